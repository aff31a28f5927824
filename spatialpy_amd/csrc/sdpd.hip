// SDPD substeps on gfx950: one thread per particle, neighbour lists read from
// the ELL layout (coalesced across threads at each list position).
//
//   sdpd_predict  take_step1     (ENG/src/simulate.cpp:56-109) + NaN check
//                                (ENG/src/particle.cpp:88-134)
//   sdpd_forces   pairwiseForce  (ENG/src/model.cpp:39-191)
//   sdpd_correct  take_step2 corrector + filterDensity (simulate.cpp:136-155,
//                                model.cpp:194-233)
//   sdpd_bvf      computeBoundaryVolumeFraction + applyBoundaryVolumeFraction
//                 + chem half step + boundary conditions (simulate.cpp:158-171,
//                 model.cpp:236-329)
//
// Neighbour rho inside the boundary volume fraction is read after every
// particle's density update of the same substep (snapshot order); the
// reference with one thread reads it in index order (model.cpp:285-293).
#include <hip/hip_runtime.h>

#include <cmath>

#include "engine.hpp"
#include "vm.hpp"

namespace ssa {

__device__ __forceinline__ void apply_bcs(const SdpdArgs &a, int i) {
    // generated applyBoundaryConditions (core/boundarycondition.py:121-167)
    for (int b = 0; b < a.nbc; b++) {
        const BcRule &r = a.bc[b];
        double x0 = a.x[i], x1 = a.x[(size_t)a.n + i], x2 = a.x[2 * (size_t)a.n + i];
        if (!((x0 >= r.lo[0]) && (x0 <= r.hi[0]) && (x1 >= r.lo[1]) && (x1 <= r.hi[1]) && (x2 >= r.lo[2]) &&
              (x2 <= r.hi[2])))
            continue;
        if (r.type >= 0 && a.type[i] != r.type) continue;
        switch (r.target) {
        case SSASDPD_BC_CHEM:
            if (r.species < a.Sc) a.C[(size_t)i * a.Sc + r.species] = r.value[0];
            break;
        case SSASDPD_BC_V:
            for (int k = 0; k < 3; k++) a.v[(size_t)k * a.n + i] = r.value[k];
            break;
        case SSASDPD_BC_NU: a.nu[i] = r.value[0]; break;
        case SSASDPD_BC_RHO: a.rho[i] = r.value[0]; break;
        }
    }
}

__device__ __forceinline__ bool bad(double x) { return isnan(x) || !isfinite(x); }

// Particle i's gather record from the arrays (eight 16-byte stores).
__device__ __forceinline__ void store_rec_to(const SdpdArgs &a, int i, double2 *p) {
    const size_t n = a.n;
    p[0] = make_double2(a.x[i], a.x[n + i]);
    p[1] = make_double2(a.x[2 * n + i], a.v[i]);
    p[2] = make_double2(a.v[n + i], a.v[2 * n + i]);
    p[3] = make_double2(a.vt[i], a.vt[n + i]);
    p[4] = make_double2(a.vt[2 * n + i], a.nu[i]);
    p[5] = make_double2(a.rho[i], a.mass[i]);
    ((int4 *)p)[6] = make_int4(a.type[i], a.solid[i], 0, 0);
    // what every pair with this particle as j forms from j alone, by the
    // same operations as the pair loop: P_j / rho_j^2 and m_j / rho_j
    const double rho = a.rho[i], P = a.P0 * (rho / a.rho0 - 1.0);
    p[7] = make_double2(P / (rho * rho), a.mass[i] / rho);
}
__device__ __forceinline__ void store_rec(const SdpdArgs &a, int i) { store_rec_to(a, i, (double2 *)(a.rec + i)); }

// Neighbour j's record: x, v, vt, nu, rho, mass (six 16-byte loads of one
// 128-byte line).
__device__ __forceinline__ void load_rec(const PRec *r, double x[3], double v[3], double vt[3], double &nu,
                                         double &rho, double &mass) {
    const double2 *p = (const double2 *)r;
    const double2 a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3], a4 = p[4], a5 = p[5];
    x[0] = a0.x; x[1] = a0.y; x[2] = a1.x;
    v[0] = a1.y; v[1] = a2.x; v[2] = a2.y;
    vt[0] = a3.x; vt[1] = a3.y; vt[2] = a4.x;
    nu = a4.y; rho = a5.x; mass = a5.y;
}

__device__ __forceinline__ double2 load_rho_mass(const PRec *r) { return *(const double2 *)&r->rho; }
// P_j / rho_j^2 and m_j / rho_j of the record (see store_rec_to)
__device__ __forceinline__ double2 load_rec_pm(const PRec *r) { return ((const double2 *)r)[7]; }

__global__ void k_pack(SdpdArgs a) {
    const int chunk = xcd_chunk((a.n + 255) / 256);
    if (chunk < 0) return;
    int i = chunk * blockDim.x + threadIdx.x;
    if (i < a.n) store_rec(a, i);
}

__device__ __forceinline__ void predict_one(const SdpdArgs &a, int i, double2 *rec_out) {
    const size_t n = a.n;
    bool nan = bad(a.rho[i]);
    for (int k = 0; k < 3; k++) nan = nan || bad(a.x[k * n + i]) || bad(a.v[k * n + i]);
    if (nan) {
        atomicCAS(&a.err->code, 0, (int)SSASDPD_ERR_NAN_STATE);
        atomicMin(&a.err->particle, i);
    }
    if (a.solid[i] == 0 && a.static_domain == 0) {
        for (int k = 0; k < 3; k++) {
            double v = a.v[k * n + i] + 0.5 * a.dt * a.F[k * n + i];
            double vt = v + 0.5 * a.dt * a.Fbp[k * n + i];
            a.v[k * n + i] = v;
            a.vt[k * n + i] = vt;
            a.x[k * n + i] = a.x[k * n + i] + a.dt * vt;
        }
        a.rho[i] = a.rho[i] + 0.5 * a.dt * a.Frho[i];
    }
    if (a.step > 0)
        for (int s = 0; s < a.Sc; s++) a.C[(size_t)i * a.Sc + s] += a.Q[(size_t)i * a.Sc + s] * a.dt * 0.5;
    apply_bcs(a, i);
    for (int k = 0; k < 3; k++) {
        a.F[k * n + i] = a.g[k];
        a.Fbp[k * n + i] = 0.0;
    }
    a.Frho[i] = 0.0;
    for (int s = 0; s < a.Sc; s++) a.Q[(size_t)i * a.Sc + s] = 0.0;
    a.old_rho[i] = a.rho[i];
    store_rec_to(a, i, rec_out);
}

// The gather records are staged in LDS and written as one contiguous range
// per block (a lane storing its own 128-byte record writes 16 bytes to each
// of 64 lines per instruction).
__global__ void __launch_bounds__(256) k_predict(SdpdArgs a) {
    __shared__ double2 srec[256 * 8];
    const int chunk = xcd_chunk((a.n + 255) / 256);
    if (chunk < 0) return;
    const int base = chunk * 256;
    const int i = base + threadIdx.x;
    if (i < a.n) predict_one(a, i, srec + 8 * threadIdx.x);
    __syncthreads();
    const int nrec = min(256, a.n - base);
    double2 *dst = (double2 *)(a.rec + base);
    for (int q = threadIdx.x; q < nrec * 8; q += 256) dst[q] = srec[q];
}

// generated det<name> functions (solver.py:160-190), called with
// vol = mass/rho and t = current_step*dt (model.cpp:181-184); VM: kind-4
// rows through the expression VM (only k_chem instantiates it)
template <bool VM>
__device__ __forceinline__ double chem_rhs(const SdpdArgs &a, int i, int r) {
    const Rxn &re = a.rx[r];
    if (re.mask && !((re.mask >> a.type[i]) & 1ULL)) return 0.0;
    const double *c = a.C + (size_t)i * a.Sc;
    if constexpr (VM) {
        if (re.kind == 4)
            return vm_eval_f64(a.vm_code, re.b, a.vm_k, c, a.mass[i] / a.rho[i], (double)a.step * a.dt,
                               a.data_fn ? a.data_fn + (size_t)i * a.n_df : nullptr);
    }
    switch (re.kind) {
    case 0: return re.k;
    case 1: return re.k * c[re.a];
    case 2: return (re.k * c[re.a]) * c[re.b];
    default: return (re.k * c[re.a]) * c[re.a];
    }
}

// pairwiseForce (model.cpp:39-191), list order.  Real = double is the
// reference's arithmetic operation for operation; Real = float evaluates the
// pair terms (kernel gradient, pressure, viscosity, background pressure,
// transport tensor, continuity) in FP32 from FP64-formed differences dx, dv
// and sums them into FP64 accumulators.  The species flux Q stays FP64.  For
// static domains only Q is observable, so the momentum part is skipped.
// CHEM: the reaction fluxes are added here; otherwise k_chem adds them
// afterwards (models with custom deterministic rates), to the same Q in the
// same order.
// SPECIES: the model has deterministic species (without them the species
// flux and its registers drop out: more wavefronts hide the gathers).
template <typename Real, bool CHEM, bool SPECIES = true>
__global__ void __launch_bounds__(256) k_forces(SdpdArgs a) {
    const int chunk = xcd_chunk((a.n + 255) / 256);
    if (chunk < 0) return;
    int i = chunk * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const size_t n = a.n;
    const double h = a.h;
    const Real hR = (Real)a.h, P0 = (Real)a.P0, rho0 = (Real)a.rho0, alpha = (Real)a.alpha;
    const double xi[3] = {a.x[i], a.x[n + i], a.x[2 * n + i]};
    const double vi[3] = {a.v[i], a.v[n + i], a.v[2 * n + i]};
    const double vti[3] = {a.vt[i], a.vt[n + i], a.vt[2 * n + i]};
    const Real viR[3] = {(Real)vi[0], (Real)vi[1], (Real)vi[2]};
    const Real dvti[3] = {(Real)(vi[0] - vti[0]), (Real)(vi[1] - vti[1]), (Real)(vi[2] - vti[2])};
    const Real rhoi = (Real)a.rho[i], mi = (Real)a.mass[i], nui = (Real)a.nu[i];
    const Real Pi = P0 * (rhoi / rho0 - (Real)1.0);
    // the sums start where the predictor resets them (F = g, Fbp, Frho and
    // Q = 0: simulate.cpp:92-105), not from the arrays, so a second launch
    // on the same lists (after a re-search) gives the same result
    double F[3] = {a.g[0], a.g[1], a.g[2]};
    double Fbp[3] = {0.0, 0.0, 0.0};
    double Frho = 0.0;
    constexpr int MAXSC = 16;
    double Q[SPECIES ? MAXSC : 1];
    const int Sc = SPECIES ? (a.Sc < MAXSC ? a.Sc : MAXSC) : 0;
    for (int s = 0; s < Sc; s++) Q[s] = 0.0;
    const int ti = a.type[i];
    const bool momentum = a.static_domain == 0;
    const int dg = a.deg[i];
    Real dx[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
    for (int k = 0; k < dg; k++) {
        const size_t e = (size_t)k * n + i;
        const int j = a.nidx[e];
        const double rd = a.nr[e];
        if (rd / h > 1.0) continue;
        if (rd == 0.0) continue;
        const Real r = (Real)rd;
        const Real R = r / hR;
        const Real dWdr = alpha * ((Real)-12 * r / (hR * hR)) * (((Real)1 - R) * ((Real)1 - R));
        double xj[3], vj[3], vtj[3], nujd, rhojd, mjd;
        double2 pmj = make_double2(0.0, 0.0);
        if (momentum) {
            load_rec(a.rec + j, xj, vj, vtj, nujd, rhojd, mjd);
            pmj = load_rec_pm(a.rec + j);
        } else {
            const double2 rm = load_rho_mass(a.rec + j);
            rhojd = rm.x;
            mjd = rm.y;
        }
        const Real rhoj = (Real)rhojd, mj = (Real)mjd;
        const Real inv_rh = (Real)1.0 / (r + (Real)0.001 * hR);
        if (momentum) {
            const Real nuj = (Real)nujd;
            Real dv_dx = 0;
            for (int d = 0; d < a.dim; d++) {
                dx[d] = (Real)(xi[d] - xj[d]);
                dv[d] = (Real)(vi[d] - vj[d]);
                dv_dx += dv[d] * dx[d];
            }
            const Real vjR[3] = {(Real)vj[0], (Real)vj[1], (Real)vj[2]};
            const Real dvtj[3] = {(Real)(vj[0] - vtj[0]), (Real)(vj[1] - vtj[1]), (Real)(vj[2] - vtj[2])};
            // P_j / rho_j^2 and m_j / rho_j from the record: the same
            // operations on the same values as here, done once per particle
            const Real Pj_r = (Real)pmj.x, mrj = (Real)pmj.y;
            Real pg = Pi / (rhoi * rhoi) + Pj_r;
            if (pg < 0) pg = -Pi / (rhoi * rhoi) + Pj_r;
            const Real fp = (Real)-1.0 * mj * pg * dWdr / (r + (Real)0.001 * hR);
            const Real fv = mj * ((Real)2.0 * (nui * nuj) / (nui + nuj)) * ((Real)1 / (r + (Real)0.001 * hR)) * dWdr /
                            ((rhoi * rhoj));
            const Real mri = mi / rhoi;
            const Real vsum = mri * mri + mrj * mrj;
            const Real fbp = (Real)-10.0 * P0 * ((Real)1.0 / mi) * vsum * dWdr / (r + (Real)0.001 * hR);
            Real ft[3];
            for (int p = 0; p < 3; p++) {
                Real acc = 0;
                Real Tp[3];
                for (int q = 0; q < 3; q++)
                    Tp[q] = (Real)0.5 * ((rhoi * viR[p] * (-dvti[q])) + (rhoj * vjR[p] * (-dvtj[q])));
                acc = Tp[0] * dx[0] + Tp[1] * dx[1] + Tp[2] * dx[2];
                ft[p] = ((Real)1.0 / mi) * vsum * acc * dWdr / (r + (Real)0.001 * hR);
            }
            for (int d = 0; d < a.dim; d++) {
                if (!a.skip_F) F[d] += (double)(fp * dx[d] + fv * dv[d] + ft[d]);
                Fbp[d] += (double)(fbp * dx[d]);
            }
            // model.cpp:143-146; its middle term (density diffusion) is
            // multiplied by 0.0 there: for finite states it adds a signed
            // zero, which leaves the sum (which starts at +0) unchanged, so
            // it is not formed
            Frho = Frho + (double)(rhoi * mrj * dv_dx * ((Real)1 / (r + (Real)0.001 * hR)) * dWdr -
                                   mrj *
                                       (rhoi * (dvti[0] * dx[0] + dvti[1] * dx[1] + dvti[2] * dx[2]) +
                                        rhoj * (dvtj[0] * dx[0] + dvtj[1] * dx[1] + dvtj[2] * dx[2])) *
                                       inv_rh * dWdr);
        }
        if (Sc > 0) {
            const double Rd = rd / h;
            const double dWdrd = a.alpha * (-12 * rd / (h * h)) * ((1 - Rd) * (1 - Rd));
            const double rhoid = a.rho[i], mid = a.mass[i];
            const double wfd = (1.0 / (rd + 0.001 * h)) * dWdrd;
            const double dQc_base = 2.0 * ((mid * mjd) / (mid + mjd)) * ((rhoid + rhojd) / (rhoid * rhojd)) *
                                    (rd * rd) * wfd / ((rd * rd) + 0.01 * h * h);
            for (int s = 0; s < Sc; s++) {
                const long long kk = (long long)a.Sc * (ti - 1) + s;   // type-major, model.cpp:163
                const double D = (kk < (long long)a.Sc * a.T) ? a.Dmat[kk] : 0.0;
                Q[s] += D * (a.C[(size_t)i * a.Sc + s] - a.C[(size_t)j * a.Sc + s]) * dQc_base;
            }
        }
    }
    if constexpr (CHEM) {
        for (int r = 0; r < a.Rc; r++) {
            const double flux = chem_rhs<false>(a, i, r);
            for (int s = 0; s < Sc; s++) {
                const long long kk = (long long)a.Rc * r + s;             // model.cpp:186
                const int nk = kk < (long long)a.Sc * a.Rc ? a.Ndense[kk] : 0;
                Q[s] += nk * flux;
            }
        }
    }
    if (momentum) {
        for (int d = 0; d < 3; d++) { a.F[d * n + i] = F[d]; a.Fbp[d * n + i] = Fbp[d]; }
        a.Frho[i] = Frho;
    }
    for (int s = 0; s < Sc; s++) a.Q[(size_t)i * a.Sc + s] = Q[s];
}

__device__ __forceinline__ double cube(double x) { return x * x * x; }

__global__ void k_correct(SdpdArgs a) {
    const int chunk = xcd_chunk((a.n + 255) / 256);
    if (chunk < 0) return;
    int i = chunk * blockDim.x + threadIdx.x;
    if (i >= a.n || a.static_domain) return;
    const size_t n = a.n;
    const bool filt = (a.step % 20) == 0;
    auto filter = [&]() {
        // filterDensity (model.cpp:194-233): no self term
        const double h = a.h;
        double num = 0.0, den = 0.0;
        const int dg = a.deg[i];
        for (int k = 0; k < dg; k++) {
            const size_t e = (size_t)k * n + i;
            const double r = a.nr[e], R = r / h;
            if (R > 1.0) continue;
            if (r == 0.0) continue;
            const double W = a.alpha * ((1 + 3 * R) * cube(1 - R));
            num += a.rec[a.nidx[e]].rho * W;     // old_rho of j (the record is written with it)
            den += W;
        }
        a.rho[i] = num / den;
    };
    if (a.solid[i] == 0) {
        for (int k = 0; k < 3; k++) a.v[k * n + i] = a.v[k * n + i] + 0.5 * a.dt * a.F[k * n + i];
        if (filt) filter();
        a.rho[i] = a.rho[i] + 0.5 * a.dt * a.Frho[i];
    } else if (a.solid[i] == 1) {
        if (filt) filter();
    }
}

__device__ __forceinline__ void bvf_one(const SdpdArgs &a, int i) {
    const size_t n = a.n;
    if (a.solid[i] == 0 && a.static_domain == 0) {
        const double h = a.h;
        double vos = 0.0, vtot = 0.0, nw[3] = {0.0, 0.0, 0.0}, dx[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < a.dim; d++) { a.normal[d * n + i] = 0.0; a.vt[d * n + i] = 0.0; }
        const double xi[3] = {a.x[i], a.x[n + i], a.x[2 * n + i]};
        const int dg = a.deg[i];
        if (a.no_solids) {
            // no solid neighbour: vos and nw stay 0, so phi = |0 / vtot| is 0
            // when some neighbour has W > 0 (0 < r < h, every (m/rho)^2 > 0)
            // and 0/0 = NaN otherwise; the normal is -0/0 = NaN and the
            // velocity is never reflected (phi >= 0.5 is false for both)
            bool pos = false;
            for (int k = 0; k < dg && !pos; k++) {
                const double r = a.nr[(size_t)k * n + i], R = r / h;
                pos = !(R > 1.0) && r != 0.0 && (1 + 3 * R) * cube(1 - R) > 0.0;
            }
            const double qnan = __longlong_as_double(0x7ff8000000000000LL);
            for (int d = 0; d < 3; d++) a.normal[d * n + i] = -qnan;
            a.bvf[i] = pos ? 0.0 : qnan;
            for (int s = 0; s < a.Sc; s++) a.C[(size_t)i * a.Sc + s] += a.Q[(size_t)i * a.Sc + s] * a.dt * 0.5;
            apply_bcs(a, i);
            return;
        }
        for (int k = 0; k < dg; k++) {
            const size_t e = (size_t)k * n + i;
            const int j = a.nidx[e];
            const double r = a.nr[e], R = r / h;
            if (R > 1.0) continue;
            if (r == 0.0) continue;
            const double W = a.alpha * ((1 + 3 * R) * cube(1 - R));
            const double dWdr = a.alpha * (-12 * r / (h * h)) * ((1 - R) * (1 - R));
            // x, mass and solid of j from its record (unchanged since the
            // predictor); rho[j] after the corrector from the array
            const PRec *pj = a.rec + j;
            const double2 x01 = *(const double2 *)&pj->x[0];
            const double xj[3] = {x01.x, x01.y, pj->x[2]};
            for (int d = 0; d < a.dim; d++) dx[d] = xi[d] - xj[d];
            const double mr = pj->mass / a.rho[j];
            const double mv = mr * mr;
            const bool sj = pj->solid != 0;
            if (sj) vos += mv * W;
            vtot += mv * W;
            if (sj)
                for (int d = 0; d < 3; d++) nw[d] += mv * dx[d] * dWdr / (r + 0.001 * h);
        }
        for (int d = 0; d < 3; d++) nw[d] = nw[d] / vtot;
        const double nn = sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
        double nm[3];
        for (int d = 0; d < 3; d++) { nm[d] = -nw[d] / nn; a.normal[d * n + i] = nm[d]; }
        const double phi = fabs(vos / vtot);
        a.bvf[i] = phi;
        double v[3] = {a.v[i], a.v[n + i], a.v[2 * n + i]};
        const double vdn = v[0] * nm[0] + v[1] * nm[1] + v[2] * nm[2];
        if (phi >= 0.5)
            for (int d = 0; d < a.dim; d++) a.v[d * n + i] = -v[d] + 2.0 * fmax(0.0, vdn) * nm[d];
    }
    for (int s = 0; s < a.Sc; s++) a.C[(size_t)i * a.Sc + s] += a.Q[(size_t)i * a.Sc + s] * a.dt * 0.5;
    apply_bcs(a, i);
}
// The BVF, the corrector's species update and the BCs; then vol = mass/rho
// for the RDME of the step, from the final rho (the BCs may set it)
__global__ void k_bvf(SdpdArgs a) {
    const int chunk = xcd_chunk((a.n + 255) / 256);
    if (chunk < 0) return;
    const int i = chunk * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    bvf_one(a, i);
    if (a.vol) a.vol[i] = a.mass[i] / a.rho[i];
}

static dim3 grid1(int n) { return xcd_grid((n + 255) / 256); }

void sdpd_predict(hipStream_t s, const SdpdArgs &a) { hipLaunchKernelGGL(k_predict, grid1(a.n), dim3(256), 0, s, a); }
// The reaction fluxes of models with custom deterministic rates
// (model.cpp:179-189), after k_forces has summed the pair fluxes into Q.
__global__ void __launch_bounds__(256) k_chem(SdpdArgs a) {
    const int chunk = xcd_chunk((a.n + 255) / 256);
    if (chunk < 0) return;
    const int i = chunk * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int Sc = a.Sc;
    double *Q = a.Q + (size_t)i * Sc;
    for (int r = 0; r < a.Rc; r++) {
        const double flux = chem_rhs<true>(a, i, r);
        for (int s = 0; s < Sc; s++) {
            const long long kk = (long long)a.Rc * r + s;             // model.cpp:186
            const int nk = kk < (long long)a.Sc * a.Rc ? a.Ndense[kk] : 0;
            Q[s] += nk * flux;
        }
    }
}

void sdpd_forces(hipStream_t s, const SdpdArgs &a) {
    const bool vm = a.vm_code != nullptr && a.Rc > 0 && a.Sc > 0;
    if (vm) {
        if (a.fp32) hipLaunchKernelGGL((k_forces<float, false>), grid1(a.n), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_forces<double, false>), grid1(a.n), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_chem, grid1(a.n), dim3(256), 0, s, a);
        return;
    }
    if (a.Sc == 0) {
        if (a.fp32) hipLaunchKernelGGL((k_forces<float, false, false>), grid1(a.n), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_forces<double, false, false>), grid1(a.n), dim3(256), 0, s, a);
        return;
    }
    if (a.fp32) hipLaunchKernelGGL((k_forces<float, true>), grid1(a.n), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_forces<double, true>), grid1(a.n), dim3(256), 0, s, a);
}
void sdpd_correct(hipStream_t s, const SdpdArgs &a) {
    if (!a.static_domain) hipLaunchKernelGGL(k_correct, grid1(a.n), dim3(256), 0, s, a);
}
void sdpd_bvf(hipStream_t s, const SdpdArgs &a) { hipLaunchKernelGGL(k_bvf, grid1(a.n), dim3(256), 0, s, a); }
void sdpd_pack(hipStream_t s, const SdpdArgs &a) { hipLaunchKernelGGL(k_pack, grid1(a.n), dim3(256), 0, s, a); }


// Row permutation: gather dst[p][c] = src[map[p]][c], or scatter
// dst[map[p]][c] = src[p][c]; rows of `cols` elements of T.
template <class T>
__global__ void k_permute_rows(int n, int cols, const int *__restrict__ map, const T *__restrict__ src,
                               T *__restrict__ dst, int scatter) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)n * cols) return;
    const int p = (int)(e / cols), c = (int)(e % cols);
    const size_t q = (size_t)map[p] * cols + c;
    if (scatter) dst[q] = src[e];
    else dst[e] = src[q];
}
void permute_rows(hipStream_t s, int n, int cols, int esz, const int *map, const void *src, void *dst, bool scatter) {
    if (n <= 0 || cols <= 0) return;
    const size_t m = (size_t)n * cols;
    const dim3 g((unsigned)((m + 255) / 256));
    if (esz == 8)
        hipLaunchKernelGGL(k_permute_rows<uint64_t>, g, dim3(256), 0, s, n, cols, map, (const uint64_t *)src,
                           (uint64_t *)dst, scatter ? 1 : 0);
    else
        hipLaunchKernelGGL(k_permute_rows<uint32_t>, g, dim3(256), 0, s, n, cols, map, (const uint32_t *)src,
                           (uint32_t *)dst, scatter ? 1 : 0);
}
}  // namespace ssa
