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

__global__ void k_predict(SdpdArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
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
}

__device__ __forceinline__ double chem_rhs(const SdpdArgs &a, int i, int r) {
    // generated det<name> functions (solver.py:160-190)
    const Rxn &re = a.rx[r];
    if (re.mask && !((re.mask >> a.type[i]) & 1ULL)) return 0.0;
    const double *c = a.C + (size_t)i * a.Sc;
    switch (re.kind) {
    case 0: return re.k;
    case 1: return re.k * c[re.a];
    case 2: return (re.k * c[re.a]) * c[re.b];
    default: return (re.k * c[re.a]) * c[re.a];
    }
}

// pairwiseForce (model.cpp:39-191) in FP64, list order.  For static domains
// only the species flux Q is observable, so the momentum part is skipped.
__global__ void __launch_bounds__(256) k_forces(SdpdArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const size_t n = a.n;
    const double h = a.h, P0 = a.P0, rho0 = a.rho0;
    const double xi[3] = {a.x[i], a.x[n + i], a.x[2 * n + i]};
    const double vi[3] = {a.v[i], a.v[n + i], a.v[2 * n + i]};
    const double vti[3] = {a.vt[i], a.vt[n + i], a.vt[2 * n + i]};
    const double rhoi = a.rho[i], mi = a.mass[i], nui = a.nu[i];
    const double Pi = P0 * (rhoi / rho0 - 1.0);
    double F[3] = {a.F[i], a.F[n + i], a.F[2 * n + i]};
    double Fbp[3] = {a.Fbp[i], a.Fbp[n + i], a.Fbp[2 * n + i]};
    double Frho = a.Frho[i];
    constexpr int MAXSC = 16;
    double Q[MAXSC];
    const int Sc = a.Sc < MAXSC ? a.Sc : MAXSC;
    for (int s = 0; s < Sc; s++) Q[s] = a.Q[(size_t)i * a.Sc + s];
    const int ti = a.type[i];
    const bool momentum = a.static_domain == 0;
    const int dg = a.deg[i];
    double dx[3] = {0.0, 0.0, 0.0}, dv[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < dg; k++) {
        const size_t e = (size_t)k * n + i;
        const int j = a.nidx[e];
        const double r = a.nr[e];
        const double R = r / h;
        if (R > 1.0) continue;
        if (r == 0.0) continue;
        const double dWdr = a.alpha * (-12 * r / (h * h)) * ((1 - R) * (1 - R));
        const double rhoj = a.rho[j], mj = a.mass[j];
        const double inv_rh = 1.0 / (r + 0.001 * h);
        if (momentum) {
            const double xj[3] = {a.x[j], a.x[n + j], a.x[2 * n + j]};
            const double vj[3] = {a.v[j], a.v[n + j], a.v[2 * n + j]};
            const double vtj[3] = {a.vt[j], a.vt[n + j], a.vt[2 * n + j]};
            const double nuj = a.nu[j];
            double dv_dx = 0.0;
            for (int d = 0; d < a.dim; d++) {
                dx[d] = xi[d] - xj[d];
                dv[d] = vi[d] - vj[d];
                dv_dx += dv[d] * dx[d];
            }
            const double Pj = P0 * (rhoj / rho0 - 1.0);
            double pg = Pi / (rhoi * rhoi) + Pj / (rhoj * rhoj);
            if (pg < 0) pg = -Pi / (rhoi * rhoi) + Pj / (rhoj * rhoj);
            const double fp = -1.0 * mj * pg * dWdr / (r + 0.001 * h);
            const double fv = mj * (2.0 * (nui * nuj) / (nui + nuj)) * (1 / (r + 0.001 * h)) * dWdr / ((rhoi * rhoj));
            const double mri = mi / rhoi, mrj = mj / rhoj;
            const double vsum = mri * mri + mrj * mrj;
            const double fbp = -10.0 * P0 * (1.0 / mi) * vsum * dWdr / (r + 0.001 * h);
            double ft[3];
            for (int p = 0; p < 3; p++) {
                double acc = 0.0;
                double Tp[3];
                for (int q = 0; q < 3; q++)
                    Tp[q] = 0.5 * ((rhoi * vi[p] * (vti[q] - vi[q])) + (rhoj * vj[p] * (vtj[q] - vj[q])));
                acc = Tp[0] * dx[0] + Tp[1] * dx[1] + Tp[2] * dx[2];
                ft[p] = (1.0 / mi) * vsum * acc * dWdr / (r + 0.001 * h);
            }
            for (int d = 0; d < a.dim; d++) {
                F[d] += fp * dx[d] + fv * dv[d] + ft[d];
                Fbp[d] += fbp * dx[d];
            }
            Frho = Frho + rhoi * (mj / rhoj) * dv_dx * (1 / (r + 0.001 * h)) * dWdr -
                   0.0 * h * rhoi * a.c0 * mj * 2.0 * (rhoj / rhoi - 1.0) * (r * r / (r * r + 0.01 * h * h)) *
                       (1.0 / (r + 0.001 * h)) * dWdr / rhoj -
                   (mj / rhoj) *
                       (rhoi * ((vi[0] - vti[0]) * dx[0] + (vi[1] - vti[1]) * dx[1] + (vi[2] - vti[2]) * dx[2]) +
                        rhoj * ((vj[0] - vtj[0]) * dx[0] + (vj[1] - vtj[1]) * dx[1] + (vj[2] - vtj[2]) * dx[2])) *
                       inv_rh * dWdr;
        }
        if (Sc > 0) {
            const double wfd = (1.0 / (r + 0.001 * h)) * dWdr;
            const double dQc_base = 2.0 * ((mi * mj) / (mi + mj)) * ((rhoi + rhoj) / (rhoi * rhoj)) * (r * r) * wfd /
                                    ((r * r) + 0.01 * h * h);
            for (int s = 0; s < Sc; s++) {
                const long long kk = (long long)a.Sc * (ti - 1) + s;   // type-major, model.cpp:163
                const double D = (kk < (long long)a.Sc * a.T) ? a.Dmat[kk] : 0.0;
                Q[s] += D * (a.C[(size_t)i * a.Sc + s] - a.C[(size_t)j * a.Sc + s]) * dQc_base;
            }
        }
    }
    for (int r = 0; r < a.Rc; r++) {
        const double flux = chem_rhs(a, i, r);
        for (int s = 0; s < Sc; s++) {
            const long long kk = (long long)a.Rc * r + s;             // model.cpp:186
            const int nk = kk < (long long)a.Sc * a.Rc ? a.Ndense[kk] : 0;
            Q[s] += nk * flux;
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
    int i = blockIdx.x * blockDim.x + threadIdx.x;
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
            num += a.old_rho[a.nidx[e]] * W;
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

__global__ void k_bvf(SdpdArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const size_t n = a.n;
    if (a.solid[i] == 0 && a.static_domain == 0) {
        const double h = a.h;
        double vos = 0.0, vtot = 0.0, nw[3] = {0.0, 0.0, 0.0}, dx[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < a.dim; d++) { a.normal[d * n + i] = 0.0; a.vt[d * n + i] = 0.0; }
        const double xi[3] = {a.x[i], a.x[n + i], a.x[2 * n + i]};
        const int dg = a.deg[i];
        for (int k = 0; k < dg; k++) {
            const size_t e = (size_t)k * n + i;
            const int j = a.nidx[e];
            const double r = a.nr[e], R = r / h;
            if (R > 1.0) continue;
            if (r == 0.0) continue;
            const double W = a.alpha * ((1 + 3 * R) * cube(1 - R));
            const double dWdr = a.alpha * (-12 * r / (h * h)) * ((1 - R) * (1 - R));
            for (int d = 0; d < a.dim; d++) dx[d] = xi[d] - a.x[d * n + j];
            const double mr = a.mass[j] / a.rho[j];
            const double mv = mr * mr;
            const bool sj = a.solid[j] != 0;
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

static dim3 grid1(int n) { return dim3((unsigned)((n + 255) / 256)); }

void sdpd_predict(hipStream_t s, const SdpdArgs &a) { hipLaunchKernelGGL(k_predict, grid1(a.n), dim3(256), 0, s, a); }
void sdpd_forces(hipStream_t s, const SdpdArgs &a) { hipLaunchKernelGGL(k_forces, grid1(a.n), dim3(256), 0, s, a); }
void sdpd_correct(hipStream_t s, const SdpdArgs &a) {
    if (!a.static_domain) hipLaunchKernelGGL(k_correct, grid1(a.n), dim3(256), 0, s, a);
}
void sdpd_bvf(hipStream_t s, const SdpdArgs &a) { hipLaunchKernelGGL(k_bvf, grid1(a.n), dim3(256), 0, s, a); }

}  // namespace ssa
