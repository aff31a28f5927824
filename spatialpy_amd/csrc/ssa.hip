// GPU-resident NSM (next-subvolume method) for the sSSA part.
//
//   rdme_rates        nsm_core__initialize_rxn_propensities and
//                     nsm_core__initialize_diff_propensities
//                     (ENG/src/simulate_rdme.cpp:113-152)
//   rdme_queue_build  nsm_core__initialize_heap + NRMConstant_v5::build
//                     (simulate_rdme.cpp:155-195, NRMConstant_v5.cpp:27-85)
//   rdme_take_step    nsm_core__take_step (simulate_rdme.cpp:211-472)
//
// The event queue is an indexed 64-ary min-tree over the voxels' firing
// times (one wavefront reduces one node).  The reference's time-binned hash
// table always yields the global minimum among firing times inside its
// current window [L, U] and rebuilds the window when that set is empty
// (NRMConstant_v5.cpp:156-281); the window scalars are kept here with the
// same update rules, so the min-tree returns the same event sequence (up to
// exactly equal firing times, which the table breaks by bin order).
//
// One 64-lane workgroup runs one trajectory's event loop; every lane
// executes the scalar logic (uniform control flow, broadcast loads), lanes
// split the min-tree node reductions and the neighbour-row loads.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>

#include "engine.hpp"

namespace ssa {

constexpr int MAXS = 16;
constexpr int MAXR = 16;

// ---------------------------------------------------------------------------
// Random numbers.
// Philox4x32-10 (Salmon et al., SC'11) for the seeded stream.
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
        uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += W0; k1 += W1;
    }
}
__device__ __forceinline__ uint64_t philox64(uint64_t seed, uint64_t stream, uint64_t ctr) {
    uint32_t c[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return (uint64_t)c[0] | ((uint64_t)c[1] << 32);
}
// rng()*1.0/rng.max() (simulate_rdme.cpp:253, 353)
__device__ __forceinline__ double word_uniform(uint64_t w) {
    return (double)w * 1.0 / (double)0xFFFFFFFFFFFFFFFFULL;
}
// -log(1 - generate_canonical) (libstdc++ exponential_distribution)
__device__ __forceinline__ double word_expo(uint64_t w) {
    double c = (double)w * 1.0 / 18446744073709551616.0;
    if (c >= 1.0) c = 0x1.fffffffffffffp-1;
    return -log(1.0 - c) / 1.0;
}

struct Rng {
    int kind;
    uint64_t seed, stream;
    const uint64_t *words;
    const double *expo;
    unsigned long long n_words;
    unsigned long long pos;
    bool exhausted;
    __device__ double uniform() {
        if (kind == 1) {
            if (pos >= n_words) { exhausted = true; return 0.5; }
            return word_uniform(words[pos++]);
        }
        return word_uniform(philox64(seed, stream, pos++));
    }
    __device__ double exponential() {
        if (kind == 1) {
            if (pos >= n_words) { exhausted = true; return 1.0; }
            return expo[pos++];
        }
        return word_expo(philox64(seed, stream, pos++));
    }
};

// Mass-action propensities (generated functions, solver.py:344-373).
__device__ __forceinline__ double propensity(const Rxn &re, const uint32_t *xv, double vol, int tp) {
    if (re.mask && !((re.mask >> tp) & 1ULL)) return 0.0;
    const int *xs = (const int *)xv;
    switch (re.kind) {
    case 0: return re.k * vol;
    case 1: return re.k * xs[re.a];
    case 2: return ((re.k * xs[re.a]) * xs[re.b]) / vol;
    default: return ((re.k * xs[re.a]) * (xs[re.a] - 1)) / vol;
    }
}

// ---------------------------------------------------------------------------
// Rates (simulate_rdme.cpp:113-152), one thread per voxel.
__global__ void k_rates(RdmeArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int S = a.S, R = a.R;
    uint32_t xv[MAXS];
    for (int s = 0; s < S; s++) xv[s] = a.xx[(size_t)i * S + s];
    double sr = 0.0;
    for (int j = 0; j < R; j++) {
        double vol = a.mass[i] / a.rho[i];
        double p = propensity(a.rx[j], xv, vol, a.type[i]);
        a.rrate[(size_t)i * R + j] = p;
        sr += p;
    }
    a.srrate[i] = sr;
    double sd = 0.0;
    const int dg = a.deg[i];
    for (int s = 0; s < S; s++) {
        double dd = 0.0;
        for (int k = 0; k < dg; k++) {
            size_t e = (size_t)k * a.n + i;
            double dc = a.Dmat[(size_t)s * a.T + (a.type[a.nidx[e]] - 1)];
            dd += dc * a.Dij[e];
        }
        a.Ddiag[(size_t)i * S + s] = dd;
        sd += dd * xv[s];
    }
    a.sdrate[i] = sd;
}
void rdme_rates(hipStream_t s, const RdmeArgs &a) {
    hipLaunchKernelGGL(k_rates, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// Time window (NRMConstant_v5::setBinNumberAndBounds, NRMConstant_v5.cpp:245-281).
__device__ bool set_bins(QueueState &q, double newL, double psum, long long active) {
    if (active == 0) return false;
    q.L = newL;
    if (q.L > q.endTime) return false;
    double w = 16.0 / psum;
    unsigned long long maxb = (unsigned long long)((double)20 * sqrt((double)(int)active));
    unsigned long long endb = ULLONG_MAX, nb;
    if (q.endTime < DBL_MAX) {
        endb = (unsigned long long)ceil((q.endTime - q.L) / w);
        nb = endb < maxb ? endb : maxb;
    } else {
        nb = maxb;
    }
    q.U = q.L + w * (double)nb;
    if (q.endTime < DBL_MAX) {
        if (q.U < q.endTime && ((double)(endb - maxb) / ((double)nb)) < .2) {
            nb = endb;
            q.U = q.L + w * (double)nb;
        }
    }
    q.nbins = nb;
    return true;
}

// prop, nonzero flags
__global__ void k_qprop(RdmeArgs a, double *prop, int *nz, int *act) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double p = a.srrate[i] + a.sdrate[i];
    prop[i] = p;
    nz[i] = p != 0.0;
    act[i] = p > 0;
}

// propensitySum in index order (simulate_rdme.cpp:167-179) by one wavefront:
// each 64-value chunk is loaded by the lanes, then added in order.
__global__ void k_qsum(RdmeArgs a, const double *prop, const int *act_scan, int exact) {
    __shared__ double part[1];
    const int lane = threadIdx.x;
    double s = 0.0;
    if (exact) {
        for (int base = 0; base < a.n; base += WAVE) {
            double v = base + lane < a.n ? prop[base + lane] : 0.0;
            int m = min(WAVE, a.n - base);
            for (int k = 0; k < m; k++) s += __shfl(v, k, WAVE);
        }
    } else {
        for (int i = lane; i < a.n; i += WAVE) s += prop[i];
        for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, WAVE);
    }
    if (lane == 0) {
        part[0] = s;
        QueueState q = *a.qs;
        q.endTime = a.endTime;
        long long active = act_scan[a.n];
        q.active = (unsigned long long)active;
        int ok = 0;
        if (a.S > 0 && s > 0.0) ok = set_bins(q, a.t0, s, active) ? 1 : 0;
        q.initialized = ok ? 2 : 0;   // 2: draws pending
        q.rxnCount = 0;
        *a.qs = q;
        *a.psum_out = s;
    }
}

// Firing times in voxel order (NRMConstant_v5.cpp:52-71): the k-th voxel with
// non-zero propensity takes the k-th exponential of the build.
__global__ void k_qtimes(RdmeArgs a, const double *prop, const int *nz_scan) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const QueueState *q = a.qs;
    if (q->initialized != 2) return;
    double p = prop[i];
    double t;
    if (p == 0.0) {
        t = INFINITY;
    } else {
        unsigned long long idx = q->rng_pos + (unsigned long long)nz_scan[i];
        double e;
        if (a.rng_kind == 1) e = idx < a.n_words ? a.expo[idx] : NAN;
        else e = word_expo(philox64(a.seed, a.stream_id, idx));
        t = e / p + a.t0;
    }
    a.qtime[i] = t;
}

__global__ void k_qfinish(RdmeArgs a, const int *nz_scan) {
    if (threadIdx.x != 0) return;
    QueueState q = *a.qs;
    if (q.initialized == 2) {
        unsigned long long cnt = (unsigned long long)nz_scan[a.n];
        if (a.rng_kind == 1 && q.rng_pos + cnt > a.n_words) q.err = SSASDPD_ERR_REPLAY_EXHAUSTED;
        q.rng_pos += cnt;
        q.initialized = (cnt == q.active) ? 1 : 0;
    }
    *a.qs = q;
}

// One level of the min-tree: node j of the upper level = argmin of 64
// children (ties -> lower voxel index).
__device__ __forceinline__ void argmin2(double &t, int &v, double t2, int v2) {
    if (t2 < t || (t2 == t && v2 < v)) { t = t2; v = v2; }
}
__device__ __forceinline__ void wave_argmin(double &t, int &v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        double t2 = __shfl_xor(t, o, WAVE);
        int v2 = __shfl_xor(v, o, WAVE);
        argmin2(t, v, t2, v2);
    }
}
__global__ void k_tree_level(const double *ct, const int *ca, int nchild, double *pt, int *pa, int nparent) {
    int node = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    int lane = threadIdx.x & (WAVE - 1);
    if (node >= nparent) return;
    int c = node * WAVE + lane;
    double t = INFINITY;
    int v = INT_MAX;
    if (c < nchild) { t = ct[c]; v = ca ? ca[c] : c; }
    wave_argmin(t, v);
    if (lane == 0) { pt[node] = t; pa[node] = v; }
}

void rdme_queue_build(hipStream_t s, const RdmeArgs &a, void *scan_tmp, int *work) {
    // work: prop (n doubles) | nz (n+1) | act (n+1) | nz_scan (n+1) | act_scan (n+1)
    double *prop = (double *)work;
    int *nz = (int *)(prop + a.n);
    int *act = nz + a.n + 1;
    int *nz_scan = act + a.n + 1;
    int *act_scan = nz_scan + a.n + 1;
    unsigned g = (unsigned)((a.n + 255) / 256);
    hipLaunchKernelGGL(k_qprop, dim3(g), dim3(256), 0, s, a, prop, nz, act);
    exclusive_scan_i32(s, nz, nz_scan, a.n, scan_tmp);
    exclusive_scan_i32(s, act, act_scan, a.n, scan_tmp);
    hipLaunchKernelGGL(k_qsum, dim3(1), dim3(WAVE), 0, s, a, prop, act_scan, a.rng_kind == 1 ? 1 : 0);
    hipLaunchKernelGGL(k_qtimes, dim3(g), dim3(256), 0, s, a, prop, nz_scan);
    hipLaunchKernelGGL(k_qfinish, dim3(1), dim3(64), 0, s, a, nz_scan);
    // min-tree levels
    const double *ct = a.qtime;
    const int *ca = nullptr;
    int nchild = a.n;
    for (int l = 0; l < a.nlev; l++) {
        int np = a.lev_n[l + 1];
        double *pt = a.tmin + a.lev_off[l];
        int *pa = a.targ + a.lev_off[l];
        hipLaunchKernelGGL(k_tree_level, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, s, ct, ca, nchild, pt, pa, np);
        ct = pt;
        ca = pa;
        nchild = np;
    }
}

// ---------------------------------------------------------------------------
// The event loop.
struct Tree {
    double *qtime, *tmin;
    int *targ;
    int n, nlev;
    int lev_off[8], lev_n[8];

    __device__ void root(double &t, int &v) const {
        if (nlev == 0) { t = qtime[0]; v = 0; return; }
        t = tmin[lev_off[nlev - 1]];
        v = targ[lev_off[nlev - 1]];
    }
    // set the firing time of voxel v and repair the path to the root
    __device__ void update(int v, double t, int lane) {
        if (lane == 0) qtime[v] = t;
        double ct = t;
        int cv = v;
        int idx = v;
        for (int l = 0; l < nlev; l++) {
            int node = idx / WAVE;
            int c = node * WAVE + lane;
            int nchild = lev_n[l];
            double tt = INFINITY;
            int vv = INT_MAX;
            if (c < nchild) {
                if (c == idx) { tt = ct; vv = cv; }
                else if (l == 0) { tt = qtime[c]; vv = c; }
                else { tt = tmin[lev_off[l - 1] + c]; vv = targ[lev_off[l - 1] + c]; }
            }
            wave_argmin(tt, vv);
            if (lane == 0) { tmin[lev_off[l] + node] = tt; targ[lev_off[l] + node] = vv; }
            ct = tt;
            cv = vv;
            idx = node;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
};

// NRMConstant_v5::rebuild (NRMConstant_v5.cpp:194-242): returns false when
// the window cannot move (selectReaction then reports an empty queue).
__device__ bool rebuild(QueueState &q) {
    double est;
    if (q.rxnCount > 0) est = (double)q.rxnCount / (q.U - q.L);
    else est = 1.0 / ((q.U - q.L) / (double)q.nbins);
    if (!set_bins(q, q.U, est, (long long)q.active)) return false;
    q.rxnCount = 0;
    return q.nbins > 0;
}

__global__ void __launch_bounds__(WAVE) k_ssa(RdmeArgs a) {
    const int lane = threadIdx.x;
    QueueState q = *a.qs;
    if (q.initialized != 1 || q.err) return;
    const int S = a.S, R = a.R, n = a.n, T = a.T;
    Tree tr;
    tr.qtime = a.qtime; tr.tmin = a.tmin; tr.targ = a.targ; tr.n = n; tr.nlev = a.nlev;
    for (int l = 0; l < 8; l++) { tr.lev_off[l] = a.lev_off[l]; tr.lev_n[l] = a.lev_n[l]; }
    Rng rng;
    rng.kind = a.rng_kind; rng.seed = a.seed; rng.stream = a.stream_id; rng.words = a.words; rng.expo = a.expo;
    rng.n_words = a.n_words; rng.pos = q.rng_pos; rng.exhausted = false;
    const bool textbook = a.textbook != 0;
    const bool lead = lane == 0;

    double tt = a.t0;
    const double end_time = a.t0 + a.dt;
    int dest = -1;
    int err = 0;
    unsigned long long guard = 0;
    unsigned long long n_react = 0, n_diff = 0;
    const double inf = INFINITY;

    while (tt <= end_time) {
        if (++guard > a.max_events) { err = SSASDPD_ERR_CAPACITY; break; }
        // ---- selectReaction
        double t;
        int sv;
        bool found = false;
        for (int it = 0; it < 4096; it++) {
            tr.root(t, sv);
            if (sv != INT_MAX && t <= q.U) { found = true; break; }
            if (!rebuild(q)) break;
            if (!(t <= q.U)) break;   // window moved but still empty: empty queue this call
        }
        if (!found) break;
        q.rxnCount++;
        tt = t;
        const double vol = a.mass[sv] / a.rho[sv];
        double sr = a.srrate[sv], sd = a.sdrate[sv];
        double totrate = sr + sd;
        const int tsv = a.type[sv];
        uint32_t xs[MAXS];
        double Dd[MAXS], rr[MAXR];
        for (int s = 0; s < S; s++) { xs[s] = a.xx[(size_t)sv * S + s]; Dd[s] = a.Ddiag[(size_t)sv * S + s]; }
        for (int j = 0; j < R; j++) rr[j] = a.rrate[(size_t)sv * R + j];
        const double rand1 = rng.uniform();
        if (rng.exhausted) { err = SSASDPD_ERR_REPLAY_EXHAUSTED; break; }
        int event;
        int which;
        double dest_sr = 0.0, dest_sd = 0.0;
        if (rand1 <= sr / totrate) {
            // ---- reaction (simulate_rdme.cpp:255-310)
            event = 0;
            double rv = textbook ? rand1 * totrate : rand1 * sr;
            int re = 0;
            double cum = R > 0 ? rr[0] : 0.0;
            while (re < R && rv > cum) { re++; cum += re < R ? rr[re] : 0.0; }
            if (re >= R) {
                if (cum != sr) {
                    double rd = 0.0;
                    for (int j = 0; j < R; j++) { rr[j] = propensity(a.rx[j], xs, vol, tsv); rd += rr[j]; }
                    sr = rd;
                    if (lead) {
                        for (int j = 0; j < R; j++) a.rrate[(size_t)sv * R + j] = rr[j];
                        a.srrate[sv] = sr;
                    }
                }
                if (sr == 0.0) continue;
                double rv2 = rand1 * sr;
                re = 0;
                cum = rr[0];
                while (re < R && rv2 > cum) { re++; cum += re < R ? rr[re] : 0.0; }
                if (re >= R) { err = SSASDPD_ERR_PROPENSITY_OVERFLOW; break; }
            }
            for (int k = a.jcN[re]; k < a.jcN[re + 1]; k++) {
                int sp = a.irN[k], d = a.prN[k];
                xs[sp] += (uint32_t)d;
                sd += Dd[sp] * d;
            }
            double rd = 0.0;
            for (int k = a.jcG[S + re]; k < a.jcG[S + re + 1]; k++) {
                int j = a.irG[k];
                double old = rr[j];
                rr[j] = propensity(a.rx[j], xs, vol, tsv);
                rd += rr[j] - old;
            }
            sr += rd;
            if (lead) {
                for (int s = 0; s < S; s++) a.xx[(size_t)sv * S + s] = xs[s];
                for (int j = 0; j < R; j++) a.rrate[(size_t)sv * R + j] = rr[j];
                a.srrate[sv] = sr;
                a.sdrate[sv] = sd;
            }
            n_react++;
            which = re;
        } else {
            // ---- diffusion (simulate_rdme.cpp:312-445)
            event = 1;
            double drand = textbook ? (rand1 * totrate - sr) : rand1 * sd;
            int spec = 0;
            double cum = S > 0 ? Dd[0] * xs[0] : 0.0;
            while (spec < S && drand > cum) { spec++; cum += spec < S ? Dd[spec] * xs[spec] : 0.0; }
            if (spec >= S) {
                if (cum != sd) {
                    double rd = 0.0;
                    for (int j = 0; j < S; j++) rd += Dd[j] * xs[j];
                    sd = rd;
                    if (lead) a.sdrate[sv] = sd;
                }
                if (sd == 0.0) continue;
                drand = cum * rand1;
                spec = 0;
                cum = Dd[0] * xs[0];
                while (spec < S && drand > cum) { spec++; cum += spec < S ? Dd[spec] * xs[spec] : 0.0; }
                if (spec >= S) {
                    spec--;
                    bool bad = false;
                    while (xs[spec] <= 0) {
                        spec--;
                        if (spec <= 0) { bad = true; break; }
                    }
                    if (bad) { err = SSASDPD_ERR_DIFFUSION_OVERFLOW; break; }
                }
            }
            const double r2 = rng.uniform();
            if (rng.exhausted) { err = SSASDPD_ERR_REPLAY_EXHAUSTED; break; }
            double rand2 = r2 * Dd[spec];
            // destination: walk the neighbour list in order
            const int dg = a.deg[sv];
            bool got = false;
            for (int pass = 0; pass < 2 && !got; pass++) {
                double cum2 = 0.0;
                for (int base = 0; base < dg && !got; base += WAVE) {
                    int k = base + lane;
                    double c = 0.0;
                    int jj = -1;
                    if (k < dg) {
                        size_t e = (size_t)k * n + sv;
                        jj = a.nidx[e];
                        int tj = T == 1 ? 1 : a.type[jj];
                        c = a.Dij[e] * a.Dmat[(size_t)spec * T + (tj - 1)];
                    }
                    int m = min(WAVE, dg - base);
                    for (int kk = 0; kk < m; kk++) {
                        cum2 += __shfl(c, kk, WAVE);
                        if (cum2 > rand2) { dest = __shfl(jj, kk, WAVE); got = true; break; }
                    }
                }
                if (!got) {
                    if (dest >= 0) break;   // stale destination of an earlier event (simulate_rdme.cpp:368)
                    rand2 = r2 * cum2;
                }
            }
            if (!got && dest < 0) { err = SSASDPD_ERR_DIFFUSION_OVERFLOW; break; }
            // move one molecule
            xs[spec]--;
            if (lead) a.xx[(size_t)sv * S + spec] = xs[spec];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            uint32_t xd[MAXS];
            double rrd[MAXR];
            const bool self = dest == sv;
            for (int s = 0; s < S; s++) xd[s] = self ? xs[s] : a.xx[(size_t)dest * S + s];
            xd[spec]++;
            if (self) xs[spec] = xd[spec];
            const int tds = a.type[dest];
            double srd = self ? sr : a.srrate[dest];
            double sdd;
            if (R > 0) {
                for (int j = 0; j < R; j++) rrd[j] = self ? rr[j] : a.rrate[(size_t)dest * R + j];
                double rd = 0.0, rrdel = 0.0;
                for (int k = a.jcG[spec]; k < a.jcG[spec + 1]; k++) {
                    int j = a.irG[k];
                    double old = rr[j];
                    rr[j] = propensity(a.rx[j], xs, vol, tsv);
                    rd += rr[j] - old;
                    if (self) rrd[j] = rr[j];
                    old = rrd[j];
                    rrd[j] = propensity(a.rx[j], xd, vol, tds);
                    rrdel += rrd[j] - old;
                    if (self) rr[j] = rrd[j];
                }
                sr += rd;
                if (self) srd = sr;
                srd += rrdel;
                if (self) sr = srd;
            }
            sd -= Dd[spec];
            sdd = self ? sd : a.sdrate[dest];
            sdd += a.Ddiag[(size_t)dest * S + spec];
            if (self) sd = sdd;
            if (lead) {
                if (!self) {
                    for (int j = 0; j < R; j++) a.rrate[(size_t)sv * R + j] = rr[j];
                    a.srrate[sv] = sr;
                    a.sdrate[sv] = sd;
                }
                a.xx[(size_t)dest * S + spec] = xd[spec];
                for (int j = 0; j < R; j++) a.rrate[(size_t)dest * R + j] = rrd[j];
                a.srrate[dest] = srd;
                a.sdrate[dest] = sdd;
            }
            n_diff++;
            which = spec;
            dest_sr = srd;
            dest_sd = sdd;
        }
        // ---- queue updates (NRMConstant_v5::update, NRMConstant_v5.cpp:89-132)
        {
            double tot = sr + sd;
            double old = a.qtime[sv];
            double nt;
            if (tot <= 0.0) {
                nt = inf;
                if (old != inf) q.active--;
            } else {
                nt = rng.exponential() / tot + tt;
                if (old == inf) q.active++;
            }
            tr.update(sv, nt, lane);
        }
        if (event) {
            double tot = dest_sr + dest_sd;
            double old = a.qtime[dest];
            double nt;
            if (tot <= 0.0) {
                nt = inf;
                if (old != inf) q.active--;
            } else {
                nt = rng.exponential() / tot + tt;
                if (old == inf) q.active++;
            }
            tr.update(dest, nt, lane);
        }
        if (rng.exhausted) { err = SSASDPD_ERR_REPLAY_EXHAUSTED; break; }
        if (a.trace && lead) {
            unsigned long long ix = q.events_traced;
            if (ix < a.trace_cap) {
                ssasdpd_event ev;
                ev.t = tt; ev.voxel = sv; ev.kind = event; ev.which = which; ev.dest = event ? dest : -1;
                a.trace[ix] = ev;
            }
        }
        q.events_traced++;
    }
    q.rng_pos = rng.pos;
    q.reactions += n_react;
    q.diffusions += n_diff;
    if (err) q.err = err;
    if (lead) *a.qs = q;
}

void rdme_take_step(hipStream_t s, const RdmeArgs &a) {
    hipLaunchKernelGGL(k_ssa, dim3(1), dim3(WAVE), 0, s, a);
}

}  // namespace ssa
