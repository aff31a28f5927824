// GPU-resident NSM (next-subvolume method) for the sSSA part.
//
//   rdme_rates        nsm_core__initialize_rxn_propensities and
//                     nsm_core__initialize_diff_propensities
//                     (ENG/src/simulate_rdme.cpp:113-152)
//   rdme_queue_build  nsm_core__initialize_heap + NRMConstant_v5::build
//                     (simulate_rdme.cpp:155-195, NRMConstant_v5.cpp:27-85)
//   rdme_take_step    nsm_core__take_step (simulate_rdme.cpp:211-472)
//
// The event queue is an indexed min-tree over the voxels' firing times:
// F0 = 64*PL voxels under each level-1 node, 64 nodes under each node above
// (one wavefront reduces one node).  The levels above the voxels live in LDS
// while the event loop runs.  The reference's time-binned hash
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
#include "vm.hpp"

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

// The event loop's stream, drawn 64 words at a time: lane i holds the
// uniform and the exponential variate of word base + i, computed in
// parallel (Philox + log for the seeded stream; the supplied words and
// host-computed exponentials in replay mode); a draw is one v_readlane pair.
// Every word is consumed either as a uniform or as an exponential, in the
// reference's draw order, so the values are those of Rng above.
// The stream's constant parameters, in LDS for the whole launch (read only
// when a new batch of 64 words is drawn, so no registers are held for them).
struct RngParams {
    const uint64_t *words;
    const double *expo;
    unsigned long long n_words;
    uint64_t seed, stream;
    int kind;
};
__shared__ RngParams g_rngp;

// Lane i's uniform (x) and exponential (y) variates of word base + i.
__device__ __noinline__ double2 rng_fill(unsigned long long base, int lane) {
    const RngParams p = g_rngp;
    const unsigned long long k = base + (unsigned long long)lane;
    double2 r;
    if (p.kind == 1) {
        const bool ok = k < p.n_words;
        r.x = ok ? word_uniform(p.words[k]) : 0.5;
        r.y = ok ? p.expo[k] : 1.0;
    } else {
        const uint64_t w = philox64(p.seed, p.stream, k);
        r.x = word_uniform(w);
        r.y = word_expo(w);
    }
    return r;
}

struct WaveRng {
    unsigned long long n_words;  // replay: words available (kind 1), else ~0
    unsigned long long pos;      // next word
    unsigned long long base;     // word held by lane 0
    double u_l, e_l;
    bool exhausted;
    __device__ __forceinline__ void refill(int lane) {
        base = pos & ~63ULL;
        const double2 r = rng_fill(base, lane);
        u_l = r.x;
        e_l = r.y;
    }
    __device__ __forceinline__ void init(int lane) { refill(lane); }
    __device__ __forceinline__ int slot(int lane) {
        if (pos - base >= 64) refill(lane);
        return (int)(pos - base);
    }
    __device__ __forceinline__ double pick(double x, int s) {
        int lo = __builtin_amdgcn_readlane(__double2loint(x), s);
        int hi = __builtin_amdgcn_readlane(__double2hiint(x), s);
        return __hiloint2double(hi, lo);
    }
    __device__ __forceinline__ double uniform(int lane) {
        if (pos >= n_words) { exhausted = true; return 0.5; }
        const int s = slot(lane);
        pos++;
        return pick(u_l, s);
    }
    __device__ __forceinline__ double exponential(int lane) {
        if (pos >= n_words) { exhausted = true; return 1.0; }
        const int s = slot(lane);
        pos++;
        return pick(e_l, s);
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
// Rates (simulate_rdme.cpp:113-152), one thread per voxel.  VM: kind-4
// reactions through the expression VM.
template <bool VM>
__global__ void k_rates(RdmeArgs a) {
    const int chunk = xcd_chunk((a.n + 255) / 256);
    if (chunk < 0) return;
    int i = chunk * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int b = blockIdx.y;
    uint32_t *xx = a.xx + b * a.bs_xx;
    double *rrate = a.rrate + b * a.bs_rr, *srrate = a.srrate + b * a.bs_n, *sdrate = a.sdrate + b * a.bs_n;
    const int S = a.S, R = a.R;
    uint32_t xv[MAXS];
    for (int s = 0; s < S; s++) xv[s] = xx[(size_t)i * S + s];
    double sr = 0.0;
    const double *df = a.data_fn ? a.data_fn + (size_t)i * a.n_df : nullptr;
    // mass / rho of this step, formed by the particle phase (simulate_rdme.cpp:123)
    const double vol = a.vol[i];
    for (int j = 0; j < R; j++) {
        const Rxn re = a.rx[j];
        double p;
        if (VM && re.kind == 4)   // simulate_rdme.cpp:124 evaluates at t = 0.0
            p = (re.mask && !((re.mask >> a.type[i]) & 1ULL))
                    ? 0.0 : vm_eval_u32(a.vm_code, re.a, a.vm_k, xx + (size_t)i * S, vol, 0.0, df);
        else
            p = propensity(re, xv, vol, a.type[i]);
        rrate[(size_t)i * R + j] = p;
        sr += p;
    }
    srrate[i] = sr;
    double sd = 0.0;
    for (int s = 0; s < S; s++) sd += a.Ddiag[(size_t)i * S + s] * xv[s];
    sdrate[i] = sd;
}
// The per-species diffusion sums over the neighbour row (Ddiag) of models
// with more species than the search's finish forms, shared by every
// trajectory, from this step's lists; the particle phase runs it (beside
// the previous step's event loop when a look-ahead runs it).
__global__ void k_diag(RdmeArgs a) {
    const int chunk = xcd_chunk((a.n + 255) / 256);
    if (chunk < 0) return;
    const int i = chunk * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int dg = a.deg[i];
    for (int s = 0; s < a.S; s++) {
        double dd = 0.0;
        for (int k = 0; k < dg; k++) {
            const size_t e = (size_t)k * a.n + i;
            const double dc = a.Dmat[(size_t)s * a.T + (a.ntype[e] - 1)];   // ntype[e] = type of neighbour e
            dd += dc * a.Dij[e];
        }
        a.Ddiag[(size_t)i * a.S + s] = dd;
    }
}
void rdme_diag(hipStream_t s, const RdmeArgs &a) {
    hipLaunchKernelGGL(k_diag, xcd_grid((a.n + 255) / 256), dim3(256), 0, s, a);
}
void rdme_rates(hipStream_t s, const RdmeArgs &a, int n_traj) {
    const dim3 g(xcd_grid((a.n + 255) / 256).x, n_traj);
    if (a.vm_ncode > 0) hipLaunchKernelGGL(k_rates<true>, g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_rates<false>, g, dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// Time window (NRMConstant_v5::setBinNumberAndBounds, NRMConstant_v5.cpp:245-281).
__device__ bool set_bins(QueueState &q, double newL, double psum, long long active) {
    if (active == 0) return false;
    q.L = newL;
    if (q.L > q.endTime) return false;
    double w = 16.0 / psum;
    // The empty asm keeps the square root here: without it the compiler
    // hoists it out of the rarely taken window rebuild into every event of
    // the event loop, since `active` does not change inside the rebuild.
    double act = (double)(int)active;
    asm volatile("" : "+v"(act));
    unsigned long long maxb = (unsigned long long)((double)20 * sqrt(act));
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

// Partial sums for the fast (seeded) mode: fixed block/lane order, so the
// result is deterministic.
constexpr int QSUM_BLOCKS = 256;
__global__ void k_qsum_part(const double *prop, int n, double *part) {
    __shared__ double w[4];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += prop[i];
    for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0) w[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((w[0] + w[1]) + w[2]) + w[3];
}

// propensitySum (simulate_rdme.cpp:167-179) and the first time window.
// exact: added in voxel order by one wavefront (bit-identical to the
// reference's loop); otherwise the block partials are added in order.
__global__ void k_qsum(RdmeArgs a, const double *prop, const int *act_scan, int exact, const double *part) {
    const int lane = threadIdx.x;
    double s = 0.0;
    if (exact) {
        for (int base = 0; base < a.n; base += WAVE) {
            double v = base + lane < a.n ? prop[base + lane] : 0.0;
            int m = min(WAVE, a.n - base);
            for (int k = 0; k < m; k++) s += __shfl(v, k, WAVE);
        }
    } else {
        for (int k = 0; k < QSUM_BLOCKS; k++) s += part[k];
    }
    if (lane == 0) {
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

// Argmin of (time, voxel) pairs: ties go to the lower voxel index.
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
// level 1 from the firing times: one wavefront per node of F0 = 64*pl voxels
__global__ void k_tree_level1(const double *qt, int n, int pl, double *pt, int *pa, int nparent) {
    int node = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    int lane = threadIdx.x & (WAVE - 1);
    if (node >= nparent) return;
    double t = INFINITY;
    int v = INT_MAX;
    int base = node * WAVE * pl + lane * pl;
    for (int p = 0; p < pl; p++) {
        int c = base + p;
        if (c < n) argmin2(t, v, qt[c], c);
    }
    wave_argmin(t, v);
    if (lane == 0) { pt[node] = t; pa[node] = v; }
}
// upper level: 64 children per node
__global__ void k_tree_level(const double *ct, const int *ca, int nchild, double *pt, int *pa, int nparent) {
    int node = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    int lane = threadIdx.x & (WAVE - 1);
    if (node >= nparent) return;
    int c = node * WAVE + lane;
    double t = INFINITY;
    int v = INT_MAX;
    if (c < nchild) { t = ct[c]; v = ca[c]; }
    wave_argmin(t, v);
    if (lane == 0) { pt[node] = t; pa[node] = v; }
}

void rdme_queue_build(hipStream_t s, const RdmeArgs &a, void *scan_tmp, int *work) {
    // work: prop (n doubles) | nz (n+1) | act (n+1) | nz_scan (n+1) | act_scan (n+1) | part
    double *prop = (double *)work;
    int *nz = (int *)(prop + a.n);
    int *act = nz + a.n + 1;
    int *nz_scan = act + a.n + 1;
    int *act_scan = nz_scan + a.n + 1;
    double *part = (double *)(((uintptr_t)(act_scan + a.n + 1) + 15) & ~(uintptr_t)15);
    unsigned g = (unsigned)((a.n + 255) / 256);
    const bool exact = a.rng_kind == 1;
    hipLaunchKernelGGL(k_qprop, dim3(g), dim3(256), 0, s, a, prop, nz, act);
    exclusive_scan_i32(s, nz, nz_scan, a.n, scan_tmp);
    exclusive_scan_i32(s, act, act_scan, a.n, scan_tmp);
    if (!exact) hipLaunchKernelGGL(k_qsum_part, dim3(QSUM_BLOCKS), dim3(256), 0, s, prop, a.n, part);
    hipLaunchKernelGGL(k_qsum, dim3(1), dim3(WAVE), 0, s, a, prop, act_scan, exact ? 1 : 0, part);
    hipLaunchKernelGGL(k_qtimes, dim3(g), dim3(256), 0, s, a, prop, nz_scan);
    hipLaunchKernelGGL(k_qfinish, dim3(1), dim3(64), 0, s, a, nz_scan);
    tree_rebuild(s, a);
}

void tree_rebuild(hipStream_t s, const RdmeArgs &a) {
    int n1 = a.lev_n[1];
    hipLaunchKernelGGL(k_tree_level1, dim3((unsigned)((n1 + 3) / 4)), dim3(256), 0, s, a.qtime, a.n, a.pl,
                       a.tmin + a.lev_off[1], a.targ + a.lev_off[1], n1);
    for (int k = 1; k < a.nlev; k++) {
        int np = a.lev_n[k + 1];
        hipLaunchKernelGGL(k_tree_level, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, s, a.tmin + a.lev_off[k],
                           a.targ + a.lev_off[k], a.lev_n[k], a.tmin + a.lev_off[k + 1], a.targ + a.lev_off[k + 1], np);
    }
}

// ---------------------------------------------------------------------------
// The event loop.

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

// LDS image of the model constants and the upper tree levels.
struct Lds {
    double *lt;     // upper-level times (offsets lev_off)
    int *lv;        // upper-level voxels
    Rxn *rx;
    int *jcN, *irN, *prN, *jcG, *irG;
    double *Dmat;
};

__device__ __forceinline__ size_t al16(size_t x) { return (x + 15) & ~(size_t)15; }

// Cross-lane moves through DPP (no LDS round trip).  Lanes whose source is
// outside the pattern keep their own value.
template <int CTRL, int RMASK>
__device__ __forceinline__ int dpp_i(int x) {
    return __builtin_amdgcn_update_dpp(x, x, CTRL, RMASK, 0xF, false);
}
template <int CTRL, int RMASK>
__device__ __forceinline__ double dpp_d(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    if (RMASK == 0xF)   // every lane is written: no copy of the old value
        return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false),
                                __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false));
    return __hiloint2double(dpp_i<CTRL, RMASK>(hi), dpp_i<CTRL, RMASK>(lo));
}
template <int CTRL, int RMASK>
__device__ __forceinline__ void dpp_step(double &t, int &v) {
    double t2 = dpp_d<CTRL, RMASK>(t);
    int v2 = dpp_i<CTRL, RMASK>(v);
    argmin2(t, v, t2, v2);
}
__device__ __forceinline__ double rl_d(double x, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}
template <int CTRL, int RMASK>
__device__ __forceinline__ double dpp_min(double t) {
    return fmin(t, dpp_d<CTRL, RMASK>(t));
}
// Minimum over the wavefront (DPP + v_min_f64), returned to every lane.
__device__ __forceinline__ double wave_min_dpp(double m) {
    m = dpp_min<0xB1, 0xF>(m);    // quad_perm [1,0,3,2]
    m = dpp_min<0x4E, 0xF>(m);    // quad_perm [2,3,0,1]
    m = dpp_min<0x141, 0xF>(m);   // row_half_mirror
    m = dpp_min<0x140, 0xF>(m);   // row_mirror
    m = dpp_min<0x142, 0xA>(m);   // row_bcast15 into rows 1, 3
    m = dpp_min<0x143, 0xC>(m);   // row_bcast31 into rows 2, 3
    return rl_d(m, 63);
}
// argmin over the wavefront; every lane gets the result.  The minimum time
// is reduced alone (DPP + v_min_f64); the winner is the lowest lane holding
// it, and lanes hold increasing voxel ranges, so ties go to the lower voxel.
__device__ __forceinline__ void wave_argmin_dpp(double &t, int &v) {
    const double m = wave_min_dpp(t);
    const unsigned long long hit = __ballot(t == m);
    const int w = hit ? (int)__builtin_ctzll(hit) : 0;
    v = __builtin_amdgcn_readlane(v, w);
    t = m;
}

// Small per-voxel vectors are held one element per lane ("lane s holds
// species s"); an element at a uniform index is read with v_readlane.
__device__ __forceinline__ uint32_t rl_u(uint32_t x, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}

// Mass-action propensity with the counts distributed over the lanes.
__device__ __forceinline__ double propensity_l(const Rxn &re, uint32_t x_l, double vol, int tp) {
    if (re.mask && !((re.mask >> tp) & 1ULL)) return 0.0;
    const int xa = (int)rl_u(x_l, re.a);
    switch (re.kind) {
    case 0: return re.k * vol;
    case 1: return re.k * xa;
    case 2: return ((re.k * xa) * (int)rl_u(x_l, re.b)) / vol;
    default: return ((re.k * xa) * (xa - 1)) / vol;
    }
}

// The model's small tables held one row per lane for the whole run: lane j
// has reaction j's kind, reactants, rate and type mask; lane c < S+R has
// the dependency-graph column c as a bitmask of reactions (irG is ascending
// within a column); lane s has species s's stoichiometric change under each
// reaction.
template <int MR> struct LaneModel {
    int kind, ra, rb, pc;   // pc: program start of a kind-4 reaction
    double k;
    uint64_t mask;
    uint32_t dep;
    int dN[MR];
};
// Propensity of reaction `lane` from lane-distributed counts: every lane
// fetches its reactants' counts with one bpermute each.  Same operations as
// propensity_l, so the same bits.
template <int MR>
__device__ __forceinline__ double propensity_par(const LaneModel<MR> &m, uint32_t x_l, double vol, int tp) {
    const int xa = __shfl((int)x_l, m.ra, WAVE);
    const int xb = __shfl((int)x_l, m.rb, WAVE);
    // every kind's operations on every lane, then a select: lanes of
    // different kinds share one division instead of diverging
    const double t1 = m.k * xa;                                  // kinds 1-3
    const double t2 = t1 * (m.kind == 2 ? xb : xa - 1);          // kinds 2, 3
    const double t3 = t2 / vol;
    double p = m.kind == 0 ? m.k * vol : (m.kind == 1 ? t1 : t3);
    if (m.mask && !((m.mask >> tp) & 1ULL)) p = 0.0;
    return p;
}
// propensity_par with kind-4 rows (custom propensities) from the expression
// VM, for the lanes in `want`: the counts are staged in LDS (vmx) so that
// every lane's program can read any species; arguments as the generated
// functions get them (counts, t, vol, data_fn, type).  Lanes outside `want`
// hold values the callers mask off.
template <int MR>
__device__ __forceinline__ double propensity_vm(const LaneModel<MR> &m, uint32_t x_l, double vol, int tp, double tnow,
                                                const double *df, uint64_t want, uint32_t *vmx,
                                                const uint32_t *vcode, const double *vk, int lane) {
    double p = propensity_par(m, x_l, vol, tp);
    vmx[lane] = x_l;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    if (m.kind == 4 && ((want >> lane) & 1ULL))
        p = (m.mask && !((m.mask >> tp) & 1ULL)) ? 0.0 : vm_eval_u32(vcode, m.pc, vk, vmx, vol, tnow, df);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    return p;
}
// Sum over the set bits of `bits` in ascending lane order of lane-held
// values, added one by one (the reference's sequential order).
__device__ __forceinline__ double ordered_sum(double acc, double x, uint64_t bits) {
    while (bits) {
        const int j = (int)__builtin_ctzll(bits);
        bits &= bits - 1;
        acc += rl_d(x, j);
    }
    return acc;
}

// Level-0 children of one level-1 node held in registers: lane l owns the
// voxels base + l*PL .. base + l*PL + PL-1.
template <int PL> struct Sib {
    double t[PL];
    int base;
    __device__ __forceinline__ void load(const double *qt, int n, int node, int lane) {
        base = node * WAVE * PL;
        // unconditional loads (clamped index) keep the round trip in one
        // basic block; children past n read as +inf
        double raw[PL];
#pragma unroll
        for (int p = 0; p < PL; p++) {
            const int c = base + lane * PL + p;
            raw[p] = qt[c < n ? c : n - 1];
        }
#pragma unroll
        for (int p = 0; p < PL; p++) t[p] = base + lane * PL + p < n ? raw[p] : INFINITY;
    }
    __device__ __forceinline__ bool holds(int v) const { return v >= base && v < base + WAVE * PL; }
    __device__ __forceinline__ void set(int v, double tv, int lane) {
        int off = v - base;
#pragma unroll
        for (int p = 0; p < PL; p++)
            if (lane * PL + p == off) t[p] = tv;
    }
    // Minimum of the node and the lowest voxel holding it: a pairwise min
    // tree over the lane's PL times, the wave minimum by DPP, then each
    // lane's first slot holding it and one ballot for the first such lane
    // (voxel base + lane*PL + p; times past n are +inf, and an all-+inf node
    // reports its first voxel).
    __device__ __forceinline__ void argmin(double &m, int &am, int lane, int n) const {
        double u[PL];
#pragma unroll
        for (int p = 0; p < PL; p++) u[p] = t[p];
#pragma unroll
        for (int st = 1; st < PL; st <<= 1)
#pragma unroll
            for (int i = 0; i + st < PL; i += 2 * st) u[i] = fmin(u[i], u[i + st]);
        m = wave_min_dpp(u[0]);
        int slot = PL;
#pragma unroll
        for (int p = PL - 1; p >= 0; p--)
            if (t[p] == m) slot = p;
        const unsigned long long h = __ballot(slot < PL);
        const int w = h ? (int)__builtin_ctzll(h) : 0;
        am = base + w * PL + (h ? __builtin_amdgcn_readlane(slot, w) : 0);
        (void)lane;
        (void)n;
    }
};

// The event tree's shape for the loop: lane k holds the offset (lo_l) and
// node count (ln_l) of level k, read with v_readlane (no scalar registers
// held for the tables).
struct TreeShape {
    int n, nlev;
    int lo_l, ln_l;
    __device__ __forceinline__ int off(int k) const { return __builtin_amdgcn_readlane(lo_l, k); }
    __device__ __forceinline__ int cnt(int k) const { return __builtin_amdgcn_readlane(ln_l, k); }
};

// The top of the event tree also lives in registers: lane l holds node l
// of level regl, the level just below the root (at most 64 nodes; level 1
// when there are only one or two levels), and the root is uniform.  LDS
// keeps every level (written back at the end of a launch); reads of these
// two levels come from registers.
struct TreeTop {
    int regl, root;          // register level, LDS index of the root
    double rt;               // lane l: node l of level regl (+inf past the count)
    int rv;
    double root_t;           // the root: the global minimum and its voxel
    int root_v;
};

// Node idx of level regl changed from (old_m, old_am) to (m, am): update its
// lane and the root, rescanning the register level only when the root's
// minimum came from this node and got later.
__device__ __forceinline__ void top_set(TreeTop &tp, Lds &L, int idx, double m, int am, double old_m, int old_am,
                                        int lane) {
    (void)old_m;
    if (lane == idx) { tp.rt = m; tp.rv = am; }
    double nm;
    int nam;
    if (tp.root_v == old_am && (m > tp.root_t || (m == tp.root_t && am != tp.root_v))) {
        nm = tp.rt;
        nam = tp.rv;
        wave_argmin_dpp(nm, nam);
    } else {
        nm = tp.root_t;
        nam = tp.root_v;
        argmin2(nm, nam, m, am);
    }
    tp.root_t = nm;
    tp.root_v = nam;
    if (lane == 0) { L.lt[tp.root] = nm; L.lv[tp.root] = nam; }
}

// Repair the levels above a changed level-1 node value (idx -> (m, am)),
// rescanning a node only when its stored minimum may have increased.
template <int PL>
__device__ void tree_fix_up(const TreeShape &a, Lds &L, TreeTop &tp, int idx, double m, int am, double old_m,
                            int old_am, int lane) {
    for (int k = 1; k < tp.regl; k++) {
        const int p = idx / WAVE;
        const int poff = a.off(k + 1) + p;
        double pm;
        int pam;
        if (k + 1 == tp.regl) {
            pm = rl_d(tp.rt, p);
            pam = __builtin_amdgcn_readlane(tp.rv, p);
        } else {
            pm = L.lt[poff];
            pam = L.lv[poff];
        }
        double nm;
        int nam;
        if (pam == old_am && (m > pm || (m == pm && am != pam))) {
            // the parent's argmin came from this child and got worse: rescan
            int c = p * WAVE + lane;
            double tt = INFINITY;
            int vv = INT_MAX;
            if (c < a.cnt(k)) {
                if (c == idx) { tt = m; vv = am; }
                else { tt = L.lt[a.off(k) + c]; vv = L.lv[a.off(k) + c]; }
            }
            wave_argmin_dpp(tt, vv);
            nm = tt;
            nam = vv;
        } else {
            nm = pm;
            nam = pam;
            argmin2(nm, nam, m, am);
        }
        if (nm == pm && nam == pam) return;   // nothing above changes
        if (lane == 0) { L.lt[poff] = nm; L.lv[poff] = nam; }
        old_m = pm;
        old_am = pam;
        m = nm;
        am = nam;
        idx = p;
    }
    top_set(tp, L, idx, m, am, old_m, old_am, lane);
}

// Level-1 node idx's stored value: registers when level 1 is the register
// level, else LDS.
__device__ __forceinline__ void level1_get(const TreeShape &a, const Lds &L, const TreeTop &tp, int idx, double &om,
                                           int &oam) {
    if (tp.regl == 1) {
        om = rl_d(tp.rt, idx);
        oam = __builtin_amdgcn_readlane(tp.rv, idx);
    } else {
        om = L.lt[a.off(1) + idx];
        oam = L.lv[a.off(1) + idx];
    }
}

// Voxel v gets time tv; its level-0 node is in sib (already holding tv).
template <int PL>
__device__ void tree_update_rescan(const TreeShape &a, double *qtime, Lds &L, TreeTop &tp, const Sib<PL> &sib, int v,
                                   double tv, int lane) {
    if (lane == 0) qtime[v] = tv;
    double m;
    int am;
    sib.argmin(m, am, lane, a.n);
    const int idx = v / (WAVE * PL);
    const int off = a.off(1) + idx;
    double om;
    int oam;
    level1_get(a, L, tp, idx, om, oam);
    if (lane == 0) { L.lt[off] = m; L.lv[off] = am; }
    if (!(m == om && am == oam)) tree_fix_up<PL>(a, L, tp, idx, m, am, om, oam, lane);
}

// The selected voxel v is the global minimum, so every ancestor's minimum
// comes from it, and its new time tv rescans all of them.  Each ancestor's
// new minimum is min((v's new node minimum), (the ancestor's minimum over
// its other children)), and those other-children minima read only LDS and
// registers: pop_prepare computes them while the event's first memory round
// trip is in flight, pop_finish merges them once the level-0 node is in.
// (time, voxel) pairs merge lexicographically, which is the lowest-lane rule
// of the wave reductions, since lanes hold increasing voxel ranges.  Level 1
// holds at most 4096 nodes, so the register level is level 1 or 2.
struct PopPre {
    double t1;   // the level-1 node's other children (register level 2 only)
    int v1;
    double t2;   // the register level's other nodes
    int v2;
};

template <int PL>
__device__ __forceinline__ PopPre pop_prepare(const TreeShape &a, const Lds &L, const TreeTop &tp, int v, int lane) {
    PopPre r;
    const int idx = v / (WAVE * PL);
    int top = idx;   // v's ancestor at the register level
    r.t1 = INFINITY;
    r.v1 = INT_MAX;
    if (tp.regl != 1) {
        top = idx / WAVE;
        const int c = top * WAVE + lane;
        double t = INFINITY;
        int vv = INT_MAX;
        if (c < a.cnt(1) && c != idx) { t = L.lt[a.off(1) + c]; vv = L.lv[a.off(1) + c]; }
        wave_argmin_dpp(t, vv);
        r.t1 = t;
        r.v1 = vv;
    }
    double t = lane == top ? INFINITY : tp.rt;
    int vv = lane == top ? INT_MAX : tp.rv;
    wave_argmin_dpp(t, vv);
    r.t2 = t;
    r.v2 = vv;
    return r;
}

template <int PL>
__device__ __forceinline__ void pop_finish(const TreeShape &a, double *qtime, Lds &L, TreeTop &tp, const Sib<PL> &sib,
                                           int v, double tv, const PopPre &pre, int lane) {
    if (lane == 0) qtime[v] = tv;
    double m;
    int am;
    sib.argmin(m, am, lane, a.n);
    const int idx = v / (WAVE * PL);
    if (lane == 0) { L.lt[a.off(1) + idx] = m; L.lv[a.off(1) + idx] = am; }
    int top = idx;
    if (tp.regl != 1) {
        argmin2(m, am, pre.t1, pre.v1);   // the level-1 node's new minimum
        top = idx / WAVE;
        if (lane == 0) { L.lt[a.off(2) + top] = m; L.lv[a.off(2) + top] = am; }
    }
    if (lane == top) { tp.rt = m; tp.rv = am; }
    argmin2(m, am, pre.t2, pre.v2);       // the root
    tp.root_t = m;
    tp.root_v = am;
    if (lane == 0) { L.lt[tp.root] = m; L.lv[tp.root] = am; }
}

// Voxel v gets time tv without its level-0 node in registers: the node's
// minimum is updated in place unless v was the node's argmin and got later,
// in which case the node is reloaded and rescanned.
template <int PL>
__device__ void tree_update_point(const TreeShape &a, double *qtime, Lds &L, TreeTop &tp, int v, double tv, int lane) {
    if (lane == 0) qtime[v] = tv;
    const int idx = v / (WAVE * PL);
    const int off = a.off(1) + idx;
    double om;
    int oam;
    level1_get(a, L, tp, idx, om, oam);
    double m;
    int am;
    if (oam == v && tv > om) {
        Sib<PL> sb;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        sb.load(qtime, a.n, idx, lane);
        sb.set(v, tv, lane);
        sb.argmin(m, am, lane, a.n);
    } else if (oam == v) {
        m = tv;
        am = v;
    } else {
        m = om;
        am = oam;
        argmin2(m, am, tv, v);
    }
    if (m == om && am == oam) return;
    if (lane == 0) { L.lt[off] = m; L.lv[off] = am; }
    tree_fix_up<PL>(a, L, tp, idx, m, am, om, oam, lane);
}

// Snapshot k of one trajectory: its counts [n][S] and its species totals
// (lane s holds total s).
__device__ __noinline__ void snapshot_counts(const uint32_t *xx, uint32_t *snap_xx, uint32_t *snap_tot, size_t nsS,
                                             int S, int k, uint32_t tot, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    if (snap_xx) {
        uint32_t *dst = snap_xx + (size_t)k * nsS;
        for (size_t i = lane; i < nsS; i += WAVE) dst[i] = xx[i];
    }
    if (snap_tot && lane < S) snap_tot[(size_t)k * S + lane] = tot;
}

// The event loop of one trajectory on one wavefront.  b: its index in a
// batch that shares one argument block (strides bs_*), else 0.
// VM: kind-4 reactions (custom propensities) through the expression VM; the
// kernels without it keep the mass-action table's register allocation.
template <int PL, int MS, int MR, bool VM>
__device__ __forceinline__ void ssa_run(const RdmeArgs &a, const size_t b) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    // this trajectory's arrays (the argument block itself stays read-only,
    // so it is never copied to private memory)
    uint32_t *const xx_ = a.xx + b * a.bs_xx;
    double *const rrate_ = a.rrate + b * a.bs_rr;
    double *const srrate_ = a.srrate + b * a.bs_n;
    double *const sdrate_ = a.sdrate + b * a.bs_n;
    double *const qtime_ = a.qtime + b * a.bs_n;
    double *const tmin_ = a.tmin + b * a.bs_tree;
    int *const targ_ = a.targ + b * a.bs_tree;
    QueueState *const qs_ = a.qs + b;
    const unsigned long long seed_ = a.seeds ? a.seeds[b] : a.seed;
    QueueState q = *qs_;
    if (q.initialized != 1 || q.err) return;
    const int S = a.S, R = a.R, n = a.n, T = a.T;

    // ---- LDS image
    Lds L;
    uint32_t *vcode = nullptr, *vmx = nullptr;   // VM programs, staged counts
    double *vk = nullptr;                        // VM constants
    {
        size_t off = 0;
        L.lt = (double *)(smem + off); off = al16(off + sizeof(double) * a.tree_nodes);
        L.lv = (int *)(smem + off); off = al16(off + sizeof(int) * a.tree_nodes);
        L.rx = (Rxn *)(smem + off); off = al16(off + sizeof(Rxn) * R);
        L.Dmat = (double *)(smem + off); off = al16(off + sizeof(double) * S * T);
        int nnzN = R > 0 ? a.jcN[R] : 0, nnzG = a.jcG[S + R];
        L.jcN = (int *)(smem + off); off = al16(off + sizeof(int) * (R + 1));
        L.irN = (int *)(smem + off); off = al16(off + sizeof(int) * nnzN);
        L.prN = (int *)(smem + off); off = al16(off + sizeof(int) * nnzN);
        L.jcG = (int *)(smem + off); off = al16(off + sizeof(int) * (S + R + 1));
        L.irG = (int *)(smem + off); off = al16(off + sizeof(int) * nnzG);
        if (VM && a.vm_ncode > 0) {
            vcode = (uint32_t *)(smem + off); off = al16(off + sizeof(uint32_t) * a.vm_ncode);
            vk = (double *)(smem + off); off = al16(off + sizeof(double) * a.vm_nk);
            vmx = (uint32_t *)(smem + off); off = al16(off + sizeof(uint32_t) * WAVE);
            for (int i = lane; i < a.vm_ncode; i += WAVE) vcode[i] = a.vm_code[i];
            for (int i = lane; i < a.vm_nk; i += WAVE) vk[i] = a.vm_k[i];
        }
        for (int i = lane; i < a.tree_nodes; i += WAVE) { L.lt[i] = tmin_[i]; L.lv[i] = targ_[i]; }
        for (int i = lane; i < R; i += WAVE) L.rx[i] = a.rx[i];
        for (int i = lane; i < S * T; i += WAVE) L.Dmat[i] = a.Dmat[i];
        for (int i = lane; i <= R && R > 0; i += WAVE) L.jcN[i] = a.jcN[i];
        for (int i = lane; i < nnzN; i += WAVE) { L.irN[i] = a.irN[i]; L.prN[i] = a.prN[i]; }
        for (int i = lane; i <= S + R; i += WAVE) L.jcG[i] = a.jcG[i];
        for (int i = lane; i < nnzG; i += WAVE) L.irG[i] = a.irG[i];
        __syncthreads();
    }
    const int root = a.lev_off[a.nlev];
    const int W = a.W;     // row stride of erow
    const int RW = a.row_w;   // lanes that load a row (no row is longer)
    TreeShape tsh;
    tsh.n = n;
    tsh.nlev = a.nlev;
    tsh.lo_l = 0;
    tsh.ln_l = 0;
#pragma unroll
    for (int k = 0; k < 10; k++)
        if (lane == k) { tsh.lo_l = a.lev_off[k]; tsh.ln_l = a.lev_n[k]; }
    TreeTop tp;
    {
        tp.regl = a.nlev > 1 ? a.nlev - 1 : 1;
        tp.root = root;
        const int cnt = tsh.cnt(tp.regl), off = tsh.off(tp.regl);
        tp.rt = lane < cnt ? L.lt[off + lane] : INFINITY;
        tp.rv = lane < cnt ? L.lv[off + lane] : INT_MAX;
        tp.root_t = L.lt[root];
        tp.root_v = L.lv[root];
    }
    LaneModel<MR> lm;
    {
        lm.kind = 0; lm.ra = 0; lm.rb = 0; lm.pc = 0; lm.k = 0.0; lm.mask = 0; lm.dep = 0;
        if (lane < R) {
            const Rxn r = L.rx[lane];
            lm.kind = r.kind; lm.ra = r.a; lm.rb = r.b; lm.k = r.k; lm.mask = r.mask;
            if (r.kind == 4) { lm.pc = r.a; lm.ra = 0; lm.rb = 0; }
        }
        if (lane < S + R)
            for (int k = L.jcG[lane]; k < L.jcG[lane + 1]; k++) lm.dep |= 1u << L.irG[k];
#pragma unroll
        for (int r = 0; r < MR; r++) {
            int d = 0;
            if (r < R)
                for (int k = L.jcN[r]; k < L.jcN[r + 1]; k++)
                    if (L.irN[k] == lane) d = L.prN[k];
            lm.dN[r] = lane < S ? d : 0;
        }
    }

    if (lane == 0) {
        g_rngp.words = a.words; g_rngp.expo = a.expo; g_rngp.n_words = a.n_words;
        g_rngp.seed = seed_; g_rngp.stream = a.stream_id; g_rngp.kind = a.rng_kind;
    }
    __syncthreads();
    WaveRng rng;
    rng.n_words = a.rng_kind == 1 ? a.n_words : ~0ULL;
    rng.pos = q.rng_pos;
    rng.exhausted = false;
    rng.init(lane);
    const bool textbook = a.textbook != 0;
    const bool has_prog = VM && a.vm_ncode > 0;
    const bool lead = lane == 0;
    const bool s_lane = lane < S, r_lane = lane < R;

    int err = 0;
    unsigned long long n_react = 0, n_diff = 0;
    const double inf = INFINITY;

    // species totals, lane s holds species s (persistent mode snapshots)
    uint32_t tot = 0;
    if (a.snap_tot) {
        uint32_t part[MS];
#pragma unroll
        for (int u = 0; u < MS; u++) part[u] = 0;
        const size_t ns = (size_t)n * S;
        for (size_t i = lane; i < ns; i += WAVE) {
            const int sp = (int)(i % (size_t)S);
            const uint32_t xv = xx_[i];
#pragma unroll
            for (int u = 0; u < MS; u++)
                if (u == sp) part[u] += xv;
        }
#pragma unroll
        for (int u = 0; u < MS; u++) {
            uint32_t v = part[u];
            for (int o = WAVE / 2; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, WAVE);
            if (lane == u) tot = v;
        }
    }
    int ks = 0;
    const size_t nsS = (size_t)n * S;
    uint32_t *const snap_xx = a.snap_xx ? a.snap_xx + b * a.bs_snap : nullptr;
    uint32_t *const snap_tot = a.snap_tot ? a.snap_tot + b * a.n_snap * S : nullptr;
    const uint32_t *const xx_all = xx_;

    const unsigned step_end = a.step_end > a.step ? a.step_end : a.step + 1;
    for (unsigned stp = a.step; stp < step_end; stp++) {
    while (ks < a.n_snap && a.snap_step[ks] <= stp) snapshot_counts(xx_all, snap_xx, snap_tot, nsS, S, ks++, tot, lane);
    // one step of simulate_rdme (simulate_rdme.cpp:53-68): events while the
    // last one fired at or before the step's end; the destination carried
    // between diffusion events restarts each step
    const double t0 = a.step_end > a.step ? (double)stp * a.dt : a.t0;
    double tt = t0;
    const double end_time = t0 + a.dt;
    int dest = -1;
    unsigned long long guard = 0;
    while (tt <= end_time) {
        if (++guard > a.max_events) { err = SSASDPD_ERR_CAPACITY; break; }
        // ---- selectReaction: global minimum inside the window [L, U]
        double t = tp.root_t;
        int sv = tp.root_v;
        bool found = false;
        for (int it = 0; it < 4096; it++) {
            if (sv != INT_MAX && t <= q.U) { found = true; break; }
            if (!rebuild(q)) break;
            if (!(t <= q.U)) break;   // window moved but still empty: empty queue this call
        }
        if (!found) break;
        q.rxnCount++;
        tt = t;
        // ---- round trip 1: the voxel's state, its neighbour row and its
        // level-0 node
        Sib<PL> sib;
        sib.load(qtime_, n, sv / (WAVE * PL), lane);
        const double vol = a.vol[sv];
        const double *const df_sv = a.data_fn ? a.data_fn + (size_t)sv * a.n_df : nullptr;
        double sr = srrate_[sv], sd = sdrate_[sv];
        const int tsv = a.type[sv];
        const int dg_l = a.deg[sv];
        // lane s: count of species s; lane j: rate of reaction j; lanes past
        // S / R / the longest row read a valid element and drop it
        const int ls = lane < S ? lane : S - 1, lr = lane < R ? lane : (R > 0 ? R - 1 : 0);
        const int lw = lane < RW ? lane : RW - 1;
        const uint32_t xs_raw = xx_[(size_t)sv * S + ls];
        const double Dd_raw = a.Ddiag[(size_t)sv * S + ls];
        const double rr_raw = R > 0 ? rrate_[(size_t)sv * R + lr] : 0.0;
        const EdgeRec er = a.erow[(size_t)sv * W + lw];
        const int nbj_raw = er.j;
        const double nbD_raw = er.D;
        const int nbt_raw = er.type;
        // while the loads above are in flight: the selected voxel's ancestors
        // without it, and the event's first random draw
        const PopPre pre = pop_prepare<PL>(tsh, L, tp, sv, lane);
        const double rand1 = rng.uniform(lane);   // the event's first draw
        if (rng.exhausted) { err = SSASDPD_ERR_REPLAY_EXHAUSTED; break; }
        uint32_t xs = s_lane ? xs_raw : 0u;
        const double Dd = s_lane ? Dd_raw : 0.0;
        double rr = r_lane ? rr_raw : 0.0;
        const int nb_j = lane < RW ? nbj_raw : -1;
        const double nb_D = lane < RW ? nbD_raw : 0.0;
        const int nb_t = (lane < RW && T > 1) ? nbt_raw : 1;
        double totrate = sr + sd;
        int event;
        int which;
        if (rand1 <= sr / totrate) {
            // ---- reaction (simulate_rdme.cpp:255-310)
            event = 0;
            double rv = textbook ? rand1 * totrate : rand1 * sr;
            int re = 0;
            double cum = R > 0 ? rl_d(rr, 0) : 0.0;
            while (re < R && rv > cum) { re++; cum += re < R ? rl_d(rr, re) : 0.0; }
            if (re >= R) {
                if (cum != sr) {
                    double rd = 0.0;
                    if (has_prog) {
                        const uint64_t all_r = (1ULL << R) - 1ULL;
                        const double p = propensity_vm(lm, xs, vol, tsv, tt, df_sv, all_r, vmx, vcode, vk, lane);
                        rd = ordered_sum(0.0, p, all_r);
                        if (r_lane) rr = p;
                    } else {
                        for (int j = 0; j < R; j++) {
                            double p = propensity_l(L.rx[j], xs, vol, tsv);
                            if (lane == j) rr = p;
                            rd += p;
                        }
                    }
                    sr = rd;
                    if (r_lane) rrate_[(size_t)sv * R + lane] = rr;
                    if (lead) srrate_[sv] = sr;
                }
                if (sr == 0.0) continue;
                double rv2 = rand1 * sr;
                re = 0;
                cum = rl_d(rr, 0);
                while (re < R && rv2 > cum) { re++; cum += re < R ? rl_d(rr, re) : 0.0; }
                if (re >= R) { err = SSASDPD_ERR_PROPENSITY_OVERFLOW; break; }
            }
            {
                // stoichiometry column re, ascending species (CSC order)
                int d_l = 0;
#pragma unroll
                for (int r = 0; r < MR; r++)
                    if (r == re) d_l = lm.dN[r];
                uint64_t nz = __ballot(d_l != 0);
                xs += (uint32_t)d_l;
                tot += (uint32_t)d_l;
                while (nz) {
                    const int sp = (int)__builtin_ctzll(nz);
                    nz &= nz - 1;
                    sd += rl_d(Dd, sp) * __builtin_amdgcn_readlane(d_l, sp);
                }
            }
            // dependency-graph column S+re: all propensities at once, the
            // changes added in the column's order
            const uint64_t dep = (uint32_t)__builtin_amdgcn_readlane((int)lm.dep, S + re);
            const double p_new = has_prog ? propensity_vm(lm, xs, vol, tsv, tt, df_sv, dep, vmx, vcode, vk, lane)
                                          : propensity_par(lm, xs, vol, tsv);
            const double rd = ordered_sum(0.0, p_new - rr, dep);
            if ((dep >> lane) & 1ULL) rr = p_new;
            sr += rd;
            if (s_lane) xx_[(size_t)sv * S + lane] = xs;
            if (r_lane) rrate_[(size_t)sv * R + lane] = rr;
            if (lead) { srrate_[sv] = sr; sdrate_[sv] = sd; }
            n_react++;
            which = re;
            // ---- queue update (NRMConstant_v5::update, NRMConstant_v5.cpp:89-132)
            const double tot = sr + sd;
            double ntv;
            if (tot <= 0.0) { ntv = inf; q.active--; }   // the selected voxel's time was finite
            else ntv = rng.exponential(lane) / tot + tt;
            sib.set(sv, ntv, lane);
            pop_finish<PL>(tsh, qtime_, L, tp, sib, sv, ntv, pre, lane);
        } else {
            // ---- diffusion (simulate_rdme.cpp:312-445)
            event = 1;
            double drand = textbook ? (rand1 * totrate - sr) : rand1 * sd;
            int spec = 0;
            double cum = S > 0 ? rl_d(Dd, 0) * rl_u(xs, 0) : 0.0;
            while (spec < S && drand > cum) { spec++; cum += spec < S ? rl_d(Dd, spec) * rl_u(xs, spec) : 0.0; }
            if (spec >= S) {
                if (cum != sd) {
                    double rd = 0.0;
                    for (int j = 0; j < S; j++) rd += rl_d(Dd, j) * rl_u(xs, j);
                    sd = rd;
                    if (lead) sdrate_[sv] = sd;
                }
                if (sd == 0.0) continue;
                drand = cum * rand1;
                spec = 0;
                cum = rl_d(Dd, 0) * rl_u(xs, 0);
                while (spec < S && drand > cum) { spec++; cum += spec < S ? rl_d(Dd, spec) * rl_u(xs, spec) : 0.0; }
                if (spec >= S) {
                    spec--;
                    bool bad = false;
                    while (rl_u(xs, spec) <= 0) {
                        spec--;
                        if (spec <= 0) { bad = true; break; }
                    }
                    if (bad) { err = SSASDPD_ERR_DIFFUSION_OVERFLOW; break; }
                }
            }
            const int dg = __builtin_amdgcn_readfirstlane(dg_l);
            const double r2 = rng.uniform(lane);
            if (rng.exhausted) { err = SSASDPD_ERR_REPLAY_EXHAUSTED; break; }
            double rand2 = r2 * rl_d(Dd, spec);
            int kk = -1;             // chosen entry of the neighbour row
            if (dg <= WAVE) {
                // destination: walk the neighbour list (held in registers) in
                // order; its state is loaded after (round trip 2).  Fetching
                // every neighbour's state while the walk runs measured slower,
                // for rows of 6 neighbours and of 30.
                const bool have = lane < dg && nb_j >= 0;
                const double c = have ? nb_D * L.Dmat[spec * T + (nb_t - 1)] : 0.0;
                for (int pass = 0; pass < 2 && kk < 0; pass++) {
                    double cum2 = 0.0;
                    for (int k = 0; k < dg; k++) {
                        cum2 += rl_d(c, k);
                        if (cum2 > rand2) { kk = k; break; }
                    }
                    if (kk < 0) {
                        if (dest >= 0) break;   // stale destination of an earlier event (simulate_rdme.cpp:368)
                        rand2 = r2 * cum2;
                    }
                }
            } else {
                // rows longer than a wavefront: the same in-order walk, 64
                // entries at a time
                for (int pass = 0; pass < 2 && kk < 0; pass++) {
                    double cum2 = 0.0;
                    for (int c0 = 0; c0 < dg && kk < 0; c0 += WAVE) {
                        const int k = c0 + lane;
                        double cl = 0.0;
                        if (k < dg) {
                            const EdgeRec ek = a.erow[(size_t)sv * W + k];
                            const int tk = T > 1 ? ek.type : 1;
                            cl = ek.D * L.Dmat[spec * T + (tk - 1)];
                        }
                        const int mm = dg - c0 < WAVE ? dg - c0 : WAVE;
                        for (int q2 = 0; q2 < mm; q2++) {
                            cum2 += rl_d(cl, q2);
                            if (cum2 > rand2) { kk = c0 + q2; break; }
                        }
                    }
                    if (kk < 0) {
                        if (dest >= 0) break;
                        rand2 = r2 * cum2;
                    }
                }
            }
            if (kk < 0 && dest < 0) { err = SSASDPD_ERR_DIFFUSION_OVERFLOW; break; }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            // destination state (lane s / lane j layout): the chosen entry, or
            // the stale destination of an earlier event when the walk chose
            // none
            uint32_t xd;
            double rrd, srd, sdd, Ddd, qtd;
            int tds;
            {
                if (kk >= 0)
                    dest = dg <= WAVE ? __builtin_amdgcn_readlane(nb_j, kk)
                                      : __builtin_amdgcn_readfirstlane(a.erow[(size_t)sv * W + kk].j);
                const size_t j = (size_t)dest;
                xd = s_lane ? xx_[j * S + lane] : 0u;
                rrd = r_lane ? rrate_[j * R + lane] : 0.0;
                srd = srrate_[j];
                sdd = sdrate_[j];
                Ddd = a.Ddiag[j * S + spec];
                qtd = qtime_[j];
                tds = a.type[j];
            }
            const bool self = dest == sv;
            if (self) { xd = xs; rrd = rr; srd = sr; sdd = sd; Ddd = rl_d(Dd, spec); tds = tsv; }
            // move one molecule
            if (lane == spec) xs -= 1u;
            if (self) xd = xs;
            if (lane == spec) xd += 1u;
            if (self) xs = xd;
            if (R > 0 && !self) {
                const uint64_t dep = (uint32_t)__builtin_amdgcn_readlane((int)lm.dep, spec);
                double p_s, p_d;
                if (has_prog) {   // destination: its counts, data_fn and type, the source's vol
                    const double *df_d = a.data_fn ? a.data_fn + (size_t)dest * a.n_df : nullptr;
                    p_s = propensity_vm(lm, xs, vol, tsv, tt, df_sv, dep, vmx, vcode, vk, lane);
                    p_d = propensity_vm(lm, xd, vol, tds, tt, df_d, dep, vmx, vcode, vk, lane);
                } else {
                    p_s = propensity_par(lm, xs, vol, tsv);
                    p_d = propensity_par(lm, xd, vol, tds);
                }
                const double rd = ordered_sum(0.0, p_s - rr, dep);
                const double rrdel = ordered_sum(0.0, p_d - rrd, dep);
                if ((dep >> lane) & 1ULL) { rr = p_s; rrd = p_d; }
                sr += rd;
                srd += rrdel;
            } else if (R > 0) {
                double rd = 0.0, rrdel = 0.0;
                // self: source and destination are one voxel, so one set of
                // propensities serves both evaluations
                double pv = 0.0;
                if (has_prog) {
                    const uint64_t dep = (uint32_t)__builtin_amdgcn_readlane((int)lm.dep, spec);
                    pv = propensity_vm(lm, xs, vol, tsv, tt, df_sv, dep, vmx, vcode, vk, lane);
                }
                for (int k = L.jcG[spec]; k < L.jcG[spec + 1]; k++) {
                    const int j = L.irG[k];
                    double old = rl_d(rr, j);
                    const double p = has_prog ? rl_d(pv, j) : propensity_l(L.rx[j], xs, vol, tsv);
                    if (lane == j) rr = p;
                    rd += p - old;
                    if (self) rrd = rr;
                    old = rl_d(rrd, j);
                    const double pd = has_prog ? rl_d(pv, j) : propensity_l(L.rx[j], xd, vol, tds);
                    if (lane == j) rrd = pd;
                    rrdel += pd - old;
                    if (self) rr = rrd;
                }
                sr += rd;
                if (self) srd = sr;
                srd += rrdel;
                if (self) sr = srd;
            }
            sd -= rl_d(Dd, spec);
            if (self) sdd = sd;
            sdd += Ddd;
            if (self) sd = sdd;
            if (!self) {
                if (lane == spec) xx_[(size_t)sv * S + spec] = xs;
                if (r_lane) rrate_[(size_t)sv * R + lane] = rr;
                if (lead) { srrate_[sv] = sr; sdrate_[sv] = sd; }
            }
            if (lane == spec) xx_[(size_t)dest * S + spec] = xd;
            if (r_lane) rrate_[(size_t)dest * R + lane] = rrd;
            if (lead) { srrate_[dest] = srd; sdrate_[dest] = sdd; }
            n_diff++;
            which = spec;
            // ---- queue updates (NRMConstant_v5::update, NRMConstant_v5.cpp:89-132)
            const double tot = sr + sd;
            double ntv;
            if (tot <= 0.0) { ntv = inf; q.active--; }   // the selected voxel's time was finite
            else ntv = rng.exponential(lane) / tot + tt;
            sib.set(sv, ntv, lane);
            pop_finish<PL>(tsh, qtime_, L, tp, sib, sv, ntv, pre, lane);
            const double dest_tot = srd + sdd;
            const double dold = self ? ntv : qtd;
            double ntd;
            if (dest_tot <= 0.0) {
                ntd = inf;
                if (dold != inf) q.active--;
            } else {
                ntd = rng.exponential(lane) / dest_tot + tt;
                if (dold == inf) q.active++;
            }
            if (sib.holds(dest)) {
                sib.set(dest, ntd, lane);
                tree_update_rescan<PL>(tsh, qtime_, L, tp, sib, dest, ntd, lane);
            } else {
                tree_update_point<PL>(tsh, qtime_, L, tp, dest, ntd, lane);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
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
    if (err) { q.err_step = stp; break; }
    }
    if (!err)
        while (ks < a.n_snap) snapshot_counts(xx_all, snap_xx, snap_tot, nsS, S, ks++, tot, lane);
    // write the upper levels back for the next launch
    __syncthreads();
    for (int i = lane; i < a.tree_nodes; i += WAVE) { tmin_[i] = L.lt[i]; targ_[i] = L.lv[i]; }
    q.rng_pos = rng.pos;
    q.reactions += n_react;
    q.diffusions += n_diff;
    if (err) q.err = err;
    if (lead) *qs_ = q;
}

template <int PL, int MS, int MR, bool VM>
__global__ void __launch_bounds__(WAVE) k_ssa(RdmeArgs a) {
    ssa_run<PL, MS, MR, VM>(a, blockIdx.x);
}

// Ensembles of different models: workgroup g runs the trajectory described
// by args[g] (its own domain, chemistry, queue and stream).
template <int PL, int MS, int MR, bool VM>
__global__ void __launch_bounds__(WAVE) k_ssa_multi(const RdmeArgs *__restrict__ args) {
    ssa_run<PL, MS, MR, VM>(args[blockIdx.x], 0);
}

template <int MS, int MR, bool VM>
static void launch_multi_ms(hipStream_t s, const RdmeArgs *d_args, int n_models, int pl, size_t lds) {
    const dim3 g(n_models);
    switch (pl) {
    case 1: hipLaunchKernelGGL((k_ssa_multi<1, MS, MR, VM>), g, dim3(WAVE), lds, s, d_args); break;
    case 2: hipLaunchKernelGGL((k_ssa_multi<2, MS, MR, VM>), g, dim3(WAVE), lds, s, d_args); break;
    case 4: hipLaunchKernelGGL((k_ssa_multi<4, MS, MR, VM>), g, dim3(WAVE), lds, s, d_args); break;
    case 8: hipLaunchKernelGGL((k_ssa_multi<8, MS, MR, VM>), g, dim3(WAVE), lds, s, d_args); break;
    case 16: hipLaunchKernelGGL((k_ssa_multi<16, MS, MR, VM>), g, dim3(WAVE), lds, s, d_args); break;
    default: hipLaunchKernelGGL((k_ssa_multi<32, MS, MR, VM>), g, dim3(WAVE), lds, s, d_args); break;
    }
}

// Models with custom propensities use the widest per-lane tables (S, R <= 16)
// and the kernels with the expression VM.
void rdme_take_step_multi(hipStream_t s, const RdmeArgs *d_args, int n_models, int pl, int max_sr, size_t lds,
                          bool any_prog) {
    if (any_prog) launch_multi_ms<16, 16, true>(s, d_args, n_models, pl, lds);
    else if (max_sr <= 4) launch_multi_ms<4, 4, false>(s, d_args, n_models, pl, lds);
    else if (max_sr <= 8) launch_multi_ms<8, 8, false>(s, d_args, n_models, pl, lds);
    else launch_multi_ms<16, 16, false>(s, d_args, n_models, pl, lds);
}

template <int MS, int MR, bool VM> static void launch_ms(hipStream_t s, const RdmeArgs &a, int nt) {
    const dim3 g(nt);
    switch (a.pl) {
    case 1: hipLaunchKernelGGL((k_ssa<1, MS, MR, VM>), g, dim3(WAVE), a.lds_bytes, s, a); break;
    case 2: hipLaunchKernelGGL((k_ssa<2, MS, MR, VM>), g, dim3(WAVE), a.lds_bytes, s, a); break;
    case 4: hipLaunchKernelGGL((k_ssa<4, MS, MR, VM>), g, dim3(WAVE), a.lds_bytes, s, a); break;
    case 8: hipLaunchKernelGGL((k_ssa<8, MS, MR, VM>), g, dim3(WAVE), a.lds_bytes, s, a); break;
    case 16: hipLaunchKernelGGL((k_ssa<16, MS, MR, VM>), g, dim3(WAVE), a.lds_bytes, s, a); break;
    default: hipLaunchKernelGGL((k_ssa<32, MS, MR, VM>), g, dim3(WAVE), a.lds_bytes, s, a); break;
    }
}

void rdme_take_step(hipStream_t s, const RdmeArgs &a, int n_traj) {
    const int m = a.S > a.R ? a.S : a.R;
    if (a.vm_ncode > 0) launch_ms<16, 16, true>(s, a, n_traj);
    else if (m <= 4) launch_ms<4, 4, false>(s, a, n_traj);
    else if (m <= 8) launch_ms<8, 8, false>(s, a, n_traj);
    else launch_ms<16, 16, false>(s, a, n_traj);
}

// The expression VM over m argument sets (ssasdpd_vm_eval).
__global__ void k_vm_eval(const uint32_t *code, int start, const double *K, int x_is_double, int n_x, int n_df,
                          int m, const double *xd, const uint32_t *xu, const double *vol, const double *t,
                          const double *df, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double *dfi = n_df > 0 ? df + (size_t)i * n_df : nullptr;
    out[i] = x_is_double ? vm_eval_f64(code, start, K, xd + (size_t)i * n_x, vol[i], t[i], dfi)
                         : vm_eval_u32(code, start, K, xu + (size_t)i * n_x, vol[i], t[i], dfi);
}
void vm_eval_batch(hipStream_t s, const uint32_t *code, int start, const double *K, int x_is_double, int n_x,
                   int n_df, int m, const double *xd, const uint32_t *xu, const double *vol, const double *t,
                   const double *df, double *out) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_vm_eval, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, code, start, K, x_is_double,
                       n_x, n_df, m, xd, xu, vol, t, df, out);
}

// LDS bytes of the event-loop kernel for a given model.
size_t ssa_lds_bytes(int tree_nodes, int S, int R, int T, int nnzN, int nnzG, int n_code, int n_consts) {
    auto a16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
    size_t off = 0;
    off = a16(off + sizeof(double) * tree_nodes);
    off = a16(off + sizeof(int) * tree_nodes);
    off = a16(off + sizeof(Rxn) * R);
    off = a16(off + sizeof(double) * S * T);
    off = a16(off + sizeof(int) * (R + 1));
    off = a16(off + sizeof(int) * nnzN);
    off = a16(off + sizeof(int) * nnzN);
    off = a16(off + sizeof(int) * (S + R + 1));
    off = a16(off + sizeof(int) * nnzG);
    if (n_code > 0) {
        off = a16(off + sizeof(uint32_t) * n_code);
        off = a16(off + sizeof(double) * n_consts);
        off = a16(off + sizeof(uint32_t) * 64);
    }
    return off;
}

}  // namespace ssa
