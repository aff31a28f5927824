// Neighbour search for the sSSA-SDPD engine on gfx950.
//
// Replaces buildKDTree + Particle::find_neighbors + add_to_neighbor_list
// (ENG/src/simulate_threads.cpp:80-108, ENG/src/particle.cpp:150-294) and the
// ANN kd-tree fixed-radius search they call (ANN/src/kd_fix_rad_search.cpp).
//
//  * The neighbour SET and the r^2 bits come from counting-sort cell binning:
//    tree points are binned into cells of side >= h, candidates in the 3^d
//    surrounding cells are tested with ANN's exact acceptance rule
//    (r2 = ((0 + t0^2) + t1^2) + t2^2 with t = q - p in FP64, r2 <= h^2,
//    r2 != 0 (no self match, ANN.h:235), sqrt(r2) <= h (particle.cpp:162)).
//  * The list ORDER is ANN's: ascending r2, ties in the kd-tree's depth-first
//    visit order (pr_queue_k.h:100-116, kd_fix_rad_search.cpp:100-140).  Ties
//    are ranked through the lowest common ancestor of the two leaves in a
//    sliding-midpoint kd-tree built on the GPU level by level
//    (kd_tree.cpp:314-405, kd_split.cpp:146-213, kd_util.cpp:291-319): the
//    point on the query's side of that node's cut is visited first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "engine.hpp"

namespace ssa {

// ---------------------------------------------------------------------------
// Order-preserving map of doubles to unsigned 64-bit keys (for atomic min/max).
__device__ __forceinline__ unsigned long long dkey(double x) {
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__device__ __forceinline__ double dkey_inv(unsigned long long k) {
    unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k;
    return __longlong_as_double((long long)b);
}

// ---------------------------------------------------------------------------
// Exclusive scan of int32 (3-phase: block reduce, scan of block sums, add).
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ int block_exclusive_scan(int v, int *lds, int &total) {
    // v: thread partial; returns exclusive prefix within the block
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    int x = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        int y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    if (lane == WAVE - 1) lds[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < SCAN_BLOCK / WAVE; i++) { int t = lds[i]; lds[i] = acc; acc += t; }
        lds[SCAN_BLOCK / WAVE] = acc;
    }
    __syncthreads();
    total = lds[SCAN_BLOCK / WAVE];
    int r = x - v + lds[w];
    __syncthreads();
    return r;
}

__global__ void k_scan_tiles(const int *in, int *out, long long n, int *tile_sums) {
    __shared__ int lds[SCAN_BLOCK / WAVE + 1];
    long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int total;
    int pre = block_exclusive_scan(s, lds, total);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void k_scan_add(int *out, long long n, const int *tile_pre) {
    long long i = (long long)blockIdx.x * SCAN_TILE + threadIdx.x;
    int add = tile_pre[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; k++, i += SCAN_BLOCK)
        if (i < n) out[i] += add;
}

__global__ void k_scan_fix(const int *in, int *out, long long n) {
    if (threadIdx.x == 0) out[n] = n > 0 ? out[n - 1] + in[n - 1] : 0;
}

size_t scan_tmp_bytes(long long n) {
    size_t b = 0;
    long long m = n + 1;
    for (;;) {
        long long tiles = (m + SCAN_TILE - 1) / SCAN_TILE;
        b += (size_t)(2 * tiles + 4) * sizeof(int);
        if (tiles <= 1) break;
        m = tiles + 1;
    }
    return b + 256;
}

// out[i] = sum(in[0..i-1]); out must have n+1 entries (out[n] = total).
void exclusive_scan_i32(hipStream_t s, const int *in, int *out, long long n, void *tmp) {
    // scan n+1 elements where in[n] is treated as 0
    long long m = n + 1;
    long long tiles = (m + SCAN_TILE - 1) / SCAN_TILE;
    int *sums = (int *)tmp;
    int *pre = sums + tiles + 1;
    // in[n] may be out of bounds: run the tile scan on n elements and fix the
    // last entry with the tile total
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, in, out, n, sums);
    if (tiles > 1) {
        exclusive_scan_i32(s, sums, pre, tiles, (void *)(pre + tiles + 1));
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, out, n, pre);
    }
    // out[n] = out[n-1] + in[n-1]
    hipLaunchKernelGGL(k_scan_fix, dim3(1), dim3(1), 0, s, in, out, n);
}

// ---------------------------------------------------------------------------
// Bounding box of the points (annEnclRect, ANN/src/kd_util.cpp:73-90).
__global__ void k_bbox_init(unsigned long long *t) {
    int i = threadIdx.x;
    if (i < 3) { t[i] = ~0ULL; t[3 + i] = 0ULL; }
}
__global__ void k_bbox(const double *pts, int n, int dim, unsigned long long *t) {
    __shared__ unsigned long long smin[3][4], smax[3][4];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    for (int d = 0; d < dim; d++) {
        unsigned long long kmin = ~0ULL, kmax = 0;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            unsigned long long k = dkey(pts[(size_t)d * n + i]);
            kmin = k < kmin ? k : kmin;
            kmax = k > kmax ? k : kmax;
        }
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            unsigned long long a = __shfl_down(kmin, o, WAVE), b = __shfl_down(kmax, o, WAVE);
            kmin = a < kmin ? a : kmin;
            kmax = b > kmax ? b : kmax;
        }
        if (lane == 0) { smin[d][w] = kmin; smax[d][w] = kmax; }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)dim) {
        int d = threadIdx.x;
        unsigned long long a = smin[d][0], b = smax[d][0];
        for (int k = 1; k < 4; k++) { a = smin[d][k] < a ? smin[d][k] : a; b = smax[d][k] > b ? smax[d][k] : b; }
        atomicMin(&t[d], a);
        atomicMax(&t[3 + d], b);
    }
}
__global__ void k_bbox_fin(const unsigned long long *t, double *bb, int dim) {
    int d = threadIdx.x;
    if (d < 3) {
        bb[2 * d] = d < dim ? dkey_inv(t[d]) : 0.0;
        bb[2 * d + 1] = d < dim ? dkey_inv(t[3 + d]) : 0.0;
    }
}
void compute_bbox(hipStream_t s, const double *pts, int n, int dim, double *d_bbox, unsigned long long *tmp) {
    hipLaunchKernelGGL(k_bbox_init, dim3(1), dim3(64), 0, s, tmp);
    unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 512);
    hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(256), 0, s, pts, n, dim, tmp);
    hipLaunchKernelGGL(k_bbox_fin, dim3(1), dim3(64), 0, s, tmp, d_bbox, dim);
}

// Cell grid with side >= h over the bounding box bb[6] (host values).
// Counts are formed in double: a point far from the rest (a bounding box of
// 1e7 cells per side) must coarsen the grid, not wrap an integer product.
// A non-finite extent (NaN/Inf positions, reported by the state check) gets
// one cell.
void grid_setup(const double *bb, int n, int dim, double h, Grid &g) {
    const double limit = (double)std::max<long long>(4LL * n, 1 << 20);
    double cell = h * 1.01;
    for (;;) {
        double tot = 1.0, c[3] = {1.0, 1.0, 1.0};
        for (int d = 0; d < dim && d < 3; d++) {
            const double ext = bb[2 * d + 1] - bb[2 * d];
            c[d] = (std::isfinite(ext) && ext > 0.0) ? std::floor(ext / cell) + 1.0 : 1.0;
            tot *= c[d];
        }
        if (tot <= limit || !std::isfinite(cell)) {
            long long t = 1;
            for (int d = 0; d < 3; d++) {
                g.dims[d] = d < dim ? (int)std::min(c[d], limit) : 1;
                g.origin[d] = d < dim ? bb[2 * d] : 0.0;
                g.tile[d] = g.ntile[d] = 0;
                t *= g.dims[d];
            }
            g.ncell = t;
            break;
        }
        cell *= 1.5;
    }
    g.inv_cell = 1.0 / cell;
}

__device__ __forceinline__ int cell_coord(double x, double o, double inv, int dimc) {
    double c = floor((x - o) * inv);
    if (c < -1.0) c = -1.0;
    if (c > (double)dimc) c = (double)dimc;
    return (int)c;
}

// Tiles of 4 x 4 x 4 cells in 3D, 8 x 8 in 2D, 64 in 1D (a few hundred
// points each at the search's cell size), narrower where the grid is.
void grid_tile(Grid &g, int dim) {
    const int T = dim >= 3 ? 4 : (dim == 2 ? 8 : 64);
    long long k = 1;
    for (int d = 0; d < 3; d++) {
        g.tile[d] = d < dim ? std::max(1, std::min(T, g.dims[d])) : 1;
        g.ntile[d] = (g.dims[d] + g.tile[d] - 1) / g.tile[d];
        k *= (long long)g.ntile[d] * g.tile[d];
    }
    g.ncell = k;
}

__device__ __forceinline__ int cell_key(const int c[3], const Grid &g) {
    if (g.tile[0] == 0) return (c[2] * g.dims[1] + c[1]) * g.dims[0] + c[0];
    const int t = (c[2] / g.tile[2] * g.ntile[1] + c[1] / g.tile[1]) * g.ntile[0] + c[0] / g.tile[0];
    const int w = ((c[2] % g.tile[2]) * g.tile[1] + c[1] % g.tile[1]) * g.tile[0] + c[0] % g.tile[0];
    return t * (g.tile[0] * g.tile[1] * g.tile[2]) + w;
}

// crowded: raised to a cell's count when it passes FR_MAX_CELL
__global__ void k_cell_count(const double *pts, int n, int dim, Grid g, int *cell_of, int *count, int *crowded) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
        int v = cell_coord(pts[(size_t)d * n + i], g.origin[d], g.inv_cell, g.dims[d]);
        c[d] = min(max(v, 0), g.dims[d] - 1);
    }
    const int cell = cell_key(c, g);
    cell_of[i] = cell;
    const int k = atomicAdd(&count[cell], 1) + 1;
    if (crowded && k > FR_MAX_CELL) atomicMax(crowded, k);
}

__global__ void k_cell_scatter(const double *pts, int n, int dim, const int *cell_of, const int *start,
                               int *fill, int *sorted, double *spts) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = cell_of[i];
    int p = start[c] + atomicAdd(&fill[c], 1);
    sorted[p] = i;
    // one 32-byte record per point: a candidate is two 16-byte loads
    double2 *o = (double2 *)(spts + 4 * (size_t)p);
    o[0] = make_double2(pts[i], dim > 1 ? pts[(size_t)n + i] : 0.0);
    o[1] = make_double2(dim > 2 ? pts[2 * (size_t)n + i] : 0.0, 0.0);
}

void bin_points(hipStream_t s, const double *pts, int n, int dim, const Grid &g, int *cell_of, int *cell_count,
                int *cell_start, int *cell_fill, int *sorted, double *spts, void *scan_tmp, int *crowded) {
    HIP_OK(hipMemsetAsync(cell_count, 0, sizeof(int) * g.ncell, s));
    HIP_OK(hipMemsetAsync(cell_fill, 0, sizeof(int) * g.ncell, s));
    hipLaunchKernelGGL(k_cell_count, dim3((n + 255) / 256), dim3(256), 0, s, pts, n, dim, g, cell_of, cell_count,
                       crowded);
    exclusive_scan_i32(s, cell_count, cell_start, g.ncell, scan_tmp);
    hipLaunchKernelGGL(k_cell_scatter, dim3((n + 255) / 256), dim3(256), 0, s, pts, n, dim, cell_of, cell_start,
                       cell_fill, sorted, spts);
}

// One thread per cell: insertion sort of the cell's indices (cells hold a
// few tens of points; this runs only when particles are renumbered).  A cell
// past FR_MAX_CELL is left as it is: the search of that step reports it.
__global__ void k_cell_order(const int *start, long long ncell, int *sorted) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int b = start[c], e = start[c + 1];
    if (e - b > FR_MAX_CELL) return;
    for (int i = b + 1; i < e; i++) {
        const int v = sorted[i];
        int j = i - 1;
        while (j >= b && sorted[j] > v) { sorted[j + 1] = sorted[j]; j--; }
        sorted[j + 1] = v;
    }
}

void cell_order_by_index(hipStream_t s, const int *cell_start, long long ncell, int *sorted) {
    if (ncell <= 0) return;
    hipLaunchKernelGGL(k_cell_order, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, cell_start, ncell, sorted);
}

// ---------------------------------------------------------------------------
// GPU kd-tree (ANN sliding-midpoint rule, bucket size 1), built level by
// level.  The split of each cell follows sl_midpt_split exactly, and the
// two-pass partition reproduces annPlaneSplit's permutation: in each pass the
// k-th misplaced element from the left is swapped with the k-th misplaced
// element from the right, which is what the two-pointer loop does
// (ANN/src/kd_util.cpp:291-319).
struct KdWork {
    int *perm, *seg, *flags, *scan, *WL, *WR;
    int *s_start, *s_n, *s_node, *s_cd, *s_mode, *s_br1, *s_nlo, *s_clo, *s_chi;
    double *s_cv, *s_lo, *s_hi;                 // boxes [3][cap]
    unsigned long long *s_min, *s_max;          // [3][cap]
    int *n_start, *n_n, *n_node;                // next level
    double *n_lo, *n_hi;                        // [3][cap]
    int *counters;                              // [0] next segs, [1] nodes
    void *scan_tmp;
};

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

size_t kd_work_bytes(int n) {
    size_t cap = (size_t)n + 1;
    size_t b = 0;
    b += align_up(sizeof(int) * cap) * 6;                      // perm seg flags WL WR + spare
    b += align_up(sizeof(int) * (cap + 1));                    // scan
    b += align_up(sizeof(int) * cap) * 9;                      // s_* ints
    b += align_up(sizeof(double) * cap) * 7;                   // s_cv, s_lo[3], s_hi[3]
    b += align_up(sizeof(unsigned long long) * cap) * 6;       // s_min, s_max
    b += align_up(sizeof(int) * cap) * 3;                      // n_start n_n n_node
    b += align_up(sizeof(double) * cap) * 6;                   // n_lo n_hi
    b += align_up(sizeof(int) * 16);
    b += align_up(scan_tmp_bytes(cap + 1));
    return b + 4096;
}

static KdWork carve(void *base, int n) {
    size_t cap = (size_t)n + 1;
    char *p = (char *)base;
    auto take = [&](size_t bytes) { void *r = p; p += align_up(bytes); return r; };
    KdWork w;
    w.perm = (int *)take(sizeof(int) * cap);
    w.seg = (int *)take(sizeof(int) * cap);
    w.flags = (int *)take(sizeof(int) * cap);
    w.WL = (int *)take(sizeof(int) * cap);
    w.WR = (int *)take(sizeof(int) * cap);
    take(sizeof(int) * cap);
    w.scan = (int *)take(sizeof(int) * (cap + 1));
    w.s_start = (int *)take(sizeof(int) * cap);
    w.s_n = (int *)take(sizeof(int) * cap);
    w.s_node = (int *)take(sizeof(int) * cap);
    w.s_cd = (int *)take(sizeof(int) * cap);
    w.s_mode = (int *)take(sizeof(int) * cap);
    w.s_br1 = (int *)take(sizeof(int) * cap);
    w.s_nlo = (int *)take(sizeof(int) * cap);
    w.s_clo = (int *)take(sizeof(int) * cap);
    w.s_chi = (int *)take(sizeof(int) * cap);
    w.s_cv = (double *)take(sizeof(double) * cap);
    w.s_lo = (double *)take(sizeof(double) * cap * 3);
    w.s_hi = (double *)take(sizeof(double) * cap * 3);
    w.s_min = (unsigned long long *)take(sizeof(unsigned long long) * cap * 3);
    w.s_max = (unsigned long long *)take(sizeof(unsigned long long) * cap * 3);
    w.n_start = (int *)take(sizeof(int) * cap);
    w.n_n = (int *)take(sizeof(int) * cap);
    w.n_node = (int *)take(sizeof(int) * cap);
    w.n_lo = (double *)take(sizeof(double) * cap * 3);
    w.n_hi = (double *)take(sizeof(double) * cap * 3);
    w.counters = (int *)take(sizeof(int) * 16);
    w.scan_tmp = take(scan_tmp_bytes(cap + 1));
    return w;
}

__global__ void k_kd_init(int n, int *perm, int *seg, KdDev kd) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { perm[i] = i; seg[i] = 0; }
}

__global__ void k_kd_root(const double *bb, int n, int dim, KdWork w, KdDev kd, int cap) {
    if (threadIdx.x != 0) return;
    w.s_start[0] = 0;
    w.s_n[0] = n;
    w.s_node[0] = 0;
    for (int d = 0; d < 3; d++) {
        w.s_lo[(size_t)d * cap] = bb[2 * d];
        w.s_hi[(size_t)d * cap] = bb[2 * d + 1];
    }
    kd.parent[0] = -1;
    kd.depth[0] = 0;
    kd.is_hi[0] = 0;
    kd.cd[0] = 0;
    kd.cv[0] = 0.0;
    w.counters[1] = 1;
    if (n == 1) kd.leaf_of[0] = 0;
}

__global__ void k_kd_mm_init(int nseg, unsigned long long *mn, unsigned long long *mx, int cap) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nseg)
        for (int d = 0; d < 3; d++) { mn[(size_t)d * cap + i] = ~0ULL; mx[(size_t)d * cap + i] = 0ULL; }
}

// Segmented min/max over positions: a wave combines lanes of the same
// segment before one lane per segment run issues the atomics.
__global__ void k_kd_minmax(const double *pts, int n, int dim, const int *perm, const int *seg,
                            unsigned long long *mn, unsigned long long *mx, int cap) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int lane = threadIdx.x & (WAVE - 1);
    int s = p < n ? seg[p] : -1;
    int sp = __shfl_up(s, 1, WAVE);
    bool head = (lane == 0) || (sp != s);
    for (int d = 0; d < dim; d++) {
        unsigned long long a = ~0ULL, b = 0ULL;
        if (s >= 0) a = b = dkey(pts[(size_t)d * n + perm[p]]);
        for (int o = 1; o < WAVE; o <<= 1) {
            unsigned long long oa = __shfl_down(a, o, WAVE), ob = __shfl_down(b, o, WAVE);
            int os = __shfl_down(s, o, WAVE);
            if (lane + o < WAVE && os == s) { a = oa < a ? oa : a; b = ob > b ? ob : b; }
        }
        if (head && s >= 0) {
            atomicMin(&mn[(size_t)d * cap + s], a);
            atomicMax(&mx[(size_t)d * cap + s], b);
        }
    }
}

// sl_midpt_split's choice of cut dimension and value (kd_split.cpp:146-188).
__global__ void k_kd_choose(int nseg, int dim, KdWork w, int cap) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const double ERR = 0.001;
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++) { lo[d] = w.s_lo[(size_t)d * cap + s]; hi[d] = w.s_hi[(size_t)d * cap + s]; }
    double max_length = hi[0] - lo[0];
    for (int d = 1; d < dim; d++) {
        double length = hi[d] - lo[d];
        if (length > max_length) max_length = length;
    }
    double max_spread = -1;
    int cut_dim = 0;
    for (int d = 0; d < dim; d++) {
        if ((hi[d] - lo[d]) >= (1 - ERR) * max_length) {
            double spr = dkey_inv(w.s_max[(size_t)d * cap + s]) - dkey_inv(w.s_min[(size_t)d * cap + s]);
            if (spr > max_spread) { max_spread = spr; cut_dim = d; }
        }
    }
    double ideal = (lo[cut_dim] + hi[cut_dim]) / 2;
    double mn = dkey_inv(w.s_min[(size_t)cut_dim * cap + s]);
    double mx = dkey_inv(w.s_max[(size_t)cut_dim * cap + s]);
    double cv;
    int mode = 0;
    if (ideal < mn) { cv = mn; mode = 1; }
    else if (ideal > mx) { cv = mx; mode = 2; }
    else cv = ideal;
    w.s_cd[s] = cut_dim;
    w.s_cv[s] = cv;
    w.s_mode[s] = mode;
}

// pass 0: flag v < cv on the whole segment; pass 1: flag v <= cv on
// [start+br1, end).
__global__ void k_kd_flags(const double *pts, int n, const int *perm, const int *seg, KdWork w, int pass) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int s = seg[p];
    int f = 0;
    if (s >= 0) {
        double v = pts[(size_t)w.s_cd[s] * n + perm[p]];
        double cv = w.s_cv[s];
        if (pass == 0) f = v < cv;
        else f = (p >= w.s_start[s] + w.s_br1[s]) && (v <= cv);
    }
    w.flags[p] = f;
}

// Ranks of misplaced elements and the swap (both passes).  Range of the pass:
// [b, e) with b = start (+br1 for pass 1); C = number of flagged elements.
__global__ void k_kd_rank(int n, const int *perm, const int *seg, KdWork w, int pass) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int s = seg[p];
    if (s < 0) return;
    int b = w.s_start[s] + (pass ? w.s_br1[s] : 0);
    int e = w.s_start[s] + w.s_n[s];
    if (p < b) return;
    const int *A = w.scan;
    int C = A[e] - A[b];
    int f = w.flags[p];
    if (p < b + C && !f) {
        int k = (p - b) - (A[p] - A[b]);
        w.WL[b + k] = perm[p];
    } else if (p >= b + C && f) {
        int k = A[e] - A[p + 1];
        w.WR[b + k] = perm[p];
    }
}
__global__ void k_kd_swap(int n, int *perm, const int *seg, KdWork w, int pass) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int s = seg[p];
    if (s < 0) return;
    int b = w.s_start[s] + (pass ? w.s_br1[s] : 0);
    int e = w.s_start[s] + w.s_n[s];
    if (p < b) return;
    const int *A = w.scan;
    int C = A[e] - A[b];
    int f = w.flags[p];
    if (p < b + C && !f) {
        int k = (p - b) - (A[p] - A[b]);
        perm[p] = w.WR[b + k];
    } else if (p >= b + C && f) {
        int k = A[e] - A[p + 1];
        perm[p] = w.WL[b + k];
    }
}
__global__ void k_kd_br1(int nseg, KdWork w) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    int b = w.s_start[s], e = b + w.s_n[s];
    w.s_br1[s] = w.scan[e] - w.scan[b];
}

// n_lo rule (kd_split.cpp:208-212), node records and the next level.
__global__ void k_kd_children(int nseg, KdWork w, KdDev kd, const int *perm, int cap) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    int b = w.s_start[s], cnt = w.s_n[s];
    int br1 = w.s_br1[s];
    int br2 = br1 + (w.scan[b + cnt] - w.scan[b + br1]);
    int mode = w.s_mode[s];
    int n_lo;
    if (mode == 1) n_lo = 1;
    else if (mode == 2) n_lo = cnt - 1;
    else if (br1 > cnt / 2) n_lo = br1;
    else if (br2 < cnt / 2) n_lo = br2;
    else n_lo = cnt / 2;
    w.s_nlo[s] = n_lo;
    int node = w.s_node[s];
    int cd = w.s_cd[s];
    double cv = w.s_cv[s];
    kd.cd[node] = cd;
    kd.cv[node] = cv;
    int c0 = atomicAdd(&w.counters[1], 2);
    int dep = kd.depth[node] + 1;
    for (int side = 0; side < 2; side++) {
        int child = c0 + side;
        int cstart = side ? b + n_lo : b;
        int cn = side ? cnt - n_lo : n_lo;
        kd.parent[child] = node;
        kd.depth[child] = dep;
        kd.is_hi[child] = side;
        kd.cd[child] = 0;
        kd.cv[child] = 0.0;
        int nseg_id = -1;
        if (cn == 1) {
            kd.leaf_of[perm[cstart]] = child;
        } else if (cn > 1) {
            nseg_id = atomicAdd(&w.counters[0], 1);
            w.n_start[nseg_id] = cstart;
            w.n_n[nseg_id] = cn;
            w.n_node[nseg_id] = child;
            for (int d = 0; d < 3; d++) {
                double lo = w.s_lo[(size_t)d * cap + s], hi = w.s_hi[(size_t)d * cap + s];
                if (d == cd) {
                    if (side == 0) hi = cv;
                    else lo = cv;
                }
                w.n_lo[(size_t)d * cap + nseg_id] = lo;
                w.n_hi[(size_t)d * cap + nseg_id] = hi;
            }
        }
        if (side == 0) w.s_clo[s] = nseg_id;
        else w.s_chi[s] = nseg_id;
    }
}

__global__ void k_kd_reseg(int n, int *seg, KdWork w) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int s = seg[p];
    if (s < 0) return;
    seg[p] = (p - w.s_start[s] < w.s_nlo[s]) ? w.s_clo[s] : w.s_chi[s];
}

__global__ void k_kd_promote(int nseg, KdWork w, int cap) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    w.s_start[s] = w.n_start[s];
    w.s_n[s] = w.n_n[s];
    w.s_node[s] = w.n_node[s];
    for (int d = 0; d < 3; d++) {
        w.s_lo[(size_t)d * cap + s] = w.n_lo[(size_t)d * cap + s];
        w.s_hi[(size_t)d * cap + s] = w.n_hi[(size_t)d * cap + s];
    }
}

void kd_build(hipStream_t st, const double *pts, int n, int dim, KdDev &kd, void *work, size_t work_bytes) {
    (void)work_bytes;
    KdWork w = carve(work, n);
    int cap = n + 1;
    const int B = 256;
    unsigned gn = (unsigned)((n + B - 1) / B);
    // root box = bounding rectangle of the points
    double *d_bb = (double *)w.n_lo;   // scratch for the 6 bbox values
    compute_bbox(st, pts, n, dim, d_bb, (unsigned long long *)w.s_min);
    hipLaunchKernelGGL(k_kd_init, dim3(gn), dim3(B), 0, st, n, w.perm, w.seg, kd);
    HIP_OK(hipMemsetAsync(w.counters, 0, sizeof(int) * 16, st));
    hipLaunchKernelGGL(k_kd_root, dim3(1), dim3(64), 0, st, d_bb, n, dim, w, kd, cap);
    int nseg = n > 1 ? 1 : 0;
    if (n <= 1) {
        HIP_OK(hipMemsetAsync(w.seg, 0xFF, sizeof(int) * n, st));
    }
    int levels = 0;
    while (nseg > 0) {
        unsigned gs = (unsigned)((nseg + B - 1) / B);
        hipLaunchKernelGGL(k_kd_mm_init, dim3(gs), dim3(B), 0, st, nseg, w.s_min, w.s_max, cap);
        hipLaunchKernelGGL(k_kd_minmax, dim3(gn), dim3(B), 0, st, pts, n, dim, w.perm, w.seg, w.s_min, w.s_max, cap);
        hipLaunchKernelGGL(k_kd_choose, dim3(gs), dim3(B), 0, st, nseg, dim, w, cap);
        // pass 1: < cv
        hipLaunchKernelGGL(k_kd_flags, dim3(gn), dim3(B), 0, st, pts, n, w.perm, w.seg, w, 0);
        exclusive_scan_i32(st, w.flags, w.scan, n, w.scan_tmp);
        hipLaunchKernelGGL(k_kd_rank, dim3(gn), dim3(B), 0, st, n, w.perm, w.seg, w, 0);
        hipLaunchKernelGGL(k_kd_swap, dim3(gn), dim3(B), 0, st, n, w.perm, w.seg, w, 0);
        hipLaunchKernelGGL(k_kd_br1, dim3(gs), dim3(B), 0, st, nseg, w);
        // pass 2: <= cv on [start+br1, end)
        hipLaunchKernelGGL(k_kd_flags, dim3(gn), dim3(B), 0, st, pts, n, w.perm, w.seg, w, 1);
        exclusive_scan_i32(st, w.flags, w.scan, n, w.scan_tmp);
        hipLaunchKernelGGL(k_kd_rank, dim3(gn), dim3(B), 0, st, n, w.perm, w.seg, w, 1);
        hipLaunchKernelGGL(k_kd_swap, dim3(gn), dim3(B), 0, st, n, w.perm, w.seg, w, 1);
        HIP_OK(hipMemsetAsync(w.counters, 0, sizeof(int), st));
        hipLaunchKernelGGL(k_kd_children, dim3(gs), dim3(B), 0, st, nseg, w, kd, w.perm, cap);
        hipLaunchKernelGGL(k_kd_reseg, dim3(gn), dim3(B), 0, st, n, w.seg, w);
        int next = 0;
        HIP_OK(hipMemcpyAsync(&next, w.counters, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (next > 0) {
            hipLaunchKernelGGL(k_kd_promote, dim3((unsigned)((next + B - 1) / B)), dim3(B), 0, st, next, w, cap);
        }
        nseg = next;
        if (++levels > n + 2) throw HipError("kd_build: no progress", hipErrorUnknown, __FILE__, __LINE__);
    }
    int nodes = 0;
    HIP_OK(hipMemcpyAsync(&nodes, w.counters + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    kd.n_nodes = nodes;
    kd.valid = true;
}

// ---------------------------------------------------------------------------
// Fixed-radius search.

// Each lane of the search builds and sorts its list in SLOTS LDS slots
// (slot-major, one column per lane of a one-wavefront block) and writes it
// out once.  A list that outgrows them moves to the output arrays (entries
// in visit order, then the later hits) and is sorted there, one global
// round trip per step of the insertion sort, which stalls the whole
// wavefront: the launch picks the smallest of 12, 16, 24 or 32 slots that
// holds the longest list of the previous search (fewer slots keep more
// wavefronts resident for the gathers; a lattice at h = 1.3 spacing holds 6
// neighbours per list, a flowing one 9 on average and up to ~20).
constexpr int FR_SLOT_CHOICES[4] = {12, 16, 24, 32};

struct KdView {
    const int *parent, *depth, *cd, *is_hi, *leaf_of;
    const double *cv;
    int valid;
};

// true when tree point a is visited before tree point b by ANN's search from
// query q (nearer child first: cut_diff < 0 visits the low child first,
// kd_fix_rad_search.cpp:106-136).
__device__ bool kd_before(const KdView &kd, int a, int b, const double *q) {
    if (!kd.valid) return a < b;
    int na = kd.leaf_of[a], nb = kd.leaf_of[b];
    int da = kd.depth[na], db = kd.depth[nb];
    while (da > db) { na = kd.parent[na]; da--; }
    while (db > da) { nb = kd.parent[nb]; db--; }
    while (kd.parent[na] != kd.parent[nb]) { na = kd.parent[na]; nb = kd.parent[nb]; }
    int p = kd.parent[na];
    double cut_diff = q[kd.cd[p]] - kd.cv[p];
    int near_hi = cut_diff < 0 ? 0 : 1;
    return kd.is_hi[na] == near_hi;
}

// Insertion sort of a list of cnt entries by (r2, ANN visit order); I and D
// return references to entry k's index and r2.
template <class IA, class DA>
__device__ __forceinline__ void sort_list(IA I, DA D, int cnt, const KdView &kd, const double *qp) {
    for (int i = 1; i < cnt; i++) {
        int ji = I(i);
        double ri = D(i);
        int k = i - 1;
        while (k >= 0) {
            double rk = D(k);
            int jk = I(k);
            bool after = rk > ri || (rk == ri && kd_before(kd, ji, jk, qp));
            if (!after) break;
            I(k + 1) = jk;
            D(k + 1) = rk;
            k--;
        }
        I(k + 1) = ji;
        D(k + 1) = ri;
    }
}

// D_ij of one edge (add_to_neighbor_list, particle.cpp:160-187): the 3-D
// constant of wfd in every dimension, operation for operation.
__device__ __forceinline__ double edge_dij(double mi, double ri, double mj, double rj, double rr2, double rr,
                                           double h) {
    double ih = 1.0 / h;
    double ihsq = ih * ih;
    double dhr = h - rr;
    double wfd = -25.066903536973515383e0 * dhr * dhr * ihsq * ihsq * ihsq * ih;
    return -2.0 * (mi * mj) / (mi + mj) * (ri + rj) / (ri * rj) * rr2 * wfd / (rr2 + 0.01 * h * h);
}

template <int SLOTS>
__global__ void __launch_bounds__(WAVE)
k_fr_search(const double *__restrict__ q, int m, const double *__restrict__ spts, const int *__restrict__ sorted,
            const int *__restrict__ cell_start, int n, int dim, double h, Grid g, KdView kd, int W,
            int *__restrict__ out_idx, double *__restrict__ out_r2, int *__restrict__ deg, DevErr *err,
            const int *__restrict__ order, FinishArgs fin, int has_fin) {
    __shared__ int l_idx[SLOTS * WAVE];
    __shared__ double l_r2[SLOTS * WAVE];
    __shared__ double s_dm[FIN_MAX_DMAT];
    const int tl = (int)threadIdx.x;
    if (has_fin && fin.Ddiag) {
        for (int k = tl; k < fin.S * fin.T; k += WAVE) s_dm[k] = fin.Dmat[k];
        __syncthreads();
    }
    const int chunk = xcd_chunk((m + WAVE - 1) / WAVE);
    if (chunk < 0) return;
    int t = chunk * WAVE + tl;
    // lanes past the end stay for the wave's reduction of the degree
    const bool live = t < m;
    int qi = live ? (order ? order[t] : t) : 0;
    int cnt = 0;
    if (err->crowded > 0) {
        // the binning put more than FR_MAX_CELL points in one cell (a point
        // far from the rest coarsened the grid): no lists, and the step
        // reports it instead of scanning those cells for every query
        if (live) deg[qi] = 0;
        if (t == 0) atomicCAS(&err->code, 0, (int)SSASDPD_ERR_CAPACITY);
        return;
    }
    if (live) {
        auto LI = [&](int k) -> int & { return l_idx[k * WAVE + tl]; };
        auto LD = [&](int k) -> double & { return l_r2[k * WAVE + tl]; };
        auto GI = [&](int k) -> int & { return out_idx[(size_t)k * m + qi]; };
        auto GD = [&](int k) -> double & { return out_r2[(size_t)k * m + qi]; };
        const int L = W < SLOTS ? W : SLOTS;   // LDS slots in use
        double qp[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < dim; d++) qp[d] = q[(size_t)d * m + qi];
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int d = 0; d < 3; d++) {
            if (d < dim) {
                const int c = cell_coord(qp[d], g.origin[d], g.inv_cell, g.dims[d]);
                lo[d] = max(c - 1, 0);
                hi[d] = min(c + 1, g.dims[d] - 1);
            }
        }
        const double sq = h * h;
        // the candidate rows' bounds, all loaded before the first is used
        int rb[9], re[9], nrow = 0;
        for (int cz = lo[2]; cz <= hi[2]; cz++)
            for (int cy = lo[1]; cy <= hi[1]; cy++) {
                const int row = (cz * g.dims[1] + cy) * g.dims[0];
                rb[nrow] = cell_start[row + lo[0]];
                re[nrow] = cell_start[row + hi[0] + 1];
                nrow++;
            }
        auto accept = [&](int j, double dist) {
            if (cnt < L) {
                LI(cnt) = j;
                LD(cnt) = dist;
            } else {
                if (cnt == L)   // the list moves to the output arrays
                    for (int k = 0; k < L; k++) { GI(k) = LI(k); GD(k) = LD(k); }
                if (cnt < W) { GI(cnt) = j; GD(cnt) = dist; }
            }
            cnt++;
        };
        // candidates in visit order, fetched CB at a time (positions and
        // tree indices together) so one round trip serves CB of them
        constexpr int CB = 4;
        for (int r = 0; r < nrow; r++) {
            const int pe = re[r];
            for (int p0 = rb[r]; p0 < pe; p0 += CB) {
                double2 a[CB], b[CB];
                int jj[CB];
#pragma unroll
                for (int u = 0; u < CB; u++) {
                    const int p = p0 + u < pe ? p0 + u : p0;
                    const double2 *sp = (const double2 *)(spts + 4 * (size_t)p);
                    a[u] = sp[0];
                    b[u] = sp[1];
                    jj[u] = sorted[p];
                }
#pragma unroll
                for (int u = 0; u < CB; u++) {
                    if (p0 + u >= pe) break;
                    double dist = 0.0;
                    double t0 = qp[0] - a[u].x;
                    dist = dist + t0 * t0;
                    if (dim > 1) { double t1 = qp[1] - a[u].y; dist = dist + t1 * t1; }
                    if (dim > 2) { double t2 = qp[2] - b[u].x; dist = dist + t2 * t2; }
                    if (dist <= sq && dist != 0.0 && sqrt(dist) <= h) accept(jj[u], dist);
                }
            }
        }
        if (cnt > W) {
            atomicExch(&err->overflow, 1);
            deg[qi] = 0;
        } else {
            deg[qi] = cnt;
            if (cnt <= L) {
                sort_list(LI, LD, cnt, kd, qp);
                // with the finish, r2 is only the spill area of long lists
                for (int k = 0; k < cnt; k++) {
                    GI(k) = LI(k);
                    if (!has_fin) GD(k) = LD(k);
                }
            } else {
                sort_list(GI, GD, cnt, kd, qp);
            }
            if (has_fin) {
                // r, D_ij, the neighbour's type and the voxel-major row, in
                // list order; the neighbours' records are fetched FB at a
                // time ahead of the stores (the compiler cannot move a load
                // above a store through these pointers)
                const double mi = fin.mass[qi], ri = fin.rho[qi];
                double dd[FIN_MAX_SPECIES] = {};
                constexpr int FB = 4;
                for (int k0 = 0; k0 < cnt; k0 += FB) {
                    int jj[FB], tj[FB];
                    double rr2[FB], rj[FB], mj[FB];
#pragma unroll
                    for (int u = 0; u < FB; u++) {
                        const int k = k0 + u < cnt ? k0 + u : k0;
                        jj[u] = cnt <= L ? LI(k) : GI(k);
                        rr2[u] = cnt <= L ? LD(k) : GD(k);
                        if (fin.rec) {
                            const PRec *pj = fin.rec + jj[u];
                            const double2 rm = *(const double2 *)&pj->rho;
                            rj[u] = rm.x;
                            mj[u] = rm.y;
                            tj[u] = pj->type;
                        } else {
                            mj[u] = fin.mass[jj[u]];
                            rj[u] = fin.rho[jj[u]];
                            tj[u] = fin.type[jj[u]];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < FB; u++) {
                        const int k = k0 + u;
                        if (k >= cnt) break;
                        const double rr = sqrt(rr2[u]);
                        const double D = edge_dij(mi, ri, mj[u], rj[u], rr2[u], rr, h);
                        if (D != D) {
                            atomicCAS(&err->code, 0, (int)SSASDPD_ERR_NAN_DIJ);
                            atomicMin(&err->particle, qi);
                        }
                        const size_t e = (size_t)k * m + qi;
                        fin.r[e] = rr;
                        if (fin.Dij) fin.Dij[e] = D;
                        if (fin.ntype) fin.ntype[e] = tj[u];
                        fin.erow[(size_t)qi * W + k] = EdgeRec{jj[u], tj[u], D};
                        if (fin.Ddiag) {
#pragma unroll
                            for (int s2 = 0; s2 < FIN_MAX_SPECIES; s2++)
                                if (s2 < fin.S) dd[s2] += s_dm[s2 * fin.T + (tj[u] - 1)] * D;
                        }
                    }
                }
                if (fin.Ddiag) {
#pragma unroll
                    for (int s2 = 0; s2 < FIN_MAX_SPECIES; s2++)
                        if (s2 < fin.S) fin.Ddiag[(size_t)qi * fin.S + s2] = dd[s2];
                }
            }
        }
    }
    // one atomic per wavefront for the longest list
    int mx = cnt;
    for (int o = WAVE / 2; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, WAVE));
    if (tl == 0) atomicMax(&err->max_deg, mx);
}

void fr_search(hipStream_t s, const double *q, int m, const double *spts, const int *sorted, const int *cell_start,
               int n, int dim, double h, const Grid &g, const KdDev *kd, int W, int *out_idx, double *out_r2,
               int *deg, DevErr *err, const int *order, const FinishArgs *fin, int slots_hint) {
    KdView v{};
    if (kd && kd->valid) {
        v.parent = kd->parent; v.depth = kd->depth; v.cd = kd->cd; v.is_hi = kd->is_hi;
        v.leaf_of = kd->leaf_of; v.cv = kd->cv; v.valid = 1;
    }
    const FinishArgs f = fin ? *fin : FinishArgs{};
    const dim3 grid = xcd_grid((m + WAVE - 1) / WAVE);
    const int hf = fin ? 1 : 0;
    if (slots_hint <= FR_SLOT_CHOICES[0])
        hipLaunchKernelGGL(k_fr_search<12>, grid, dim3(WAVE), 0, s, q, m, spts, sorted, cell_start, n, dim, h, g, v, W,
                           out_idx, out_r2, deg, err, order, f, hf);
    else if (slots_hint <= FR_SLOT_CHOICES[1])
        hipLaunchKernelGGL(k_fr_search<16>, grid, dim3(WAVE), 0, s, q, m, spts, sorted, cell_start, n, dim, h, g, v, W,
                           out_idx, out_r2, deg, err, order, f, hf);
    else if (slots_hint <= FR_SLOT_CHOICES[2])
        hipLaunchKernelGGL(k_fr_search<24>, grid, dim3(WAVE), 0, s, q, m, spts, sorted, cell_start, n, dim, h, g, v, W,
                           out_idx, out_r2, deg, err, order, f, hf);
    else
        hipLaunchKernelGGL(k_fr_search<32>, grid, dim3(WAVE), 0, s, q, m, spts, sorted, cell_start, n, dim, h, g, v, W,
                           out_idx, out_r2, deg, err, order, f, hf);
}

}  // namespace ssa
