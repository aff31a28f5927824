// ssasdpd_fr_search_noalloc: the stand-alone fixed-radius search
// (ssasdpd_fr_search) with every device buffer -- staging copies of the
// points and queries, the cell binning, the kd-tree and its build work
// space, the ELL lists, the CSR output and the error record -- carved from
// one file-scope device array, so that a call makes no device allocation.
// For callers that cannot allocate (tight loops, harnesses that forbid it).
//
// Each GPU holds its own copy of the array at its own address; the address
// is looked up once per device with hipGetSymbolAddress, with that device
// current, and only that address is handed to kernels and copies.  Calls
// that overlap on one GPU share the array and race: callers serialise them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "engine.hpp"

namespace ssa {
namespace {

constexpr int NA_MAX_POINTS = SSASDPD_NOALLOC_MAX_POINTS;
constexpr int NA_MAX_QUERIES = SSASDPD_NOALLOC_MAX_QUERIES;
constexpr int NA_MAX_WIDTH = SSASDPD_NOALLOC_MAX_WIDTH;
constexpr long long NA_MAX_CELLS = 131072;
constexpr int NA_MAX_DEVICES = 64;

constexpr size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte layout of the scratch array.
constexpr size_t CAP_P = (size_t)NA_MAX_POINTS + 1, CAP_Q = (size_t)NA_MAX_QUERIES + 1;
constexpr size_t O_STAGE = 0;                                                 // AoS upload [3][max(n, m)]
constexpr size_t O_PTS = O_STAGE + al(sizeof(double) * 3 * std::max(CAP_P, CAP_Q));
constexpr size_t O_SPTS = O_PTS + al(sizeof(double) * 3 * CAP_P);           // sorted tree points [n][4]
constexpr size_t O_Q = O_SPTS + al(sizeof(double) * 4 * CAP_P);
constexpr size_t O_SORTED = O_Q + al(sizeof(double) * 3 * CAP_Q);
constexpr size_t O_CELL_OF = O_SORTED + al(sizeof(int) * CAP_P);
constexpr size_t O_CC = O_CELL_OF + al(sizeof(int) * CAP_P);
constexpr size_t O_CS = O_CC + al(sizeof(int) * (NA_MAX_CELLS + 1));
constexpr size_t O_CF = O_CS + al(sizeof(int) * (NA_MAX_CELLS + 2));
constexpr size_t O_KD = O_CF + al(sizeof(int) * (NA_MAX_CELLS + 1));
constexpr size_t KD_NODES = 2 * CAP_P + 2;
constexpr size_t KD_BYTES = al(sizeof(int) * KD_NODES) * 4 + al(sizeof(double) * KD_NODES) + al(sizeof(int) * CAP_P);
constexpr size_t O_KDW = O_KD + KD_BYTES;
// kd_build work space: kd_work_bytes(n) is below 256 bytes per point plus
// the scan space (checked at set-up)
constexpr size_t KDW_BYTES = 256 * CAP_P + 65536;
constexpr size_t O_SCAN = O_KDW + KDW_BYTES;
constexpr size_t SCAN_BYTES = 4 * (std::max<size_t>(NA_MAX_CELLS, CAP_Q) + 16) + 65536;
constexpr size_t O_OI = O_SCAN + SCAN_BYTES;                                 // ELL [W][m]
constexpr size_t O_OR = O_OI + al(sizeof(int) * NA_MAX_WIDTH * CAP_Q);
constexpr size_t O_DEG = O_OR + al(sizeof(double) * NA_MAX_WIDTH * CAP_Q);
constexpr size_t O_ROW32 = O_DEG + al(sizeof(int) * CAP_Q);
constexpr size_t O_ROW64 = O_ROW32 + al(sizeof(int) * (CAP_Q + 1));
constexpr size_t O_CI = O_ROW64 + al(sizeof(int64_t) * (CAP_Q + 1));         // CSR out
constexpr size_t O_CR = O_CI + al(sizeof(int) * NA_MAX_WIDTH * CAP_Q);
constexpr size_t O_MISC = O_CR + al(sizeof(double) * NA_MAX_WIDTH * CAP_Q);  // bbox, bbox tmp, DevErr
constexpr size_t NA_BYTES = O_MISC + 4096;

}  // namespace
}  // namespace ssa

// The scratch array (one copy per GPU, held until the process ends).
static __device__ __align__(256) unsigned char scratch[ssa::NA_BYTES];

namespace ssa {
namespace {

std::once_flag g_once[NA_MAX_DEVICES];
unsigned char *g_ptr[NA_MAX_DEVICES] = {};
hipError_t g_err[NA_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once).
unsigned char *scratch_ptr() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) throw HipError("noalloc: device index", hipErrorInvalidDevice, __FILE__, __LINE__);
    std::call_once(g_once[dev], [dev] {
        void *p = nullptr;
        g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[dev] = (unsigned char *)p;
    });
    if (g_err[dev] != hipSuccess || !g_ptr[dev])
        throw HipError("hipGetSymbolAddress(scratch)", g_err[dev] != hipSuccess ? g_err[dev] : hipErrorUnknown,
                       __FILE__, __LINE__);
    return g_ptr[dev];
}

__global__ void k_na_soa(const double *aos, int n, double *soa) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < 3; d++) soa[(size_t)d * n + i] = aos[3 * (size_t)i + d];
}

__global__ void k_na_row64(const int *row32, int m, int64_t *row64) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= m) row64[i] = row32[i];
}

// ELL [W][m] -> CSR by the row offsets
__global__ void k_na_csr(int m, int W, const int *deg, const int *row32, const int *oi, const double *orr, int *ci,
                         double *cr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int d = deg[i], o = row32[i];
    for (int k = 0; k < d && k < W; k++) {
        ci[o + k] = oi[(size_t)k * m + i];
        cr[o + k] = orr[(size_t)k * m + i];
    }
}

// Grid over the tree points' box with cell side >= h (every accepted pair
// lies in neighbouring cells) and at most NA_MAX_CELLS cells.  The lists do
// not depend on the grid: acceptance is exact and the order is (r^2, ANN
// visit order).
// Counts in double, as grid_setup: points far apart coarsen the grid
// instead of wrapping an integer product past the scratch array's cells.
void grid_capped(const double *bb, int dim, double h, Grid &g) {
    double cell = h * 1.01;
    for (;;) {
        double tot = 1.0, c[3] = {1.0, 1.0, 1.0};
        for (int d = 0; d < dim && d < 3; d++) {
            const double ext = bb[2 * d + 1] - bb[2 * d];
            c[d] = (std::isfinite(ext) && ext > 0.0) ? std::floor(ext / cell) + 1.0 : 1.0;
            tot *= c[d];
        }
        if (tot <= (double)NA_MAX_CELLS || !std::isfinite(cell)) {
            long long t = 1;
            for (int d = 0; d < 3; d++) {
                g.dims[d] = d < dim ? (int)std::min(c[d], (double)NA_MAX_CELLS) : 1;
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

}  // namespace
}  // namespace ssa

using namespace ssa;

#define NA_GUARD_BEGIN try {
#define NA_GUARD_END                                                         \
    }                                                                        \
    catch (const HipError &e) {                                              \
        return report_error(SSASDPD_ERR_HIP, e.msg);                            \
    }

extern "C" int ssasdpd_fr_search_noalloc(const double *pts, int32_t n, const double *q, int32_t m, int32_t dim,
                                         double h, int64_t *row_ptr, int32_t *idx, double *r2) {
    NA_GUARD_BEGIN
    if (n <= 0 || m < 0 || dim < 1 || dim > 3 || !(h > 0) || !pts || (m > 0 && !q) || !row_ptr)
        return report_error(SSASDPD_ERR_BAD_DESC, "fr_search_noalloc: bad arguments");
    if (n > NA_MAX_POINTS || m > NA_MAX_QUERIES)
        return report_error(SSASDPD_ERR_CAPACITY, "fr_search_noalloc: at most " + std::to_string(NA_MAX_POINTS) +
                                                   " points and " + std::to_string(NA_MAX_QUERIES) + " queries");
    if (kd_work_bytes(NA_MAX_POINTS) > KDW_BYTES || scan_tmp_bytes(std::max<long long>(NA_MAX_CELLS, CAP_Q) + 8) > SCAN_BYTES)
        return report_error(SSASDPD_ERR_CAPACITY, "fr_search_noalloc: scratch layout too small");
    unsigned char *base = scratch_ptr();
    hipStream_t st = 0;   // the null stream: nothing created per call
    double *stage = (double *)(base + O_STAGE), *dp = (double *)(base + O_PTS), *sp = (double *)(base + O_SPTS);
    double *dq = (double *)(base + O_Q);
    int *srt = (int *)(base + O_SORTED), *cof = (int *)(base + O_CELL_OF);
    int *cc = (int *)(base + O_CC), *cs = (int *)(base + O_CS), *cf = (int *)(base + O_CF);
    void *kw = base + O_KDW, *stmp = base + O_SCAN;
    int *oi = (int *)(base + O_OI), *dg = (int *)(base + O_DEG);
    double *orr = (double *)(base + O_OR);
    int *row32 = (int *)(base + O_ROW32);
    int64_t *row64 = (int64_t *)(base + O_ROW64);
    int *ci = (int *)(base + O_CI);
    double *cr = (double *)(base + O_CR);
    double *bb = (double *)(base + O_MISC);
    unsigned long long *bbt = (unsigned long long *)(base + O_MISC + 256);
    DevErr *de = (DevErr *)(base + O_MISC + 1024);
    KdDev kd;
    {
        unsigned char *k = base + O_KD;
        kd.parent = (int *)k; k += al(sizeof(int) * KD_NODES);
        kd.depth = (int *)k; k += al(sizeof(int) * KD_NODES);
        kd.cd = (int *)k; k += al(sizeof(int) * KD_NODES);
        kd.is_hi = (int *)k; k += al(sizeof(int) * KD_NODES);
        kd.cv = (double *)k; k += al(sizeof(double) * KD_NODES);
        kd.leaf_of = (int *)k;
    }
    const unsigned B = 256;
    // points: host AoS -> staging -> SoA
    HIP_OK(hipMemcpyAsync(stage, pts, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_na_soa, dim3((n + B - 1) / B), dim3(B), 0, st, stage, n, dp);
    if (m > 0) {
        HIP_OK(hipMemcpyAsync(stage, q, sizeof(double) * 3 * (size_t)m, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_na_soa, dim3((m + B - 1) / B), dim3(B), 0, st, stage, m, dq);
    }
    compute_bbox(st, dp, n, dim, bb, bbt);
    double hb[6];
    HIP_OK(hipMemcpyAsync(hb, bb, sizeof hb, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    Grid g;
    grid_capped(hb, dim, h, g);
    bin_points(st, dp, n, dim, g, cof, cc, cs, cf, srt, sp, stmp);
    kd_build(st, dp, n, dim, kd, kw, KDW_BYTES);
    int W = 16;
    DevErr e{};
    for (;;) {
        const DevErr e0{0, INT32_MAX, 0, 0, 0};
        HIP_OK(hipMemcpyAsync(de, &e0, sizeof e0, hipMemcpyHostToDevice, st));
        if (m > 0) fr_search(st, dq, m, sp, srt, cs, n, dim, h, g, &kd, W, oi, orr, dg, de, nullptr);
        HIP_OK(hipMemcpyAsync(&e, de, sizeof e, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (!e.overflow) break;
        const int want = ((e.max_deg + 7) / 8) * 8 + 8;
        if (want > NA_MAX_WIDTH && W >= NA_MAX_WIDTH)
            return report_error(SSASDPD_ERR_CAPACITY, "fr_search_noalloc: a list holds " + std::to_string(e.max_deg) +
                                                       " neighbours, more than " + std::to_string(NA_MAX_WIDTH));
        W = std::min(want, NA_MAX_WIDTH);
    }
    if (m == 0) {
        row_ptr[0] = 0;
        return 0;
    }
    exclusive_scan_i32(st, dg, row32, m, stmp);
    hipLaunchKernelGGL(k_na_row64, dim3((m + 1 + B - 1) / B), dim3(B), 0, st, row32, m, row64);
    HIP_OK(hipMemcpyAsync(row_ptr, row64, sizeof(int64_t) * ((size_t)m + 1), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (idx) {
        const size_t tot = (size_t)row_ptr[m];
        hipLaunchKernelGGL(k_na_csr, dim3((m + B - 1) / B), dim3(B), 0, st, m, W, dg, row32, oi, orr, ci, cr);
        HIP_OK(hipMemcpyAsync(idx, ci, sizeof(int) * tot, hipMemcpyDeviceToHost, st));
        if (r2) HIP_OK(hipMemcpyAsync(r2, cr, sizeof(double) * tot, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
    }
    return 0;
    NA_GUARD_END
}
