// Internal declarations shared by the engine's translation units.
// Device memory layout (all arrays live in HBM for the whole run):
//   particles: component-major FP64 [3][n] for x, v, vt, F, Fbp, normal;
//              FP64 [n] rho, old_rho, Frho, mass, nu, bvf; int32 type, solid
//   deterministic species: FP64 C[n][Sc], Q[n][Sc]
//   neighbour lists: ELL, column-major [W][n] for idx (int32), r (FP64 |x_i-x_j|)
//              and Dij (FP64); deg[n]
//   RDME: xx uint32 [n][S], rrate FP64 [n][R], srrate/sdrate FP64 [n],
//         Ddiag FP64 [n][S], firing times FP64 [n] + a 64-ary min tree
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ssasdpd.h"

namespace ssa {

#define HIP_OK(expr)                                                              \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) throw ::ssa::HipError(#expr, _e, __FILE__, __LINE__); \
    } while (0)

struct HipError {
    std::string msg;
    HipError(const char *what, hipError_t e, const char *file, int line) {
        msg = std::string(what) + ": " + hipGetErrorString(e) + " at " + file + ":" + std::to_string(line);
    }
};

// Record msg as the calling thread's ssasdpd_last_error() and return code.
int report_error(int code, const std::string &msg);

constexpr int WAVE = 64;
constexpr int TREE_FAN = 64;

// XCD-aware block order for particle-order kernels: the hardware deals
// consecutive workgroups round-robin to the 8 XCDs, each with its own L2;
// launching round_up(nb, 8) workgroups and mapping workgroup b to chunk
// (b % 8) * ceil(nb/8) + b / 8 gives every XCD one contiguous range of
// particles (neighbours in space share its L2).  Returns -1 past the end.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_chunk(int nb) {
    const int per = (nb + N_XCD - 1) / N_XCD;
    const int c = (int)(blockIdx.x % N_XCD) * per + (int)(blockIdx.x / N_XCD);
    return c < nb ? c : -1;
}
inline dim3 xcd_grid(int nb) { return dim3((unsigned)(((nb + N_XCD - 1) / N_XCD) * N_XCD)); }

// Device-side error record.
struct DevErr {
    int code;          // ssasdpd_status
    int particle;      // offending particle (NaN), or -1
    unsigned step;
    int max_deg;       // largest neighbour count seen by the last search
    int overflow;      // a list did not fit in the ELL width
    int crowded;       // a binned cell's count when it passed FR_MAX_CELL, else 0
};

// Mass-action reaction (same as ssasdpd_reaction, device copy).
struct Rxn {
    int kind, a, b, pad;
    double k;
    uint64_t mask;
};

struct BcRule {
    double lo[3], hi[3];
    int type, target, species, pad;
    double value[3];
};

// Per-trajectory SSA queue/time-window state (restates the scalar state of
// NRMConstant_v5, ENG/include/NRMConstant_v5.hpp:58-107).
struct QueueState {
    double L, U, endTime;
    unsigned long long nbins, rxnCount, active;
    unsigned long long rng_pos;      // replay: next word; seeded: draw counter
    unsigned long long reactions, diffusions, events_traced;
    int initialized;
    int err;
    unsigned err_step;       // step at which err was set
    int pad;
};

// Kd-tree of the tree points (ANN sliding-midpoint, bucket 1), flattened.
struct KdDev {
    int *parent = nullptr;   // [2n]
    int *depth = nullptr;    // [2n]
    int *cd = nullptr;       // [2n]
    double *cv = nullptr;    // [2n]
    int *is_hi = nullptr;    // [2n]
    int *leaf_of = nullptr;  // [n] leaf node of each point
    int n_nodes = 0;
    bool valid = false;
};

// Cell grid over the tree points.
struct Grid {
    double origin[3];
    double inv_cell;
    int dims[3];
    long long ncell;
    // tile[0] == 0: cell (cx, cy, cz) is (cz * dims[1] + cy) * dims[0] + cx,
    // as the search reads it.  Otherwise (a binning for the renumbering
    // only) cells are numbered tile by tile, tiles of tile[d] cells per
    // side, x fastest within a tile and among tiles; ncell counts the key
    // space, partial tiles included.
    int tile[3];
    int ntile[3];
};
// The tiled numbering of g's cells (see Grid).
void grid_tile(Grid &g, int dim);

// What the neighbour loops gather from particle j, packed after the
// predictor: one 128-byte line per neighbour instead of a load from each of
// twelve component arrays.  Written by k_predict (and k_pack); the force,
// D_ij, density-filter and boundary-volume-fraction kernels read it.
struct __align__(16) PRec {
    double x[3], v[3], vt[3];
    double nu, rho, mass;        // rho, mass: one aligned 16-byte pair
    int type, solid, pad0, pad1;
    double pad2[2];
};
static_assert(sizeof(PRec) == 128, "PRec is one 128-byte line");
static_assert(offsetof(PRec, rho) % 16 == 0, "rho, mass load as one double2");

struct Bufs;   // defined in engine.hip

// ---- launch wrappers (implemented in nbr.hip / sdpd.hip / ssa.hip) -------
// Neighbour search: bins tree points pts[3][n] into cells, then for each
// query q[3][m] collects accepted points, sorts them in ANN order and writes
// ELL rows (idx, r2) of width W.  Returns via err->overflow when W is too
// small.
void grid_setup(const double *pts_host_bbox, int n, int dim, double h, Grid &g);
// Most points one search cell may hold: past it the grid is taken as broken
// (one point far from the rest coarsens every cell until the bulk of the
// points shares a few), and the engine's search reports an error instead of
// scanning them for every query.
constexpr int FR_MAX_CELL = 4096;
void bin_points(hipStream_t s, const double *pts, int n, int dim, const Grid &g, int *cell_of,
                int *cell_count, int *cell_start, int *cell_fill, int *sorted, double *spts,
                void *scan_tmp,    // spts: sorted points [n][4] (x, y, z, 0)
                int *crowded = nullptr);   // raised to a cell's count past FR_MAX_CELL
// Within each cell of a binning, the point indices of sorted[] in ascending
// order (the scatter's atomics leave them in arrival order).  spts is not
// reordered: the result is an ordering, not a binning to search.
void cell_order_by_index(hipStream_t s, const int *cell_start, long long ncell, int *sorted);
void compute_bbox(hipStream_t s, const double *pts, int n, int dim, double *d_bbox /*[6]*/,
                  unsigned long long *tmp);
// One edge of a voxel's neighbour row as the event loop reads it: a row is
// contiguous (erow[i*W + k]), so one event's row load touches a few cache
// lines instead of one per slot of the neighbour-major ELL arrays.
struct EdgeRec {
    int j;        // neighbour
    int type;     // its type
    double D;     // D_ij
};

// The per-edge finish of add_to_neighbor_list (r, D_ij; ENG/src/particle.cpp:
// 150-210) plus the neighbour's type and the event loop's voxel-major row,
// done by the search itself once a query's list is sorted.  Own mass and rho
// from the arrays; the neighbour's from rec when non-null, else the arrays.
// Ddiag (optional, models of at most FIN_MAX_SPECIES species and
// FIN_MAX_DMAT species x type entries): the finish also forms
// Ddiag[i][s] = sum over the list, in list order, of Dmat[s][type_j - 1] *
// D_ij (nsm_core__initialize_diff_propensities, simulate_rdme.cpp:131-152),
// which then needs no pass of its own; Dmat is read from an LDS copy.
constexpr int FIN_MAX_SPECIES = 4;
constexpr int FIN_MAX_DMAT = 64;
struct FinishArgs {
    double *r, *Dij;         // Dij, ntype: column-major per edge, or null (not formed)
    int *ntype;
    EdgeRec *erow;
    const double *mass, *rho;
    const int *type;
    const PRec *rec;
    double *Ddiag;
    const double *Dmat;
    int S, T;
};
void fr_search(hipStream_t s, const double *q, int m, const double *spts, const int *sorted,
               const int *cell_start, int n, int dim, double h, const Grid &g, const KdDev *kd,
               int W, int *out_idx, double *out_r2, int *deg, DevErr *err, const int *query_order,
               const FinishArgs *fin = nullptr, int slots_hint = 0);
void kd_build(hipStream_t s, const double *pts, int n, int dim, KdDev &kd, void *work, size_t work_bytes);
size_t kd_work_bytes(int n);
size_t scan_tmp_bytes(long long n);
void exclusive_scan_i32(hipStream_t s, const int *in, int *out, long long n, void *tmp);


// SDPD substeps.
struct SdpdArgs {
    int n, dim, Sc, Rc, T, static_domain, nbc, W;
    int fp32;                // pair-force terms in FP32 (FP64 accumulation)
    int skip_F;              // leave F at gravity (the surrogate supplies it)
    int no_solids;           // no particle is solid (the BVF then needs no neighbour data)
    unsigned step;
    double dt, h, rho0, c0, P0, alpha;
    double g[3];
    double *x, *v, *vt, *F, *Fbp, *rho, *old_rho, *Frho, *bvf, *normal;
    double *vol;             // mass / rho after the step's last update (RDME), or null
    const double *mass;
    double *nu;
    const int *type, *solid;
    double *C, *Q;
    const double *Dmat;      // [S][T] species-major
    const int *Ndense;       // [Sc][Rc]
    const Rxn *rx;
    const BcRule *bc;
    const int *deg, *nidx;
    const double *nr;
    PRec *rec;               // [n] packed gather records (after the predictor)
    DevErr *err;
    // expression-VM programs (kind-4 reactions) and data functions [n][n_df]
    const uint32_t *vm_code;
    const double *vm_k;
    const double *data_fn;
    int n_df;
};
void sdpd_predict(hipStream_t s, const SdpdArgs &a);
void sdpd_forces(hipStream_t s, const SdpdArgs &a);
void sdpd_correct(hipStream_t s, const SdpdArgs &a);
void sdpd_bvf(hipStream_t s, const SdpdArgs &a);
// rec from the current particle arrays (for force evaluations outside a step)
void sdpd_pack(hipStream_t s, const SdpdArgs &a);

// RDME.
struct RdmeArgs {
    int n, S, R, T, W;
    const double *mass, *rho;
    const int *type;
    uint32_t *xx;
    double *rrate, *srrate, *sdrate, *Ddiag;
    const double *Dmat;
    const Rxn *rx;
    const int *jcN, *irN, *prN, *jcG, *irG;
    const int *deg, *nidx, *ntype;
    const double *Dij;
    const EdgeRec *erow;     // the same rows voxel-major (event loop)
    int row_w;               // no row is longer (<= W)
    double *vol;             // mass/rho per voxel, set by the rates kernel
    // queue: level 0 = firing times of the voxels; level k >= 1 holds the
    // (time, voxel) minimum of its children: F0 = 64*pl voxels per level-1
    // node, 64 nodes per node above.  Levels 1..nlev live in tmin/targ at
    // lev_off[k]; level nlev is the root.
    double *qtime;
    double *tmin;
    int *targ;
    int pl;                  // level-0 children per lane (F0 = 64*pl)
    int nlev;                // number of levels above level 0 (>= 1)
    int lev_off[10];
    int lev_n[10];           // nodes per level (lev_n[0] = n)
    int tree_nodes;          // sum of lev_n[1..nlev]
    size_t lds_bytes;        // dynamic LDS of the event-loop kernel
    QueueState *qs;
    // random stream
    int rng_kind;            // 0 seeded (Philox), 1 replay
    unsigned long long seed;
    unsigned long long stream_id;
    const uint64_t *words;
    const double *expo;
    unsigned long long n_words;
    // options
    int textbook;
    ssasdpd_event *trace;
    unsigned long long trace_cap;
    unsigned long long max_events;
    double t0, dt, endTime;
    unsigned step;
    double *psum_out;
    int *active_out;
    // batched ensembles: trajectory b = blockIdx.x (rates: blockIdx.y) uses
    // its own counts / rates / queue at these element strides; neighbour
    // lists, Ddiag and vol are shared
    size_t bs_xx, bs_rr, bs_n, bs_tree;
    const unsigned long long *seeds;   // per-trajectory seeds, or null
    // persistent mode: steps [step, step_end) in one launch (static domains
    // whose particle phase is constant after step 0).  Snapshot k of this
    // launch is taken before step snap_step[k] (after the last step when
    // snap_step[k] == step_end): counts into snap_xx + k*n*S (+ b*bs_snap),
    // species totals into snap_tot[(b*n_snap + k)*S + s]; either may be null.
    unsigned step_end;
    const unsigned *snap_step;
    int n_snap;
    uint32_t *snap_xx;
    size_t bs_snap;
    uint32_t *snap_tot;
    // expression-VM programs of kind-4 reactions (copied to LDS by the event
    // loop) and data functions [n][n_df]
    const uint32_t *vm_code;
    const double *vm_k;
    int vm_ncode, vm_nk;
    const double *data_fn;
    int n_df;
};
// Ddiag from this step's lists, for models with more than FIN_MAX_SPECIES
// species (the search's finish forms it otherwise; k_bvf forms vol)
void rdme_diag(hipStream_t s, const RdmeArgs &a);
void rdme_rates(hipStream_t s, const RdmeArgs &a, int n_traj = 1);
void rdme_queue_build(hipStream_t s, const RdmeArgs &a, void *scan_tmp, int *rank_tmp);
void rdme_take_step(hipStream_t s, const RdmeArgs &a, int n_traj = 1);
void tree_rebuild(hipStream_t s, const RdmeArgs &a);
// one launch over n_models argument blocks (device array), common PL and
// MS/MR bucket (max over the models of max(S, R)), LDS = max over models
void rdme_take_step_multi(hipStream_t s, const RdmeArgs *d_args, int n_models, int pl, int max_sr, size_t lds,
                          bool any_prog);
void vm_eval_batch(hipStream_t s, const uint32_t *code, int start, const double *K, int x_is_double, int n_x,
                   int n_df, int m, const double *xd, const uint32_t *xu, const double *vol, const double *t,
                   const double *df, double *out);
// Row permutation of an [n][cols] array of esz-byte elements (4 or 8):
// gather dst[p] = src[map[p]], or scatter dst[map[p]] = src[p] (sdpd.hip).
void permute_rows(hipStream_t s, int n, int cols, int esz, const int *map, const void *src, void *dst, bool scatter);
size_t ssa_lds_bytes(int tree_nodes, int S, int R, int T, int nnzN, int nnzG, int n_code = 0, int n_consts = 0);

}  // namespace ssa
