// Host side of the engine: device memory, the step loop and the C ABI
// (include/ssasdpd.h).
//
// One engine = one model on one GPU.  Work is issued on one HIP stream; the
// step loop follows run_simulation (ENG/src/simulate_threads.cpp:171-312)
// with the three worker substeps (ENG/src/simulate.cpp:33-173) and the serial
// RDME step (ENG/src/simulate_rdme.cpp:53-68) as kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <random>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "engine.hpp"
#include "vm.hpp"
#include "mpn.hpp"

using namespace ssa;

namespace {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

template <class T> T *dalloc(size_t count) {
    T *p = nullptr;
    if (count == 0) count = 1;
    HIP_OK(hipMalloc((void **)&p, sizeof(T) * count));
    return p;
}
template <class T> void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}
template <class T> void h2d(hipStream_t s, T *dst, const T *src, size_t count) {
    if (count) HIP_OK(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
}
template <class T> void d2h(hipStream_t s, T *dst, const T *src, size_t count) {
    if (count) HIP_OK(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost, s));
}

// component-major [3][n] <-> particle-major [n][3]
std::vector<double> to_soa3(const double *aos, int n) {
    std::vector<double> o(3 * (size_t)n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) o[(size_t)k * n + i] = aos ? aos[3 * (size_t)i + k] : 0.0;
    return o;
}
void from_soa3(const std::vector<double> &soa, double *aos, int n) {
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) aos[3 * (size_t)i + k] = soa[(size_t)k * n + i];
}

double kernel_alpha(int dim, double h) {
    // ENG/src/particle.cpp:169-175 (1-D: 5 / 4 * h is integer 1 times h)
    if (dim == 3) return 105 / (16 * M_PI * h * h * h);
    if (dim == 2) return 5 / (M_PI * h * h);
    return 5 / 4 * h;
}

}  // namespace

// The calling thread's error message, for the other translation units.
int ssa::report_error(int code, const std::string &msg) { return set_error(code, msg); }

// With a time limit set, a fused static-domain run launches at most this
// many steps at a time and checks the limit between launches.
constexpr uint32_t FUSED_SEGMENT = 64;

struct ssasdpd_engine {
    // configuration
    int n = 0, dim = 3, T = 1, S = 0, R = 0, Sc = 0, Rc = 0, static_domain = 1, nbc = 0;
    uint32_t nt = 0, flags = 0;
    double dt = 0, h = 0, rho0 = 1, c0 = 10, P0 = 100, alpha = 0;
    double g[3] = {0, 0, 0};
    std::vector<uint32_t> out_steps;
    // initial state (host)
    std::vector<double> x0, v0, rho0v, C0;
    std::vector<uint32_t> u0;
    std::vector<double> nu0;

    hipStream_t st = nullptr;
    // particles
    double *x = nullptr, *v = nullptr, *vt = nullptr, *F = nullptr, *Fbp = nullptr, *normal = nullptr;
    double *rho = nullptr, *old_rho = nullptr, *Frho = nullptr, *mass = nullptr, *nu = nullptr, *bvf = nullptr;
    int *type = nullptr, *solid = nullptr;
    double *C = nullptr, *Q = nullptr;
    // neighbour lists
    int W = 0;
    int *nidx = nullptr, *deg = nullptr;
    double *nr2 = nullptr, *nr = nullptr, *Dij = nullptr;
    EdgeRec *erow = nullptr;   // the rows voxel-major, for the event loop
    // Look-ahead (advance on moving domains): the next step's particle phase
    // runs on its own stream while this step's event loop runs.  The event
    // loop reads only deg and erow of the lists, so those two are double
    // buffered; it keeps the previous step's pair.
    hipStream_t ps = nullptr;
    hipEvent_t la_done = nullptr;
    bool la_pending = false;
    hipEvent_t ev_bbox = nullptr;   // the bounding box has reached the host
    hipEvent_t ev_srch = nullptr;   // a search's error record has reached the host
    // a static domain's queue state, read into pinned memory behind the
    // event loop's launch (ev_q: it has arrived)
    QueueState *h_q = nullptr;
    hipEvent_t ev_q = nullptr;
    // a step's error record and queue state, read back together after its
    // event loop (the same pinned block)
    DevErr *h_err = nullptr;
    QueueState *h_qc = nullptr;
    // A look-ahead starts behind ev_rt, recorded on the engine stream after
    // this step's particle phase (which may have run there).  A renumber
    // inside the look-ahead: the event loop beside it reads the types from
    // rtype (a copy made before ev_rt) while the particle arrays are
    // permuted; the counts are permuted after it on the engine stream, by
    // the order kept in rmap (ev_map), through xbuf.
    int *rtype = nullptr, *rmap = nullptr;
    uint32_t *xbuf = nullptr;
    hipEvent_t ev_rt = nullptr, ev_map = nullptr;
    // The particle phase also forms vol and Ddiag (the particle part of the
    // rates); the event loop reads them, so a look-ahead writes the other
    // pair, swapped in with deg and erow.
    double *Ddiag_alt = nullptr, *vol_alt = nullptr;
    // Wall-clock limit of one run / run_batch / advance call (0: none),
    // checked between steps and between the segments of a fused static run.
    double timeout_s = 0;
    bool deadline_on = false;
    std::chrono::steady_clock::time_point deadline{};
    void arm_deadline() {
        deadline_on = timeout_s > 0;
        if (deadline_on)
            deadline = std::chrono::steady_clock::now() +
                       std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                           std::chrono::duration<double>(timeout_s));
    }
    // SSASDPD_ERR_TIMEOUT once the deadline has passed (the state is that
    // after step cur - 1), else 0.
    int deadline_rc() const {
        if (!deadline_on || std::chrono::steady_clock::now() < deadline) return 0;
        char buf[96];
        snprintf(buf, sizeof buf, "timeout after %.3f s: stopped before step %u", timeout_s, cur);
        return set_error(SSASDPD_ERR_TIMEOUT, buf);
    }
    int *deg_alt = nullptr;
    EdgeRec *erow_alt = nullptr;
    // binning / tree
    double *xtree = nullptr, *spts = nullptr, *d_bbox = nullptr;
    PRec *rec = nullptr;        // packed gather records, written by the predictor
    bool any_solid = true;      // some particle is fixed (solidTag, static)
    bool rec_ready = false;     // rec holds this step's post-predictor state
    // pinned landing slot for the per-step bbox readback: a pageable copy goes
    // through the runtime's shared staging buffer, which can hold the particle
    // stream behind a running event loop
    double *h_bbox = nullptr;
    unsigned long long *d_bbtmp = nullptr;
    int *sorted = nullptr, *cell_of = nullptr, *cell_count = nullptr, *cell_start = nullptr, *cell_fill = nullptr;
    long long cell_cap = 0;
    Grid grid{};
    KdDev kd;
    void *kd_work = nullptr;
    size_t kd_work_sz = 0;
    void *scan_tmp = nullptr;
    void *scan_tmp_q = nullptr;   // the queue build's (it can run beside a binning)
    // chemistry constants
    Rxn *rx = nullptr;
    // expression-VM programs of kind-4 reactions, data functions [n][n_df]
    uint32_t *vm_code = nullptr;
    double *vm_k = nullptr, *dfn = nullptr;
    int n_code = 0, n_consts = 0, n_df = 0;
    int *jcN = nullptr, *irN = nullptr, *prN = nullptr, *jcG = nullptr, *irG = nullptr, *Ndense = nullptr;
    double *Dmat = nullptr;
    BcRule *bc = nullptr;
    // RDME state
    uint32_t *xx = nullptr;
    double *rrate = nullptr, *srrate = nullptr, *sdrate = nullptr, *Ddiag = nullptr, *qtime = nullptr, *tmin = nullptr;
    double *vol = nullptr;
    int *targ = nullptr, *ntype = nullptr;
    int pl = 1, nlev = 0, lev_off[10] = {0}, lev_n[10] = {0}, tree_nodes = 0;
    size_t lds_bytes = 0;
    QueueState *qs = nullptr;
    int *qwork = nullptr;
    double *d_psum = nullptr;
    // random stream
    int rng_kind = 0;
    uint64_t seed = 0;
    uint64_t *d_words = nullptr;
    double *d_expo = nullptr;
    uint64_t n_words = 0, words_cap = 0;
    // trace
    ssasdpd_event *d_trace = nullptr;
    uint64_t trace_cap = 0;
    // errors
    // the particle phase's error record; a look-ahead writes the other one
    // (swapped in with deg and erow), so each step's check reads its own
    DevErr *derr = nullptr, *derr_alt = nullptr;
    // progress
    uint32_t cur = 0;
    bool lists_built = false;
    bool kd_built = false;
    bool rdme_initialized = false;
    ssasdpd_stats stats{};
    int max_deg_seen = 0;
    int last_max_deg = 0;       // longest list of the last search
    // Particle renumbering (moving domains, seeded mode): every
    // RENUMBER_EVERY steps the particles are put in cell order, so that
    // neighbours in space are neighbours in memory.  perm[p] is the caller's
    // index of internal particle p (the identity until the first renumber
    // and after every reset); outputs go through it back to that numbering.
    static constexpr unsigned RENUMBER_EVERY = 20;
    bool renumbered = false;
    int *perm = nullptr, *perm_tmp = nullptr;
    // host copies, refreshed on first use after a renumber (the run itself
    // needs neither)
    std::vector<int> perm_h;
    bool perm_h_stale = false, movers_stale = false;
    const std::vector<int> &perm_host() {
        if (perm_h_stale) {
            d2h(st, perm_h.data(), perm, n);
            HIP_OK(hipStreamSynchronize(st));
            perm_h_stale = false;
        }
        return perm_h;
    }
    void *pbuf = nullptr, *stage = nullptr;
    size_t pbuf_cap = 0, stage_cap = 0;
    std::vector<double> mass_h, dfn_h;
    std::vector<int> type_h, solid_h;
    // GPU time per phase, HIP events on the engine stream: 0 binning + tree,
    // 1 search + D_ij, 2 SDPD substeps, 3 rates + queue build, 4 SSA loop,
    // 5 force kernel alone
    static constexpr int NPH = 6;
    hipEvent_t ev_a[NPH][2] = {}, ev_b[NPH][2] = {};
    int ev_n[NPH] = {0};
    double ph_ms[NPH] = {0};
    uint64_t ph_launch[NPH] = {0};

    void ph_begin(int p) {
        if (!ev_a[p][0]) return;
        int k = ev_n[p] & 1;
        if (pending[p][k]) ph_take(p, k);   // a phase timed twice in one step
        HIP_OK(hipEventRecord(ev_a[p][k], st));
    }
    void ph_end(int p) {
        if (!ev_a[p][0]) return;
        int k = ev_n[p] & 1;
        HIP_OK(hipEventRecord(ev_b[p][k], st));
        ev_n[p]++;
        pending[p][k] = true;
        ahead[p][k] = in_lookahead;
    }
    bool pending[NPH][2] = {};
    // ahead: recorded by a look-ahead (the next step's particle phase on the
    // second stream), which the host does not wait for within its step
    bool ahead[NPH][2] = {};
    bool in_lookahead = false;
    void ph_take(int p, int k) {
        float ms = 0;
        HIP_OK(hipEventSynchronize(ev_b[p][k]));
        HIP_OK(hipEventElapsedTime(&ms, ev_a[p][k], ev_b[p][k]));
        ph_ms[p] += ms;
        ph_launch[p]++;
        pending[p][k] = false;
    }
    // skip_ahead: leave the phases of a look-ahead launched in this step for
    // the next call (by then the engine stream has waited for them), so the
    // host does not wait here for the next step's particle phase
    void ph_collect(bool skip_ahead = false) {
        for (int p = 0; p < NPH; p++)
            for (int k = 0; k < 2; k++)
                if (pending[p][k]) {
                    if (skip_ahead && ahead[p][k]) {
                        ahead[p][k] = false;
                        continue;
                    }
                    ph_take(p, k);
                }
    }

    ~ssasdpd_engine() { release(); }

    void release() {
        dfree(nn_F); dfree(nn_acc);
        dfree(perm); dfree(perm_tmp);
        if (pbuf) (void)hipFree(pbuf);
        if (stage) (void)hipFree(stage);
        pbuf = stage = nullptr;
        pbuf_cap = stage_cap = 0;
        dfree(b_xx); dfree(b_rr); dfree(b_sr); dfree(b_sd); dfree(b_qt); dfree(b_tmin); dfree(b_targ);
        dfree(b_qs); dfree(b_seeds);
        dfree(x); dfree(v); dfree(vt); dfree(F); dfree(Fbp); dfree(normal);
        dfree(rho); dfree(old_rho); dfree(Frho); dfree(mass); dfree(nu); dfree(bvf);
        dfree(type); dfree(solid); dfree(C); dfree(Q);
        if (ps) (void)hipStreamSynchronize(ps);
        dfree(nidx); dfree(deg); dfree(nr2); dfree(nr); dfree(Dij); dfree(ntype); dfree(erow); dfree(vol);
        dfree(deg_alt); dfree(erow_alt);
        if (la_done) (void)hipEventDestroy(la_done);
        la_done = nullptr;
        if (ev_bbox) (void)hipEventDestroy(ev_bbox);
        ev_bbox = nullptr;
        if (ev_srch) (void)hipEventDestroy(ev_srch);
        ev_srch = nullptr;
        if (ev_q) (void)hipEventDestroy(ev_q);
        ev_q = nullptr;
        h_q = h_qc = nullptr;   // inside h_bbox's block
        h_err = nullptr;
        if (ev_rt) (void)hipEventDestroy(ev_rt);
        if (ev_map) (void)hipEventDestroy(ev_map);
        ev_rt = ev_map = nullptr;
        dfree(rtype); dfree(rmap); dfree(xbuf); dfree(Ddiag_alt); dfree(vol_alt);
        if (ps) (void)hipStreamDestroy(ps);
        ps = nullptr;
        dfree(xtree); dfree(spts); dfree(d_bbox); dfree(d_bbtmp); dfree(rec);
        if (h_bbox) (void)hipHostFree(h_bbox);
        h_bbox = nullptr;
        dfree(sorted); dfree(cell_of); dfree(cell_count); dfree(cell_start); dfree(cell_fill);
        dfree(kd.parent); dfree(kd.depth); dfree(kd.cd); dfree(kd.cv); dfree(kd.is_hi); dfree(kd.leaf_of);
        if (kd_work) (void)hipFree(kd_work);
        kd_work = nullptr;
        if (scan_tmp) (void)hipFree(scan_tmp);
        scan_tmp = nullptr;
        if (scan_tmp_q) (void)hipFree(scan_tmp_q);
        scan_tmp_q = nullptr;
        dfree(derr_alt);
        dfree(vm_code); dfree(vm_k); dfree(dfn);
        dfree(rx); dfree(jcN); dfree(irN); dfree(prN); dfree(jcG); dfree(irG); dfree(Ndense); dfree(Dmat); dfree(bc);
        dfree(xx); dfree(rrate); dfree(srrate); dfree(sdrate); dfree(Ddiag); dfree(qtime); dfree(tmin); dfree(targ);
        dfree(qs); dfree(qwork); dfree(d_psum); dfree(d_words); dfree(d_expo); dfree(d_trace); dfree(derr);
        for (int p = 0; p < NPH; p++)
            for (int k = 0; k < 2; k++) {
                if (ev_a[p][k]) (void)hipEventDestroy(ev_a[p][k]);
                if (ev_b[p][k]) (void)hipEventDestroy(ev_b[p][k]);
                ev_a[p][k] = ev_b[p][k] = nullptr;
            }
        if (st && !ext_stream) (void)hipStreamDestroy(st);
        st = nullptr;
    }

    void alloc_lists(int w) {
        // the event loop of a step may still read the other row buffer
        if (ps) HIP_OK(hipDeviceSynchronize());
        dfree(nidx); dfree(nr2); dfree(nr); dfree(Dij); dfree(ntype); dfree(erow); dfree(erow_alt);
        W = w;
        size_t cap = (size_t)W * n;
        nidx = dalloc<int>(cap);
        ntype = dalloc<int>(cap);
        nr2 = dalloc<double>(cap);
        nr = dalloc<double>(cap);
        Dij = dalloc<double>(cap);
        erow = dalloc<EdgeRec>(cap);
        if (ps) erow_alt = dalloc<EdgeRec>(cap);
    }

    // options set before setup (ensembles): a common event-tree width and a
    // shared stream
    int pl_override = 0;
    hipStream_t ext_stream = nullptr;
    int nnzN_ = 0, nnzG_ = 0;
    // surrogate for the pair accelerations (null: plain SDPD)
    Mpn *nn = nullptr;
    int nn_check_every = 50;
    double nn_tol = 0.02, nn_last_err = -1.0, nn_max_err = 0.0;
    bool nn_active = false;
    uint64_t nn_steps = 0, nn_checks = 0;
    int nn_refresh_iters = 0, nn_refresh_keep = 32;
    float nn_refresh_lr = 1e-4f;
    uint64_t nn_refreshes = 0, nn_fails = 0;
    bool nn_rearm = false;
    Mpn *nn_rec = nullptr;         // record training samples during plain steps
    int nn_rec_every = 0, nn_rec_max = 32768;
    uint64_t nn_rec_seed = 1;
    // nn_acc: the check's sums [0], [1], its count of moving particles [2]
    // and the count of capped predictions [3] (uint64)
    double *nn_F = nullptr, *nn_acc = nullptr;
    // Between checks the network's pair acceleration of a particle is capped
    // at nn_cap_factor times the RMS plain pair acceleration of the last
    // check (off until the first check, or with a factor <= 0): a particle
    // in a state the network has not seen cannot run away before the next
    // check catches it.
    double nn_cap_factor = 20.0;
    double nn_acc_cap = 0.0;
    void nn_clear_capped() {
        if (nn_acc) HIP_OK(hipMemsetAsync(nn_acc + 3, 0, sizeof(double), st));
    }
    int nn_cap = 0;
    std::vector<int> movers;       // particles that are not fixed (internal numbering)
    const std::vector<int> &moving() {
        if (movers_stale) rebuild_movers();
        return movers;
    }

    // the network's view of the current state: the gather records of the
    // predictor (or of eval_forces) and this step's lists
    EcnGraph nn_graph() {
        EcnGraph gr;
        gr.n = n; gr.W = W; gr.dim = dim; gr.h = h; gr.rec = rec; gr.deg = deg; gr.nidx = nidx; gr.nr = nr;
        return gr;
    }
    void nn_ensure() {
        if (nn_cap >= n) return;
        dfree(nn_F); dfree(nn_acc);
        nn_F = dalloc<double>((size_t)n * 3);
        nn_acc = dalloc<double>(4);
        HIP_OK(hipMemsetAsync(nn_acc, 0, sizeof(double) * 4, st));
        nn_cap = n;
    }
    // up to maxp moving particles, a uniform random subset in index order
    std::vector<int> pick_particles(int maxp, uint64_t seed) {
        std::vector<int> sel = moving();
        if ((int)sel.size() > maxp) {
            std::mt19937_64 rng(seed);
            for (int k = 0; k < maxp; k++) {
                std::uniform_int_distribution<size_t> u(k, sel.size() - 1);
                std::swap(sel[k], sel[u(rng)]);
            }
            sel.resize(maxp);
            std::sort(sel.begin(), sel.end());
        }
        return sel;
    }
    // A training sample from the state the force kernel just saw (after the
    // predictor, lists of this step) with its accelerations F as the target.
    int record_sample(Mpn *M, int maxp, uint64_t seed) {
        std::vector<int> dg(n);
        d2h(st, dg.data(), deg, n);
        HIP_OK(hipStreamSynchronize(st));
        const std::vector<int> sel = pick_particles(maxp, seed);
        if (sel.empty()) return -1;
        return M->add_sample(st, nn_graph(), dg.data(), sel.data(), (int)sel.size(), nullptr, F, g);
    }

    // the network's accelerations for the current state into Fout (SoA
    // [3][n])
    // capped: the network drives this step (the acceleration cap applies)
    void nn_predict(double *Fout, bool capped = false) {
        if (!Fout) throw HipError("surrogate output buffer missing", hipErrorInvalidValue, __FILE__, __LINE__);
        nn->infer(st, nn_graph(), g, Fout, capped ? nn_acc_cap : 0.0,
                  capped ? (unsigned long long *)(nn_acc + 3) : nullptr);
    }

    // Level-1 nodes held in the event loop's LDS (with the levels above:
    // about 48 KiB at the limit).  The narrowest width that fits is the
    // fastest: each event loads one level-0 node of 64*pl firing times.
    static constexpr int L1_MAX = 4096;
    static int auto_pl(int nv) {
        int p = 1;
        while (p < 32 && (nv + 64 * p - 1) / (64 * p) > L1_MAX) p *= 2;
        return p;
    }
    static bool pl_valid(int p, int nv) {
        return (p == 1 || p == 2 || p == 4 || p == 8 || p == 16 || p == 32) && (nv + 64 * p - 1) / (64 * p) <= L1_MAX;
    }
    // Level sizes/offsets for width p, tree storage and the event loop's LDS.
    void layout_tree(int p) {
        if (!pl_valid(p, n))
            throw HipError("too many voxels for the event tree (max 8.3M)", hipErrorUnknown, __FILE__, __LINE__);
        pl = p;
        lev_n[0] = n;
        lev_n[1] = (n + 64 * pl - 1) / (64 * pl);
        lev_off[1] = 0;
        nlev = 1;
        tree_nodes = lev_n[1];
        while (lev_n[nlev] > 1) {
            lev_n[nlev + 1] = (lev_n[nlev] + TREE_FAN - 1) / TREE_FAN;
            lev_off[nlev + 1] = tree_nodes;
            tree_nodes += lev_n[nlev + 1];
            nlev++;
        }
        // the event loop holds level nlev-1 in registers and repairs the
        // levels below it with at most one level in LDS
        if (nlev > 3)
            throw HipError("event tree deeper than three levels", hipErrorUnknown, __FILE__, __LINE__);
        dfree(tmin); dfree(targ);
        tmin = dalloc<double>(tree_nodes);
        targ = dalloc<int>(tree_nodes);
        lds_bytes = ssa_lds_bytes(tree_nodes, S, R, T, nnzN_, nnzG_, n_code, n_consts);
        if (lds_bytes > 65536)
            throw HipError("event-loop constants exceed 64 KiB of LDS", hipErrorUnknown, __FILE__, __LINE__);
    }
    // Change the tree width of a built queue: the levels are rebuilt from
    // the firing times, so the event sequence is unchanged.
    void set_pl(int p) {
        if (p == pl) return;
        layout_tree(p);
        if (rdme_initialized) tree_rebuild(st, rdme_args(cur));
    }

    void setup(const ssasdpd_desc *d) {
        n = d->n; dim = d->dimension; T = d->n_types; S = d->n_species; R = d->n_reactions;
        Sc = d->n_chem_species; Rc = d->n_chem_reactions; static_domain = d->static_domain;
        nbc = d->n_bcs; nt = d->nt; flags = d->flags;
        dt = d->dt; h = d->h; rho0 = d->rho0; c0 = d->c0; P0 = d->P0;
        for (int k = 0; k < 3; k++) g[k] = d->gravity[k];
        alpha = kernel_alpha(dim, h);
        out_steps.assign(d->output_steps, d->output_steps + d->n_output_steps);
        if (ext_stream) st = ext_stream;
        else {
            HIP_OK(hipStreamCreate(&st));
            HIP_OK(hipStreamCreate(&ps));
            HIP_OK(hipEventCreateWithFlags(&la_done, hipEventDisableTiming));
        }
        HIP_OK(hipEventCreateWithFlags(&ev_bbox, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&ev_srch, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&ev_q, hipEventDisableTiming));
        for (int p = 0; p < NPH; p++)
            for (int k = 0; k < 2; k++) {
                HIP_OK(hipEventCreate(&ev_a[p][k]));
                HIP_OK(hipEventCreate(&ev_b[p][k]));
            }
        size_t n3 = 3 * (size_t)n;
        x = dalloc<double>(n3); v = dalloc<double>(n3); vt = dalloc<double>(n3); F = dalloc<double>(n3);
        Fbp = dalloc<double>(n3); normal = dalloc<double>(n3); xtree = dalloc<double>(n3);
        spts = dalloc<double>(4 * (size_t)n);
        rec = dalloc<PRec>(n);
        rho = dalloc<double>(n); old_rho = dalloc<double>(n); Frho = dalloc<double>(n); mass = dalloc<double>(n);
        nu = dalloc<double>(n); bvf = dalloc<double>(n);
        type = dalloc<int>(n); solid = dalloc<int>(n);
        C = dalloc<double>((size_t)n * Sc); Q = dalloc<double>((size_t)n * Sc);
        deg = dalloc<int>(n);
        if (ps) deg_alt = dalloc<int>(n);
        sorted = dalloc<int>(n); cell_of = dalloc<int>(n);
        d_bbox = dalloc<double>(8); d_bbtmp = dalloc<unsigned long long>(8);
        {
            // one pinned block: the bounding box, the queue state and the
            // error record read back each step
            constexpr size_t qsz = ((sizeof(QueueState) + 63) / 64) * 64;
            constexpr size_t o_q = 64, o_qc = o_q + qsz, o_err = o_qc + qsz;
            void *blk = nullptr;
            HIP_OK(hipHostMalloc(&blk, o_err + ((sizeof(DevErr) + 63) / 64) * 64, hipHostMallocDefault));
            h_bbox = (double *)blk;
            h_q = (QueueState *)((char *)blk + o_q);
            h_qc = (QueueState *)((char *)blk + o_qc);
            h_err = (DevErr *)((char *)blk + o_err);
        }
        derr = dalloc<DevErr>(1);
        if (ps) derr_alt = dalloc<DevErr>(1);
        d_psum = dalloc<double>(2);
        kd.parent = dalloc<int>(2 * (size_t)n + 2); kd.depth = dalloc<int>(2 * (size_t)n + 2);
        kd.cd = dalloc<int>(2 * (size_t)n + 2); kd.cv = dalloc<double>(2 * (size_t)n + 2);
        kd.is_hi = dalloc<int>(2 * (size_t)n + 2); kd.leaf_of = dalloc<int>(n);
        kd_work_sz = kd_work_bytes(n);
        HIP_OK(hipMalloc(&kd_work, kd_work_sz));
        size_t scan_sz = std::max(scan_tmp_bytes(4LL * n + (1 << 20) + 8), scan_tmp_bytes(n + 8));
        HIP_OK(hipMalloc(&scan_tmp, scan_sz));
        HIP_OK(hipMalloc(&scan_tmp_q, scan_tmp_bytes(n + 8)));
        alloc_lists(dim == 3 ? 40 : (dim == 2 ? 16 : 4));

        // constants
        std::vector<int> ty(d->type, d->type + n), so(d->fixed, d->fixed + n);
        h2d(st, type, ty.data(), n);
        h2d(st, solid, so.data(), n);
        type_h = ty;
        solid_h = so;
        mass_h.assign(d->mass, d->mass + n);
        perm = dalloc<int>(n);
        perm_tmp = dalloc<int>(n);
        perm_h.resize(n);
        for (int i = 0; i < n; i++) perm_h[i] = i;
        h2d(st, perm, perm_h.data(), n);
        {
            // the renumber's staging buffer at its largest row size, so no
            // renumber (one beside a running event loop included) reallocates
            const size_t row = std::max<size_t>({8, (size_t)8 * std::max(d->n_chem_species, 0),
                                                 (size_t)8 * std::max(d->n_data_fn, 0),
                                                 (size_t)4 * std::max(d->n_species, 0)});
            grow(pbuf, pbuf_cap, (size_t)n * row);
        }
        if (ps && d->n_species > 0 && !d->static_domain) {
            rtype = dalloc<int>(n);
            rmap = dalloc<int>(n);
            xbuf = dalloc<uint32_t>((size_t)n * d->n_species);
            Ddiag_alt = dalloc<double>((size_t)n * d->n_species);
            vol_alt = dalloc<double>(n);
            HIP_OK(hipEventCreateWithFlags(&ev_rt, hipEventDisableTiming));
            HIP_OK(hipEventCreateWithFlags(&ev_map, hipEventDisableTiming));
        }
        any_solid = std::any_of(so.begin(), so.end(), [](int v) { return v != 0; });
        movers.clear();
        for (int i = 0; i < n; i++)
            if (!so[i]) movers.push_back(i);
        h2d(st, mass, d->mass, n);
        x0 = to_soa3(d->x, n);
        v0 = to_soa3(d->v, n);
        rho0v.assign(d->rho, d->rho + n);
        nu0.assign(d->nu, d->nu + n);
        if (S > 0) u0.assign(d->u0, d->u0 + (size_t)n * S);
        C0.assign((size_t)n * Sc, 0.0);
        for (size_t k = 0; k < (size_t)n * Sc; k++)
            C0[k] = d->C0 ? d->C0[k] : (S == Sc && d->u0 ? (double)d->u0[k] : 0.0);
        int Smax = std::max(S, Sc), Rmax = std::max(R, Rc);
        if (Smax > 0 && d->diffusion_matrix) {
            Dmat = dalloc<double>((size_t)Smax * T);
            h2d(st, Dmat, d->diffusion_matrix, (size_t)Smax * T);
        }
        if (Rmax > 0) {
            std::vector<Rxn> r(Rmax);
            for (int i = 0; i < Rmax; i++) {
                const ssasdpd_reaction &q = d->reactions[i];
                r[i] = Rxn{q.kind, q.a, q.b, 0, q.k, q.type_mask};
            }
            rx = dalloc<Rxn>(Rmax);
            h2d(st, rx, r.data(), Rmax);
        }
        if (d->n_code > 0) {
            n_code = d->n_code;
            n_consts = d->n_consts;
            vm_code = dalloc<uint32_t>(n_code);
            h2d(st, vm_code, d->code, n_code);
            vm_k = dalloc<double>(std::max(n_consts, 1));
            if (n_consts > 0) h2d(st, vm_k, d->consts, n_consts);
        }
        if (d->n_data_fn > 0) {
            n_df = d->n_data_fn;
            dfn = dalloc<double>((size_t)n * n_df);
            h2d(st, dfn, d->data_fn, (size_t)n * n_df);
            dfn_h.assign(d->data_fn, d->data_fn + (size_t)n * n_df);
        }
        if (R > 0) {
            int nnz = d->jcN[R];
            jcN = dalloc<int>(R + 1); h2d(st, jcN, d->jcN, R + 1);
            irN = dalloc<int>(nnz); h2d(st, irN, d->irN, nnz);
            prN = dalloc<int>(nnz); h2d(st, prN, d->prN, nnz);
        }
        if (S > 0) {
            int nnz = d->jcG[S + R];
            jcG = dalloc<int>(S + R + 1); h2d(st, jcG, d->jcG, S + R + 1);
            irG = dalloc<int>(nnz); h2d(st, irG, d->irG, nnz);
            xx = dalloc<uint32_t>((size_t)n * S);
            rrate = dalloc<double>((size_t)n * std::max(R, 1));
            srrate = dalloc<double>(n); sdrate = dalloc<double>(n);
            Ddiag = dalloc<double>((size_t)n * S);
            qtime = dalloc<double>(n);
            // event tree: level 1 groups F0 = 64*pl voxels (pl chosen so the
            // levels above the voxels fit the event loop's LDS), 64 per node above
            nnzN_ = R > 0 ? d->jcN[R] : 0;
            nnzG_ = d->jcG[S + R];
            int p0 = pl_override ? pl_override : auto_pl(n);
            if (const char *e = getenv("SSASDPD_EVENT_TREE_PL")) {
                int want = atoi(e);
                if (!pl_override && pl_valid(want, n)) p0 = want;
            }
            layout_tree(p0);
            vol = dalloc<double>(n);
            qs = dalloc<QueueState>(1);
            qwork = (int *)dalloc<char>(sizeof(double) * (size_t)n + sizeof(int) * 4 * ((size_t)n + 1) + 64 +
                                        sizeof(double) * 512);
        }
        if (Sc > 0 && Rc > 0 && d->N_dense) {
            Ndense = dalloc<int>((size_t)Sc * Rc);
            h2d(st, Ndense, d->N_dense, (size_t)Sc * Rc);
        }
        if (nbc > 0) {
            std::vector<BcRule> b(nbc);
            for (int i = 0; i < nbc; i++) {
                const ssasdpd_bc &q = d->bcs[i];
                for (int k = 0; k < 3; k++) { b[i].lo[k] = q.lo[k]; b[i].hi[k] = q.hi[k]; b[i].value[k] = q.value[k]; }
                b[i].type = q.type; b[i].target = q.target; b[i].species = q.species; b[i].pad = 0;
            }
            bc = dalloc<BcRule>(nbc);
            h2d(st, bc, b.data(), nbc);
        }
        HIP_OK(hipStreamSynchronize(st));
    }

    void reset(const ssasdpd_rng *r) {
        if (ps) HIP_OK(hipStreamSynchronize(ps));
        la_pending = false;
        size_t n3 = 3 * (size_t)n;
        if (renumbered) {
            // the constant per-particle arrays back in the caller's numbering
            h2d(st, mass, mass_h.data(), n);
            h2d(st, type, type_h.data(), n);
            h2d(st, solid, solid_h.data(), n);
            if (n_df > 0) h2d(st, dfn, dfn_h.data(), (size_t)n * n_df);
            for (int i = 0; i < n; i++) perm_h[i] = i;
            h2d(st, perm, perm_h.data(), n);
            renumbered = false;
            perm_h_stale = false;
            rebuild_movers();
        }
        h2d(st, x, x0.data(), n3);
        h2d(st, v, v0.data(), n3);
        HIP_OK(hipMemsetAsync(vt, 0, sizeof(double) * n3, st));
        HIP_OK(hipMemsetAsync(F, 0, sizeof(double) * n3, st));
        HIP_OK(hipMemsetAsync(Fbp, 0, sizeof(double) * n3, st));
        HIP_OK(hipMemsetAsync(normal, 0, sizeof(double) * n3, st));
        h2d(st, rho, rho0v.data(), n);
        h2d(st, old_rho, rho0v.data(), n);
        h2d(st, nu, nu0.data(), n);
        HIP_OK(hipMemsetAsync(Frho, 0, sizeof(double) * n, st));
        HIP_OK(hipMemsetAsync(bvf, 0, sizeof(double) * n, st));
        if (Sc > 0) {
            h2d(st, C, C0.data(), (size_t)n * Sc);
            HIP_OK(hipMemsetAsync(Q, 0, sizeof(double) * n * Sc, st));
        }
        HIP_OK(hipMemsetAsync(deg, 0, sizeof(int) * n, st));
        if (S > 0) {
            h2d(st, xx, u0.data(), (size_t)n * S);
            HIP_OK(hipMemsetAsync(srrate, 0, sizeof(double) * n, st));
            HIP_OK(hipMemsetAsync(sdrate, 0, sizeof(double) * n, st));
            QueueState q{};
            h2d(st, qs, &q, 1);
        }
        DevErr e0{0, INT32_MAX, 0, 0, 0};
        h2d(st, derr, &e0, 1);
        if (derr_alt) h2d(st, derr_alt, &e0, 1);
        rng_kind = (r && r->kind == 1) ? 1 : 0;
        seed = r ? r->seed : 0;
        if (rng_kind == 1) {
            if (r->n_words > words_cap) {
                dfree(d_words); dfree(d_expo);
                words_cap = r->n_words;
                d_words = dalloc<uint64_t>(words_cap);
                d_expo = dalloc<double>(words_cap);
            }
            n_words = r->n_words;
            h2d(st, d_words, r->words, n_words);
            h2d(st, d_expo, r->expo, n_words);
        } else {
            n_words = 0;
        }
        HIP_OK(hipStreamSynchronize(st));
        cur = 0;
        lists_built = false;
        kd_built = false;
        rdme_initialized = false;
        max_deg_seen = 0;
        memset(&stats, 0, sizeof(stats));
        stats.bad_particle = -1;
        nn_active = nn != nullptr;
        nn_steps = nn_checks = 0;
        nn_last_err = -1.0;
        nn_max_err = 0.0;
        nn_refreshes = 0;
        nn_fails = 0;
        nn_acc_cap = 0.0;
        nn_clear_capped();
        for (int p = 0; p < NPH; p++) { ph_ms[p] = 0; ph_launch[p] = 0; pending[p][0] = pending[p][1] = false; }
    }

    DevErr read_err(const DevErr *from = nullptr) {
        d2h(st, h_err, from ? from : derr, 1);
        HIP_OK(hipStreamSynchronize(st));
        return *h_err;
    }

    // Bin the tree points (= current x), and build the kd-tree when list
    // order needs it.
    void rebuild_movers() {
        const std::vector<int> &ph = perm_host();
        movers.clear();
        for (int p = 0; p < n; p++)
            if (!solid_h[ph[p]]) movers.push_back(p);
        movers_stale = false;
    }

    bool rdme_on() const { return S > 0 && !(flags & SSASDPD_FLAG_NO_RDME); }

    bool renumber_allowed() const {
        return static_domain == 0 && rng_kind == 0 && trace_cap == 0 &&
               !(flags & (SSASDPD_FLAG_ANN_ORDER_MOVING | SSASDPD_FLAG_NO_RENUMBER));
    }
    bool renumber_due(unsigned step) const { return step > 0 && step % RENUMBER_EVERY == 0 && renumber_allowed(); }

    static void *grow(void *&buf, size_t &cap, size_t bytes) {
        if (bytes > cap) {
            if (buf) HIP_OK(hipFree(buf));
            HIP_OK(hipMalloc(&buf, bytes));
            cap = bytes;
        }
        return buf;
    }
    // an [n][cols] array of esz-byte elements gathered into the new order
    void permute_inplace(void *arr, int cols, int esz, const int *map) {
        if (!arr || cols <= 0) return;
        const size_t bytes = (size_t)n * cols * esz;
        void *tmp = grow(pbuf, pbuf_cap, bytes);
        permute_rows(st, n, cols, esz, map, arr, tmp, false);
        HIP_OK(hipMemcpyAsync(arr, tmp, bytes, hipMemcpyDeviceToDevice, st));
    }
    // The new order: cell order of the current positions, cells taken tile
    // by tile (4 x 4 x 4 in 3D), ties inside a cell by index, so the
    // numbering (which the random decisions of the chemistry see) is a
    // function of the positions.  m[p]: the particle
    // that becomes p; valid until the next binning.
    const int *renumber_map() {
        bin_tree_points(true);
        cell_order_by_index(st, cell_start, grid.ncell, sorted);
        return sorted;
    }
    // Put the particles in cell order of their current positions: every
    // array that carries state from one step to the next, the counts of
    // n_traj batch trajectories, and the map to the caller's numbering.
    // Lists, rates and the queue are rebuilt by the next step.
    void renumber(int n_traj = 0) {
        const int *m = renumber_map();
        permute_particles(m);
        if (S > 0) permute_inplace(xx, S, 4, m);
        for (int b = 0; b < n_traj; b++) permute_inplace(b_xx + (size_t)b * n * S, S, 4, m);
    }
    // every per-particle array but the counts, and the map to the caller's
    // numbering
    void permute_particles(const int *m) {
        for (double *a : {x, v, vt, F, Fbp, normal})
            for (int k = 0; k < 3; k++) permute_inplace(a + (size_t)k * n, 1, 8, m);
        for (double *a : {rho, old_rho, Frho, mass, nu, bvf}) permute_inplace(a, 1, 8, m);
        permute_inplace(type, 1, 4, m);
        permute_inplace(solid, 1, 4, m);
        if (Sc > 0) {
            permute_inplace(C, Sc, 8, m);
            permute_inplace(Q, Sc, 8, m);
        }
        if (n_df > 0) permute_inplace(dfn, n_df, 8, m);
        permute_rows(st, n, 1, 4, m, perm, perm_tmp, false);
        std::swap(perm, perm_tmp);
        renumbered = true;
        rec_ready = false;
        perm_h_stale = movers_stale = true;
    }
    // an [n][cols] device array in the caller's numbering (the array itself
    // while no renumber has happened); the result lives until the next call
    template <class T> const T *orig(const T *src, int cols) {
        static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte elements");
        if (!renumbered || !src || cols <= 0) return src;
        void *out = grow(stage, stage_cap, (size_t)n * cols * sizeof(T));
        permute_rows(st, n, cols, (int)sizeof(T), perm, src, out, true);
        return (const T *)out;
    }
    // a component-major [3][n] array in the caller's numbering
    const double *orig3(const double *src) {
        if (!renumbered || !src) return src;
        double *out = (double *)grow(stage, stage_cap, 3 * (size_t)n * sizeof(double));
        for (int k = 0; k < 3; k++) permute_rows(st, n, 1, 8, perm, src + (size_t)k * n, out + (size_t)k * n, true);
        return out;
    }
    int orig_index(int p) { return (renumbered && p >= 0 && p < n) ? perm_host()[p] : p; }

    // tiled: cells numbered tile by tile (a binning for the renumbering
    // only; the search needs the plain numbering)
    void bin_tree_points(bool tiled = false) {
        bin_begin();
        bin_finish(tiled);
    }
    // The tree points (= current x) and their bounding box on its way to
    // the host; work enqueued before bin_finish runs while the host waits.
    void bin_begin() {
        HIP_OK(hipMemcpyAsync(xtree, x, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, st));
        compute_bbox(st, xtree, n, dim, d_bbox, d_bbtmp);
        d2h(st, h_bbox, d_bbox, 6);
        HIP_OK(hipEventRecord(ev_bbox, st));
    }
    // The cell grid from the bounding box, then the binning (phase 0 times
    // this part).
    void bin_finish(bool tiled = false) {
        HIP_OK(hipEventSynchronize(ev_bbox));
        ph_begin(0);
        grid_setup(h_bbox, n, dim, h, grid);
        if (tiled) {
            Grid tg = grid;
            grid_tile(tg, dim);
            // within what the scan's work space was sized for
            if (tg.ncell <= 4LL * n + (1 << 20)) grid = tg;
        }
        if (grid.ncell + 1 > cell_cap) {
            // hipFree waits for the whole device (a running event loop too), so
            // grow with a quarter of headroom: a slowly widening bounding box
            // then reallocates rarely instead of every few steps
            dfree(cell_count); dfree(cell_start); dfree(cell_fill);
            cell_cap = grid.ncell + 1 + grid.ncell / 4 + 64;
            cell_count = dalloc<int>(cell_cap);
            cell_start = dalloc<int>(cell_cap + 1);
            cell_fill = dalloc<int>(cell_cap);
        }
        bin_points(st, xtree, n, dim, grid, cell_of, cell_count, cell_start, cell_fill, sorted, spts, scan_tmp,
                   &derr->crowded);
        bool want_kd = static_domain || (flags & SSASDPD_FLAG_ANN_ORDER_MOVING);
        if (want_kd && !(static_domain && kd_built)) {
            kd_build(st, xtree, n, dim, kd, kd_work, kd_work_sz);
            kd_built = true;
        }
        if (!want_kd) kd.valid = false;
        ph_end(0);
    }

    // Fixed-radius search from query = current x against the binned tree
    // points, then r and D_ij with the current rho (find_neighbors).
    // forces_step >= 0: the step's plain pair forces are launched behind
    // each search, before the host reads its overflow flag (the pair kernel
    // starts its sums from the predictor's values, so the launch after a
    // re-search gives the result of a single one)
    void search(int forces_step = -1) {
        for (int attempt = 0; attempt < 4; attempt++) {
            ph_begin(1);
            HIP_OK(hipMemsetAsync(&derr->overflow, 0, sizeof(int), st));
            HIP_OK(hipMemsetAsync(&derr->max_deg, 0, sizeof(int), st));
            // queries in particle order: each wave's list writes are
            // coalesced (cell order shares candidate cells but scatters the
            // writes, and measured slower); each sorted list is finished (r,
            // D_ij with the current rho, rows) in the same kernel
            // (with the RDME on and few species the finish also forms Ddiag;
            // then nothing reads the per-edge D_ij and neighbour types but
            // the event loop, which has them in its rows)
            const bool fin_diag = rdme_on() && S <= FIN_MAX_SPECIES && S * T <= FIN_MAX_DMAT;
            const FinishArgs fa{nr, fin_diag ? nullptr : Dij, fin_diag ? nullptr : ntype, erow, mass, rho, type,
                                rec_ready ? rec : nullptr, fin_diag ? Ddiag : nullptr, Dmat, S, T};
            // LDS list slots sized by the previous search's longest list
            fr_search(st, x, n, spts, sorted, cell_start, n, dim, h, grid, kd.valid ? &kd : nullptr, W, nidx, nr2,
                      deg, derr, nullptr, &fa, last_max_deg);
            ph_end(1);
            // the host waits for the search's overflow flag only, not for
            // the pair kernel behind it
            d2h(st, h_err, derr, 1);
            HIP_OK(hipEventRecord(ev_srch, st));
            if (forces_step >= 0) {
                ph_begin(5);
                sdpd_forces(st, sdpd_args((unsigned)forces_step));
                ph_end(5);
            }
            HIP_OK(hipEventSynchronize(ev_srch));
            const DevErr e = *h_err;
            last_max_deg = e.max_deg;
            max_deg_seen = std::max(max_deg_seen, e.max_deg);
            if (!e.overflow) break;
            HIP_OK(hipStreamSynchronize(st));   // the pair kernel reads the lists freed here
            alloc_lists(((e.max_deg + 7) / 8) * 8 + 8);
        }
        lists_built = true;
    }

    SdpdArgs sdpd_args(unsigned step) {
        SdpdArgs a{};
        a.n = n; a.dim = dim; a.Sc = Sc; a.Rc = Rc; a.T = T; a.static_domain = static_domain; a.nbc = nbc; a.W = W;
        a.step = step; a.dt = dt; a.h = h; a.rho0 = rho0; a.c0 = c0; a.P0 = P0; a.alpha = alpha;
        for (int k = 0; k < 3; k++) a.g[k] = g[k];
        a.x = x; a.v = v; a.vt = vt; a.F = F; a.Fbp = Fbp; a.rho = rho; a.old_rho = old_rho; a.Frho = Frho;
        a.bvf = bvf; a.normal = normal; a.mass = mass; a.nu = nu; a.type = type; a.solid = solid; a.C = C; a.Q = Q;
        a.vol = rdme_on() ? vol : nullptr;
        a.Dmat = Dmat; a.Ndense = Ndense; a.rx = rx; a.bc = bc; a.deg = deg; a.nidx = nidx; a.nr = nr; a.err = derr;
        a.rec = rec;
        a.no_solids = any_solid ? 0 : 1;
        a.fp32 = (flags & SSASDPD_FLAG_FP32_FORCES) ? 1 : 0;
        a.vm_code = vm_code; a.vm_k = vm_k; a.data_fn = dfn; a.n_df = n_df;
        return a;
    }

    RdmeArgs rdme_args(unsigned step) {
        RdmeArgs a{};
        a.n = n; a.S = S; a.R = R; a.T = T; a.W = W;
        a.mass = mass; a.rho = rho; a.type = type; a.xx = xx; a.rrate = rrate; a.srrate = srrate; a.sdrate = sdrate;
        a.Ddiag = Ddiag; a.Dmat = Dmat; a.rx = rx; a.jcN = jcN; a.irN = irN; a.prN = prN; a.jcG = jcG; a.irG = irG;
        a.deg = deg; a.nidx = nidx; a.ntype = ntype; a.Dij = Dij; a.vol = vol;
        a.erow = erow;
        a.row_w = std::max(1, std::min(W, max_deg_seen));
        a.qtime = qtime; a.tmin = tmin; a.targ = targ; a.nlev = nlev; a.pl = pl; a.tree_nodes = tree_nodes;
        a.lds_bytes = lds_bytes;
        for (int l = 0; l < 10; l++) { a.lev_off[l] = lev_off[l]; a.lev_n[l] = lev_n[l]; }
        a.qs = qs;
        a.rng_kind = rng_kind; a.seed = seed; a.stream_id = 0; a.words = d_words; a.expo = d_expo; a.n_words = n_words;
        a.textbook = (flags & SSASDPD_FLAG_TEXTBOOK_SELECTION) ? 1 : 0;
        a.trace = trace_cap > 0 ? d_trace : nullptr;
        a.trace_cap = trace_cap;
        a.max_events = 1ULL << 26;   // per step launch; beyond it the run stops with CAPACITY
        a.t0 = step * dt;
        a.dt = dt;
        a.endTime = static_domain ? nt * dt : a.t0 + dt;
        a.step = step;
        a.psum_out = d_psum;
        a.vm_code = vm_code; a.vm_k = vm_k; a.vm_ncode = n_code; a.vm_nk = n_consts;
        a.data_fn = dfn; a.n_df = n_df;
        return a;
    }

    // The step's particle-phase errors (from its error record: `eb`, else
    // the current one), then its event loop's.
    int check(unsigned step, const DevErr *eb = nullptr) {
        const bool q_on = S > 0 && !(flags & SSASDPD_FLAG_NO_RDME);
        d2h(st, h_err, eb ? eb : derr, 1);
        if (q_on) d2h(st, h_qc, qs, 1);
        HIP_OK(hipStreamSynchronize(st));
        const DevErr e = *h_err;
        if (e.code == SSASDPD_ERR_CAPACITY && e.crowded > 0) {
            stats.bad_particle = -1;
            char buf[320];
            snprintf(buf, sizeof buf,
                     "step %u: %d particles share one cell of the neighbour-search grid (at most %d): the bounding "
                     "box [%g, %g] x [%g, %g] x [%g, %g] is too wide for %d particles at h = %g (a particle far from "
                     "the others?)",
                     step, e.crowded, FR_MAX_CELL, h_bbox[0], h_bbox[1], h_bbox[2], h_bbox[3], h_bbox[4], h_bbox[5],
                     n, h);
            return set_error(e.code, buf);
        }
        if (e.code) {
            stats.bad_particle = e.particle == INT32_MAX ? -1 : orig_index(e.particle);
            char buf[160];
            snprintf(buf, sizeof buf, "step %u: particle %d", step, stats.bad_particle);
            return set_error(e.code, e.code == SSASDPD_ERR_NAN_STATE ? std::string("nan/inf state at ") + buf
                                                                        : std::string("NaN D_ij at ") + buf);
        }
        if (q_on) {
            const QueueState q = *h_qc;
            stats.reactions = q.reactions;
            stats.diffusions = q.diffusions;
            stats.rng_words_used = q.rng_pos;
            if (q.err) {
                char buf[96];
                snprintf(buf, sizeof buf, "SSA failed at step %u", q.err_step);
                return set_error(q.err, buf);
            }
        }
        return 0;
    }

    // Neighbour search and the three SDPD substeps of one step; the particle
    // state does not depend on the molecule counts, so a batch of
    // trajectories shares it.
    void particle_phase(unsigned step) {
        const bool moving = static_domain == 0;
        rec_ready = false;
        // after step 0 on a moving domain the tree points are copied before
        // the predictor moves the particles, and the predictor runs while
        // the host waits for their bounding box
        const bool overlap = moving && step > 0;
        if (overlap) {
            bin_begin();
        } else if (step == 0 || moving) {
            bin_tree_points();
            // substep 0 lists (simulate.cpp:61-63).  On a moving domain the
            // reference rebuilds them at every step, but after step 0 nothing
            // reads them before substep 1 rebuilds them again
            // (simulate.cpp:121-123), so they are built at step 0 only.
            if (step == 0) search();
        }
        SdpdArgs sa = sdpd_args(step);
        ph_begin(2);
        sdpd_predict(st, sa);                      // substep 0 (+ the gather records)
        ph_end(2);
        if (overlap) bin_finish();
        rec_ready = true;
        // attached: checks run while the network is on, or always when a
        // failed check only pauses it until the next passing one (re-arm)
        const bool nn_on = nn && step > 0 && static_domain == 0 && (nn_active || nn_rearm);
        const bool nn_step = nn_on && nn_active && step % nn_check_every != 0;
        bool forces_done = false;
        if (step > 0 && moving) {
            // substep 1 lists (simulate.cpp:121-123); a plain step's pair
            // forces go behind the search
            search(nn_step ? -1 : (int)step);
            sa = sdpd_args(step);
            forces_done = !nn_step;
        }
        if (nn_on) nn_ensure();
        if (nn_step) {
            // surrogate step: the pair kernel supplies Fbp, Frho and Q only,
            // the network the accelerations
            ph_begin(5);
            sa.skip_F = 1;
            sdpd_forces(st, sa);
            sa.skip_F = 0;
            nn_predict(F, true);
            ph_end(5);
            nn_steps++;
        } else {
            if (!forces_done) {
                ph_begin(5);
                sdpd_forces(st, sa);               // substep 1
                ph_end(5);
            }
            if (nn_rec && step > 0 && static_domain == 0 && step % nn_rec_every == 0)
                record_sample(nn_rec, nn_rec_max, nn_rec_seed++);
            if (nn_on && step % nn_check_every == 0) {
                // periodic check against the plain SDPD: above the budget the
                // network is switched off (for the rest of the run, or until
                // a check passes again with re-arm)
                nn_predict(nn_F);
                HIP_OK(hipMemsetAsync(nn_acc, 0, sizeof(double) * 3, st));
                ecn_err(st, n, F, nn_F, solid, g, nn_acc);
                double acc[3];
                d2h(st, acc, nn_acc, 3);
                HIP_OK(hipStreamSynchronize(st));
                nn_last_err = acc[1] > 0 ? std::sqrt(acc[0] / acc[1]) : 0.0;
                nn_acc_cap = (nn_cap_factor > 0.0 && acc[2] > 0.0) ? nn_cap_factor * std::sqrt(acc[1] / acc[2]) : 0.0;
                nn_checks++;
                nn_active = nn_last_err <= nn_tol;
                if (nn_active)
                    nn_max_err = std::max(nn_max_err, nn_last_err);
                else
                    nn_fails++;
                if (nn_refresh_iters > 0) {
                    // online refresh: this plain step becomes a sample (the
                    // oldest beyond the window are dropped) and the network
                    // takes a few Adam steps
                    record_sample(nn, nn_rec_max, nn_rec_seed++);
                    while ((int)nn->samples.size() > nn_refresh_keep) nn->drop_oldest();
                    nn->train_steps(nn_refresh_iters, nn_refresh_lr, nn_rec_seed++);
                    nn_refreshes++;
                }
            }
        }
        ph_begin(2);
        sdpd_correct(st, sa);                      // substep 2
        sdpd_bvf(st, sa);
        ph_end(2);
        if (rdme_on() && (S > FIN_MAX_SPECIES || S * T > FIN_MAX_DMAT) && (step == 0 || moving)) {
            // Ddiag of a model with more species than the search's finish forms
            ph_begin(3);
            rdme_diag(st, rdme_args(step));
            ph_end(3);
        }
    }

    void end_step(bool skip_ahead = false) {
        ph_collect(skip_ahead);
        stats.ms_neighbors = ph_ms[0] + ph_ms[1];
        stats.ms_sdpd = ph_ms[2] + ph_ms[5];
        stats.ms_rates = ph_ms[3];
        stats.ms_ssa = ph_ms[4];
        cur++;
        stats.steps = cur;
        stats.particle_steps = (uint64_t)cur * n;
        stats.max_degree = max_deg_seen;
    }

    // Static domains: after step 0 the event loops of all remaining steps
    // run in one launch with in-kernel snapshots (see particle_steps).
    bool fusable_static() const {
        return static_domain && S > 0 && !(flags & SSASDPD_FLAG_NO_RDME) && cur > 0;
    }

    // Static domains after step 0: the particle phases of steps [s0, s1)
    // alone.  The event loop neither reads nor writes anything they change
    // (rates and vol are built once, rho does not evolve), so they run ahead
    // of the fused event loop; without deterministic species they change
    // nothing at all and are skipped.  snap(j) takes the field snapshot
    // scheduled before step sst[j] (sst[j] == s1: after the last step).
    template <class Snap>
    int particle_steps(unsigned s0, unsigned s1, const std::vector<unsigned> &sst, Snap snap) {
        size_t j = 0;
        for (unsigned s = s0; s < s1; s++) {
            while (j < sst.size() && sst[j] <= s) snap(j++);
            if (Sc > 0) particle_phase(s);
        }
        while (j < sst.size()) snap(j++);
        DevErr e = read_err();
        if (e.code) {
            stats.bad_particle = e.particle == INT32_MAX ? -1 : e.particle;
            return set_error(e.code, "non-finite particle state in a static-domain step");
        }
        return 0;
    }
    bool fusable() const { return fusable_static() && rdme_initialized; }

    // Steps [cur, s_end) of the event loop in one launch for n_traj
    // trajectories (ra: single-trajectory or batch arguments).  snap: the
    // step before which each snapshot is taken (s_end: after the last step);
    // d_xx / d_tot: device snapshot buffers or null.
    void ssa_fused(RdmeArgs ra, int n_traj, unsigned s_end, const std::vector<unsigned> &snap, uint32_t *d_xx,
                   size_t bs_snap, uint32_t *d_tot) {
        unsigned *d_steps = nullptr;
        if (!snap.empty()) {
            d_steps = dalloc<unsigned>(snap.size());
            h2d(st, d_steps, snap.data(), snap.size());
        }
        ra.step = cur;
        ra.step_end = s_end;
        ra.t0 = cur * dt;
        ra.snap_step = d_steps;
        ra.n_snap = (int)snap.size();
        ra.snap_xx = d_xx;
        ra.bs_snap = bs_snap;
        ra.snap_tot = d_tot;
        ph_begin(4);
        rdme_take_step(st, ra, n_traj);
        ph_end(4);
        HIP_OK(hipStreamSynchronize(st));
        dfree(d_steps);
        ph_collect();
        stats.ms_ssa = ph_ms[4];
        cur = s_end;
        stats.steps = cur;
        stats.particle_steps = (uint64_t)cur * n;
    }

    // Step 0 of a static domain up to the event loop: neighbour lists, the
    // step-0 particle phase, rates and the queue.  Returns whether the queue
    // holds an event (the event loop then runs from step 0).
    bool prepare_static() {
        particle_phase(0);
        RdmeArgs ra = rdme_args(0);
        rdme_rates(st, ra);
        rdme_queue_build(st, ra, scan_tmp_q, qwork);
        QueueState q;
        d2h(st, &q, qs, 1);
        HIP_OK(hipStreamSynchronize(st));
        rdme_initialized = q.initialized == 1;
        return rdme_initialized;
    }

    // One step.  lookahead: the caller runs the next step right after, so
    // its particle phase may start now, beside this step's event loop.
    int do_step(bool lookahead = false) {
        const unsigned step = cur;
        // (a renumber step prepared by a look-ahead was renumbered there)
        if (!la_pending && renumber_due(step)) renumber();
        if (la_pending) {
            HIP_OK(hipStreamWaitEvent(st, la_done, 0));
            la_pending = false;
        } else {
            particle_phase(step);
        }
        bool la = false;
        if (S > 0 && !(flags & SSASDPD_FLAG_NO_RDME)) {
            RdmeArgs ra = rdme_args(step);
            const bool build = static_domain == 0 || !rdme_initialized;
            // A look-ahead runs the next step's particle phase beside this
            // step's rates, queue build and event loop: they read vol, Ddiag,
            // deg and the rows of this step (the look-ahead writes the other
            // buffers), the types from a copy when it renumbers, and not the
            // particle state; its errors go to the other error record.  A
            // renumber step is prepared too, unless the event loop reads
            // data-function values (those are permuted with the particles).
            const bool want = build && lookahead && ps && Ddiag_alt && derr_alt && static_domain == 0 && !nn &&
                              step + 1 < nt && (!renumber_due(step + 1) || (rtype && n_df == 0));
            if (want) {
                if (renumber_due(step + 1)) {
                    HIP_OK(hipMemcpyAsync(rtype, type, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
                    ra.type = rtype;
                }
                HIP_OK(hipEventRecord(ev_rt, st));
            }
            bool qpend = false;
            if (build) {
                ph_begin(3);
                rdme_rates(st, ra);
                rdme_queue_build(st, ra, scan_tmp_q, qwork);
                ph_end(3);
                if (static_domain) {
                    // whether the queue holds an event (a static domain then
                    // runs its event loops from here on); the event loop,
                    // which checks the queue state itself, is launched first
                    d2h(st, h_q, qs, 1);
                    HIP_OK(hipEventRecord(ev_q, st));
                    qpend = true;
                }
            }
            ph_begin(4);
            rdme_take_step(st, ra);
            ph_end(4);
            if (qpend) {
                HIP_OK(hipEventSynchronize(ev_q));
                if (h_q->initialized == 1) rdme_initialized = true;
            }
            la = want;
        }
        if (la) {
            const bool ren = renumber_due(step + 1);
            std::swap(deg, deg_alt);
            std::swap(erow, erow_alt);
            std::swap(Ddiag, Ddiag_alt);
            std::swap(vol, vol_alt);
            std::swap(derr, derr_alt);
            std::swap(st, ps);
            in_lookahead = true;
            HIP_OK(hipStreamWaitEvent(st, ev_rt, 0));
            if (ren) {
                const int *m = renumber_map();
                HIP_OK(hipMemcpyAsync(rmap, m, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
                HIP_OK(hipEventRecord(ev_map, st));
                permute_particles(m);
            }
            particle_phase(step + 1);
            HIP_OK(hipEventRecord(la_done, st));
            in_lookahead = false;
            std::swap(st, ps);
            if (ren) {
                // the counts, once this step's event loop is done with them
                HIP_OK(hipStreamWaitEvent(st, ev_map, 0));
                permute_rows(st, n, S, 4, rmap, xx, xbuf, false);
                HIP_OK(hipMemcpyAsync(xx, xbuf, sizeof(uint32_t) * n * S, hipMemcpyDeviceToDevice, st));
            }
            la_pending = true;
        }
        // this step's particle phase wrote the error record now in derr_alt
        // when a look-ahead took derr
        int rc = check(step, la ? derr_alt : nullptr);
        if (rc && la) {
            // the next step's particle phase is still running on ps: let it
            // finish so that state reads on st do not race it
            HIP_OK(hipStreamSynchronize(ps));
        }
        end_step(la);
        return rc;
    }

    // ---- batched ensembles: B trajectories of the same model, seeded
    // streams, one event-loop workgroup per trajectory.  Per trajectory:
    // counts [n][S], rates [n][R], srrate / sdrate / firing times [n], the
    // event tree and the queue scalars.
    int bcap = 0;
    uint32_t *b_xx = nullptr;
    double *b_rr = nullptr, *b_sr = nullptr, *b_sd = nullptr, *b_qt = nullptr, *b_tmin = nullptr;
    int *b_targ = nullptr;
    QueueState *b_qs = nullptr;
    unsigned long long *b_seeds = nullptr;
    std::vector<uint64_t> b_seed_h;
    std::vector<char> b_init;

    void batch_alloc(int B) {
        if (B <= bcap) return;
        dfree(b_xx); dfree(b_rr); dfree(b_sr); dfree(b_sd); dfree(b_qt); dfree(b_tmin); dfree(b_targ);
        dfree(b_qs); dfree(b_seeds);
        const size_t nn = n;
        b_xx = dalloc<uint32_t>(nn * S * B);
        b_rr = dalloc<double>(nn * std::max(R, 1) * B);
        b_sr = dalloc<double>(nn * B);
        b_sd = dalloc<double>(nn * B);
        b_qt = dalloc<double>(nn * B);
        b_tmin = dalloc<double>((size_t)std::max(tree_nodes, 1) * B);
        b_targ = dalloc<int>((size_t)std::max(tree_nodes, 1) * B);
        b_qs = dalloc<QueueState>(B);
        b_seeds = dalloc<unsigned long long>(B);
        bcap = B;
    }

    // The arguments of trajectory b alone (queue build), or of the whole
    // batch (b < 0: base pointers plus strides for the event loop).
    RdmeArgs batch_args(unsigned step, int b) {
        RdmeArgs a = rdme_args(step);
        const size_t nn = n;
        a.bs_xx = nn * S;
        a.bs_rr = nn * std::max(R, 1);
        a.bs_n = nn;
        a.bs_tree = (size_t)std::max(tree_nodes, 1);
        a.rng_kind = 0;
        a.words = nullptr; a.expo = nullptr; a.n_words = 0;
        a.trace = nullptr; a.trace_cap = 0;
        const size_t o = b < 0 ? 0 : (size_t)b;
        a.xx = b_xx + o * a.bs_xx;
        a.rrate = b_rr + o * a.bs_rr;
        a.srrate = b_sr + o * a.bs_n;
        a.sdrate = b_sd + o * a.bs_n;
        a.qtime = b_qt + o * a.bs_n;
        a.tmin = b_tmin + o * a.bs_tree;
        a.targ = b_targ + o * a.bs_tree;
        a.qs = b_qs + o;
        if (b < 0) a.seeds = b_seeds;
        else a.seed = b_seed_h[b];
        return a;
    }

    void batch_reset(int B, const uint64_t *seeds) {
        ssasdpd_rng r{};
        r.kind = 0;
        r.seed = seeds[0];
        reset(&r);
        batch_alloc(B);
        b_seed_h.assign(seeds, seeds + B);
        b_init.assign(B, 0);
        h2d(st, b_seeds, (const unsigned long long *)seeds, B);
        const size_t nn = n;
        for (int b = 0; b < B; b++) h2d(st, b_xx + (size_t)b * nn * S, u0.data(), nn * S);
        HIP_OK(hipMemsetAsync(b_sr, 0, sizeof(double) * nn * B, st));
        HIP_OK(hipMemsetAsync(b_sd, 0, sizeof(double) * nn * B, st));
        std::vector<QueueState> q0(B);
        memset(q0.data(), 0, sizeof(QueueState) * B);
        h2d(st, b_qs, q0.data(), B);
        HIP_OK(hipStreamSynchronize(st));
    }

    int batch_step(int B) {
        const unsigned step = cur;
        if (renumber_due(step)) renumber(B);
        particle_phase(step);
        int rc = check(step);
        if (rc) return rc;
        if (S > 0 && !(flags & SSASDPD_FLAG_NO_RDME)) {
            bool any = false;
            ph_begin(3);
            for (int b = 0; b < B; b++) {
                if (static_domain && b_init[b]) continue;
                RdmeArgs ra = batch_args(step, b);
                rdme_rates(st, ra);
                rdme_queue_build(st, ra, scan_tmp_q, qwork);
                any = true;
            }
            ph_end(3);
            std::vector<QueueState> q(B);
            if (any) {
                d2h(st, q.data(), b_qs, B);
                HIP_OK(hipStreamSynchronize(st));
                for (int b = 0; b < B; b++) if (q[b].initialized == 1) b_init[b] = 1;
            }
            ph_begin(4);
            rdme_take_step(st, batch_args(step, -1), B);
            ph_end(4);
            rc = batch_check(B, step);
        }
        end_step();
        return rc;
    }

    bool batch_all_init() const {
        for (char c : b_init)
            if (!c) return false;
        return !b_init.empty();
    }

    // Sum the trajectories' counters into stats; first error wins.
    int batch_check(int B, unsigned step) {
        std::vector<QueueState> q(B);
        d2h(st, q.data(), b_qs, B);
        HIP_OK(hipStreamSynchronize(st));
        int rc = 0;
        stats.reactions = stats.diffusions = stats.rng_words_used = 0;
        for (int b = 0; b < B; b++) {
            stats.reactions += q[b].reactions;
            stats.diffusions += q[b].diffusions;
            stats.rng_words_used += q[b].rng_pos;
            if (q[b].err && !rc) {
                char buf[128];
                snprintf(buf, sizeof buf, "SSA failed at step %u in trajectory %d", q[b].err ? q[b].err_step : step, b);
                rc = set_error(q[b].err, buf);
            }
        }
        return rc;
    }

    void snapshot(ssasdpd_outputs *o, int k) {
        size_t nn = n;
        std::vector<double> tmp;
        if (o->xx && S > 0) d2h(st, o->xx + (size_t)k * nn * S, orig(xx, S), nn * S);
        if (o->C && Sc > 0) d2h(st, o->C + (size_t)k * nn * Sc, orig(C, Sc), nn * Sc);
        if (o->rho) d2h(st, o->rho + (size_t)k * nn, orig(rho, 1), nn);
        if (o->bvf_phi) d2h(st, o->bvf_phi + (size_t)k * nn, orig(bvf, 1), nn);
        if (o->nu) d2h(st, o->nu + (size_t)k * nn, orig(nu, 1), nn);
        if (o->x) {
            tmp.resize(3 * nn);
            d2h(st, tmp.data(), orig3(x), 3 * nn);
            HIP_OK(hipStreamSynchronize(st));
            from_soa3(tmp, o->x + (size_t)k * nn * 3, n);
        }
        if (o->v) {
            tmp.resize(3 * nn);
            d2h(st, tmp.data(), orig3(v), 3 * nn);
            HIP_OK(hipStreamSynchronize(st));
            from_soa3(tmp, o->v + (size_t)k * nn * 3, n);
        }
        if (o->steps) o->steps[k] = cur;
        HIP_OK(hipStreamSynchronize(st));
    }
};

// ---------------------------------------------------------------------------
// Ensembles of different static models (uneven domains, own chemistry), one
// trajectory each, all event loops in one persistent launch.
struct ssasdpd_ensemble {
    std::vector<ssasdpd_engine *> eng;
    hipStream_t st = nullptr;
    int M = 0, Smax = 0, max_sr = 0, pl = 1;
    uint32_t nt = 0;
    std::vector<unsigned> snap_steps;   // snapshot k is taken before step snap_steps[k]
    std::vector<int> order;             // launch order: largest domain first
    RdmeArgs *d_args = nullptr;
    unsigned *d_snap = nullptr;
    uint32_t *d_tot = nullptr;
    ssasdpd_stats stats{};
    double ms_launch = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    ~ssasdpd_ensemble() {
        for (auto *e : eng) delete e;
        dfree(d_args); dfree(d_snap); dfree(d_tot);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    }
    size_t lds_max() const {
        size_t m = 0;
        for (auto *e : eng) m = std::max(m, e->lds_bytes);
        return m;
    }
    bool pl_ok(int p) const {
        for (auto *e : eng)
            if (!ssasdpd_engine::pl_valid(p, e->n)) return false;
        for (auto *e : eng)
            if (ssa_lds_bytes_for(e, p) > 65536) return false;
        return true;
    }
    static size_t ssa_lds_bytes_for(const ssasdpd_engine *e, int p) {
        int n1 = (e->n + 64 * p - 1) / (64 * p), nodes = n1;
        while (n1 > 1) { n1 = (n1 + TREE_FAN - 1) / TREE_FAN; nodes += n1; }
        return ssa_lds_bytes(nodes, e->S, e->R, e->T, e->nnzN_, e->nnzG_, e->n_code, e->n_consts);
    }
    void set_pl(int p) {
        for (auto *e : eng) e->set_pl(p);
        pl = p;
    }
    // Reset every model with its seed and build step 0 up to the event loop.
    // init[m] = 0 for a model whose queue holds no event.
    void prepare(const uint64_t *seeds, std::vector<char> &init) {
        init.assign(M, 0);
        for (int m = 0; m < M; m++) {
            ssasdpd_rng r{};
            r.kind = 0;
            r.seed = seeds[m];
            eng[m]->reset(&r);
            init[m] = eng[m]->prepare_static() ? 1 : 0;
        }
    }
    // One launch: steps [0, s_end) of every initialised model.
    float launch(unsigned s_end, bool snapshots, std::vector<uint32_t *> *d_counts, const std::vector<char> &init) {
        std::vector<RdmeArgs> args;
        args.reserve(M);
        const int K = (int)snap_steps.size();
        for (int m : order) {
            if (!init[m]) continue;
            ssasdpd_engine *e = eng[m];
            RdmeArgs a = e->rdme_args(0);
            a.step = 0;
            a.step_end = s_end;
            a.t0 = 0.0;
            a.trace = nullptr;
            a.trace_cap = 0;
            if (snapshots) {
                a.snap_step = d_snap;
                a.n_snap = K;
                a.snap_tot = d_tot + (size_t)m * K * Smax;
                a.snap_xx = d_counts ? (*d_counts)[m] : nullptr;
            }
            args.push_back(a);
        }
        if (args.empty()) return 0.f;
        h2d(st, d_args, args.data(), args.size());
        HIP_OK(hipEventRecord(e0, st));
        bool any_prog = false;
        for (auto *e : eng) any_prog = any_prog || e->n_code > 0;
        rdme_take_step_multi(st, d_args, (int)args.size(), pl, max_sr, lds_max(), any_prog);
        HIP_OK(hipEventRecord(e1, st));
        HIP_OK(hipStreamSynchronize(st));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        return ms;
    }
    int collect(const std::vector<char> &init, unsigned s_end) {
        stats.reactions = stats.diffusions = stats.rng_words_used = 0;
        int rc = 0;
        for (int m = 0; m < M; m++) {
            if (!init[m]) continue;
            QueueState q;
            d2h(st, &q, eng[m]->qs, 1);
            HIP_OK(hipStreamSynchronize(st));
            stats.reactions += q.reactions;
            stats.diffusions += q.diffusions;
            stats.rng_words_used += q.rng_pos;
            if (q.err && !rc) {
                char buf[128];
                snprintf(buf, sizeof buf, "SSA failed at step %u in model %d", q.err_step, m);
                rc = set_error(q.err, buf);
            }
            eng[m]->cur = s_end;
        }
        return rc;
    }
};

// ---------------------------------------------------------------------------
// C ABI
#define GUARD_BEGIN try {
#define GUARD_END                                                            \
    }                                                                        \
    catch (const HipError &e) {                                              \
        return set_error(SSASDPD_ERR_HIP, e.msg);                            \
    }                                                                        \
    catch (const std::exception &e) {                                        \
        return set_error(SSASDPD_ERR_HIP, e.what());                         \
    }

static int validate(const ssasdpd_desc *d) {
    if (!d) return set_error(SSASDPD_ERR_BAD_DESC, "null descriptor");
    if (d->abi_version != SSASDPD_ABI_VERSION) return set_error(SSASDPD_ERR_BAD_DESC, "ABI version mismatch");
    if (d->n <= 0) return set_error(SSASDPD_ERR_BAD_DESC, "n must be positive");
    if (d->dimension < 1 || d->dimension > 3) return set_error(SSASDPD_ERR_BAD_DESC, "dimension must be 1, 2 or 3");
    if (d->n_types < 1 || d->n_types > 63) return set_error(SSASDPD_ERR_BAD_DESC, "n_types must be in 1..63");
    if (!d->x || !d->type || !d->mass || !d->rho || !d->nu || !d->fixed)
        return set_error(SSASDPD_ERR_BAD_DESC, "missing particle arrays");
    if (!(d->h > 0)) return set_error(SSASDPD_ERR_BAD_DESC, "h (basis function width) must be positive");
    if (d->n_species > 16 || d->n_reactions > 16 || d->n_chem_species > 16)
        return set_error(SSASDPD_ERR_UNSUPPORTED, "at most 16 species and 16 reactions");
    if (d->n_species > 0 && (!d->u0 || !d->diffusion_matrix || !d->jcG))
        return set_error(SSASDPD_ERR_BAD_DESC, "missing chemistry arrays");
    if (d->n_reactions > 0 && (!d->reactions || !d->jcN))
        return set_error(SSASDPD_ERR_BAD_DESC, "missing reaction arrays");
    for (int i = 0; i < d->n; i++)
        if (d->type[i] < 1 || d->type[i] > d->n_types)
            return set_error(SSASDPD_ERR_BAD_DESC, "particle type out of range 1..n_types");
    int Rmax = std::max(d->n_reactions, d->n_chem_reactions);
    int Smax = std::max(d->n_species, d->n_chem_species);
    if (d->n_code < 0 || d->n_consts < 0 || d->n_data_fn < 0 || (d->n_code > 0 && !d->code) ||
        (d->n_consts > 0 && !d->consts) || (d->n_data_fn > 0 && !d->data_fn))
        return set_error(SSASDPD_ERR_BAD_DESC, "bad program / data-function arrays");
    if (d->n_code > (1 << 16) || d->n_consts > (1 << 16))
        return set_error(SSASDPD_ERR_UNSUPPORTED, "at most 65536 program words and constants");
    for (int r = 0; r < Rmax; r++) {
        const ssasdpd_reaction &q = d->reactions[r];
        if (q.kind == 4) {
            // programs: the propensity (species = stochastic counts) and the
            // deterministic rate (species = concentrations) must both check
            if (d->n_reactions > 0 &&
                (q.a < 0 || vm_check(d->code, d->n_code, d->n_consts, q.a, d->n_species, d->n_data_fn) <= 0))
                return set_error(SSASDPD_ERR_BAD_DESC, "reaction " + std::to_string(r) + ": bad propensity program");
            if (d->n_chem_reactions > 0 &&
                (q.b < 0 || vm_check(d->code, d->n_code, d->n_consts, q.b, d->n_chem_species, d->n_data_fn) <= 0))
                return set_error(SSASDPD_ERR_BAD_DESC,
                                 "reaction " + std::to_string(r) + ": bad deterministic-rate program");
            continue;
        }
        if (q.kind < 0 || q.kind > 3 || q.a < 0 || q.a >= std::max(Smax, 1) || q.b < 0 || q.b >= std::max(Smax, 1))
            return set_error(SSASDPD_ERR_BAD_DESC, "bad reaction table entry");
    }
    // CSC columns: row indices strictly ascending and in range (the event
    // loop applies a column as a bitmask in ascending order, as the
    // reference's loops over these columns do for the matrices model.py builds)
    auto csc_ok = [](const int32_t *jc, const int32_t *ir, int cols, int rows) {
        if (jc[0] != 0) return false;
        for (int c = 0; c < cols; c++) {
            if (jc[c + 1] < jc[c]) return false;
            for (int k = jc[c]; k < jc[c + 1]; k++) {
                if (ir[k] < 0 || ir[k] >= rows) return false;
                if (k > jc[c] && ir[k] <= ir[k - 1]) return false;
            }
        }
        return true;
    };
    if (d->n_species > 0) {
        if (!csc_ok(d->jcG, d->irG, d->n_species + d->n_reactions, std::max(d->n_reactions, 1)))
            return set_error(SSASDPD_ERR_BAD_DESC, "dependency graph CSC: rows must be ascending and < n_reactions");
        if (d->n_reactions > 0 && !csc_ok(d->jcN, d->irN, d->n_reactions, d->n_species))
            return set_error(SSASDPD_ERR_BAD_DESC, "stoichiometry CSC: rows must be ascending and < n_species");
    }
    return 0;
}

extern "C" {

const char *ssasdpd_last_error(void) { return g_last_error.c_str(); }
int ssasdpd_abi_version(void) { return SSASDPD_ABI_VERSION; }

int ssasdpd_create(const ssasdpd_desc *d, ssasdpd_engine **out) {
    if (!out) return set_error(SSASDPD_ERR_BAD_DESC, "null out");
    *out = nullptr;
    int rc = validate(d);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(SSASDPD_ERR_HIP, "no HIP device");
    ssasdpd_engine *e = new ssasdpd_engine();
    try {
        e->setup(d);
        e->reset(nullptr);
    } catch (const HipError &err) {
        delete e;
        return set_error(SSASDPD_ERR_HIP, err.msg);
    }
    *out = e;
    return 0;
}

void ssasdpd_destroy(ssasdpd_engine *e) { delete e; }

int ssasdpd_num_snapshots(const ssasdpd_desc *d, int32_t *n_snap) {
    uint32_t next = 0;
    int idx = 0, k = 0;
    for (uint32_t step = 0; step < d->nt; step++) {
        if (step >= next) {
            k++;
            next = idx < d->n_output_steps ? d->output_steps[idx] : 0xFFFFFFFFu;
            idx++;
        }
    }
    *n_snap = k + 1;
    return 0;
}

int ssasdpd_reset(ssasdpd_engine *e, const ssasdpd_rng *r) {
    GUARD_BEGIN
    e->reset(r);
    return 0;
    GUARD_END
}

int ssasdpd_advance(ssasdpd_engine *e, uint32_t n_steps) {
    GUARD_BEGIN
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t target = (uint32_t)std::min<uint64_t>((uint64_t)e->cur + n_steps, e->nt);
    e->arm_deadline();
    while (e->cur < target) {
        if (int rc = e->deadline_rc()) return rc;
        if (e->fusable()) {
            // with a time limit, fused launches of at most FUSED_SEGMENT
            // steps, the limit checked between them
            const uint32_t seg = e->deadline_on ? std::min<uint32_t>(target, e->cur + FUSED_SEGMENT) : target;
            int rc = e->particle_steps(e->cur, seg, {}, [](size_t) {});
            if (rc) return rc;
            e->ssa_fused(e->rdme_args(e->cur), 1, seg, {}, nullptr, 0, nullptr);
            rc = e->check(seg - 1);
            if (rc) return rc;
            continue;
        }
        int rc = e->do_step(e->cur + 1 < target);
        if (rc) return rc;
    }
    e->stats.ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
    GUARD_END
}

// The reference's output schedule (run_simulation, simulate_threads.cpp:
// 238-247, 284-289) continued from step s0 to the end: the snapshot taken
// before each step that reaches the next output step, plus the final one.
// Returns the steps; ks receives the snapshot indices.
static std::vector<unsigned> plan_snapshots(const ssasdpd_engine *e, uint32_t s0, uint32_t &next, size_t &idx, int &k,
                                            std::vector<int> &ks) {
    std::vector<unsigned> steps;
    for (uint32_t step = s0; step < e->nt; step++) {
        if (step >= next) {
            steps.push_back(step);
            ks.push_back(k++);
            next = idx < e->out_steps.size() ? e->out_steps[idx] : 0xFFFFFFFFu;
            idx++;
        }
    }
    steps.push_back(e->nt);
    ks.push_back(k++);
    return steps;
}

int ssasdpd_run(ssasdpd_engine *e, const ssasdpd_rng *r, ssasdpd_outputs *o) {
    GUARD_BEGIN
    e->reset(r);
    e->arm_deadline();
    auto t0 = std::chrono::steady_clock::now();
    uint32_t next = 0;
    size_t idx = 0;
    int k = 0;
    for (uint32_t step = 0; step < e->nt; step++) {
        if (int rc = e->deadline_rc()) return rc;
        if (e->fusable() && e->deadline_on) {
            // with a time limit: the snapshot due before this step, then the
            // steps up to the next snapshot (at most FUSED_SEGMENT) in one
            // launch, the limit checked between launches
            if (step >= next) {
                if (o && k < o->n_snapshots) e->snapshot(o, k);
                k++;
                next = idx < e->out_steps.size() ? e->out_steps[idx] : 0xFFFFFFFFu;
                idx++;
            }
            // the next snapshot is due before step max(next, step + 1)
            const uint32_t seg = std::min<uint32_t>({e->nt, std::max<uint32_t>(next, step + 1), step + FUSED_SEGMENT});
            int rc = e->particle_steps(step, seg, {}, [](size_t) {});
            if (rc) return rc;
            e->ssa_fused(e->rdme_args(step), 1, seg, {}, nullptr, 0, nullptr);
            rc = e->check(seg - 1);
            if (rc) return rc;
            step = seg - 1;
            continue;
        }
        if (e->fusable()) {
            // all remaining steps in one launch, count snapshots in-kernel
            std::vector<int> ks;
            std::vector<unsigned> sst = plan_snapshots(e, step, next, idx, k, ks);
            const int K = o ? (int)std::count_if(ks.begin(), ks.end(), [&](int q) { return q < o->n_snapshots; }) : 0;
            sst.resize(K);
            const size_t per = (size_t)e->n * e->S;
            // particle fields first (independent of the event loop here)
            int rc = e->particle_steps(step, e->nt, sst, [&](size_t j) {
                uint32_t *keep = o->xx;
                o->xx = nullptr;
                e->snapshot(o, ks[j]);
                o->xx = keep;
                if (o->steps) o->steps[ks[j]] = sst[j];
            });
            if (rc) return rc;
            uint32_t *d_xx = (K > 0 && o->xx) ? dalloc<uint32_t>(per * K) : nullptr;
            e->ssa_fused(e->rdme_args(step), 1, e->nt, sst, d_xx, 0, nullptr);
            rc = e->check(e->nt - 1);
            for (int j = 0; j < K && !rc; j++)
                if (d_xx) d2h(e->st, o->xx + (size_t)ks[j] * per, d_xx + (size_t)j * per, per);
            HIP_OK(hipStreamSynchronize(e->st));
            dfree(d_xx);
            if (rc) return rc;
            e->stats.ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            return 0;
        }
        if (step >= next) {
            if (o && k < o->n_snapshots) e->snapshot(o, k);
            k++;
            next = idx < e->out_steps.size() ? e->out_steps[idx] : 0xFFFFFFFFu;
            idx++;
        }
        int rc = e->do_step();
        if (rc) return rc;
    }
    if (o && k < o->n_snapshots) e->snapshot(o, k);
    e->stats.ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
    GUARD_END
}

int ssasdpd_run_batch(ssasdpd_engine *e, int32_t n_traj, const uint64_t *seeds, ssasdpd_outputs *shared,
                      uint32_t *counts) {
    GUARD_BEGIN
    if (n_traj <= 0 || !seeds) return set_error(SSASDPD_ERR_BAD_DESC, "run_batch: n_traj must be positive and seeds given");
    if (e->S > 0 && e->lds_bytes == 0) return set_error(SSASDPD_ERR_BAD_DESC, "run_batch: no RDME queue");
    if (e->trace_cap > 0) return set_error(SSASDPD_ERR_UNSUPPORTED, "run_batch: event trace is single-trajectory only");
    const int B = n_traj;
    int n_snap = 0;
    {
        uint32_t nx = 0;
        size_t ix = 0;
        for (uint32_t step = 0; step < e->nt; step++)
            if (step >= nx) { n_snap++; nx = ix < e->out_steps.size() ? e->out_steps[ix] : 0xFFFFFFFFu; ix++; }
        n_snap++;
    }

    e->batch_reset(B, seeds);
    e->arm_deadline();
    auto t0 = std::chrono::steady_clock::now();
    const size_t per = (size_t)e->n * e->S;
    auto snap_shared = [&](int k, uint32_t step) {
        if (shared && k < shared->n_snapshots) {
            uint32_t *keep = shared->xx;
            shared->xx = nullptr;
            e->snapshot(shared, k);
            shared->xx = keep;
            if (shared->steps) shared->steps[k] = step;
        }
    };
    auto snap = [&](int k) {
        snap_shared(k, e->cur);
        if (counts && e->S > 0) {
            for (int b = 0; b < B; b++)
                d2h(e->st, counts + ((size_t)b * n_snap + k) * per, e->orig(e->b_xx + (size_t)b * per, e->S), per);
            HIP_OK(hipStreamSynchronize(e->st));
        }
    };
    uint32_t next = 0;
    size_t idx = 0;
    int k = 0;
    for (uint32_t step = 0; step < e->nt; step++) {
        if (int rc = e->deadline_rc()) return rc;
        if (e->fusable_static() && e->batch_all_init() && e->deadline_on) {
            // as in ssasdpd_run: segments between snapshots, limit checked
            // between launches
            if (step >= next) {
                if (k < n_snap) snap(k);
                k++;
                next = idx < e->out_steps.size() ? e->out_steps[idx] : 0xFFFFFFFFu;
                idx++;
            }
            // the next snapshot is due before step max(next, step + 1)
            const uint32_t seg = std::min<uint32_t>({e->nt, std::max<uint32_t>(next, step + 1), step + FUSED_SEGMENT});
            int rc = e->particle_steps(step, seg, {}, [](size_t) {});
            if (rc) return rc;
            e->ssa_fused(e->batch_args(step, -1), B, seg, {}, nullptr, 0, nullptr);
            rc = e->batch_check(B, seg - 1);
            if (rc) return rc;
            step = seg - 1;
            continue;
        }
        if (e->fusable_static() && e->batch_all_init()) {
            std::vector<int> ks;
            std::vector<unsigned> sst = plan_snapshots(e, step, next, idx, k, ks);
            const int K = (int)sst.size();
            int rc = e->particle_steps(step, e->nt, sst, [&](size_t j) { snap_shared(ks[j], sst[j]); });
            if (rc) return rc;
            uint32_t *d_xx = (counts && e->S > 0) ? dalloc<uint32_t>(per * K * B) : nullptr;
            e->ssa_fused(e->batch_args(step, -1), B, e->nt, sst, d_xx, per * K, nullptr);
            rc = e->batch_check(B, e->nt - 1);
            for (int j = 0; j < K && !rc; j++) {
                if (d_xx)
                    for (int b = 0; b < B; b++)
                        d2h(e->st, counts + ((size_t)b * n_snap + ks[j]) * per, d_xx + ((size_t)b * K + j) * per, per);
            }
            HIP_OK(hipStreamSynchronize(e->st));
            dfree(d_xx);
            if (rc) return rc;
            e->stats.ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            return 0;
        }
        if (step >= next) {
            if (k < n_snap) snap(k);
            k++;
            next = idx < e->out_steps.size() ? e->out_steps[idx] : 0xFFFFFFFFu;
            idx++;
        }
        int rc = e->batch_step(B);
        if (rc) return rc;
    }
    if (k < n_snap) snap(k);
    e->stats.ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
    GUARD_END
}

int ssasdpd_eval_forces(ssasdpd_engine *e, int32_t fp32, double *F, double *Fbp, double *Frho) {
    GUARD_BEGIN
    if (!e->lists_built) return set_error(SSASDPD_ERR_BAD_DESC, "eval_forces: no neighbour lists yet (advance first)");
    const size_t n = e->n, n3 = 3 * n;
    double *dF = dalloc<double>(n3), *dB = dalloc<double>(n3), *dR = dalloc<double>(n);
    double *dQ = e->Sc > 0 ? dalloc<double>(n * e->Sc) : nullptr;
    std::vector<double> g0(n3);
    for (int k = 0; k < 3; k++) std::fill(g0.begin() + k * n, g0.begin() + (k + 1) * n, e->g[k]);
    h2d(e->st, dF, g0.data(), n3);
    HIP_OK(hipMemsetAsync(dB, 0, sizeof(double) * n3, e->st));
    HIP_OK(hipMemsetAsync(dR, 0, sizeof(double) * n, e->st));
    if (dQ) HIP_OK(hipMemsetAsync(dQ, 0, sizeof(double) * n * e->Sc, e->st));
    SdpdArgs a = e->sdpd_args(e->cur);
    a.F = dF; a.Fbp = dB; a.Frho = dR; a.Q = dQ ? dQ : a.Q;
    a.Sc = dQ ? a.Sc : 0;
    a.Rc = dQ ? a.Rc : 0;
    a.fp32 = fp32 ? 1 : 0;
    sdpd_pack(e->st, a);      // gather records of the current state
    sdpd_forces(e->st, a);
    std::vector<double> tmp(n3);
    if (F) { d2h(e->st, tmp.data(), e->orig3(dF), n3); HIP_OK(hipStreamSynchronize(e->st)); from_soa3(tmp, F, (int)n); }
    if (Fbp) { d2h(e->st, tmp.data(), e->orig3(dB), n3); HIP_OK(hipStreamSynchronize(e->st)); from_soa3(tmp, Fbp, (int)n); }
    if (Frho) d2h(e->st, Frho, e->orig(dR, 1), n);
    HIP_OK(hipStreamSynchronize(e->st));
    dfree(dF); dfree(dB); dfree(dR); dfree(dQ);
    return 0;
    GUARD_END
}

int ssasdpd_vm_eval(const uint32_t *code, int32_t n_code, const double *consts, int32_t n_consts, int32_t start,
                    int32_t x_is_double, int32_t n_x, int32_t n_df, int32_t m, const double *x, const double *vol,
                    const double *t, const double *df, double *out) {
    GUARD_BEGIN
    if (!code || n_code <= 0 || n_consts < 0 || (n_consts > 0 && !consts) || n_x < 0 || n_df < 0 || m < 0 ||
        (m > 0 && (!vol || !t || !out || (n_x > 0 && !x) || (n_df > 0 && !df))))
        return set_error(SSASDPD_ERR_BAD_DESC, "vm_eval: bad arguments");
    if (start < 0 || vm_check(code, n_code, n_consts, start, n_x, n_df) <= 0)
        return set_error(SSASDPD_ERR_BAD_DESC, "vm_eval: program fails the check");
    if (m == 0) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(SSASDPD_ERR_HIP, "no HIP device");
    hipStream_t s;
    HIP_OK(hipStreamCreate(&s));
    const size_t nx = (size_t)m * n_x;
    uint32_t *dc = dalloc<uint32_t>(n_code);
    double *dk = dalloc<double>(std::max(n_consts, 1));
    double *dxd = x_is_double ? dalloc<double>(std::max<size_t>(nx, 1)) : nullptr;
    uint32_t *dxu = x_is_double ? nullptr : dalloc<uint32_t>(std::max<size_t>(nx, 1));
    double *dv = dalloc<double>(m), *dt = dalloc<double>(m), *dout = dalloc<double>(m);
    double *ddf = n_df > 0 ? dalloc<double>((size_t)m * n_df) : nullptr;
    h2d(s, dc, code, n_code);
    if (n_consts > 0) h2d(s, dk, consts, n_consts);
    std::vector<uint32_t> xu;
    if (x_is_double) {
        if (nx) h2d(s, dxd, x, nx);
    } else if (nx) {
        xu.resize(nx);
        for (size_t i = 0; i < nx; i++) xu[i] = (uint32_t)(int32_t)x[i];
        h2d(s, dxu, xu.data(), nx);
    }
    h2d(s, dv, vol, m);
    h2d(s, dt, t, m);
    if (ddf) h2d(s, ddf, df, (size_t)m * n_df);
    vm_eval_batch(s, dc, start, dk, x_is_double ? 1 : 0, n_x, n_df, m, dxd, dxu, dv, dt, ddf, dout);
    HIP_OK(hipGetLastError());
    d2h(s, out, dout, m);
    HIP_OK(hipStreamSynchronize(s));
    dfree(dc); dfree(dk); dfree(dxd); dfree(dxu); dfree(dv); dfree(dt); dfree(dout); dfree(ddf);
    HIP_OK(hipStreamDestroy(s));
    return 0;
    GUARD_END
}

int ssasdpd_ensemble_create(const ssasdpd_desc *const *descs, int32_t n_models, ssasdpd_ensemble **out) {
    GUARD_BEGIN
    if (!descs || n_models <= 0 || !out) return set_error(SSASDPD_ERR_BAD_DESC, "ensemble: no models");
    for (int m = 0; m < n_models; m++) {
        int rc = validate(descs[m]);
        if (rc) return rc;
        const ssasdpd_desc *d = descs[m];
        if (!d->static_domain) return set_error(SSASDPD_ERR_UNSUPPORTED, "ensemble: static domains only");
        if (d->n_species <= 0) return set_error(SSASDPD_ERR_BAD_DESC, "ensemble: every model needs species");
        const ssasdpd_desc *d0 = descs[0];
        bool same = d->nt == d0->nt && d->dt == d0->dt && d->n_output_steps == d0->n_output_steps;
        for (int k = 0; same && k < d->n_output_steps; k++) same = d->output_steps[k] == d0->output_steps[k];
        if (!same) return set_error(SSASDPD_ERR_BAD_DESC, "ensemble: models must share nt, dt and output steps");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(SSASDPD_ERR_HIP, "no HIP device");
    auto *E = new ssasdpd_ensemble();
    try {
        E->M = n_models;
        HIP_OK(hipStreamCreate(&E->st));
        HIP_OK(hipEventCreate(&E->e0));
        HIP_OK(hipEventCreate(&E->e1));
        int p = 1;
        for (int m = 0; m < n_models; m++) p = std::max(p, ssasdpd_engine::auto_pl(descs[m]->n));
        for (int m = 0; m < n_models; m++) {
            // the deterministic species do not enter the event loop, and an
            // ensemble reports counts only, so they are left out
            ssasdpd_desc dm = *descs[m];
            dm.n_chem_species = 0;
            dm.n_chem_reactions = 0;
            auto *e = new ssasdpd_engine();
            e->pl_override = p;
            e->ext_stream = E->st;
            E->eng.push_back(e);
            e->setup(&dm);
            E->Smax = std::max(E->Smax, e->S);
            E->max_sr = std::max(E->max_sr, std::max(e->S, e->R));
        }
        E->pl = p;
        E->nt = descs[0]->nt;
        uint32_t next = 0;
        size_t idx = 0;
        int k = 0;
        std::vector<int> ks;
        E->snap_steps = plan_snapshots(E->eng[0], 0, next, idx, k, ks);
        E->order.resize(n_models);
        for (int m = 0; m < n_models; m++) E->order[m] = m;
        std::stable_sort(E->order.begin(), E->order.end(),
                         [&](int a, int b) { return E->eng[a]->n > E->eng[b]->n; });
        E->d_args = dalloc<RdmeArgs>(n_models);
        E->d_snap = dalloc<unsigned>(E->snap_steps.size());
        h2d(E->st, E->d_snap, E->snap_steps.data(), E->snap_steps.size());
        E->d_tot = dalloc<uint32_t>((size_t)n_models * E->snap_steps.size() * E->Smax);
        HIP_OK(hipStreamSynchronize(E->st));
    } catch (...) {
        delete E;
        throw;
    }
    *out = E;
    return 0;
    GUARD_END
}

void ssasdpd_ensemble_destroy(ssasdpd_ensemble *m) { delete m; }

int ssasdpd_ensemble_num_snapshots(ssasdpd_ensemble *m, int32_t *k) {
    *k = (int32_t)m->snap_steps.size();
    return 0;
}

int ssasdpd_ensemble_set_tree_width(ssasdpd_ensemble *m, int32_t pl) {
    GUARD_BEGIN
    if (!m->pl_ok(pl)) return set_error(SSASDPD_ERR_BAD_DESC, "ensemble: tree width not valid for every model");
    m->set_pl(pl);
    return 0;
    GUARD_END
}

int ssasdpd_ensemble_autotune(ssasdpd_ensemble *m, const uint64_t *seeds, int32_t probe_steps, int32_t *chosen,
                              double *ms_per_width) {
    GUARD_BEGIN
    const int cand[4] = {1, 2, 4, 8};
    const unsigned s_end = (unsigned)std::max(1, std::min<int>(probe_steps, (int)m->nt));
    float best = 3.4e38f;
    int best_pl = m->pl;
    for (int c = 0; c < 4; c++) {
        if (ms_per_width) ms_per_width[c] = -1.0;
        if (!m->pl_ok(cand[c])) continue;
        m->set_pl(cand[c]);
        std::vector<char> init;
        m->prepare(seeds, init);
        const float ms = m->launch(s_end, false, nullptr, init);
        int rc = m->collect(init, s_end);
        if (rc) return rc;
        if (ms_per_width) ms_per_width[c] = ms;
        if (ms < best) { best = ms; best_pl = cand[c]; }
    }
    m->set_pl(best_pl);
    if (chosen) *chosen = best_pl;
    return 0;
    GUARD_END
}

int ssasdpd_ensemble_run(ssasdpd_ensemble *m, const uint64_t *seeds, uint32_t *totals, uint32_t *const *counts) {
    GUARD_BEGIN
    if (!seeds) return set_error(SSASDPD_ERR_BAD_DESC, "ensemble_run: seeds required");
    auto t0 = std::chrono::steady_clock::now();
    const int K = (int)m->snap_steps.size();
    std::vector<char> init;
    m->prepare(seeds, init);
    std::vector<uint32_t *> d_counts;
    if (counts) {
        d_counts.assign(m->M, nullptr);
        for (int q = 0; q < m->M; q++)
            if (counts[q]) d_counts[q] = dalloc<uint32_t>((size_t)K * m->eng[q]->n * m->eng[q]->S);
    }
    m->ms_launch = m->launch(m->nt, true, counts ? &d_counts : nullptr, init);
    int rc = m->collect(init, m->nt);
    if (!rc) {
        std::vector<uint32_t> tot((size_t)m->M * K * m->Smax);
        d2h(m->st, tot.data(), m->d_tot, tot.size());
        HIP_OK(hipStreamSynchronize(m->st));
        for (int q = 0; q < m->M; q++) {
            ssasdpd_engine *e = m->eng[q];
            const size_t per = (size_t)e->n * e->S;
            std::vector<uint32_t> u0sum(e->S, 0);
            if (!init[q])
                for (size_t i = 0; i < per; i++) u0sum[i % e->S] += e->u0[i];
            for (int k = 0; k < K; k++)
                for (int s = 0; s < m->Smax; s++) {
                    uint32_t v = 0;
                    if (s < e->S) v = init[q] ? tot[((size_t)q * K * m->Smax) + (size_t)k * e->S + s] : u0sum[s];
                    if (totals) totals[((size_t)q * K + k) * m->Smax + s] = v;
                }
            if (counts && counts[q]) {
                if (init[q]) {
                    d2h(m->st, counts[q], d_counts[q], per * K);
                } else {
                    for (int k = 0; k < K; k++) std::copy(e->u0.begin(), e->u0.end(), counts[q] + (size_t)k * per);
                }
            }
        }
        HIP_OK(hipStreamSynchronize(m->st));
    }
    for (auto *p : d_counts) dfree(p);
    m->stats.steps = m->nt;
    uint64_t ps = 0;
    for (auto *e : m->eng) ps += (uint64_t)e->n * m->nt;
    m->stats.particle_steps = ps;
    m->stats.ms_ssa = m->ms_launch;
    m->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
    GUARD_END
}

int ssasdpd_ensemble_stats(ssasdpd_ensemble *m, ssasdpd_stats *s) {
    *s = m->stats;
    return 0;
}

struct ssasdpd_mpn {
    ssa::Mpn *m;
};

int32_t ssasdpd_mpn_num_params(void) { return Mpn::P; }

int ssasdpd_mpn_create(uint64_t seed, ssasdpd_mpn **out) {
    GUARD_BEGIN
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(SSASDPD_ERR_HIP, "no HIP device");
    auto *h = new ssasdpd_mpn{new Mpn(seed)};
    *out = h;
    return 0;
    GUARD_END
}

void ssasdpd_mpn_destroy(ssasdpd_mpn *m) {
    if (!m) return;
    delete m->m;
    delete m;
}

int ssasdpd_mpn_get_params(ssasdpd_mpn *m, float *w) {
    GUARD_BEGIN
    m->m->get_params(w);
    return 0;
    GUARD_END
}

int ssasdpd_mpn_set_params(ssasdpd_mpn *m, const float *w) {
    GUARD_BEGIN
    m->m->set_params(w);
    return 0;
    GUARD_END
}

int ssasdpd_mpn_set_scales(ssasdpd_mpn *m, const float *sc) {
    for (int k = 0; k < 6; k++)
        if (!(sc[k] > 0.f) || !std::isfinite(sc[k]))
            return set_error(SSASDPD_ERR_BAD_DESC, "mpn scales must be positive and finite");
    EcnScales &s = m->m->sc;
    s.rho0 = sc[0]; s.vs = sc[1]; s.rs = sc[2]; s.ns = sc[3]; s.acc = sc[4]; s.ts = sc[5];
    return 0;
}

// A host-supplied particle graph on the device (records and lists).
namespace {
struct DevGraph {
    PRec *rec = nullptr;
    int *deg = nullptr, *nidx = nullptr;
    double *nr = nullptr;
    EcnGraph view;
    std::vector<int> deg_h, movers;
    ~DevGraph() { dfree(rec); dfree(deg); dfree(nidx); dfree(nr); }
};
int upload_graph(const ssasdpd_mpn_graph *g, hipStream_t st, DevGraph &d) {
    if (!g || g->n <= 0 || g->W <= 0 || g->dim < 1 || g->dim > 3 || !(g->h > 0.0) || !g->x || !g->v || !g->vt ||
        !g->rho || !g->mass || !g->nu || !g->solid || !g->deg || !g->nidx || !g->nr)
        return set_error(SSASDPD_ERR_BAD_DESC, "mpn graph: bad arguments");
    const int n = g->n, W = g->W;
    for (int i = 0; i < n; i++) {
        if (g->deg[i] < 0 || g->deg[i] > W) return set_error(SSASDPD_ERR_BAD_DESC, "mpn graph: degree out of range");
        for (int k = 0; k < g->deg[i]; k++) {
            const int j = g->nidx[(size_t)k * n + i];
            if (j < 0 || j >= n) return set_error(SSASDPD_ERR_BAD_DESC, "mpn graph: neighbour index out of range");
        }
        if (!(g->mass[i] > 0.0) || !(g->rho[i] > 0.0))
            return set_error(SSASDPD_ERR_BAD_DESC, "mpn graph: mass and density must be positive");
    }
    std::vector<PRec> r(n);
    for (int i = 0; i < n; i++) {
        PRec &q = r[i];
        std::memset(&q, 0, sizeof(PRec));
        for (int c = 0; c < 3; c++) {
            q.x[c] = g->x[3 * (size_t)i + c];
            q.v[c] = g->v[3 * (size_t)i + c];
            q.vt[c] = g->vt[3 * (size_t)i + c];
        }
        q.nu = g->nu[i]; q.rho = g->rho[i]; q.mass = g->mass[i]; q.type = 1; q.solid = g->solid[i] ? 1 : 0;
        if (!q.solid) d.movers.push_back(i);
    }
    d.deg_h.assign(g->deg, g->deg + n);
    d.rec = dalloc<PRec>(n);
    d.deg = dalloc<int>(n);
    d.nidx = dalloc<int>((size_t)W * n);
    d.nr = dalloc<double>((size_t)W * n);
    h2d(st, d.rec, r.data(), n);
    h2d(st, d.deg, g->deg, n);
    h2d(st, d.nidx, g->nidx, (size_t)W * n);
    h2d(st, d.nr, g->nr, (size_t)W * n);
    HIP_OK(hipStreamSynchronize(st));
    d.view.n = n; d.view.W = W; d.view.dim = g->dim; d.view.h = g->h;
    d.view.rec = d.rec; d.view.deg = d.deg; d.view.nidx = d.nidx; d.view.nr = d.nr;
    return 0;
}
}  // namespace

int ssasdpd_mpn_add_sample(ssasdpd_mpn *m, const ssasdpd_mpn_graph *g, const double *acc, int32_t n_sel,
                           const int32_t *sel, int32_t *index) {
    GUARD_BEGIN
    ssa::Mpn *M = m->m;
    DevGraph d;
    if (int rc = upload_graph(g, M->st, d)) return rc;
    if (!acc) return set_error(SSASDPD_ERR_BAD_DESC, "mpn sample: no target accelerations");
    std::vector<int> s = sel ? std::vector<int>(sel, sel + std::max(n_sel, 0)) : d.movers;
    if (s.empty()) return set_error(SSASDPD_ERR_BAD_DESC, "mpn sample: no particles selected");
    for (int i : s)
        if (i < 0 || i >= g->n) return set_error(SSASDPD_ERR_BAD_DESC, "mpn sample: selected index out of range");
    std::vector<float> y(3 * s.size());
    for (size_t p = 0; p < s.size(); p++)
        for (int c = 0; c < 3; c++) y[3 * p + c] = (float)(acc[3 * (size_t)s[p] + c] / (double)M->sc.acc);
    const int k = M->add_sample(M->st, d.view, d.deg_h.data(), s.data(), (int)s.size(), y.data(), nullptr, nullptr);
    if (index) *index = k;
    return 0;
    GUARD_END
}

int ssasdpd_mpn_predict(ssasdpd_mpn *m, const ssasdpd_mpn_graph *g, double *acc) {
    GUARD_BEGIN
    ssa::Mpn *M = m->m;
    DevGraph d;
    if (int rc = upload_graph(g, M->st, d)) return rc;
    if (!acc) return set_error(SSASDPD_ERR_BAD_DESC, "mpn predict: no output");
    const size_t n = g->n;
    double *F = dalloc<double>(3 * n);
    const double zero[3] = {0.0, 0.0, 0.0};
    M->infer(M->st, d.view, zero, F);
    std::vector<double> tmp(3 * n);
    d2h(M->st, tmp.data(), F, 3 * n);
    HIP_OK(hipStreamSynchronize(M->st));
    dfree(F);
    from_soa3(tmp, acc, (int)n);
    return 0;
    GUARD_END
}

int ssasdpd_mpn_clear_samples(ssasdpd_mpn *m) {
    m->m->clear_samples();
    return 0;
}

int ssasdpd_mpn_sample_size(ssasdpd_mpn *m, int32_t idx, int32_t *n_particles, int32_t *n_edges) {
    if (idx < 0 || idx >= (int)m->m->samples.size()) return set_error(SSASDPD_ERR_BAD_DESC, "mpn: no such sample");
    if (n_particles) *n_particles = m->m->samples[idx].np;
    if (n_edges) *n_edges = m->m->samples[idx].ne;
    return 0;
}

int ssasdpd_mpn_forward(ssasdpd_mpn *m, int32_t idx, float *out, float *target) {
    GUARD_BEGIN
    ssa::Mpn *M = m->m;
    if (idx < 0 || idx >= (int)M->samples.size()) return set_error(SSASDPD_ERR_BAD_DESC, "mpn: no such sample");
    const EcnSample &s = M->samples[idx];
    M->forward(s, nullptr);
    std::vector<float> p(3 * (size_t)s.np);
    if (out) d2h(M->st, p.data(), M->pred(), p.size());
    std::vector<float> y(3 * (size_t)s.np);
    if (target) d2h(M->st, y.data(), s.y, y.size());
    HIP_OK(hipStreamSynchronize(M->st));
    for (size_t q = 0; q < p.size(); q++) {
        if (out) out[q] = p[q] * M->sc.acc;
        if (target) target[q] = y[q] * M->sc.acc;
    }
    return 0;
    GUARD_END
}

int ssasdpd_mpn_grad(ssasdpd_mpn *m, int32_t idx, float *grad, double *loss) {
    GUARD_BEGIN
    ssa::Mpn *M = m->m;
    if (idx < 0 || idx >= (int)M->samples.size()) return set_error(SSASDPD_ERR_BAD_DESC, "mpn: no such sample");
    const EcnSample &s = M->samples[idx];
    float *dl = dalloc<float>(1);
    HIP_OK(hipMemsetAsync(dl, 0, sizeof(float), M->st));
    M->forward(s, dl);
    M->backward(s);
    float l = 0.f;
    d2h(M->st, &l, dl, 1);
    if (grad) d2h(M->st, grad, M->g, Mpn::P);
    HIP_OK(hipStreamSynchronize(M->st));
    dfree(dl);
    if (loss) *loss = l;
    return 0;
    GUARD_END
}

int ssasdpd_mpn_train(ssasdpd_mpn *m, int32_t epochs, float lr, uint64_t shuffle_seed, double *mean_loss) {
    GUARD_BEGIN
    ssa::Mpn *M = m->m;
    const int ns = (int)M->samples.size();
    if (ns == 0) return set_error(SSASDPD_ERR_BAD_DESC, "mpn: no training samples");
    if (epochs < 1 || !(lr > 0.f)) return set_error(SSASDPD_ERR_BAD_DESC, "mpn: epochs >= 1 and lr > 0");
    float *dl = dalloc<float>(1);
    std::vector<int> order(ns);
    for (int i = 0; i < ns; i++) order[i] = i;
    std::mt19937_64 rng(shuffle_seed);
    double last = 0.0;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(order.begin(), order.end(), rng);
        HIP_OK(hipMemsetAsync(dl, 0, sizeof(float), M->st));
        for (int k = 0; k < ns; k++) {
            const EcnSample &s = M->samples[order[k]];
            M->forward(s, dl);
            M->backward(s);
            M->adam(lr);
        }
        if (ep == epochs - 1) {
            float l = 0.f;
            d2h(M->st, &l, dl, 1);
            HIP_OK(hipStreamSynchronize(M->st));
            last = l / ns;
        }
    }
    HIP_OK(hipStreamSynchronize(M->st));
    dfree(dl);
    if (mean_loss) *mean_loss = last;
    return 0;
    GUARD_END
}

int ssasdpd_mpn_add_engine_sample(ssasdpd_mpn *m, ssasdpd_engine *e, int32_t max_particles, uint64_t seed,
                                  int32_t *index) {
    GUARD_BEGIN
    if (!e->lists_built) return set_error(SSASDPD_ERR_BAD_DESC, "mpn: engine has no neighbour lists yet");
    if (max_particles < 1) return set_error(SSASDPD_ERR_BAD_DESC, "mpn: max_particles must be >= 1");
    if (e->moving().empty()) return set_error(SSASDPD_ERR_BAD_DESC, "mpn: the engine has no moving particles");
    const size_t n = e->n, n3 = 3 * n;
    // the plain pair-force kernel on the current state is the target
    double *dF = dalloc<double>(n3), *dB = dalloc<double>(n3), *dR = dalloc<double>(n);
    std::vector<double> g0(n3);
    for (int k = 0; k < 3; k++) std::fill(g0.begin() + k * n, g0.begin() + (k + 1) * n, e->g[k]);
    h2d(e->st, dF, g0.data(), n3);
    HIP_OK(hipMemsetAsync(dB, 0, sizeof(double) * n3, e->st));
    HIP_OK(hipMemsetAsync(dR, 0, sizeof(double) * n, e->st));
    SdpdArgs a = e->sdpd_args(e->cur);
    a.F = dF; a.Fbp = dB; a.Frho = dR;
    a.Sc = 0;
    a.Rc = 0;
    a.fp32 = 0;
    sdpd_pack(e->st, a);
    sdpd_forces(e->st, a);
    std::vector<int> dg(n);
    d2h(e->st, dg.data(), e->deg, n);
    HIP_OK(hipStreamSynchronize(e->st));
    const std::vector<int> sel = e->pick_particles(max_particles, seed);
    const int k = m->m->add_sample(e->st, e->nn_graph(), dg.data(), sel.data(), (int)sel.size(), nullptr, dF, e->g);
    dfree(dF); dfree(dB); dfree(dR);
    if (index) *index = k;
    return 0;
    GUARD_END
}

int ssasdpd_mpn_time_engine(ssasdpd_mpn *m, ssasdpd_engine *e, int32_t reps, double *ms_per_pass) {
    GUARD_BEGIN
    if (!e->lists_built || !e->rec_ready || reps < 1)
        return set_error(SSASDPD_ERR_BAD_DESC, "mpn: engine has not stepped yet, or reps < 1");
    e->nn_ensure();
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    ssa::Mpn *M = m->m;
    M->infer(e->st, e->nn_graph(), e->g, e->nn_F);
    HIP_OK(hipEventRecord(a, e->st));
    for (int r = 0; r < reps; r++) M->infer(e->st, e->nn_graph(), e->g, e->nn_F);
    HIP_OK(hipEventRecord(b, e->st));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_per_pass = ms / reps;
    return 0;
    GUARD_END
}

int ssasdpd_mpn_record(ssasdpd_mpn *m, ssasdpd_engine *e, int32_t every, int32_t max_particles) {
    if (m && (every < 1 || max_particles < 1))
        return set_error(SSASDPD_ERR_BAD_DESC, "mpn_record: every and max_particles must be >= 1");
    e->nn_rec = m ? m->m : nullptr;
    e->nn_rec_every = every;
    if (m) e->nn_rec_max = max_particles;
    return 0;
}

int32_t ssasdpd_mpn_num_samples(ssasdpd_mpn *m) { return (int32_t)m->m->samples.size(); }

int ssasdpd_set_surrogate(ssasdpd_engine *e, ssasdpd_mpn *m, int32_t check_every, double max_rel_err) {
    if (m && check_every < 1) return set_error(SSASDPD_ERR_BAD_DESC, "surrogate: check_every must be >= 1");
    e->nn = m ? m->m : nullptr;
    e->nn_check_every = check_every;
    e->nn_tol = max_rel_err;
    e->nn_active = e->nn != nullptr;
    e->nn_steps = e->nn_checks = 0;
    e->nn_last_err = -1.0;
    e->nn_max_err = 0.0;
    e->nn_refreshes = 0;
    e->nn_fails = 0;
    e->nn_acc_cap = 0.0;
    e->nn_clear_capped();
    return 0;
}

int ssasdpd_set_surrogate_refresh(ssasdpd_engine *e, int32_t iters, float lr, int32_t keep_samples,
                                  int32_t max_particles) {
    if (iters < 0 || (iters > 0 && (!(lr > 0.f) || keep_samples < 1 || max_particles < 1)))
        return set_error(SSASDPD_ERR_BAD_DESC, "surrogate refresh: iters >= 0, lr > 0, keep and max_particles >= 1");
    e->nn_refresh_iters = iters;
    e->nn_refresh_lr = lr;
    e->nn_refresh_keep = keep_samples;
    if (iters > 0) e->nn_rec_max = max_particles;
    return 0;
}

int ssasdpd_set_surrogate_rearm(ssasdpd_engine *e, int32_t on) {
    e->nn_rearm = on != 0;
    return 0;
}

int ssasdpd_surrogate_stats(ssasdpd_engine *e, uint64_t *nn_steps, uint64_t *checks, double *last_err,
                            double *max_err, uint64_t *refreshes, uint64_t *failed_checks, int32_t *active) {
    if (failed_checks) *failed_checks = e->nn_fails;
    if (nn_steps) *nn_steps = e->nn_steps;
    if (checks) *checks = e->nn_checks;
    if (last_err) *last_err = e->nn_last_err;
    if (max_err) *max_err = e->nn_max_err;
    if (refreshes) *refreshes = e->nn_refreshes;
    if (active) *active = e->nn_active ? 1 : 0;
    return 0;
}

int ssasdpd_set_surrogate_cap(ssasdpd_engine *e, double factor) {
    e->nn_cap_factor = factor;
    if (factor <= 0.0) e->nn_acc_cap = 0.0;
    return 0;
}

int ssasdpd_surrogate_capped(ssasdpd_engine *e, uint64_t *capped, double *cap) {
    GUARD_BEGIN
    uint64_t c = 0;
    if (e->nn_acc) {
        d2h(e->st, &c, (const uint64_t *)(e->nn_acc + 3), 1);
        HIP_OK(hipStreamSynchronize(e->st));
    }
    if (capped) *capped = c;
    if (cap) *cap = e->nn_acc_cap;
    return 0;
    GUARD_END
}

int ssasdpd_stats_get(ssasdpd_engine *e, ssasdpd_stats *s) {
    *s = e->stats;
    return 0;
}

int ssasdpd_phase_times(ssasdpd_engine *e, double *ms, uint64_t *launches, int32_t n_phases) {
    for (int p = 0; p < n_phases && p < ssasdpd_engine::NPH; p++) {
        if (ms) ms[p] = e->ph_ms[p];
        if (launches) launches[p] = e->ph_launch[p];
    }
    return ssasdpd_engine::NPH;
}

int ssasdpd_get_state(ssasdpd_engine *e, ssasdpd_state *s) {
    GUARD_BEGIN
    const int n = e->n;
    hipStream_t st = e->st;
    std::vector<double> tmp(3 * (size_t)n);
    // every array in the caller's numbering
    auto v3 = [&](double *dst, double *src) {
        if (!dst) return;
        d2h(st, tmp.data(), e->orig3(src), 3 * (size_t)n);
        HIP_OK(hipStreamSynchronize(st));
        from_soa3(tmp, dst, n);
    };
    v3(s->x, e->x); v3(s->v, e->v); v3(s->vt, e->vt); v3(s->F, e->F); v3(s->Fbp, e->Fbp);
    auto v1 = [&](double *dst, double *src, int cols) {
        if (dst && src) {
            d2h(st, dst, e->orig(src, cols), (size_t)n * cols);
            HIP_OK(hipStreamSynchronize(st));
        }
    };
    v1(s->rho, e->rho, 1); v1(s->Frho, e->Frho, 1); v1(s->bvf_phi, e->bvf, 1); v1(s->nu, e->nu, 1);
    v1(s->old_rho, e->old_rho, 1);
    if (e->Sc) { v1(s->C, e->C, e->Sc); v1(s->Q, e->Q, e->Sc); }
    if (e->S) {
        if (s->xx) d2h(st, s->xx, e->orig(e->xx, e->S), (size_t)n * e->S);
        HIP_OK(hipStreamSynchronize(st));
        v1(s->srrate, e->srrate, 1); v1(s->sdrate, e->sdrate, 1);
        if (e->R) v1(s->rrate, e->rrate, e->R);
        v1(s->Ddiag, e->Ddiag, e->S);
        v1(s->next_time, e->qtime, 1);
    }
    HIP_OK(hipStreamSynchronize(st));
    return 0;
    GUARD_END
}

int ssasdpd_get_neighbors(ssasdpd_engine *e, int64_t *row_ptr, int32_t *idx, double *dist, double *dWdr, double *Dij) {
    GUARD_BEGIN
    const int n = e->n, W = e->W;
    std::vector<int> deg(n);
    d2h(e->st, deg.data(), e->orig(e->deg, 1), n);
    HIP_OK(hipStreamSynchronize(e->st));
    int64_t off = 0;
    for (int i = 0; i < n; i++) { row_ptr[i] = off; off += deg[i]; }
    row_ptr[n] = off;
    if (!idx) return 0;
    std::vector<int> ni((size_t)W * n);
    std::vector<double> nr((size_t)W * n);
    std::vector<EdgeRec> er((size_t)W * n);
    d2h(e->st, ni.data(), e->nidx, (size_t)W * n);
    d2h(e->st, nr.data(), e->nr, (size_t)W * n);
    d2h(e->st, er.data(), e->erow, (size_t)W * n);
    HIP_OK(hipStreamSynchronize(e->st));
    const double h = e->h, a = e->alpha;
    // rows in the caller's numbering: caller's particle i is internal p
    std::vector<int> inv;
    if (e->renumbered) {
        inv.resize(n);
        const std::vector<int> &ph = e->perm_host();
        for (int p = 0; p < n; p++) inv[ph[p]] = p;
    }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < deg[i]; k++) {
            const int p = e->renumbered ? inv[i] : i;
            size_t s = (size_t)k * n + p;
            int64_t o = row_ptr[i] + k;
            idx[o] = e->orig_index(ni[s]);
            double r = nr[s];
            if (dist) dist[o] = r;
            if (dWdr) {
                double R = r / h;
                dWdr[o] = a * (-12 * r / (h * h)) * ((1 - R) * (1 - R));   // particle.cpp:178
            }
            if (Dij) Dij[o] = er[(size_t)p * W + k].D;
        }
    return 0;
    GUARD_END
}

int ssasdpd_grid_shape(const double *bbox, int32_t n, int32_t dim, double h, int32_t *dims, int64_t *ncell,
                       double *cell) {
    if (!bbox || n < 1 || dim < 1 || dim > 3 || !(h > 0.0)) return set_error(SSASDPD_ERR_BAD_DESC, "grid_shape: bad arguments");
    Grid g;
    grid_setup(bbox, n, dim, h, g);
    for (int d = 0; d < 3 && dims; d++) dims[d] = g.dims[d];
    if (ncell) *ncell = g.ncell;
    if (cell) *cell = 1.0 / g.inv_cell;
    return 0;
}

int ssasdpd_fr_search(const double *pts, int32_t n, const double *q, int32_t m, int32_t dim, double h,
                      int64_t *row_ptr, int32_t *idx, double *r2) {
    GUARD_BEGIN
    if (n <= 0 || m < 0 || dim < 1 || dim > 3 || !(h > 0)) return set_error(SSASDPD_ERR_BAD_DESC, "bad fr_search args");
    hipStream_t st;
    HIP_OK(hipStreamCreate(&st));
    std::vector<double> hp = to_soa3(pts, n), hq = to_soa3(q, m);
    double *dp = dalloc<double>(3 * (size_t)n), *dq = dalloc<double>(3 * (size_t)std::max(m, 1));
    double *sp = dalloc<double>(4 * (size_t)n), *bb = dalloc<double>(8);
    unsigned long long *bbt = dalloc<unsigned long long>(8);
    int *srt = dalloc<int>(n), *cof = dalloc<int>(n);
    h2d(st, dp, hp.data(), 3 * (size_t)n);
    h2d(st, dq, hq.data(), 3 * (size_t)m);
    compute_bbox(st, dp, n, dim, bb, bbt);
    double hb[6];
    d2h(st, hb, bb, 6);
    HIP_OK(hipStreamSynchronize(st));
    Grid g;
    grid_setup(hb, n, dim, h, g);
    int *cc = dalloc<int>(g.ncell + 1), *cs = dalloc<int>(g.ncell + 2), *cf = dalloc<int>(g.ncell + 1);
    void *stmp;
    HIP_OK(hipMalloc(&stmp, std::max(scan_tmp_bytes(g.ncell + 8), scan_tmp_bytes(n + 8))));
    bin_points(st, dp, n, dim, g, cof, cc, cs, cf, srt, sp, stmp);
    KdDev kd;
    kd.parent = dalloc<int>(2 * (size_t)n + 2); kd.depth = dalloc<int>(2 * (size_t)n + 2);
    kd.cd = dalloc<int>(2 * (size_t)n + 2); kd.cv = dalloc<double>(2 * (size_t)n + 2);
    kd.is_hi = dalloc<int>(2 * (size_t)n + 2); kd.leaf_of = dalloc<int>(n);
    size_t kws = kd_work_bytes(n);
    void *kw;
    HIP_OK(hipMalloc(&kw, kws));
    kd_build(st, dp, n, dim, kd, kw, kws);
    DevErr *de = dalloc<DevErr>(1);
    int W = 16;
    int *oi = nullptr, *dg = dalloc<int>(std::max(m, 1));
    double *orr = nullptr;
    for (int attempt = 0; attempt < 6; attempt++) {
        dfree(oi); dfree(orr);
        oi = dalloc<int>((size_t)W * std::max(m, 1));
        orr = dalloc<double>((size_t)W * std::max(m, 1));
        DevErr e0{0, INT32_MAX, 0, 0, 0};
        h2d(st, de, &e0, 1);
        if (m > 0) fr_search(st, dq, m, sp, srt, cs, n, dim, h, g, &kd, W, oi, orr, dg, de, nullptr);
        DevErr e;
        d2h(st, &e, de, 1);
        HIP_OK(hipStreamSynchronize(st));
        if (!e.overflow) break;
        W = ((e.max_deg + 7) / 8) * 8 + 8;
    }
    std::vector<int> hd(std::max(m, 1));
    d2h(st, hd.data(), dg, m);
    HIP_OK(hipStreamSynchronize(st));
    int64_t off = 0;
    for (int i = 0; i < m; i++) { row_ptr[i] = off; off += hd[i]; }
    row_ptr[m] = off;
    if (idx) {
        std::vector<int> hi((size_t)W * m);
        std::vector<double> hr((size_t)W * m);
        d2h(st, hi.data(), oi, (size_t)W * m);
        d2h(st, hr.data(), orr, (size_t)W * m);
        HIP_OK(hipStreamSynchronize(st));
        for (int i = 0; i < m; i++)
            for (int k = 0; k < hd[i]; k++) {
                idx[row_ptr[i] + k] = hi[(size_t)k * m + i];
                r2[row_ptr[i] + k] = hr[(size_t)k * m + i];
            }
    }
    dfree(dp); dfree(dq); dfree(sp); dfree(bb); dfree(bbt); dfree(srt); dfree(cof); dfree(cc); dfree(cs); dfree(cf);
    (void)hipFree(stmp); (void)hipFree(kw);
    dfree(kd.parent); dfree(kd.depth); dfree(kd.cd); dfree(kd.cv); dfree(kd.is_hi); dfree(kd.leaf_of);
    dfree(de); dfree(oi); dfree(orr); dfree(dg);
    (void)hipStreamDestroy(st);
    return 0;
    GUARD_END
}

int ssasdpd_set_timeout(ssasdpd_engine *e, double seconds) {
    GUARD_BEGIN
    if (!(seconds >= 0) || !std::isfinite(seconds))
        return set_error(SSASDPD_ERR_BAD_DESC, "set_timeout: seconds must be finite and >= 0");
    e->timeout_s = seconds;
    return 0;
    GUARD_END
}

int ssasdpd_set_trace_capacity(ssasdpd_engine *e, uint64_t cap) {
    GUARD_BEGIN
    dfree(e->d_trace);
    e->trace_cap = cap;
    if (cap) e->d_trace = dalloc<ssasdpd_event>(cap);
    return 0;
    GUARD_END
}

int ssasdpd_get_trace(ssasdpd_engine *e, ssasdpd_event *ev, uint64_t cap, uint64_t *n_events) {
    GUARD_BEGIN
    QueueState q{};
    if (e->qs) {
        d2h(e->st, &q, e->qs, 1);
        HIP_OK(hipStreamSynchronize(e->st));
    }
    *n_events = q.events_traced;
    uint64_t m = std::min<uint64_t>(std::min<uint64_t>(cap, e->trace_cap), q.events_traced);
    if (ev && m) {
        d2h(e->st, ev, e->d_trace, m);
        HIP_OK(hipStreamSynchronize(e->st));
    }
    return 0;
    GUARD_END
}

// std::mt19937_64 (Matsumoto & Nishimura 2004) and libstdc++'s
// exponential_distribution of each word, on the host.
int ssasdpd_mt19937_64_stream(uint64_t seed, uint64_t n, uint64_t *words, double *expo) {
    uint64_t mt[312];
    mt[0] = seed;
    for (int i = 1; i < 312; i++) mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
    int mti = 312;
    for (uint64_t k = 0; k < n; k++) {
        if (mti >= 312) {
            for (int i = 0; i < 312; i++) {
                uint64_t y = (mt[i] & 0xFFFFFFFF80000000ULL) | (mt[(i + 1) % 312] & 0x7FFFFFFFULL);
                uint64_t z = mt[(i + 156) % 312] ^ (y >> 1);
                if (y & 1ULL) z ^= 0xB5026F5AA96619E9ULL;
                mt[i] = z;
            }
            mti = 0;
        }
        uint64_t y = mt[mti++];
        y ^= (y >> 29) & 0x5555555555555555ULL;
        y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
        y ^= (y << 37) & 0xFFF7EEE000000000ULL;
        y ^= (y >> 43);
        if (words) words[k] = y;
        if (expo) {
            double c = (double)y * 1.0 / 18446744073709551616.0;
            if (c >= 1.0) c = std::nextafter(1.0, 0.0);
            expo[k] = -std::log(1.0 - c) / 1.0;
        }
    }
    return 0;
}

}  // extern "C"
