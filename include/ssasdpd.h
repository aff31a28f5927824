/*
 * ssasdpd.h - C ABI of the MI355X (gfx950) sSSA-SDPD engine.
 *
 * The reference engine is reached through a generated C++ translation unit
 * and a subprocess (spatialpy/solvers/solver.py:100-158, 443-507, 553-569;
 * spatialpy/solvers/c_base/ssa_sdpd-c-simulation-engine/propensity_file_template.cpp:101-142).
 * This library replaces that generated program: the model is lowered to plain
 * arrays (the same arrays the generator bakes into the program text,
 * solver.py:208-299, 312-331, 375-419) and handed over through the
 * descriptor below.  All pointers are caller-owned host memory, C-contiguous;
 * the engine copies what it needs to the GPU in ssasdpd_create.
 *
 * Conventions: every function returns 0 (SSASDPD_OK) or a status code; the
 * message for the last failure on the calling thread is ssasdpd_last_error().
 * An engine handle is used by one host thread at a time.
 */
#ifndef SSASDPD_H
#define SSASDPD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSASDPD_ABI_VERSION 3

/* Status codes.  The reference prints and calls exit(1) for each failure
 * class (particle.cpp:88-126, 189-206; simulate_rdme.cpp:262-294, 322-349,
 * 368-400); here each one is a return code, raised as SimulationError by
 * the Python Solver (solver.py:595-597). */
enum ssasdpd_status {
    SSASDPD_OK = 0,
    SSASDPD_ERR_BAD_DESC = 1,            /* descriptor fails validation       */
    SSASDPD_ERR_NAN_STATE = 2,           /* particle.cpp:88-126               */
    SSASDPD_ERR_NEGATIVE_STATE = 3,      /* simulate_rdme.cpp:288-294,395-400 */
    SSASDPD_ERR_PROPENSITY_OVERFLOW = 4, /* simulate_rdme.cpp:277-280         */
    SSASDPD_ERR_DIFFUSION_OVERFLOW = 5,  /* simulate_rdme.cpp:342-346,381-388 */
    SSASDPD_ERR_NAN_DIJ = 6,             /* particle.cpp:189-206              */
    SSASDPD_ERR_HIP = 7,                 /* HIP runtime / no device           */
    SSASDPD_ERR_TIMEOUT = 8,             /* solver.py:583-586                 */
    SSASDPD_ERR_REPLAY_EXHAUSTED = 9,    /* replay stream shorter than needed */
    SSASDPD_ERR_CAPACITY = 10,           /* a fixed-size buffer is too small  */
    SSASDPD_ERR_UNSUPPORTED = 11
};

/* Mass-action propensity kinds, evaluated left to right in FP64 exactly as
 * the generated C expressions (spatialpy/core/reaction.py:260-313):
 *   0: (k*vol)   1: (k*X)   2: (((k*X)*Y)/vol)   3: (((k*X)*(X-1))/vol)
 * X, Y are the species counts read as signed int (the generated functions
 * take `const int *x`, solver.py:345).  The deterministic (ODE) right-hand
 * side of the same reaction drops `vol` and the -1:  0: k  1: k*X
 * 2: ((k*X)*Y)  3: ((k*X)*X)  (reaction.py:297-311, solver.py:160-190).
 * Kind 4 is any other propensity (custom propensity functions, reaction.py:
 * 120-135, 325-330): a = start of its propensity program in the
 * descriptor's `code`, b = start of its deterministic-rate program, k is
 * unused.  Programs are the expression VM's bytecode (spatialpy_amd/
 * expression.py compiles the C++ text solver.py:160-190, 344-373
 * generates), evaluated with the arguments of the generated functions:
 * counts (as int) or concentrations, t, vol, data_fn[], and type (for
 * type_mask, the generated `if (sd == type_i || ...)`). */
typedef struct ssasdpd_reaction {
    int32_t kind;
    int32_t a;            /* first reactant species index (kinds 1-3) */
    int32_t b;            /* second reactant species index (kind 2)   */
    int32_t pad;
    double k;             /* rate constant                            */
    uint64_t type_mask;   /* restrict_to: bit t set => allowed in particle type t;
                             0 => no restriction (solver.py:353-368)  */
} ssasdpd_reaction;

/* Boundary-condition rule (spatialpy/core/boundarycondition.py:121-167),
 * applied in array order, twice per step (simulate.cpp:88, 171).  A bound
 * that the model leaves unset is passed as -inf / +inf. */
enum ssasdpd_bc_target { SSASDPD_BC_CHEM = 0, SSASDPD_BC_V = 1, SSASDPD_BC_NU = 2, SSASDPD_BC_RHO = 3 };
typedef struct ssasdpd_bc {
    double lo[3];
    double hi[3];
    int32_t type;         /* particle type, or -1 for any              */
    int32_t target;       /* enum ssasdpd_bc_target                    */
    int32_t species;      /* deterministic species index (target CHEM) */
    int32_t pad;
    double value[3];      /* value[0] for scalar targets                */
} ssasdpd_bc;

/* Engine behaviour flags. */
enum ssasdpd_flags {
    /* Select reaction/species with rand1*totrate instead of the reference's
     * rand1*srrate / rand1*sdrate (simulate_rdme.cpp:255-261, 317-321). */
    SSASDPD_FLAG_TEXTBOOK_SELECTION = 1u << 0,
    /* Order neighbour-list ties by ANN kd-tree visit order on moving domains
     * too (always on for static domains). */
    SSASDPD_FLAG_ANN_ORDER_MOVING = 1u << 1,
    /* Skip the sSSA part (diffusion + reactions). */
    SSASDPD_FLAG_NO_RDME = 1u << 2,
    /* Record the event trace (see ssasdpd_get_trace). */
    SSASDPD_FLAG_TRACE = 1u << 3,
    /* Pair-force terms of pairwiseForce (model.cpp:39-191) in FP32 from
     * FP64-formed dx, dv, summed in FP64; integration stays FP64.  Off: the
     * reference's FP64 arithmetic, operation for operation. */
    SSASDPD_FLAG_FP32_FORCES = 1u << 4,
    /* Keep particles in the caller's numbering in device memory.  By default
     * a moving domain in the seeded mode (not replay, not with
     * SSASDPD_FLAG_ANN_ORDER_MOVING or an event trace) renumbers its
     * particles into cell order every 20 steps, so that neighbours in space
     * are neighbours in memory for the search and every gather; outputs,
     * state and lists are always reported in the caller's numbering.  The
     * order inside a list among exactly equal distances follows the internal
     * numbering, so seeded runs with and without renumbering agree in
     * distribution, not bit for bit.  A seed still reproduces its run, and
     * a batch trajectory its single run: the new order is cell order with
     * ties inside a cell taken by the previous index. */
    SSASDPD_FLAG_NO_RENUMBER = 1u << 5
};

typedef struct ssasdpd_desc {
    int32_t abi_version;      /* SSASDPD_ABI_VERSION                           */
    int32_t n;                /* particles == voxels                           */
    int32_t dimension;        /* 1, 2 or 3 (solver.py:407-413)                 */
    int32_t n_types;          /* types 1..n_types (type 0 excluded)            */
    int32_t n_species;        /* stochastic species S (0: RDME off)            */
    int32_t n_reactions;      /* stochastic reactions R                        */
    int32_t n_chem_species;   /* deterministic species (0: PDE part off)       */
    int32_t n_chem_reactions; /* deterministic reactions                       */
    int32_t static_domain;    /* model.staticDomain                            */
    int32_t n_bcs;
    int32_t n_output_steps;
    uint32_t flags;           /* enum ssasdpd_flags                            */
    uint32_t nt;              /* number of steps                               */
    uint32_t pad0;
    double dt, h, rho0, c0, P0;
    double bbox[6];           /* xlo xhi ylo yhi zlo zhi                       */
    double gravity[3];

    /* particles (propensity_file_template.cpp:73-83) */
    const double *x;          /* [n][3]                                         */
    const double *v;          /* [n][3] or NULL (zero)                          */
    const int32_t *type;      /* [n], 1..n_types                                */
    const double *nu;         /* [n]                                            */
    const double *mass;       /* [n]                                            */
    const double *rho;        /* [n]                                            */
    const int32_t *fixed;     /* [n], solidTag                                  */
    const uint32_t *u0;       /* [n][n_species] (solver.py:211-219)             */
    const double *C0;         /* [n][n_chem_species] or NULL (= u0)             */

    /* chemistry (solver.py:208-286; core/model.py:187-268) */
    const double *diffusion_matrix;   /* [n_species][n_types] species-major     */
    const ssasdpd_reaction *reactions;/* [n_reactions]                          */
    const int32_t *jcN;       /* [R+1] stoichiometry CSC column pointers        */
    const int32_t *irN;       /* [nnz] species rows, ascending per column       */
    const int32_t *prN;       /* [nnz] net change                               */
    const int32_t *jcG;       /* [S+R+1] dependency-graph CSC column pointers   */
    const int32_t *irG;       /* [nnzG] reaction rows, ascending per column     */
    const int32_t *N_dense;   /* [S][R] row-major (chem RHS, model.cpp:183-189) */

    const ssasdpd_bc *bcs;    /* [n_bcs]                                        */
    const uint32_t *output_steps; /* [n_output_steps] (core/timespan.py:72-85)  */

    /* expression-VM programs of kind-4 reactions (may be empty)            */
    int32_t n_code, n_consts;
    const uint32_t *code;     /* [n_code] instruction words                     */
    const double *consts;     /* [n_consts] constant pool                        */
    /* data functions (solver.py:192-198, 239-249): value of function k at
     * particle i's initial position, read by programs as data_fn[k]          */
    int32_t n_data_fn, pad1;
    const double *data_fn;    /* [n][n_data_fn]                                 */
} ssasdpd_desc;

/* Output snapshot buffers, caller-allocated with ssasdpd_num_snapshots()
 * entries.  Any pointer may be NULL to skip that field. */
typedef struct ssasdpd_outputs {
    int32_t n_snapshots;      /* capacity in snapshots                          */
    int32_t pad;
    uint32_t *xx;             /* [n_snap][n][S]                                 */
    double *C;                /* [n_snap][n][Sc]                                */
    double *x;                /* [n_snap][n][3]                                 */
    double *v;                /* [n_snap][n][3]                                 */
    double *rho;              /* [n_snap][n]                                    */
    double *bvf_phi;          /* [n_snap][n]                                    */
    double *nu;               /* [n_snap][n]                                    */
    uint32_t *steps;          /* [n_snap] step index each snapshot was taken at */
} ssasdpd_outputs;

/* Random stream for a run.  kind 0: seeded (GPU counter-based generator,
 * statistically exact, not the reference's bit stream).  kind 1: replay --
 * the engine consumes words[] in exactly the reference's draw order
 * (simulate_rdme.cpp:253, 353; NRMConstant_v5.cpp:57, 98): a uniform draw
 * uses words[i] as double(words[i])/2^64, an exponential draw uses expo[i]
 * = -log(1 - canonical(words[i])) as computed by the caller's libm.
 * ssasdpd_mt19937_64_stream() fills both arrays from a seed. */
typedef struct ssasdpd_rng {
    int32_t kind;
    int32_t pad;
    uint64_t seed;
    const uint64_t *words;
    const double *expo;
    uint64_t n_words;
} ssasdpd_rng;

typedef struct ssasdpd_stats {
    uint64_t steps;
    uint64_t particle_steps;
    uint64_t reactions;       /* total_reactions (simulate_rdme.cpp:310)        */
    uint64_t diffusions;      /* total_diffusion (simulate_rdme.cpp:445)        */
    uint64_t rng_words_used;  /* replay words consumed                          */
    double ms_neighbors, ms_sdpd, ms_rates, ms_ssa, ms_total;
    int32_t bad_particle;     /* NaN particle id, or -1                         */
    int32_t max_degree;       /* widest neighbour list seen                     */
} ssasdpd_stats;

typedef struct ssasdpd_engine ssasdpd_engine;

/* ---- lifecycle (replaces main(), propensity_file_template.cpp:101-142) -- */
int ssasdpd_create(const ssasdpd_desc *d, ssasdpd_engine **out);
void ssasdpd_destroy(ssasdpd_engine *e);
const char *ssasdpd_last_error(void);
int ssasdpd_abi_version(void);

/* Number of snapshots a full run writes, following the output schedule of
 * run_simulation (simulate_threads.cpp:231-289). */
int ssasdpd_num_snapshots(const ssasdpd_desc *d, int32_t *n_snap);

/* Reset the state to the descriptor's initial state and set the random
 * stream (main(): rng seeding + init_all_particles + initialize_rdme,
 * propensity_file_template.cpp:127-138, simulate_rdme.cpp:40-48). */
int ssasdpd_reset(ssasdpd_engine *e, const ssasdpd_rng *rng);

/* Advance n_steps steps (the body of the loop at simulate_threads.cpp:232-281
 * without the output branch). */
int ssasdpd_advance(ssasdpd_engine *e, uint32_t n_steps);

/* Full trajectory: reset + the step loop with snapshots (run_simulation,
 * simulate_threads.cpp:171-312). */
int ssasdpd_run(ssasdpd_engine *e, const ssasdpd_rng *rng, ssasdpd_outputs *out);

/* Many independent trajectories of the same model (the reference runs them
 * one after another, solver.py:547-605, each a separate process).  The
 * particle state does not depend on the molecule counts, so it is computed
 * once per step and shared; every trajectory has its own counts, rates,
 * event queue and seeded GPU stream, and trajectory b is bit-identical to
 * ssasdpd_run with rng {kind 0, seed seeds[b]}.  The event loops of all
 * trajectories run in one launch per step, one workgroup each.
 *   shared: snapshots of the particle fields (its xx is ignored), may be NULL;
 *   counts: [n_traj][K][n][S] with K from ssasdpd_num_snapshots, may be NULL. */
int ssasdpd_run_batch(ssasdpd_engine *e, int32_t n_traj, const uint64_t *seeds,
                      ssasdpd_outputs *shared, uint32_t *counts);

/* Wall-clock limit, in seconds, of each later ssasdpd_run, ssasdpd_run_batch
 * and ssasdpd_advance call (0: none, the default).  The reference waits on
 * the trajectory's process with Popen.wait(timeout), kills it when the limit
 * passes and marks that result timed out (solver.py:574-586).  Here the call
 * returns SSASDPD_ERR_TIMEOUT at the first step boundary past the limit: the
 * snapshots taken before it are filled in, and the engine holds the state
 * after the last completed step.  Static domains, whose remaining steps
 * otherwise run in one launch, then run in launches of at most 64 steps
 * (ending at snapshot steps), with the limit checked between them. */
int ssasdpd_set_timeout(ssasdpd_engine *e, double seconds);

/* Evaluate the pair-force kernel (pairwiseForce, model.cpp:39-191) once on
 * the engine's current state and neighbour lists without advancing:
 * F starts at gravity, Fbp and Frho at zero.  fp32 selects the FP32 pair
 * terms.  Outputs: F, Fbp [n][3], Frho [n] (any may be NULL).  Needs lists
 * from at least one step; static domains return gravity only (the reference
 * computes no momentum there that any later step reads). */
int ssasdpd_eval_forces(ssasdpd_engine *e, int32_t fp32, double *F, double *Fbp, double *Frho);

/* Evaluate one expression-VM program on the GPU for m argument sets (the
 * device interpreter the engine uses): x [m][n_x] (species as int when
 * x_is_double is 0 -- values must be integral -- else concentrations), vol,
 * t [m], df [m][n_df] (may be NULL when n_df is 0), out [m].  For testing
 * compiled propensities against the generated C++. */
int ssasdpd_vm_eval(const uint32_t *code, int32_t n_code, const double *consts, int32_t n_consts, int32_t start,
                    int32_t x_is_double, int32_t n_x, int32_t n_df, int32_t m, const double *x, const double *vol,
                    const double *t, const double *df, double *out);

/* ---- Ensembles of different models ------------------------------------
 * n_models static-domain models (uneven domain sizes, each its own
 * chemistry), one trajectory each, with every event loop in ONE persistent
 * launch that runs all nt steps (one workgroup per model, largest domains
 * first).  The reference runs such a set as separate processes one after
 * another (solver.py:547-605).  Model m's trajectory is bit-identical to
 * ssasdpd_run of that model with rng {kind 0, seed seeds[m]}.  Models must
 * share nt, dt and output steps.  Deterministic (PDE) species do not enter
 * the event loop; an ensemble leaves them out and reports counts only. */
typedef struct ssasdpd_ensemble ssasdpd_ensemble;
int ssasdpd_ensemble_create(const ssasdpd_desc *const *descs, int32_t n_models, ssasdpd_ensemble **out);
void ssasdpd_ensemble_destroy(ssasdpd_ensemble *m);
/* Snapshot count K (the reference's output schedule). */
int ssasdpd_ensemble_num_snapshots(ssasdpd_ensemble *m, int32_t *k);
/* Event-tree width (voxels per lane under a level-1 node: 1, 2, 4, 8, 16 or
 * 32) used by every model; results do not depend on it. */
int ssasdpd_ensemble_set_tree_width(ssasdpd_ensemble *m, int32_t pl);
/* Autotuner: times probe_steps steps of the ensemble (seeds as for run) at
 * widths 1, 2, 4, 8 and keeps the fastest.  ms_per_width[4] (may be NULL)
 * receives each probe's GPU time, -1 for widths not valid for every model. */
int ssasdpd_ensemble_autotune(ssasdpd_ensemble *m, const uint64_t *seeds, int32_t probe_steps, int32_t *chosen,
                              double *ms_per_width);
/* Run all models: totals [n_models][K][S_max] species totals per snapshot
 * (species past a model's S read 0), counts[m] (each may be NULL, the
 * array itself may be NULL) [K][n_m][S_m]. */
int ssasdpd_ensemble_run(ssasdpd_ensemble *m, const uint64_t *seeds, uint32_t *totals, uint32_t *const *counts);
/* Events and times of the last run (ms_ssa: the persistent launch). */
int ssasdpd_ensemble_stats(ssasdpd_ensemble *m, ssasdpd_stats *s);

/* ---- Learned surrogate for the SDPD accelerations ------------------------
 * No counterpart in the reference.  An edge network over the neighbour
 * lists predicts each moving particle's pair acceleration F - g (the F sum
 * of the pair loop): per listed pair (i, j) a 3-layer MLP (16 rotation-
 * invariant features -> 64 -> 64 -> 4, hard-SiLU, bf16 weights, FP32
 * accumulation on the matrix cores) gives the coefficients of dx/h,
 * dv/v_scale, v_i/v_scale and v_j/v_scale, times the envelope
 * (1-R)^2 R/(R+0.001), R = r/h; the prediction is their sum over the
 * neighbours, in units of acc_scale.  Training runs on the GPU (forward,
 * backward, Adam); the engine uses the network between periodic checks
 * against the plain SDPD and falls back to it for the rest of the run when
 * the relative error exceeds the budget.  Reactions and diffusion stay
 * exact SSA. */
typedef struct ssasdpd_mpn ssasdpd_mpn;
/* A particle state and its neighbour lists, host arrays: x, v, vt [n][3];
 * rho, mass, nu [n]; solid [n] (nonzero: fixed); ELL lists deg [n],
 * nidx [W][n] and nr [W][n] (pair distances), cut-off h. */
typedef struct {
    int32_t n, W, dim;
    double h;
    const double *x, *v, *vt, *rho, *mass, *nu;
    const int32_t *solid, *deg, *nidx;
    const double *nr;
} ssasdpd_mpn_graph;
int32_t ssasdpd_mpn_num_params(void);
int ssasdpd_mpn_create(uint64_t seed, ssasdpd_mpn **out);
void ssasdpd_mpn_destroy(ssasdpd_mpn *m);
int ssasdpd_mpn_get_params(ssasdpd_mpn *m, float *params);
int ssasdpd_mpn_set_params(ssasdpd_mpn *m, const float *params);
/* scales[6] = {rho0, v_scale, rho_scale, nu_scale, acc_scale, tv_scale}
 * (tv_scale: |v - vt|); all positive.  Set before adding samples. */
int ssasdpd_mpn_set_scales(ssasdpd_mpn *m, const float *scales);
/* A training sample from a host graph: the n_sel particles sel[] (NULL: all
 * moving particles) with targets acc [n][3] (pair accelerations F - g). */
int ssasdpd_mpn_add_sample(ssasdpd_mpn *m, const ssasdpd_mpn_graph *g, const double *acc, int32_t n_sel,
                           const int32_t *sel, int32_t *index);
/* A training sample from an engine's current state: up to max_particles
 * moving particles (a seeded uniform subset), the plain pair kernel's
 * accelerations on that state as the target. */
int ssasdpd_mpn_add_engine_sample(ssasdpd_mpn *m, ssasdpd_engine *e, int32_t max_particles, uint64_t seed,
                                  int32_t *index);
int ssasdpd_mpn_clear_samples(ssasdpd_mpn *m);
/* Record a sample (up to max_particles particles) every `every` plain SDPD
 * steps of the engine (m NULL: stop): the state the pair kernel sees inside
 * the step (after the predictor) and its accelerations.  Scales first. */
int ssasdpd_mpn_record(ssasdpd_mpn *m, ssasdpd_engine *e, int32_t every, int32_t max_particles);
int32_t ssasdpd_mpn_num_samples(ssasdpd_mpn *m);
int ssasdpd_mpn_sample_size(ssasdpd_mpn *m, int32_t sample, int32_t *n_particles, int32_t *n_edges);
/* Training-path predictions out [n_particles][3] and the sample's targets
 * [n_particles][3] (either may be NULL), pair accelerations. */
int ssasdpd_mpn_forward(ssasdpd_mpn *m, int32_t sample, float *out, float *target);
/* The fused inference kernel on a host graph: acc [n][3] pair accelerations
 * (0 for solid particles and for components past dim). */
int ssasdpd_mpn_predict(ssasdpd_mpn *m, const ssasdpd_mpn_graph *g, double *acc);
/* GPU time of one fused inference pass over an engine's current state
 * (HIP events over reps passes on the engine's stream). */
int ssasdpd_mpn_time_engine(ssasdpd_mpn *m, ssasdpd_engine *e, int32_t reps, double *ms_per_pass);
/* Loss (1/n_particles) sum |pred - target|^2 in units of acc_scale^2 of a
 * stored sample and its gradient [num_params]. */
int ssasdpd_mpn_grad(ssasdpd_mpn *m, int32_t sample, float *grad, double *loss);
/* Adam, one step per stored sample per epoch (shuffled); mean loss of the
 * last epoch. */
int ssasdpd_mpn_train(ssasdpd_mpn *m, int32_t epochs, float lr, uint64_t shuffle_seed, double *mean_loss);
/* Attach (m != NULL) or detach the surrogate: on moving domains every step
 * whose index is not a multiple of check_every takes F from the network;
 * the others run the plain kernel and compare (relative L2 error of F - g
 * over moving particles), and an error above max_rel_err switches the
 * network off until the next reset or attach.  Counters restart at 0. */
int ssasdpd_set_surrogate(ssasdpd_engine *e, ssasdpd_mpn *m, int32_t check_every, double max_rel_err);
/* Online refresh while the surrogate is attached (iters 0: off, the
 * default): at every check that passes, the plain step's state becomes a
 * training sample of up to max_particles moving particles, samples beyond
 * the newest keep_samples are dropped, and the network takes `iters` Adam
 * steps (learning rate lr), alternately on the new sample and on one drawn
 * from all stored samples, so it follows the run's own states without
 * forgetting the recorded ones. */
int ssasdpd_set_surrogate_refresh(ssasdpd_engine *e, int32_t iters, float lr, int32_t keep_samples,
                                  int32_t max_particles);
/* Re-arm (on != 0; off by default): a failed check switches the network
 * off only until the next check, which still runs and switches it back on
 * if it passes, so every interval the network drives starts from a passing
 * check.  Off: the first failure ends its use for the run. */
int ssasdpd_set_surrogate_rearm(ssasdpd_engine *e, int32_t on);
/* Since the last reset or attach: steps that took F from the network,
 * checks, the last check's error, the largest error of a passing check,
 * online refreshes, failed checks, and whether the network is on. */
int ssasdpd_surrogate_stats(ssasdpd_engine *e, uint64_t *nn_steps, uint64_t *checks, double *last_err,
                            double *max_err, uint64_t *refreshes, uint64_t *failed_checks, int32_t *active);
/* No counterpart in the reference.  Between checks a particle's network
 * pair acceleration is capped at factor times the RMS plain pair
 * acceleration over the moving particles of the last check (default 20; no
 * cap before the first check, or with factor <= 0), so that a particle in a
 * state the network has not seen cannot run away before the next check. */
int ssasdpd_set_surrogate_cap(ssasdpd_engine *e, double factor);
/* Since the last reset or attach: how many predictions were capped, and the
 * current cap (0: none). */
int ssasdpd_surrogate_capped(ssasdpd_engine *e, uint64_t *capped, double *cap);

int ssasdpd_stats_get(ssasdpd_engine *e, ssasdpd_stats *s);

/* GPU time per phase since the last reset, from HIP events on the engine's
 * stream: 0 binning + kd-tree, 1 fixed-radius search + D_ij, 2 SDPD
 * substeps 0 and 2, 3 rates + event-queue build, 4 SSA event loop, 5 the
 * pairwise-force kernel.  Returns the number of phases.  (No reference
 * counterpart: the reference has only gprof hooks, solver.py:421-430.) */
int ssasdpd_phase_times(ssasdpd_engine *e, double *ms, uint64_t *launches, int32_t n_phases);

/* ---- state access (parity tests) ---------------------------------------- */
typedef struct ssasdpd_state {
    double *x, *v, *vt;       /* [n][3] each (any may be NULL)                  */
    double *F, *Fbp;          /* [n][3]                                         */
    double *rho, *Frho, *bvf_phi, *nu, *old_rho; /* [n]                          */
    double *C, *Q;            /* [n][Sc]                                        */
    uint32_t *xx;             /* [n][S]                                         */
    double *srrate, *sdrate;  /* [n]                                            */
    double *rrate;            /* [n][R]                                         */
    double *Ddiag;            /* [n][S]                                         */
    double *next_time;        /* [n] event queue firing times                   */
} ssasdpd_state;
int ssasdpd_get_state(ssasdpd_engine *e, ssasdpd_state *st);

/* Current neighbour lists in list order (particle.cpp:150-210, 240-294):
 * row_ptr [n+1]; idx/dist/dWdr/Dij [row_ptr[n]].  Pass idx == NULL to get
 * only row_ptr (and hence the total). */
int ssasdpd_get_neighbors(ssasdpd_engine *e, int64_t *row_ptr, int32_t *idx,
                          double *dist, double *dWdr, double *Dij);

/* Stand-alone fixed-radius search: tree points pts[n][3], queries q[m][3],
 * dimension dim, radius h.  Same acceptance rule and list order as
 * Particle::find_neighbors on a tree built from pts.  row_ptr [m+1];
 * idx/r2 sized by the total in row_ptr[m] (call once with idx == NULL). */
int ssasdpd_fr_search(const double *pts, int32_t n, const double *q, int32_t m,
                      int32_t dim, double h, int64_t *row_ptr, int32_t *idx,
                      double *r2);

/* No counterpart in the reference (its search is a kd-tree).  The cell grid
 * the search bins n points with in dimension dim at radius h, for the
 * bounding box bbox = {xlo, xhi, ylo, yhi, zlo, zhi}: cells per side, their
 * total and the cell side.  Host arithmetic only (no GPU); the grid never
 * exceeds max(4n, 2^20) cells, however far apart the points are. */
int ssasdpd_grid_shape(const double *bbox, int32_t n, int32_t dim, double h, int32_t *dims, int64_t *ncell,
                       double *cell);

/* ssasdpd_fr_search without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and two-call protocol as ssasdpd_fr_search; every device buffer (staging
 * of pts and q, cell binning, kd-tree and its work space, lists, error
 * record) lives in one file-scope device array of about 45 MiB that each
 * GPU holds until the process ends.  Limits: n <= SSASDPD_NOALLOC_MAX_POINTS,
 * m <= SSASDPD_NOALLOC_MAX_QUERIES (larger inputs are refused with
 * SSASDPD_ERR_CAPACITY before any GPU work), and at most
 * SSASDPD_NOALLOC_MAX_WIDTH neighbours per query (a wider list is reported
 * as SSASDPD_ERR_CAPACITY after the search).  Calls on one GPU share that
 * array and each overwrites it: calls that overlap in time race, so callers
 * serialise them. */
#define SSASDPD_NOALLOC_MAX_POINTS 32768
#define SSASDPD_NOALLOC_MAX_QUERIES 32768
#define SSASDPD_NOALLOC_MAX_WIDTH 32
int ssasdpd_fr_search_noalloc(const double *pts, int32_t n, const double *q, int32_t m,
                              int32_t dim, double h, int64_t *row_ptr, int32_t *idx,
                              double *r2);

/* Event trace of the last run / advance when SSASDPD_FLAG_TRACE is set:
 * per event (time, voxel, kind 0=reaction 1=diffusion, reaction or species,
 * destination or -1). */
typedef struct ssasdpd_event {
    double t;
    int32_t voxel;
    int32_t kind;
    int32_t which;
    int32_t dest;
} ssasdpd_event;
int ssasdpd_set_trace_capacity(ssasdpd_engine *e, uint64_t capacity);
int ssasdpd_get_trace(ssasdpd_engine *e, ssasdpd_event *ev, uint64_t cap, uint64_t *n_events);

/* ---- host utilities ------------------------------------------------------ */
/* MT19937-64 stream as std::mt19937_64(seed) produces it, plus the
 * exponential variate -log(1-canonical(w)) of every word computed with the
 * host libm (libstdc++ exponential_distribution, random.h:4761-4767). */
int ssasdpd_mt19937_64_stream(uint64_t seed, uint64_t n, uint64_t *words, double *expo);

#ifdef __cplusplus
}
#endif
#endif /* SSASDPD_H */
