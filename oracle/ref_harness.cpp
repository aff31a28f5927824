/*
 * ref_harness.cpp -- model translation unit that drives SpatialPy's own
 * C++ engine, for producing golden fixtures.  TEST INFRASTRUCTURE ONLY.
 *
 * It plays the role of the program SpatialPy generates from
 * ENG/propensity_file_template.cpp (ENG = spatialpy/solvers/c_base/
 * ssa_sdpd-c-simulation-engine): it defines the symbols the engine needs
 * (debug_flag, rng, propensity and chem-RHS tables, applyBoundaryConditions,
 * get_next_output) and, instead of run_simulation's threads, runs the same
 * step loop on one thread (simulate_threads.cpp:232-289 with -t 1):
 * buildKDTree, then take_step substeps 0..2 over the particles in index
 * order, then simulate_rdme.
 *
 * It is compiled together with the reference sources where they lie
 * (oracle/Makefile, target `ref`), into oracle/_ref/; nothing of the
 * reference is copied into this repository.
 *
 * Input: a directory with meta.txt (key value lines) and raw little-endian
 * arrays written by tests/golden/make_fixtures.py.  Output: raw arrays in
 * the same directory (neighbour lists after step 0, forces after the first
 * force substep, snapshots) and events.txt parsed from the engine's own
 * debug-level-1 event lines (simulate_rdme.cpp:282-283, 405-413).
 *
 * Models with custom propensities: meta key prop_lib names a shared library
 * built from the C++ text the reference generator writes for each reaction
 * (solver.py:160-190, 344-373; prop_<r> and det_<r>), whose functions
 * replace the table's; n_data_fn > 0 reads data_fn.f64 in the generated
 * input_data_fn layout [function][particle] (solver.py:192-198, 239-249).
 *
 * Deliberate differences from a generated program: Particle::F, Fbp, Frho
 * and vt (left uninitialised by Particle::Particle, particle.cpp:70-83) are
 * set to 0 after the particles are created; N_dense is padded with zeros so
 * the chem-RHS read at index R*rxn+s (model.cpp:186) never leaves the array.
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <ctime>
#include <dlfcn.h>
#include <unistd.h>
#include <pthread.h>
#include <random>
#include <string>
#include <vector>

#include "particle_system.hpp"
#include "propensities.hpp"
#include "simulate.hpp"
#include "simulate_rdme.hpp"
#include "output.hpp"

int debug_flag;
std::mt19937_64 rng;

namespace Spatialpy {
void buildKDTree(ParticleSystem *system);
}

namespace {

std::string g_dir;
std::map<std::string, std::string> g_meta;

template <class T> std::vector<T> load(const std::string &name, size_t count) {
    std::vector<T> v(count);
    if (count == 0) return v;
    std::string p = g_dir + "/" + name;
    FILE *f = fopen(p.c_str(), "rb");
    if (!f) { fprintf(stderr, "missing %s\n", p.c_str()); exit(2); }
    size_t got = fread(v.data(), sizeof(T), count, f);
    fclose(f);
    if (got != count) { fprintf(stderr, "short %s\n", p.c_str()); exit(2); }
    return v;
}
template <class T> void save(const std::string &name, const std::vector<T> &v) {
    std::string p = g_dir + "/" + name;
    FILE *f = fopen(p.c_str(), "wb");
    if (!f) { perror(p.c_str()); exit(2); }
    if (!v.empty()) fwrite(v.data(), sizeof(T), v.size(), f);
    fclose(f);
}
long geti(const char *k) { return atol(g_meta.at(k).c_str()); }
double getd(const char *k) { return strtod(g_meta.at(k).c_str(), nullptr); }

struct Rx { int kind, a, b; double k; uint64_t mask; };
std::vector<Rx> g_rx;
struct Bc { double lo[3], hi[3]; int type, target, species; double value[3]; };
std::vector<Bc> g_bc;
std::vector<unsigned> g_out_steps;

/* same expressions as the generated propensity functions
 * (solver.py:344-373, core/reaction.py:260-313) */
template <int r> double prop_fn(const unsigned int *xu, double t, double vol, double *data_fn, int sd) {
    (void)t; (void)data_fn;
    const int *x = (const int *)xu;
    const Rx &re = g_rx[r];
    if (re.mask && !((re.mask >> sd) & 1ULL)) return 0.0;
    switch (re.kind) {
    case 0: return (re.k * vol);
    case 1: return (re.k * x[re.a]);
    case 2: return (((re.k * x[re.a]) * x[re.b]) / vol);
    default: return (((re.k * x[re.a]) * (x[re.a] - 1)) / vol);
    }
}
/* generated det<name> functions (solver.py:160-190) */
template <int r> double chem_fn(double *x, double t, double vol, double *data_fn, int sd) {
    (void)t; (void)data_fn; (void)vol;
    const Rx &re = g_rx[r];
    if (re.mask && !((re.mask >> sd) & 1ULL)) return 0.0;
    switch (re.kind) {
    case 0: return re.k;
    case 1: return (re.k * x[re.a]);
    case 2: return ((re.k * x[re.a]) * x[re.b]);
    default: return ((re.k * x[re.a]) * x[re.a]);
    }
}
#define RX16(F) F<0>, F<1>, F<2>, F<3>, F<4>, F<5>, F<6>, F<7>, F<8>, F<9>, F<10>, F<11>, F<12>, F<13>, F<14>, F<15>
Spatialpy::PropensityFun g_prop_tab[16] = {RX16((Spatialpy::PropensityFun)prop_fn)};
Spatialpy::ChemRxnFun g_chem_tab[16] = {RX16(chem_fn)};

/* generated functions from the model's shared library, when given */
void load_prop_lib(const std::string &path, int R) {
    void *h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) { fprintf(stderr, "dlopen %s: %s\n", path.c_str(), dlerror()); exit(2); }
    for (int r = 0; r < R; r++) {
        char nm[32];
        snprintf(nm, sizeof nm, "prop_%d", r);
        if (void *f = dlsym(h, nm)) g_prop_tab[r] = (Spatialpy::PropensityFun)f;
        snprintf(nm, sizeof nm, "det_%d", r);
        if (void *f = dlsym(h, nm)) g_chem_tab[r] = (Spatialpy::ChemRxnFun)f;
    }
}

}  // namespace

namespace Spatialpy {
double *parameters = nullptr;

PropensityFun *ALLOC_propensities(void) {
    PropensityFun *p = (PropensityFun *)malloc(sizeof(PropensityFun) * (g_rx.size() + 1));
    for (size_t i = 0; i < g_rx.size(); i++) p[i] = g_prop_tab[i];
    return p;
}
void FREE_propensities(PropensityFun *ptr) { free(ptr); }
ChemRxnFun *ALLOC_ChemRxnFun(void) {
    ChemRxnFun *p = (ChemRxnFun *)malloc(sizeof(ChemRxnFun) * (g_rx.size() + 1));
    for (size_t i = 0; i < g_rx.size(); i++) p[i] = g_chem_tab[i];
    return p;
}
/* generated get_next_output (solver.py:290-299) */
unsigned int get_next_output(ParticleSystem *system) {
    (void)system;
    static size_t index = 0;
    unsigned next = index < g_out_steps.size() ? g_out_steps[index] : 0xFFFFFFFFu;
    index++;
    return next;
}
/* generated applyBoundaryConditions (core/boundarycondition.py:121-167) */
void applyBoundaryConditions(Particle *me, ParticleSystem *system) {
    (void)system;
    for (const Bc &b : g_bc) {
        if ((me->x[0] >= b.lo[0]) && (me->x[0] <= b.hi[0]) && (me->x[1] >= b.lo[1]) && (me->x[1] <= b.hi[1]) &&
            (me->x[2] >= b.lo[2]) && (me->x[2] <= b.hi[2]) && (b.type < 0 || me->type == b.type)) {
            if (b.target == 0) me->C[b.species] = b.value[0];
            else if (b.target == 1) { me->v[0] = b.value[0]; me->v[1] = b.value[1]; me->v[2] = b.value[2]; }
            else if (b.target == 2) me->nu = b.value[0];
            else if (b.target == 3) me->rho = b.value[0];
        }
    }
}
}  // namespace Spatialpy

int main(int argc, char **argv) {
    using namespace Spatialpy;
    if (argc < 2) { fprintf(stderr, "usage: %s DIR\n", argv[0]); return 2; }
    g_dir = argv[1];
    {
        FILE *f = fopen((g_dir + "/meta.txt").c_str(), "r");
        if (!f) { fprintf(stderr, "no meta.txt\n"); return 2; }
        char k[128], v[256];
        while (fscanf(f, "%127s %255s", k, v) == 2) g_meta[k] = v;
        fclose(f);
    }
    const int n = geti("n"), dim = geti("dimension"), T = geti("n_types");
    const int S = geti("n_species"), R = geti("n_reactions");
    const int Sc = geti("n_chem_species"), Rc = geti("n_chem_reactions");
    const unsigned nt = geti("nt");
    const unsigned run_steps = g_meta.count("run_steps") ? geti("run_steps") : nt;
    const long long seed = atoll(g_meta.at("seed").c_str());
    const int force_step = g_meta.count("force_step") ? geti("force_step") : 0;
    const int nbr_step = g_meta.count("nbr_step") ? geti("nbr_step") : 0;
    debug_flag = geti("debug_flag");

    auto x = load<double>("x.f64", 3 * (size_t)n);
    auto vin = load<double>("v.f64", 3 * (size_t)n);
    auto type = load<int32_t>("type.i32", n);
    auto nu = load<double>("nu.f64", n);
    auto mass = load<double>("mass.f64", n);
    auto rho = load<double>("rho.f64", n);
    auto fixed = load<int32_t>("fixed.i32", n);
    std::vector<unsigned> u0 = load<unsigned>("u0.u32", (size_t)n * S);
    int Smax = S > Sc ? S : Sc;
    auto Dm = load<double>("dmat.f64", (size_t)Smax * T);
    int Rmax = R > Rc ? R : Rc;
    if (Rmax > 16) { fprintf(stderr, "too many reactions\n"); return 2; }
    auto rxk = load<int32_t>("rx_int.i32", 3 * (size_t)Rmax);
    auto rxd = load<double>("rx_k.f64", Rmax);
    auto rxm = load<uint64_t>("rx_mask.u64", Rmax);
    for (int r = 0; r < Rmax; r++) g_rx.push_back({rxk[3 * r], rxk[3 * r + 1], rxk[3 * r + 2], rxd[r], rxm[r]});
    if (g_meta.count("prop_lib")) load_prop_lib(g_dir + "/" + g_meta.at("prop_lib"), Rmax);
    const int ndf = g_meta.count("n_data_fn") ? geti("n_data_fn") : 0;
    auto dfv = load<double>("data_fn.f64", (size_t)ndf * n);
    std::vector<int32_t> jcN32 = load<int32_t>("jcN.i32", R + 1 > 1 ? R + 1 : 0);
    size_t nnzN = R > 0 ? (size_t)jcN32[R] : 0;
    auto irN32 = load<int32_t>("irN.i32", nnzN);
    auto prN32 = load<int32_t>("prN.i32", nnzN);
    std::vector<int32_t> jcG32 = load<int32_t>("jcG.i32", S > 0 ? S + R + 1 : 0);
    size_t nnzG = S > 0 ? (size_t)jcG32[S + R] : 0;
    auto irG32 = load<int32_t>("irG.i32", nnzG);
    auto Nd = load<int32_t>("ndense.i32", (size_t)Sc * Rc);
    int nbc = geti("n_bcs");
    auto bcd = load<double>("bc.f64", 9 * (size_t)nbc);
    auto bci = load<int32_t>("bc.i32", 3 * (size_t)nbc);
    for (int b = 0; b < nbc; b++) {
        Bc c;
        for (int k = 0; k < 3; k++) { c.lo[k] = bcd[9 * b + k]; c.hi[k] = bcd[9 * b + 3 + k]; c.value[k] = bcd[9 * b + 6 + k]; }
        c.type = bci[3 * b]; c.target = bci[3 * b + 1]; c.species = bci[3 * b + 2];
        g_bc.push_back(c);
    }
    int nout = geti("n_output_steps");
    g_out_steps = load<unsigned>("out_steps.u32", nout);

    /* the generated arrays are size_t (solver.py:236-258) */
    static std::vector<size_t> irN(irN32.begin(), irN32.end()), jcN(jcN32.begin(), jcN32.end());
    static std::vector<int> prN(prN32.begin(), prN32.end());
    static std::vector<size_t> irG(irG32.begin(), irG32.end()), jcG(jcG32.begin(), jcG32.end());
    static std::vector<int> Ndense(Nd.begin(), Nd.end());
    Ndense.resize(Ndense.size() + 64 + (size_t)Rmax * Rmax, 0);
    static std::vector<double> Dmat(Dm.begin(), Dm.end());
    static const char *names[] = {"s0", "s1", "s2", "s3", "s4", "s5", "s6", "s7", 0};

    /* __SYSTEM_CONFIG__ (solver.py:375-419) */
    ParticleSystem *system = new ParticleSystem(T, Sc, Rc, S, R, ndf);
    system->static_domain = geti("static_domain");
    system->subdomain_diffusion_matrix = Dmat.data();
    system->stoichiometric_matrix = Ndense.data();
    system->chem_rxn_rhs_functions = ALLOC_ChemRxnFun();
    system->stoch_rxn_propensity_functions = ALLOC_propensities();
    system->species_names = names;
    system->dt = getd("dt");
    system->nt = nt;
    system->h = getd("h");
    system->rho0 = getd("rho0");
    system->c0 = getd("c0");
    system->P0 = getd("P0");
    system->xlo = getd("xlo"); system->xhi = getd("xhi");
    system->ylo = getd("ylo"); system->yhi = getd("yhi");
    system->zlo = getd("zlo"); system->zhi = getd("zhi");
    system->dimension = dim;
    system->gravity[0] = getd("g0"); system->gravity[1] = getd("g1"); system->gravity[2] = getd("g2");
    rng.seed((uint64_t)seed);

    /* init_all_particles (propensity_file_template.cpp:73-90) */
    system->particles.reserve(n);
    for (int i = 0; i < n; i++) {
        system->particles.emplace_back(Particle(system, i, x[3 * i], x[3 * i + 1], x[3 * i + 2], type[i], nu[i],
                                                mass[i], 0.0, rho[i], fixed[i]));
        Particle *p = &system->particles.back();
        for (int s = 0; s < Sc; s++) p->C[s] = (double)u0[(size_t)i * Sc + s];
        if (S > 0)   /* __DATA_FUNCTION_ASSIGN__ (solver.py:192-198) */
            for (int k = 0; k < ndf; k++) p->data_fn[k] = dfv[(size_t)k * n + i];
        for (int k = 0; k < 3; k++) { p->v[k] = vin[3 * i + k]; p->vt[k] = 0; p->F[k] = 0; p->Fbp[k] = 0; }
        p->Frho = 0;
        p->bvf_phi = 0;
    }
    /* continuing a run mid-way: the state after step start_step - 1 (the
     * quantities take_step1 carries over from the previous step:
     * simulate.cpp:68-83) */
    const unsigned start_step = g_meta.count("start_step") ? geti("start_step") : 0;
    if (g_meta.count("start_step")) {
        auto F = load<double>("F.f64", 3 * (size_t)n), Fbp = load<double>("Fbp.f64", 3 * (size_t)n);
        auto Frho = load<double>("Frho.f64", n);
        auto Cst = load<double>("C.f64", (size_t)n * Sc), Qst = load<double>("Q.f64", (size_t)n * Sc);
        for (int i = 0; i < n; i++) {
            Particle &p = system->particles[i];
            for (int k = 0; k < 3; k++) { p.F[k] = F[3 * i + k]; p.Fbp[k] = Fbp[3 * i + k]; }
            p.Frho = Frho[i];
            for (int s = 0; s < Sc; s++) { p.C[s] = Cst[(size_t)i * Sc + s]; p.Q[s] = Qst[(size_t)i * Sc + s]; }
        }
    }
    if (S > 0) initialize_rdme(system, irN.data(), jcN.data(), prN.data(), irG.data(), jcG.data(), u0.data());

    std::vector<uint32_t> snap_xx;
    std::vector<double> snap_x, snap_v, snap_rho, snap_C;
    std::vector<uint32_t> snap_steps;
    const bool write_vtk = g_meta.count("write_vtk") && geti("write_vtk");
    auto snapshot = [&](unsigned step) {
        if (write_vtk) {
            /* the reference's own VTK writer, synchronously, into the cwd */
            Spatialpy::output_vtk__sync_step(system, (int)step);
            Spatialpy::output_vtk__async_step(system);
        }
        for (auto &p : system->particles) {
            for (int s = 0; s < S; s++) snap_xx.push_back(p.xx[s]);
            for (int k = 0; k < 3; k++) { snap_x.push_back(p.x[k]); snap_v.push_back(p.v[k]); }
            snap_rho.push_back(p.rho);
            for (int s = 0; s < Sc; s++) snap_C.push_back(p.C[s]);
        }
        snap_steps.push_back(step);
    };
    auto dump_nbrs = [&]() {
        std::vector<int64_t> rp;
        std::vector<int32_t> idx;
        std::vector<double> dist, dwdr, dij;
        int64_t off = 0;
        Particle *base = &system->particles[0];
        for (auto &p : system->particles) {
            rp.push_back(off);
            for (auto &nb : p.neighbors) {
                idx.push_back((int32_t)(nb.data - base));
                dist.push_back(nb.dist);
                dwdr.push_back(nb.dWdr);
                dij.push_back(nb.D_i_j);
                off++;
            }
        }
        rp.push_back(off);
        save("out_nbr_rowptr.i64", rp);
        save("out_nbr_idx.i32", idx);
        save("out_nbr_dist.f64", dist);
        save("out_nbr_dwdr.f64", dwdr);
        save("out_nbr_dij.f64", dij);
    };
    auto dump_forces = [&]() {
        std::vector<double> F, Fbp, Frho, Q;
        for (auto &p : system->particles) {
            for (int k = 0; k < 3; k++) { F.push_back(p.F[k]); Fbp.push_back(p.Fbp[k]); }
            Frho.push_back(p.Frho);
            for (int s = 0; s < Sc; s++) Q.push_back(p.Q[s]);
        }
        save("out_F.f64", F);
        save("out_Fbp.f64", Fbp);
        save("out_Frho.f64", Frho);
        save("out_Q.f64", Q);
    };

    const int threads = g_meta.count("threads") ? geti("threads") : 0;
    if (threads > 0) {
        /* timing mode: the worker structure of run_simulation
         * (simulate_threads.cpp:122-281) -- contiguous particle blocks per
         * thread, barriers between substeps, kd-tree build and RDME on the
         * master -- without the VTK output thread, so only the engine's
         * compute is timed. */
        const unsigned steps = run_steps;
        pthread_barrier_t bar;
        pthread_barrier_init(&bar, nullptr, threads + 1);
        struct Ctx { ParticleSystem *sys; int first, count; pthread_barrier_t *bar; unsigned first_step, steps; };
        std::vector<Ctx> ctx(threads);
        std::vector<pthread_t> th(threads);
        int per = n / threads, left = n, first = 0;
        for (int t = 0; t < threads; t++) {
            int cnt = (t == threads - 1) ? left : per;
            ctx[t] = Ctx{system, first, cnt, &bar, start_step, steps};
            first += cnt;
            left -= cnt;
            pthread_create(&th[t], nullptr, [](void *arg) -> void * {
                Ctx *c = (Ctx *)arg;
                for (unsigned step = c->first_step; step < c->first_step + c->steps; step++) {
                    pthread_barrier_wait(c->bar);
                    for (unsigned sub = 0; sub < 3; sub++) {
                        for (int i = 0; i < c->count; i++)
                            take_step(&c->sys->particles[c->first + i], c->sys, step, sub);
                        pthread_barrier_wait(c->bar);
                    }
                }
                return nullptr;
            }, &ctx[t]);
        }
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        double kd_s = 0, rdme_s = 0;
        for (system->current_step = start_step; system->current_step < start_step + steps; system->current_step++) {
            struct timespec a, b;
            clock_gettime(CLOCK_MONOTONIC, &a);
            buildKDTree(system);
            clock_gettime(CLOCK_MONOTONIC, &b);
            kd_s += (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
            pthread_barrier_wait(&bar);
            for (unsigned sub = 0; sub < 3; sub++) pthread_barrier_wait(&bar);
            clock_gettime(CLOCK_MONOTONIC, &a);
            simulate_rdme(system, system->current_step);
            clock_gettime(CLOCK_MONOTONIC, &b);
            rdme_s += (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        for (int t = 0; t < threads; t++) pthread_join(th[t], nullptr);
        double sec = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
        printf("REF_TIMING seconds=%.6f steps=%u n=%d reactions=%ld diffusions=%ld threads=%d kd_s=%.6f rdme_s=%.6f "
               "start_step=%u\n",
               sec, steps, n, system->total_reactions, system->total_diffusion, threads, kd_s, rdme_s, start_step);
        fflush(stdout);
        _exit(0);
    }
    unsigned next_output = 0;
    for (system->current_step = start_step; system->current_step < start_step + run_steps; system->current_step++) {
        unsigned step = system->current_step;
        buildKDTree(system);
        if (step >= next_output) {
            snapshot(step);
            next_output = get_next_output(system);
        }
        for (unsigned sub = 0; sub < 3; sub++) {
            for (auto &p : system->particles) take_step(&p, system, step, sub);
            if (sub == 1 && (int)step == nbr_step) dump_nbrs();
            if (sub == 1 && (int)step == force_step) dump_forces();
        }
        simulate_rdme(system, step);
    }
    snapshot(system->current_step);
    fflush(stdout);
    save("out_xx.u32", snap_xx);
    save("out_x.f64", snap_x);
    save("out_v.f64", snap_v);
    save("out_rho.f64", snap_rho);
    save("out_C.f64", snap_C);
    save("out_steps.u32", snap_steps);
    std::vector<int64_t> counters = {system->total_reactions, system->total_diffusion};
    save("out_counters.i64", counters);
    return 0;
}
