#!/usr/bin/env python3
"""Benchmark: the 1M-particle 3D SDPD + sSSA cylinder run on one MI355X.

One step = neighbour search (cell binning), the three SDPD substeps, the
diffusion jump-rate assembly and that step's exact SSA, i.e. one iteration
of the reference's step loop (spatialpy/solvers/c_base/
ssa_sdpd-c-simulation-engine/src/simulate_threads.cpp:232-281).

Prints ONE JSON line (metric particle-steps/s, higher is better), with the
SSA event rate, the roofline of the dominant HBM-bound kernels and the
reference CPU engine timed on this host's cores beside it.

Three more workloads print the same kind of line: ``--workload static3d``
(the 40^3 static A+B<->C case: SSA events/s of a seed batch, with one
trajectory beside it), ``--workload ensemble`` (uneven gene-network models
in one persistent launch: trajectories/s) and ``--workload dambreak`` (the
2D SDPD dam break with walls: particle-steps/s, FP32 pair forces beside).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def measured_traffic():
    """HBM bytes per launch of k_ssa and per step of the particle path, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    (profiles/cylinder_1M_traffic.txt; None when it is absent)."""
    path = os.path.join(ROOT, "profiles", "cylinder_1M_traffic.txt")
    out = {"k_ssa": None, "particle_path": None}
    if not os.path.exists(path):
        return out
    for line in open(path):
        if line.startswith("k_ssa traffic_bytes_per_launch="):
            out["k_ssa"] = float(line.split("=")[1])
        if line.startswith("particle_path traffic_bytes_per_step="):
            out["particle_path"] = float(line.split("=")[1])
    return out


def algorithmic_bytes_per_particle_step(dim, S, E):
    """Compulsory HBM traffic per particle-step of the SDPD + rate path:
    read x, v, vt (3*8d), rho, mass, nu (24), type (4); write x, v, vt (3*8d),
    rho (8); per species int32 count read + FP64 Ddiag write (12 S); per
    neighbour int32 index + FP64 D_ij written once per step (12 E)."""
    return (3 * 8 * dim + 8 + 8 + 8 + 4) + (3 * 8 * dim + 8) + 12 * S + 12 * E


def build_workload(n_target, nt):
    from spatialpy_amd import workloads
    return workloads.cylinder_3d(n_target=n_target, nt=nt)


# The cylinder run: CYL_RUN_STEPS steps in all (the configured 2,000-step
# plain run).  The first CYL_SETTLE steps run untimed: the first ~100 are the
# transient in which the A and B molecules that start in the same voxel
# react (about 490k reactions in the first 50 steps, against a few thousand
# per 50 steps from step 1000 on).  The timed window starts at step
# CYL_SETTLE + warmup, in the middle of the run; the whole run's rate is
# reported beside it.
CYL_RUN_STEPS = 2000
CYL_SETTLE = 1000


def harness_exe():
    return os.path.join(ROOT, "oracle", "_ref", "ref_harness_O2")


def cpu_baseline(sa, state, start_step, steps):
    """Time the reference engine itself (oracle/_ref/ref_harness_O2, the
    reference's engine sources compiled by oracle/Makefile) on this host's
    cores, over a bounded sample of the GPU's own timed window: `steps` steps
    starting at `start_step` from the GPU's particle state and counts at
    that step.  Without that build the repository's FP64 restatement (one
    thread) is timed instead, and the line says so (kind "port")."""
    threads = min(os.cpu_count() or 1, 16)
    exe = harness_exe()
    from spatialpy_amd import lowering
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as d:
            lowering.write_harness_dir(sa, d, seed=1, threads=threads, run_steps=steps, state=state,
                                       start_step=start_step)
            r = subprocess.run([exe, d], capture_output=True, text=True, timeout=600)
            line = [l for l in r.stdout.splitlines() if l.startswith("REF_TIMING")]
            if r.returncode == 0 and line:
                kv = dict(tok.split("=") for tok in line[0].split()[1:])
                sec = float(kv["seconds"])
                ev = int(kv["reactions"]) + int(kv["diffusions"])
                return {"value": sa.n * steps / sec, "unit": "particle-steps/s", "cores": threads,
                        "kind": "reference",
                        "sample": f"reference C++ engine (-O2, {threads} threads, run_simulation worker "
                                  f"structure without VTK output) on the same {sa.n}-particle cylinder, steps "
                                  f"{start_step}..{start_step + steps - 1} (the first {steps} of the GPU's timed "
                                  f"window) from the GPU's state at step {start_step}: {sec:.2f} s",
                        "ssa_events_per_s": ev / sec,
                        "ssa_events": ev,
                        "ssa_events_per_s_in_rdme": ev / max(float(kv["rdme_s"]), 1e-9),
                        "kd_tree_s": float(kv["kd_s"]), "rdme_s": float(kv["rdme_s"])}
            print(f"warning: the reference build failed (rc={r.returncode}); timing the FP64 oracle instead",
                  file=sys.stderr)
    else:
        print(f"warning: {exe} is absent (built by `make -C oracle ref` where the reference sources are); "
              "cpu_baseline times the FP64 oracle restatement instead (kind 'port')", file=sys.stderr)
    from oracle import oracle_py
    o = oracle_py.Oracle(sa, gauss_seidel=True)
    o.reset(seed=1)
    t = time.perf_counter()
    o.advance(steps)
    sec = time.perf_counter() - t
    return {"value": sa.n * steps / sec, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle restatement, 1 thread, {sa.n} particles, steps 0..{steps - 1} "
                      f"(the reference build was absent): {sec:.2f} s"}


# matrix-core flops of the surrogate's edge network per listed pair:
# 16 features -> 64 -> 64 -> 4 (the products the fused kernel issues are
# larger: activations enter as two bf16 halves, the first layer's K is
# padded to 32 and the head's outputs to 16)
SURROGATE_FLOPS_PER_EDGE = 2 * (16 * 64 + 64 * 64 + 64 * 4)
SURROGATE_MFMA_FLOPS_PER_SLOT = 2 * 2 * (32 * 64 + 64 * 64 + 64 * 16)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def surrogate_leg(args, sa, E, plain_ms_per_step, plain_pair_ms):
    """The surrogate on the same cylinder run: trained on the plain GPU run's
    own steps (a sample of 32,768 moving particles every 25 steps of the
    2,000-step run; feature scales from the state at the settle point),
    then the run again with the network attached: a check against the plain
    pair kernel every 50 steps with the 2% budget, re-armed (a failed check
    pauses the network until the next check passes) and refreshed online
    from each check's plain step (8 Adam steps).  Reports the training cost
    and error, the checks' outcomes, the share of steps the network drove,
    the timed window's ms/step with the network beside the plain run's, the
    fused inference kernel against the plain pair kernel, and its
    matrix-core rate."""
    from spatialpy_amd.engine import Engine
    from spatialpy_amd.surrogate import Surrogate, fit_scales
    K = 20 if args.steps is None else args.steps
    W = 5 if args.warmup is None else args.warmup
    first = args.settle + W
    eng = Engine(sa)
    eng.reset(seed=12345)
    eng.advance(args.settle)
    s = Surrogate(seed=0)
    s.set_scales(*fit_scales(eng, sa))
    eng.reset(seed=12345)
    s.record(eng, 25, max_particles=32768)
    t = time.perf_counter()
    eng.advance(CYL_RUN_STEPS)
    eng.state()
    rec_s = time.perf_counter() - t
    s.stop_recording(eng)
    n_samples = s.num_samples()
    _progress(f"surrogate: {n_samples} samples recorded in {rec_s:.1f} s")
    t = time.perf_counter()
    loss = s.fit(args.surrogate_epochs, lr=3e-3, lr_end=3e-5)
    train_s = time.perf_counter() - t
    _progress(f"surrogate: {args.surrogate_epochs} epochs in {train_s:.1f} s")
    train_err = s.relative_error()
    # the run with the network attached; the same window timed
    eng.reset(seed=12345)
    s.attach(eng, check_every=50, max_rel_err=0.02, refresh_iters=8, refresh_lr=1e-4, keep_samples=400,
             max_particles=32768, rearm=True)
    eng.advance(first)
    eng.state()
    t0 = time.perf_counter()
    eng.advance(K)
    eng.state()
    win_s = time.perf_counter() - t0
    st_win = Surrogate.stats(eng)
    eng.advance(CYL_RUN_STEPS - first - K)
    st = Surrogate.stats(eng)
    xs = eng.state()["x"]
    infer_ms = s.time_engine(eng, reps=5)
    s.close()
    eng.close()
    tflops = SURROGATE_FLOPS_PER_EDGE * E * sa.n / (infer_ms * 1e-3) / 1e12
    return {"protocol": "trained on the plain 2,000-step run (a 32,768-particle sample every 25 steps), then the "
                        "run again with a check against the plain pair kernel every 50 steps, 2% budget, re-armed, "
                        "refreshed online from the checks",
            "samples": n_samples, "record_run_s": round(rec_s, 2), "train_epochs": args.surrogate_epochs,
            "train_s": round(train_s, 2), "train_loss": loss, "train_rel_err": round(train_err, 5),
            "checks": st["checks"], "failed_checks": st["failed_checks"],
            "max_passing_check_err": round(st["max_err"], 5), "last_check_err": round(st["last_err"], 5),
            "nn_steps": st["nn_steps"], "nn_share": round(st["nn_steps"] / CYL_RUN_STEPS, 4),
            "capped_predictions": st["capped"],
            "final_position_range": [round(float(xs.min()), 4), round(float(xs.max()), 4)],
            "window_ms_per_step": round(1e3 * win_s / K, 4), "window_nn_steps": st_win["nn_steps"],
            "plain_window_ms_per_step": round(plain_ms_per_step, 4),
            "speed_vs_plain_step": round(plain_ms_per_step / (1e3 * win_s / K), 4),
            "infer_ms": round(infer_ms, 4), "plain_pair_kernel_ms": round(plain_pair_ms, 4),
            "speed_vs_plain_pair_kernel": round(plain_pair_ms / infer_ms, 4),
            "mfma": {"bound": "mfma", "achieved": round(tflops, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 5),
                     "flops_per_edge": SURROGATE_FLOPS_PER_EDGE,
                     "issued_mfma_flops_per_list_slot": SURROGATE_MFMA_FLOPS_PER_SLOT}}


def ssa_bytes_per_event(S, R, E):
    """Algorithmic bytes of one event: a reaction reads and writes the
    selected voxel's state (type, degree, vol, queue time: 28 B; counts 4 S;
    Ddiag 8 S; rates 8 R; srrate/sdrate 16) and writes back its counts, rates
    and sums; a diffusion does that for the source and the destination and
    reads the source's neighbour row (int32 index + FP64 D_ij)."""
    vox = 28 + 4 * S + 8 * S + 8 * R + 16 + 4 * S + 8 * R + 16
    return vox, 2 * vox + 12 * E


def mean_degree(x, h, samples=2000, seed=0):
    """Mean neighbour count within h (self excluded) over a sample of points."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    idx = rng.choice(x.shape[0], size=min(samples, x.shape[0]), replace=False)
    cnt = cKDTree(x).query_ball_point(x[idx], r=h, return_length=True)
    return float(np.mean(cnt) - 1.0)


def ssa_roofline(S, R, E, n_re, n_di, gpu_ms, kernel):
    b_react, b_diff = ssa_bytes_per_event(S, R, E)
    by = b_react * n_re + b_diff * n_di
    gbs = by / (gpu_ms * 1e-3) / 1e9 if gpu_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(gbs, 4), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 7), "traffic": None,
            "kernel": kernel, "algorithmic_bytes_per_event": {"reaction": b_react, "diffusion": b_diff},
            "mean_degree": round(E, 2), "events": int(n_re + n_di),
            "us_per_event": round(1e3 * gpu_ms / max(n_re + n_di, 1), 3)}


def reference_timing(sa, steps):
    """The reference engine (oracle/_ref/ref_harness_O2, built from the
    reference sources) on one model for `steps` steps, with min(nproc, 16)
    threads: (seconds, reactions + diffusions, threads), or None when the
    build is absent."""
    threads = min(os.cpu_count() or 1, 16)
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness_O2")
    if not os.path.exists(exe):
        return None
    from spatialpy_amd import lowering
    with tempfile.TemporaryDirectory() as d:
        lowering.write_harness_dir(sa, d, seed=1, threads=threads, run_steps=steps)
        r = subprocess.run([exe, d], capture_output=True, text=True, timeout=600)
    line = [l for l in r.stdout.splitlines() if l.startswith("REF_TIMING")]
    if r.returncode != 0 or not line:
        return None
    kv = dict(tok.split("=") for tok in line[0].split()[1:])
    return float(kv["seconds"]), int(kv["reactions"]) + int(kv["diffusions"]), threads


def oracle_timing(sa, steps):
    """The repository's FP64 restatement, one thread, when the reference
    build is absent: (seconds, events, 1)."""
    from oracle import oracle_py
    o = oracle_py.Oracle(sa)
    o.reset(seed=1)
    t = time.perf_counter()
    o.advance(steps)
    sec = time.perf_counter() - t
    st = o.stats()
    return sec, int(st.reactions + st.diffusions), 1


def bench_static3d(args):
    """The static 3D reaction-diffusion case (40^3 = 64,000 fixed voxels,
    A+B<->C, 1e5 A + 1e5 B): one trajectory, then a batch of seeds of the
    same model (one event-loop workgroup per trajectory), over the first
    `steps` unit time steps of its t = 0..100 run."""
    from spatialpy_amd import workloads
    from spatialpy_amd.engine import Engine
    steps = args.steps if args.steps is not None else 2
    warm = args.warmup if args.warmup is not None else 0
    B = args.batch
    sa = workloads.static_rd_3d(t_end=warm + steps, num_points=warm + steps + 1)
    E = mean_degree(sa.x, sa.h)
    eng = Engine(sa)
    eng.reset(seed=12345)
    if warm:
        eng.advance(warm)
    st0, ph0 = eng.stats(), eng.phase_times()
    t0 = time.perf_counter()
    eng.advance(steps)
    sec = time.perf_counter() - t0
    st1, ph1 = eng.stats(), eng.phase_times()
    n_re, n_di = st1.reactions - st0.reactions, st1.diffusions - st0.diffusions
    ssa_ms = ph1["ssa"][0] - ph0["ssa"][0]
    single = {"events": int(n_re + n_di), "seconds": round(sec, 4), "events_per_s": (n_re + n_di) / sec,
              "ssa_gpu_ms": round(ssa_ms, 3)}
    roof = ssa_roofline(sa.n_species, sa.n_reactions, E, n_re, n_di, ssa_ms,
                        "k_ssa (event loop; latency-bound, one wavefront per trajectory)")
    eng.close()
    # the batch: B seeded trajectories of the whole window (t = 0..warm+steps)
    beng = Engine(sa)
    t0 = time.perf_counter()
    beng.run_batch(list(range(1, B + 1)), fields=("xx",))
    bsec = time.perf_counter() - t0
    bst = beng.stats()
    bph = beng.phase_times()
    bev = bst.reactions + bst.diffusions
    batch = {"trajectories": B, "sim_time": float((warm + steps) * sa.dt), "events": int(bev),
             "seconds": round(bsec, 4), "events_per_s": bev / bsec, "trajectories_per_s": B / bsec,
             "ssa_gpu_ms": round(bph["ssa"][0], 3),
             "note": "wall time includes copying every trajectory's count snapshots to the host"}
    beng.close()
    out = {
        "metric": "SSA events/s", "value": bev / bsec, "unit": "events/s", "n_gpus": 1,
        "steps": steps, "warmup": warm, "ms_per_step": 1e3 * sec / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"static 3D reaction-diffusion, 40^3 = {sa.n} fixed voxels on the unit cube, "
                               f"A+B->C k=1e-3, C->A+B k=0.1, D = 0.01/0.01/0.001, 1e5 A + 1e5 B scattered, "
                               f"dt=1, t = 0..{warm + steps}; value: {B} seeded trajectories in one batch",
                   "voxels": sa.n, "batch": B},
        "single_trajectory": single,
        "batch": batch,
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        r = reference_timing(sa, warm + steps)
        kind = "reference"
        if r is None:
            r, kind = oracle_timing(sa, warm + steps), "port"
        csec, cev, thr = r
        out["cpu_baseline"] = {"value": cev / csec, "unit": "events/s", "cores": thr, "kind": kind,
                               "sample": f"one trajectory of the same model, t = 0..{warm + steps}: {cev} events "
                                         f"in {csec:.2f} s (the reference's event loop is serial)"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        out["single_trajectory"]["speedup_vs_cpu"] = single["events_per_s"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


def bench_ensemble(args):
    """Uneven gene-network models (jittered 3D lattices, N ~ U[8k, 64k], own
    rates and initial counts each) in one persistent launch, after the
    event-tree width autotuner; t = 0..t_end."""
    from spatialpy_amd import workloads
    from spatialpy_amd.ensemble import Ensemble
    M, t_end = args.models, args.t_end
    rng = np.random.default_rng(0)
    sizes = rng.integers(8000, 64001, size=M)
    t0 = time.perf_counter()
    models = []
    for i, nv in enumerate(sizes):
        models.append(workloads.gene_network_3d(n_target=int(nv), seed=1000 + i, t_end=t_end))
        if (i + 1) % 128 == 0:
            print(f"built {i + 1} of {M} models", file=sys.stderr, flush=True)
    build_s = time.perf_counter() - t0
    print(f"built {M} models in {build_s:.1f} s", file=sys.stderr, flush=True)
    E = float(np.mean([mean_degree(m.x, m.h, samples=500, seed=i) for i, m in enumerate(models[:16])]))
    ens = Ensemble(models)
    seeds = list(range(1, M + 1))
    width, probe = ens.autotune(seeds, probe_steps=1)
    print(f"autotuned event-tree width {width}: {probe}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ens.run(seeds)
    sec = time.perf_counter() - t0
    st = ens.stats()
    ev = st.reactions + st.diffusions
    S, R = models[0].n_species, models[0].n_reactions
    roof = ssa_roofline(S, R, E, st.reactions, st.diffusions, st.ms_ssa,
                        "k_ssa_multi (all models' event loops in one launch; latency-bound)")
    roof["note"] = "achieved is the sum over concurrent trajectories of their algorithmic bytes per event"
    voxels = int(sum(m.n for m in models))
    out = {
        "metric": "trajectories/s", "value": M / sec, "unit": "trajectories/s", "n_gpus": 1,
        "steps": int(models[0].nt), "warmup": 0, "ms_per_step": 1e3 * sec / max(int(models[0].nt), 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{M} uneven static gene-network models (5 species, 6 mass-action reactions, "
                               f"jittered 3D lattices with N ~ U[8000, 64000], rates log-U[1e-3, 10], "
                               f"D log-U[1e-4, 1e-2], Poisson(50) counts), one trajectory each, t = 0..{t_end}, "
                               f"one persistent launch",
                   "models": M, "voxels": voxels, "t_end": t_end, "tree_width": width},
        "ssa_events_per_s": ev / sec,
        "ssa_events": int(ev),
        "autotune_ms": probe,
        "host_model_build_s": round(build_s, 2),
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        k = min(args.cpu_models, M)
        tot_s, tot_ev, thr, kind = 0.0, 0, 1, "reference"
        for m in models[:k]:
            r = reference_timing(m, int(m.nt))
            if r is None:
                r, kind = oracle_timing(m, int(m.nt)), "port"
            tot_s += r[0]
            tot_ev += r[1]
            thr = r[2]
        out["cpu_baseline"] = {"value": k / tot_s, "unit": "trajectories/s", "cores": thr, "kind": kind,
                               "sample": f"the first {k} of the same models run one after another, as the "
                                         f"reference runs trajectories: {tot_ev} events in {tot_s:.2f} s",
                               "events_per_s": tot_ev / tot_s}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


# The dam break's time step: at the configured 1e-4 the reference engine and
# the oracle both turn non-finite by step 12 (above the acoustic limit
# 0.25 h / c0 = 8e-5), and at 5e-5 the density triples within 200 steps;
# at 2.5e-5 it stays within 10% of rho0.
DAMBREAK_DT = 2.5e-5


def bench_dambreak(args):
    """The 2D SDPD dam break with transport: 200x200 fluid particles at
    spacing 0.0025 inside a 2-particle-thick fixed wall around [0,2]x[0,1.5],
    h = 1.3 spacing, dt = DAMBREAK_DT, A->0 on 1000 molecules.  FP64 pair
    forces (the reference's arithmetic) for `value`, the FP32 pair-force
    option timed beside it."""
    from spatialpy_amd import workloads, _abi
    from spatialpy_amd.engine import Engine
    steps = args.steps if args.steps is not None else 400
    warm = args.warmup if args.warmup is not None else 20
    total = warm + steps

    def timed(fp32):
        sa = workloads.dam_break_2d(nt=total + 1, out_every=total + 1, dt=DAMBREAK_DT)
        if fp32:
            sa.flags |= _abi.FLAG_FP32_FORCES
        eng = Engine(sa)
        eng.reset(seed=12345)
        eng.advance(warm)
        ph0, st0 = eng.phase_times(), eng.stats()
        t0 = time.perf_counter()
        eng.advance(steps)
        sec = time.perf_counter() - t0
        ph1, st1 = eng.phase_times(), eng.stats()
        eng.close()
        return sa, sec, {k: ph1[k][0] - ph0[k][0] for k in ph1}, (st1.reactions - st0.reactions,
                                                                  st1.diffusions - st0.diffusions)

    sa, sec, ph, (n_re, n_di) = timed(False)
    _, sec32, ph32, _ = timed(True)
    n = sa.n
    E = mean_degree(sa.x, sa.h)
    B = algorithmic_bytes_per_particle_step(2, sa.n_species, E)
    hbm_ms = ph["search"] + ph["sdpd"] + ph["forces"] + ph["rates_queue"] + ph["bin_tree"]
    achieved = B * n * steps / (hbm_ms * 1e-3) / 1e9 if hbm_ms > 0 else 0.0
    particle_roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernels": "binning + fixed-radius search + D_ij + SDPD substeps + rates/queue build",
                     "algorithmic_bytes_per_particle_step": round(B, 1), "mean_degree": round(E, 2),
                     "ms_per_step": round(hbm_ms / steps, 4)}
    ssa_roof = ssa_roofline(sa.n_species, sa.n_reactions, E, n_re, n_di, ph["ssa"],
                            "k_ssa (event loop; latency-bound, one wavefront per trajectory)")
    ssa_roof["ms_per_step"] = round(ph["ssa"] / steps, 4)
    out = {
        "metric": "particle-steps/s", "value": n * steps / sec, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": warm, "ms_per_step": 1e3 * sec / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D SDPD dam break with transport: 200x200 fluid particles at spacing 0.0025 and a "
                               f"2-particle-thick fixed wall around [0,2]x[0,1.5] ({n} particles), h = 1.3 spacing, "
                               f"rho0 = 1, c0 = 10, nu = 1e-3, gravity -1, dt = {DAMBREAK_DT} (1e-4 turns non-finite by "
                               f"step 12 in the reference engine too), A->0 k = 0.05 on 1000 molecules; "
                               f"FP64 pair forces", "particles": n, "dt": DAMBREAK_DT},
        "phase_ms_per_step": {k: round(v / steps, 4) for k, v in ph.items()},
        "fp32_forces": {"value": n * steps / sec32, "ms_per_step": 1e3 * sec32 / steps,
                        "forces_ms_per_step": round(ph32["forces"] / steps, 4),
                        "note": "FP32 pair terms from FP64-formed dx and dv, FP64 sums (within 1e-5 of the FP64 "
                                "oracle in tests/test_gpu_forces_fp32.py)"},
        "ssa_events_per_s": (n_re + n_di) / sec,
        "dominant": "k_ssa" if ph["ssa"] > hbm_ms else "particle_path",
        "roofline": ssa_roof if ph["ssa"] > hbm_ms else particle_roof,
        "roofline_particle_path": particle_roof,
        "roofline_ssa": ssa_roof,
    }
    if not args.no_cpu_baseline:
        k = args.cpu_steps if args.cpu_steps else 20
        csa = workloads.dam_break_2d(nt=k + 1, out_every=k + 1, dt=DAMBREAK_DT)
        r = reference_timing(csa, k)
        kind = "reference"
        if r is None:
            r, kind = oracle_timing(csa, k), "port"
        out["cpu_baseline"] = {"value": n * k / r[0], "unit": "particle-steps/s", "cores": r[2], "kind": kind,
                               "sample": f"the same dam break, {k} steps from t = 0: {r[0]:.2f} s"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


def _progress(msg):
    print(msg, file=sys.stderr, flush=True)


def bench_cylinder(args):
    """The 1M-particle 3D SDPD + sSSA cylinder: `settle` untimed steps, W
    warm-up steps, K timed steps, then the rest of the 2,000-step run."""
    from spatialpy_amd.engine import Engine
    from spatialpy_amd import _abi
    K = 20 if args.steps is None else args.steps
    W = 5 if args.warmup is None else args.warmup
    settle = args.settle
    cpu_steps = 2 if args.cpu_steps is None else args.cpu_steps
    first = settle + W
    if first + K > CYL_RUN_STEPS:
        print(f"the cylinder run is {CYL_RUN_STEPS} steps (settle {settle} + warmup {W} + steps {K} = "
              f"{first + K}); choose a window inside it", file=sys.stderr)
        sys.exit(2)
    sa = build_workload(args.n, CYL_RUN_STEPS + 1)
    if args.ann_order:
        sa.flags |= _abi.FLAG_ANN_ORDER_MOVING
    n = sa.n
    eng = Engine(sa)
    eng.reset(seed=12345)
    t_run = time.perf_counter()
    eng.advance(settle)
    eng.advance(W)
    pre_s = time.perf_counter() - t_run
    _progress(f"cylinder: {n} particles, steps 0..{first - 1} in {pre_s:.2f} s")
    # untimed: the state at the window's first step (the CPU sample starts
    # from it) and the neighbour count per particle there
    state = None if args.no_cpu_baseline else eng.state()
    rp = eng.neighbors()[0]
    E = float(np.diff(rp).mean())
    st0, ph0 = eng.stats(), eng.phase_times()
    t0 = time.perf_counter()
    eng.advance(K)
    t1 = time.perf_counter()
    st1, ph1 = eng.stats(), eng.phase_times()
    sec = t1 - t0
    value = n * K / sec
    events = (st1.reactions + st1.diffusions) - (st0.reactions + st0.diffusions)
    dph = {k: (ph1[k][0] - ph0[k][0], ph1[k][1] - ph0[k][1]) for k in ph1}
    _progress(f"timed steps {first}..{first + K - 1}: {1e3 * sec / K:.4f} ms/step")
    full = None
    if not args.no_full_run:
        t2 = time.perf_counter()
        eng.advance(CYL_RUN_STEPS - (first + K))
        rest_s = time.perf_counter() - t2
        _progress(f"rest of the run: {rest_s:.2f} s")
        st2 = eng.stats()
        run_s = pre_s + sec + rest_s
        full = {"steps": CYL_RUN_STEPS, "seconds": round(run_s, 3),
                "particle_steps_per_s": n * CYL_RUN_STEPS / run_s,
                "ssa_events": int(st2.reactions + st2.diffusions),
                "ssa_events_per_s": (st2.reactions + st2.diffusions) / run_s,
                "note": "the whole configured run, steps 0..1999 (host-side state and neighbour readbacks "
                        "before the window excluded)"}

    # roofline of the HBM-bound particle path (binning, search, SDPD
    # substeps, rates and queue build)
    B = algorithmic_bytes_per_particle_step(3, sa.n_species, E)
    hbm_ms = dph["search"][0] + dph["sdpd"][0] + dph["forces"][0] + dph["rates_queue"][0] + dph["bin_tree"][0]
    achieved = B * n * K / (hbm_ms * 1e-3) / 1e9 if hbm_ms > 0 else 0.0
    traffic = measured_traffic()
    particle_roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic["particle_path"],
                     "kernels": "binning + fixed-radius search + D_ij + SDPD substeps + rates/queue build",
                     "algorithmic_bytes_per_particle_step": round(B, 1), "mean_degree": round(E, 3),
                     "ms_per_step": round(hbm_ms / K, 4)}
    # the SSA event loop: one dependent chain per trajectory; its algorithmic
    # bytes are the voxel states an event reads and writes
    ssa_ms = dph["ssa"][0]
    n_re = st1.reactions - st0.reactions
    n_di = st1.diffusions - st0.diffusions
    ssa_roof = ssa_roofline(sa.n_species, sa.n_reactions, E, n_re, n_di, ssa_ms,
                            "k_ssa (event loop; latency-bound, one wavefront per trajectory)")
    ssa_roof["traffic"] = traffic["k_ssa"]
    ssa_roof["ms_per_step"] = round(ssa_ms / K, 4)
    dominant = "k_ssa" if ssa_ms > hbm_ms else "particle_path"
    roofline = ssa_roof if dominant == "k_ssa" else particle_roof
    out = {
        "metric": "particle-steps/s", "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": 1e3 * sec / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D SDPD+sSSA cylinder r=1 L=4, {n} particles, h=1.3*spacing, c0=20, "
                               f"nu=1e-2, inflow 1.0, A+B->C k=1e-2 with {n} A + {n} B, dt={sa.dt:g}, "
                               f"plain SDPD (no surrogate), seeded GPU stream; timed steps {first}..{first + K - 1} "
                               f"of the {CYL_RUN_STEPS}-step run ({settle} settle + {W} warm-up steps first)",
                   "particles": n, "dt": sa.dt, "run_steps": CYL_RUN_STEPS, "first_timed_step": first,
                   "ann_order_moving": bool(args.ann_order)},
        "ssa_events_per_s": events / sec,
        "ssa_events": int(events),
        "phase_ms_per_step": {k: round(v[0] / K, 4) for k, v in dph.items()},
        "dominant": dominant,
        "roofline": roofline,
        "roofline_particle_path": particle_roof,
        "roofline_ssa": ssa_roof,
    }
    if full is not None:
        out["full_run"] = full
    eng.close()
    if not args.no_fp32_leg:
        # the same window with FP32 pair terms (FP64 integration and sums):
        # the north-star form of the pair forces, beside the reference's
        # FP64 arithmetic that value is measured with
        import copy
        sa32 = copy.copy(sa)
        sa32.flags = sa.flags | _abi.FLAG_FP32_FORCES
        e32 = Engine(sa32)
        e32.reset(seed=12345)
        e32.advance(first)
        e32.state()
        p0 = e32.phase_times()
        t = time.perf_counter()
        e32.advance(K)
        s32 = time.perf_counter() - t
        p1 = e32.phase_times()
        e32.close()
        _progress(f"fp32 pair forces: {1e3 * s32 / K:.4f} ms/step")
        out["fp32_forces"] = {"value": n * K / s32, "ms_per_step": round(1e3 * s32 / K, 4),
                              "phase_ms_per_step": {k: round((p1[k][0] - p0[k][0]) / K, 4) for k in p1},
                              "note": "same run and window, FP32 pair terms from FP64-formed dx and dv with FP64 "
                                      "sums and FP64 integration (within 1e-5 of the FP64 oracle in "
                                      "tests/test_gpu_forces_fp32.py); value above keeps the reference's FP64 "
                                      "arithmetic"}
    if not args.no_cpu_baseline:
        _progress(f"cpu baseline: {cpu_steps} reference steps from step {first}")
        out["cpu_baseline"] = cpu_baseline(sa, state, first, cpu_steps)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    if not args.no_surrogate:
        from spatialpy_amd.core.errors import SimulationError
        try:
            out["surrogate"] = surrogate_leg(args, sa, E, 1e3 * sec / K, dph["forces"][0] / K)
        except SimulationError as ex:
            # the plain line stands; the leg's failure is reported with it
            out["surrogate"] = {"error": str(ex)}
        if "ssa_events_per_s" in out["cpu_baseline"]:
            out["ssa_speedup_vs_cpu"] = out["ssa_events_per_s"] / out["cpu_baseline"]["ssa_events_per_s"]
    print(json.dumps(out))


def _short_kernel(name):
    import re
    return re.sub(r"\(.*", "", name).replace("void ", "").strip()


def profile_summary(dbs, steps, calibration_mib):
    """Summarise rocprofv3 result databases (rocpd SQLite) of this script.

    A database of a cylinder run covers every dispatch of the run; the
    window is its last `steps` event-loop steps (from the end of the
    (steps+1)-th last k_ssa launch to the end of the last one), i.e. the
    timed window of `bench.py --no-full-run`.  Per kernel: dispatches in the
    window, mean duration and, for counter databases, the mean value of each
    counter per dispatch.  A database without k_ssa launches is taken as the
    copy calibration (`--copy-calibration MIB`): its copy kernel's FETCH_SIZE
    and WRITE_SIZE per KiB copied give the factors that turn those counters
    into HBM bytes, and the HBM traffic per step and per k_ssa launch is
    printed with them."""
    import collections
    import sqlite3
    tables, scale = [], {}
    for db in dbs:
        c = sqlite3.connect(db)
        views = {r[0] for r in c.execute("select name from sqlite_master where type in ('view', 'table')")}
        rows = []
        if "kernels" in views:
            rows = c.execute("select dispatch_id, name, start, end, duration from kernels").fetchall()
        if not rows and "counters_collection" in views:
            rows = c.execute("select dispatch_id, min(kernel_name), min(start), min(end), min(duration) "
                             "from counters_collection group by dispatch_id").fetchall()
        rows.sort(key=lambda r: r[2])
        counters = collections.defaultdict(dict)
        if "counters_collection" in views:
            for d, name, val in c.execute("select dispatch_id, counter_name, sum(value) from counters_collection "
                                          "group by dispatch_id, counter_name"):
                counters[d][name] = float(val)
        ssa = [r for r in rows if _short_kernel(r[1]).startswith("ssa::k_ssa")]
        calib = not ssa
        if calib:
            win = rows
        else:
            if len(ssa) <= steps:
                raise SystemExit(f"{db}: {len(ssa)} k_ssa launches, need more than {steps}")
            w0, w1 = ssa[-steps - 1][3], ssa[-1][3]
            win = [r for r in rows if r[2] >= w0 and r[3] <= w1]
        agg = collections.OrderedDict()
        for r in sorted(win, key=lambda r: -r[4]):
            a = agg.setdefault(_short_kernel(r[1]), {"n": 0, "ns": 0.0, "cnt": collections.defaultdict(float)})
            a["n"] += 1
            a["ns"] += r[4]
            for k, v in counters.get(r[0], {}).items():
                a["cnt"][k] += v
        names = sorted({k for a in agg.values() for k in a["cnt"]})
        print(f"# {db}: {len(rows)} dispatches; " +
              ("copy calibration (whole run)" if calib else f"window = last {steps} steps, {len(win)} dispatches"))
        hdr = f"{'kernel':60s} {'calls':>6s} {'avg us':>10s}" + "".join(f" {n + ' KiB':>18s}" for n in names)
        print(hdr)
        for k, a in sorted(agg.items(), key=lambda kv: -kv[1]["ns"]):
            print(f"{k[:60]:60s} {a['n']:6d} {a['ns'] / a['n'] / 1e3:10.2f}" +
                  "".join(f" {a['cnt'].get(n, 0.0) / a['n']:18.1f}" for n in names))
        if calib:
            cp = [a for k, a in agg.items() if "copyBuffer" in k or "copy" in k.lower()]
            if cp:
                a = max(cp, key=lambda a: a["ns"])
                for n in names:
                    per = a["cnt"].get(n, 0.0) / a["n"]
                    if per > 0:
                        scale[n] = (calibration_mib << 10) / per
                        print(f"calibration {n}: {per:.0f} KiB per dispatch for {calibration_mib << 10} KiB copied "
                              f"-> x{scale[n]:.3f}")
        else:
            tables.append((agg, names))
        print()
    if scale:
        tot = collections.defaultdict(float)
        for agg, names in tables:
            for k, a in agg.items():
                for n in names:
                    if n in scale:
                        tot[("ssa" if k.startswith("ssa::k_ssa") else "pp", n)] += a["cnt"].get(n, 0.0) * scale[n] * 1024
        pp = sum(v for (g, _), v in tot.items() if g == "pp") / steps
        ks = sum(v for (g, _), v in tot.items() if g == "ssa") / steps
        print(f"scaled HBM bytes per step, particle path (every kernel but k_ssa): {pp:.4g}")
        print(f"particle_path traffic_bytes_per_step={pp:.6g}")
        print(f"k_ssa traffic_bytes_per_launch={ks:.6g}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("cylinder", "static3d", "ensemble", "dambreak"), default="cylinder",
                    help="cylinder: the 1M-particle moving SDPD+sSSA run (the headline line); static3d: the "
                         "64k-voxel static A+B<->C case, one trajectory and a seed batch; ensemble: uneven "
                         "gene-network models in one launch; dambreak: the 2D SDPD dam break with walls")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (cylinder 20, static3d 2, dambreak 400)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (cylinder 5 after the settle "
                                                              "steps, static3d 0, dambreak 20)")
    ap.add_argument("--settle", type=int, default=CYL_SETTLE,
                    help="cylinder: untimed steps before the warm-up (the timed window starts at step "
                         "settle + warmup of the 2,000-step run)")
    ap.add_argument("--n", type=int, default=1_000_000, help="particles in the cylinder")
    ap.add_argument("--batch", type=int, default=256, help="static3d: trajectories in the seed batch")
    ap.add_argument("--models", type=int, default=1024, help="ensemble: number of models")
    ap.add_argument("--t-end", type=int, default=10, help="ensemble: simulated time (unit steps)")
    ap.add_argument("--cpu-models", type=int, default=8, help="ensemble: models in the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=None, help="steps of the CPU baseline sample (cylinder 2, "
                                                                   "dambreak 20)")
    ap.add_argument("--no-full-run", action="store_true",
                    help="cylinder: stop after the timed window instead of finishing the 2,000-step run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32-leg", action="store_true",
                    help="cylinder: skip timing the same window with FP32 pair terms")
    ap.add_argument("--ann-order", action="store_true", help="ANN tie order on the moving domain")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate leg (training on the plain run, the run with the network attached)")
    ap.add_argument("--surrogate-epochs", type=int, default=150, help="Adam epochs over the recorded samples")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device with torch (the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--profile-summary", nargs="+", default=None, metavar="DB",
                    help="only summarise rocprofv3 result databases of this script (per-kernel means over "
                         "the last --steps steps; a copy-calibration database scales FETCH_SIZE/WRITE_SIZE "
                         "into HBM bytes) and exit; needs no GPU")
    ap.add_argument("--calibration-mib", type=int, default=1024,
                    help="MiB copied by the --copy-calibration run summarised with --profile-summary")
    args = ap.parse_args()
    if args.profile_summary:
        return profile_summary(args.profile_summary, 20 if args.steps is None else args.steps,
                               args.calibration_mib)
    if args.copy_calibration:
        import torch
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nb, "dispatches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build has no "
              "multi-GPU path (domain slabs with halo exchange, and ensembles split by trajectory, are "
              "design only)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "static3d":
        return bench_static3d(args)
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "dambreak":
        return bench_dambreak(args)
    return bench_cylinder(args)


if __name__ == "__main__":
    main()
