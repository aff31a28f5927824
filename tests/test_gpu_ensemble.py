"""Ensembles of different static models in one persistent launch
(ssasdpd_ensemble_*): uneven domain sizes, each model its own gene-network
chemistry.  Every model's trajectory must equal the single-model engine run
with the same seed bit for bit (counts at every snapshot), for any
event-tree width, and the autotuner must not change results.  The reference
runs such a set as one process per trajectory (spatialpy/solvers/
solver.py:547-605)."""
import numpy as np
import pytest

from spatialpy_amd import workloads
from spatialpy_amd.engine import Engine
from spatialpy_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu

SIZES = [1500, 300, 2700, 800, 1200]


def _models(t_end=4):
    return [workloads.gene_network_3d(n_target=n, seed=10 + i, t_end=t_end) for i, n in enumerate(SIZES)]


def test_ensemble_matches_single_model_runs():
    models = _models()
    seeds = [101, 202, 303, 404, 505]
    ens = Ensemble(models)
    totals, counts = ens.run(seeds, counts=True)
    assert totals.shape == (len(models), ens.K, 5)
    st = ens.stats()
    assert st.reactions + st.diffusions > 0
    for m, (sa, s) in enumerate(zip(models, seeds)):
        single = Engine(sa).run(seed=s, fields=("xx",))["xx"]
        assert np.array_equal(counts[m], single), f"model {m} (n={sa.n}) differs from its single run"
        assert np.array_equal(totals[m], single.sum(axis=1)), f"model {m}: totals"


def test_ensemble_tree_width_and_autotune_do_not_change_results():
    models = _models(t_end=3)
    seeds = [7, 8, 9, 10, 11]
    ens = Ensemble(models)
    ref = ens.run(seeds)
    for w in (1, 2, 4):
        ens.set_tree_width(w)
        assert np.array_equal(ens.run(seeds), ref), f"tree width {w}"
    chosen, times = ens.autotune(seeds, probe_steps=2)
    assert chosen in times and len(times) >= 3
    assert np.array_equal(ens.run(seeds), ref)


def test_ensemble_rejects_mixed_schedules_and_moving_domains():
    from spatialpy_amd.core.errors import SimulationError
    a = workloads.gene_network_3d(n_target=300, seed=1, t_end=4)
    b = workloads.gene_network_3d(n_target=300, seed=2, t_end=5)
    with pytest.raises(SimulationError, match="share nt"):
        Ensemble([a, b])
    mv = workloads.dam_break_2d(nx=10, ny=10, spacing=0.05, n_mol=20, nt=4, out_every=2, dt=2e-4)
    with pytest.raises(SimulationError, match="static"):
        Ensemble([mv])


def test_gene_network_replay_matches_the_oracle_event_by_event(oracle):
    """A gene-network model (5 species, 6 mass-action reactions, a 2,744-voxel
    jittered lattice) in replay mode: the engine's event sequence, counters
    and count snapshots equal the FP64 oracle's (the reference's event loop,
    simulate_rdme.cpp:211-472) on the same std::mt19937_64 stream, and every
    one of the six reactions fires."""
    from spatialpy_amd.engine import mt19937_64_stream
    sa = workloads.gene_network_3d(n_target=3000, seed=21, t_end=6)
    cap = 200_000
    words, expo = mt19937_64_stream(5, 1_000_000)
    eng = Engine(sa, trace_capacity=cap)
    res = eng.run(words=words, expo=expo, fields=("xx",))
    ev, n = eng.trace(cap)
    o = oracle.Oracle(sa)
    o.set_trace(cap)
    ref = o.run(seed=5)
    oev, on = o.trace(cap)
    assert n == on and 1000 < n < cap
    assert np.array_equal(ev, oev)
    assert np.array_equal(res["xx"], ref["xx"])
    st, ost = eng.stats(), o.stats()
    assert (st.reactions, st.diffusions, st.rng_words_used) == (ost.reactions, ost.diffusions, ost.rng_words_used)
    fired = np.bincount(ev["which"][ev["kind"] == 0], minlength=6)
    assert (fired > 0).all(), fired


def test_configured_size_ensemble_equals_per_trajectory_runs():
    """The batched-ensemble configuration at its size: 64 gene-network models
    with N ~ U[8000, 64000] voxels in one launch (t = 0..2).  Per model:
    totals equal the plain single-model GPU run with the same seed (hence
    the ensemble mean and variance over trajectories equal the per-trajectory
    runs' at every output time), the first snapshot holds u0, and the gene
    count G_on + G_off -- conserved by all six reactions -- never changes."""
    rng = np.random.default_rng(7)
    sizes = rng.integers(8000, 64001, size=64)
    models = [workloads.gene_network_3d(n_target=int(nv), seed=500 + i, t_end=2) for i, nv in enumerate(sizes)]
    assert min(m.n for m in models) >= 6859 and max(m.n for m in models) <= 64000
    seeds = [9000 + i for i in range(len(models))]
    ens = Ensemble(models)
    totals = ens.run(seeds).astype(np.int64)
    st = ens.stats()
    assert st.reactions > 0 and st.diffusions > 0
    single = np.zeros_like(totals)
    for m, (sa, s) in enumerate(zip(models, seeds)):
        single[m] = Engine(sa).run(seed=s, fields=("xx",))["xx"].astype(np.int64).sum(axis=1)
        assert np.array_equal(totals[m, 0], sa.u0.astype(np.int64).sum(axis=0)), m
        genes = totals[m, :, 0] + totals[m, :, 1]
        assert (genes == genes[0]).all(), (m, genes)
    assert np.array_equal(totals, single)
    assert np.array_equal(totals.mean(axis=0), single.mean(axis=0))
    assert np.array_equal(totals.var(axis=0), single.var(axis=0))
