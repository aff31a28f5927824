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
