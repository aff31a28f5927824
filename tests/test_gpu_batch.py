"""Batched ensembles (ssasdpd_run_batch): every trajectory of a batch must be
bit-identical to the single run with the same seed, on static and moving
domains.  The reference runs the trajectories of an ensemble one after
another (spatialpy/solvers/solver.py:547-605); the batch shares the particle
state and runs all event loops in one launch per step."""
import numpy as np
import pytest

from spatialpy_amd import workloads
from spatialpy_amd.engine import Engine

pytestmark = pytest.mark.gpu

FIELDS = ("xx", "C", "x", "v", "rho")


def _check_batch(sa, seeds):
    eng = Engine(sa)
    batch = eng.run_batch(seeds, fields=FIELDS)
    for b, s in enumerate(seeds):
        single = eng.run(seed=s, fields=FIELDS)
        assert np.array_equal(batch["xx"][b], single["xx"]), f"trajectory {b} (seed {s}) differs"
        for f in ("x", "v", "rho", "C"):
            assert np.array_equal(batch[f], single[f]), f"shared field {f} differs from the single run"
    # distinct seeds give distinct trajectories
    assert not np.array_equal(batch["xx"][0, -1], batch["xx"][1, -1])
    return batch


def test_batch_birth_death_matches_single_runs():
    sa = workloads.birth_death_2d(seed=1)
    b = _check_batch(sa, [3, 4, 5, 1 << 40])
    assert b["xx"].shape == (4, 11, 121, 1)


def test_batch_static_rd_matches_single_runs():
    sa = workloads.static_rd_3d(n_side=12, nA=3000, nB=3000, t_end=3, num_points=4, seed=2)
    _check_batch(sa, [11, 12, 13])


def test_batch_moving_dam_break_matches_single_runs():
    sa = workloads.dam_break_2d(nx=20, ny=20, spacing=0.025, n_mol=100, nt=45, out_every=15, dt=2e-4)
    _check_batch(sa, [7, 8])


def test_batch_of_one_and_reuse_after_larger_batch():
    sa = workloads.birth_death_2d(seed=1)
    eng = Engine(sa)
    big = eng.run_batch([21, 22, 23], fields=("xx",))
    one = eng.run_batch([22], fields=("xx",))
    assert np.array_equal(one["xx"][0], big["xx"][1])


def test_solver_ensemble_uses_batch_and_matches_sequential():
    import spatialpy_amd as sp
    from test_solver_api import birth_death_model
    model = birth_death_model()
    res = model.run(number_of_trajectories=3, seed=40)
    assert len(res) == 3
    solo = sp.Solver(model).run(seed=41)
    assert np.array_equal(res[1].get_species("Rabbits"), solo.get_species("Rabbits"))
