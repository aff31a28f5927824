"""Solver.run's timeout semantics (spatialpy/solvers/solver.py:547-605):
the limit applies to each trajectory; a trajectory past it keeps what it
produced and is marked result.timeout; later trajectories still run.  The
engine is replaced by a stand-in that reports the time limit on the first
trajectory only, so this runs without a GPU (the engine's own limit is
checked on the GPU in tests/test_gpu_timeout.py)."""
import numpy as np

from spatialpy_amd import workloads
from spatialpy_amd.solver import Solver


class _Engine:
    def __init__(self, sa):
        self.sa = sa
        self.calls = []
        self.timeout = None

    def set_timeout(self, s):
        self.timeout = s

    def run(self, seed=None, fields=(), partial_ok=False, **kw):
        self.calls.append((seed, self.timeout, partial_ok))
        k = 3
        d = {"xx": np.zeros((k, self.sa.n, 1), np.uint32), "steps": np.zeros(k, np.uint32)}
        d["timeout"] = len(self.calls) == 1
        return d


def test_timeout_marks_the_trajectory_and_later_ones_still_run():
    sa = workloads.birth_death_2d(seed=1)
    sol = Solver(model=object())
    sol.sa, sol.engine, sol.is_compiled = sa, _Engine(sa), True
    sol.model = type("M", (), {"tspan": None, "listOfSpecies": {"A": None}})()
    sa_species = sa.n_species
    assert sa_species == 1
    sol.exact_replay = True          # one trajectory at a time, as the reference
    sol._run_replay = lambda s: sol.engine.run(seed=s, partial_ok=True)
    res = sol.run(number_of_trajectories=3, seed=10, timeout=0.5)
    assert len(res) == 3
    assert [r.timeout for r in res] == [True, False, False]
    assert [c[0] for c in sol.engine.calls] == [10, 11, 12]
    assert all(c[1] == 0.5 for c in sol.engine.calls)
