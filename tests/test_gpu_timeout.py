"""The engine's time limit (ssasdpd_set_timeout), the counterpart of the
reference's per-trajectory Popen.wait(timeout) (solver.py:574-586).

* A limit that has already passed stops a run before its first step: the
  call reports TIMEOUT, the state is the initial one, and Solver.run marks
  every trajectory timed out and still returns all of them.
* A limit that is not reached changes nothing: static domains then run in
  launches of at most 64 steps that end at snapshot steps, and their count
  snapshots equal the one-launch run's bit for bit (replay and seeded
  streams, single and batched runs); moving domains are checked too."""
import numpy as np
import pytest

from spatialpy_amd import _abi, workloads
from spatialpy_amd.core.errors import SimulationError
from spatialpy_amd.engine import Engine, mt19937_64_stream

pytestmark = pytest.mark.gpu

FIELDS = ("xx", "C", "x", "v", "rho")


def test_passed_limit_stops_before_the_first_step():
    sa = workloads.birth_death_2d(seed=1)
    eng = Engine(sa)
    eng.reset(seed=3)
    eng.set_timeout(1e-9)
    with pytest.raises(SimulationError) as e:
        eng.advance(5)
    assert e.value.code == _abi.TIMEOUT
    assert eng.stats().steps == 0
    assert np.array_equal(eng.state()["xx"], sa.u0)
    res = eng.run(seed=3, fields=FIELDS, partial_ok=True)
    assert res["timeout"]
    eng.set_timeout(0)
    eng.advance(2)
    assert eng.stats().steps == 2


def test_solver_marks_timed_out_trajectories():
    from test_solver_api import birth_death_model
    import spatialpy_amd as sp
    sol = sp.Solver(birth_death_model())
    res = sol.run(number_of_trajectories=2, seed=40, timeout=1e-9)
    assert len(res) == 2 and all(r.timeout for r in res)
    res = sol.run(number_of_trajectories=1, seed=40, timeout=1e-9)
    assert res.timeout
    res = sol.run(number_of_trajectories=1, seed=40, timeout=600)
    assert not res.timeout


CASES = {
    # 100 steps, snapshots at 0 and 100: launches [1, 65) and [65, 100)
    "birth_death_2_outputs": lambda: workloads.birth_death_2d(seed=1, t_end=10, num_points=2, dt=0.1),
    # 100 steps, a snapshot every 10 steps: launches end at snapshot steps
    "birth_death_11_outputs": lambda: workloads.birth_death_2d(seed=1, t_end=10, num_points=11, dt=0.1),
    "static_rd": lambda: workloads.static_rd_3d(n_side=8, nA=800, nB=800, t_end=3, num_points=4, seed=2),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_unreached_limit_gives_the_same_run(case):
    sa = CASES[case]()
    eng = Engine(sa)
    words, expo = mt19937_64_stream(7, 4_000_000)
    plain_r = eng.run(words=words, expo=expo, fields=FIELDS)
    plain_s = eng.run(seed=9, fields=FIELDS)
    plain_b = eng.run_batch([9, 10], fields=("xx",))
    eng.set_timeout(600)
    seg_r = eng.run(words=words, expo=expo, fields=FIELDS, partial_ok=True)
    seg_s = eng.run(seed=9, fields=FIELDS, partial_ok=True)
    seg_b = eng.run_batch([9, 10], fields=("xx",), partial_ok=True)
    assert not seg_r["timeout"] and not seg_s["timeout"] and not seg_b["timeout"]
    for a, b in ((plain_r, seg_r), (plain_s, seg_s)):
        assert np.array_equal(a["steps"], b["steps"])
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), f
    assert np.array_equal(plain_b["xx"], seg_b["xx"])
    eng.reset(seed=9)
    eng.advance(sa.nt)
    st = eng.state()
    assert np.array_equal(st["xx"], plain_s["xx"][-1])


def test_unreached_limit_on_a_moving_domain():
    sa = workloads.dam_break_2d(nx=20, ny=20, spacing=0.025, n_mol=100, nt=45, out_every=15, dt=2e-4)
    eng = Engine(sa)
    a = eng.run(seed=4, fields=FIELDS)
    eng.set_timeout(600)
    b = eng.run(seed=4, fields=FIELDS, partial_ok=True)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
