"""Time step of the 1M-particle cylinder, pinned on the oracle.

The cylinder (r = 1, L = 4, h = 1.3 x spacing, c0 = 20, nu = 1e-2, inflow
1.0) uses dt = 1.25e-4 at 1M particles, scaled with the lattice spacing at
other sizes (the same Courant number c0 dt / h).  At 2.5e-4 the run turns
non-finite after about 20 steps in the reference engine and the oracle
alike; at the chosen dt the FP64 oracle at 20k particles (the reference's
step loop, simulate_threads.cpp:232-281, for all 2,000 configured steps)
keeps every position, velocity and density finite and the density within a
factor 3 of rho0.  The 1M run itself is checked on the GPU
(tests/test_gpu_fullsize.py)."""
import numpy as np
import pytest

from spatialpy_amd import workloads


@pytest.mark.slow
def test_cylinder_dt_keeps_the_oracle_finite_for_2000_steps(oracle):
    steps = 2000
    sa = workloads.cylinder_3d(n_target=20_000, nt=steps + 1, nA=0, nB=0)
    assert sa.dt == pytest.approx(workloads.CYLINDER_DT_1M * (1_000_000 / 20_000) ** (1 / 3), rel=1e-12)
    o = oracle.Oracle(sa)
    o.reset(seed=1)
    lo, hi = np.inf, 0.0
    for _ in range(0, steps, 250):
        o.advance(250)      # raises on the engine's own NaN check (particle.cpp:88-126)
        st = o.state()
        assert np.isfinite(st["x"]).all() and np.isfinite(st["v"]).all() and np.isfinite(st["rho"]).all()
        lo, hi = min(lo, st["rho"].min()), max(hi, st["rho"].max())
    assert o.stats().steps == steps
    assert 1 / 3 < lo and hi < 3.0, (lo, hi)


def test_cylinder_dt_scales_with_spacing():
    assert workloads.cylinder_dt(1_000_000) == workloads.CYLINDER_DT_1M
    # the acoustic limit 0.25 h / c0 at 1M particles is about 3 times the step
    s = (np.pi * 4 / 1e6) ** (1 / 3)
    assert workloads.CYLINDER_DT_1M < 0.25 * 1.3 * s / 20.0 / 3
