"""The reference-engine harness (oracle/_ref, the reference's engine sources
linked with oracle/ref_harness.cpp) started mid-run.

bench.py times the reference engine over the same steps as the GPU's timed
window by handing it the state at the window's first step.  Pinned here:
the harness continued at step k from the oracle's state after step k - 1
(x, v, rho, F, Fbp, Frho, C, Q, counts) ends in the oracle's own state after
more steps, bit for bit -- positions, velocities and densities depend on
every carried quantity (take_step1, simulate.cpp:68-83)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from spatialpy_amd import lowering, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness_O0")


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="reference harness not built (needs the reference sources)")
def test_harness_continues_from_a_mid_run_state(oracle):
    sa = workloads.cylinder_3d(n_target=1500, nt=12, out_every=100, dt=1e-3, nA=300, nB=300)
    o = oracle.Oracle(sa)
    o.reset(seed=1)
    o.advance(5)
    st = o.state()
    o.advance(5)
    ref = o.state()
    with tempfile.TemporaryDirectory() as d:
        lowering.write_harness_dir(sa, d, seed=3, run_steps=5, state=st, start_step=5)
        r = subprocess.run([HARNESS, d], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-400:] + r.stderr[-400:]
        n = sa.n
        steps = np.fromfile(os.path.join(d, "out_steps.u32"), dtype=np.uint32)
        x = np.fromfile(os.path.join(d, "out_x.f64")).reshape(-1, n, 3)
        v = np.fromfile(os.path.join(d, "out_v.f64")).reshape(-1, n, 3)
        rho = np.fromfile(os.path.join(d, "out_rho.f64")).reshape(-1, n)
    assert steps[0] == 5 and steps[-1] == 10
    assert np.array_equal(x[0], st["x"])
    assert np.array_equal(x[-1], ref["x"]) and np.array_equal(v[-1], ref["v"]) and np.array_equal(rho[-1], ref["rho"])
