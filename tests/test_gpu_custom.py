"""GPU tests of custom propensities: the device expression VM against the
C++ text compiled by g++ (tests/golden/expr_cases.json), and whole runs of
models with custom propensities, data functions and type restrictions
against the reference engine that ran g++-compiled generated functions
(tests/golden/c6_*.npz, replay mode), batched and through the Solver API."""
import json
import math
import os

import numpy as np
import pytest

import spatialpy_amd as sp
from spatialpy_amd import expression as E, workloads
from spatialpy_amd.engine import Engine, mt19937_64_stream, vm_eval

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _ulp(v):
    return math.ulp(v) if math.isfinite(v) else 0.0


def test_device_vm_matches_gxx():
    """Bit-exact for + - * /, int arithmetic, comparisons, sqrt, fabs, floor,
    ceil, trunc, fmod, copysign, remainder; other math functions (device
    libm) within 16 ulp or 1e-15 absolute."""
    with open(os.path.join(GOLDEN, "expr_cases.json")) as f:
        fx = json.load(f)
    ns = E.model_namespace(fx["species"], list(fx["params"]), fx["data_functions"])
    cparams = {f"P{i}": v for i, v in enumerate(fx["params"].values())}
    n_exact = n_tol = 0
    for c in fx["cases"]:
        if c["stage"] == "cpp":
            n_x, n_df, params = len(c["args"]), 0, {}
        else:
            n_x, n_df, params = len(fx["species"]), len(fx["data_functions"]), cparams
        p = E.Program()
        start = p.add(E.parse_c(c["cpp"], params, n_x, n_df, c["x_type"]))
        code, consts = p.arrays()
        rows = np.array(c["inputs"], dtype=np.float64)
        got = vm_eval(code, consts, start, rows[:, :n_x], rows[:, n_x], rows[:, n_x + 1],
                      rows[:, n_x + 2:] if n_df else None, x_is_double=c["x_type"] == E.D)
        want = np.array([float.fromhex(h) for h in c["expect_hex"]])
        if c["exact"]:
            same = (got.view(np.uint64) == want.view(np.uint64)) | (np.isnan(got) & np.isnan(want))
            assert same.all(), (c["name"], rows[~same][:3], got[~same][:3], want[~same][:3])
            n_exact += got.size
        else:
            for g, w, r in zip(got, want, rows):
                assert (math.isnan(g) and math.isnan(w)) or abs(g - w) <= max(16 * _ulp(w), 1e-15), \
                    (c["name"], r, g, w)
            n_tol += got.size
    print(f"[vm] {n_exact} exact evaluations, {n_tol} within tolerance", flush=True)
    assert n_exact > 1500 and n_tol > 200


def _fixture(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    sa = getattr(workloads, meta["gen"])(**meta["params"])
    assert np.array_equal(sa.x, z["x"]) and np.array_equal(sa.u0, z["u0"])
    return sa, meta, z


def test_custom_propensities_replay_matches_reference_engine():
    """Integer arithmetic, a data function, type restrictions and sqrt: the
    reference engine's counters, every count snapshot and the deterministic
    concentrations, bit for bit."""
    sa, meta, z = _fixture("c6_custom_chemistry")
    words, expo = mt19937_64_stream(meta["seed"], 1_000_000)
    eng = Engine(sa)
    res = eng.run(words=words, expo=expo)
    st = eng.stats()
    assert [st.reactions, st.diffusions] == z["counters"].tolist()
    assert np.array_equal(res["xx"].ravel(), z["xx"])
    assert np.array_equal(res["C"].ravel().view(np.uint64), z["snap_C"].view(np.uint64))


def test_custom_propensities_with_device_libm_replay():
    """exp(-t/4), pow: the device libm differs from the host's by ulps; the
    event sequence (counters, counts) still matches the reference run, and
    the deterministic concentrations agree to 1e-9 relative (NaN where the
    reference has NaN)."""
    sa, meta, z = _fixture("c6_custom_chemistry_libm")
    words, expo = mt19937_64_stream(meta["seed"], 1_000_000)
    eng = Engine(sa)
    res = eng.run(words=words, expo=expo)
    st = eng.stats()
    assert [st.reactions, st.diffusions] == z["counters"].tolist()
    assert np.array_equal(res["xx"].ravel(), z["xx"])
    c, ref = res["C"].ravel(), z["snap_C"]
    assert np.array_equal(np.isnan(c), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.allclose(c[ok], ref[ok], rtol=1e-9, atol=1e-12)


def test_custom_propensities_batch_equals_single_runs():
    sa = workloads.custom_chemistry_2d(seed=5)
    eng = Engine(sa)
    seeds = [101, 202, 303]
    b = eng.run_batch(seeds, fields=("xx",))
    for i, s in enumerate(seeds):
        r = eng.run(seed=s, fields=("xx",))
        assert np.array_equal(b["xx"][i], r["xx"]), s


def test_solver_runs_a_model_with_custom_propensities_and_data_functions():
    """The Model API end to end: a Hill-type production switched by a data
    function, int-arithmetic degradation; counts stay non-negative and the
    data-function region produces more."""
    class Left(sp.DataFunction):
        def map(self, point):
            return 1.0 if point[0] < 0.5 else 0.0

    m = sp.Model("custom")
    m.add_domain(sp.Domain.create_2D_domain(xlim=(0, 1), ylim=(0, 1), numx=10, numy=10, type_id=1, fixed=True))
    m.add_species(sp.Species("P", 0.001))
    m.add_parameter([sp.Parameter("kp", expression=50.0), sp.Parameter("Kh", expression=4.0)])
    m.add_data_function(Left(name="left"))
    m.add_reaction([
        sp.Reaction("make", reactants={}, products={"P": 1}, propensity_function="kp*vol*left*Kh^2/(Kh^2+P^2)"),
        sp.Reaction("lose", reactants={"P": 1}, products={}, propensity_function="P/2"),
    ])
    m.timespan(sp.TimeSpan.linspace(t=4, num_points=5, timestep_size=0.5))
    res = m.run(seed=11)
    P = res.get_species("P", -1)
    x = m.domain.vertices[:, 0]
    assert P.min() >= 0 and P[x < 0.5].sum() > 5 * max(1, P[x >= 0.5].sum())
