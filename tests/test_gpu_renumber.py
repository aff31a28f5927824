"""Particle renumbering on moving domains (seeded mode): every 20 steps the
engine puts its particles in cell order of their positions and reports
everything in the caller's numbering.

The particle dynamics do not depend on the numbering except through the
order of exactly equal distances inside a list (the summation order of a
few pair terms), so a renumbered run follows the plain one to rounding;
neighbour sets are the same sets; snapshots, state, forces and lists come
back in the caller's numbering; and the chemistry, whose random decisions
do see the numbering, agrees in distribution (family-wise 95% bands)."""
import numpy as np
import pytest
from scipy import stats

from spatialpy_amd import _abi, workloads
from spatialpy_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _dam(nt=80, **kw):
    return workloads.dam_break_2d(nx=24, ny=24, spacing=0.025, n_mol=200, nt=nt, out_every=20, dt=2e-4, **kw)


def _cyl(nt=80):
    return workloads.cylinder_3d(n_target=20_000, nt=nt, out_every=20)


def _with_flag(sa, flag):
    import copy
    s2 = copy.copy(sa)
    s2.flags = sa.flags | flag
    return s2


@pytest.mark.parametrize("make", [_dam, _cyl])
def test_renumbered_run_follows_the_plain_run(make):
    sa = make()
    print(f"renumber: n={sa.n} nt={sa.nt}", flush=True)
    ren = Engine(sa)
    plain = Engine(_with_flag(sa, _abi.FLAG_NO_RENUMBER))
    fields = ("xx", "x", "v", "rho")
    r1 = ren.run(seed=3, fields=fields)
    r0 = plain.run(seed=3, fields=fields)
    for f in ("x", "v", "rho"):
        scale = np.abs(r0[f]).max()
        err = np.abs(r1[f] - r0[f]).max() / scale
        assert err < 1e-8, (f, err)
    # the same molecules (totals per species are numbering-free at t = 0 and
    # conserved up to the reactions; here diffusion and A+B->C alike keep
    # A + C and B + C)
    assert np.array_equal(r1["xx"][0], r0["xx"][0])
    t1 = r1["xx"].astype(np.int64).sum(axis=1)
    if sa.n_species >= 3:
        assert (t1[:, 0] + t1[:, 2] == t1[0, 0] + t1[0, 2]).all()
    # state, lists and forces in the caller's numbering after the run
    st1, st0 = ren.state(), plain.state()
    assert np.abs(st1["x"] - st0["x"]).max() <= 1e-8 * np.abs(st0["x"]).max()
    rp1, i1, d1, _, _ = ren.neighbors()
    rp0, i0, d0, _, _ = plain.neighbors()
    assert np.array_equal(rp1, rp0)
    for i in range(0, sa.n, 7):
        a = sorted(i1[rp1[i]:rp1[i + 1]])
        b = sorted(i0[rp0[i]:rp0[i + 1]])
        assert a == b, i
    F1, _, _ = ren.eval_forces()
    F0, _, _ = plain.eval_forces()
    assert np.abs(F1 - F0).max() <= 1e-6 * max(np.abs(F0).max(), 1e-300)
    # the numbering is a function of the positions (ties inside a cell go by
    # index), so a seed reproduces its renumbered trajectory bit for bit
    r2 = ren.run(seed=3, fields=fields)
    assert np.array_equal(r2["xx"], r1["xx"])
    assert np.array_equal(r2["x"], r1["x"])


@pytest.mark.parametrize("case", ["dam_break", "cylinder20k"])
def test_renumber_beside_the_event_loop_equals_the_serial_one(case):
    """advance() renumbers inside its look-ahead (the next step's particle
    phase on a second stream, beside this step's event loop: the particles
    are permuted there, the counts after the event loop); run() renumbers
    between steps.  Same seed, same state bit for bit.  advance(20) ends a
    call on the step before a renumber (no look-ahead there), so step 20 is
    renumbered between steps and step 40 beside the event loop."""
    if case == "dam_break":
        sa = workloads.dam_break_2d(nx=24, ny=24, spacing=0.025, n_mol=200, nt=45, out_every=45, dt=2e-4)
    else:
        sa = workloads.cylinder_3d(n_target=20_000, nt=45, out_every=45)
    print(f"renumber beside the event loop: {case} n={sa.n}", flush=True)
    fields = ("xx", "x", "v", "rho")
    ref = Engine(sa).run(seed=11, fields=fields)
    eng = Engine(sa)
    eng.reset(seed=11)
    eng.advance(20)
    eng.advance(25)
    st = eng.state()
    assert np.array_equal(st["xx"], ref["xx"][-1])
    for f in ("x", "v", "rho"):
        assert np.array_equal(st[f].view(np.uint64), ref[f][-1].view(np.uint64)), f
    # and in one call (renumbers at 20 and 40 beside the event loop)
    eng.reset(seed=11)
    eng.advance(45)
    st2 = eng.state()
    assert np.array_equal(st2["xx"], ref["xx"][-1])
    assert np.array_equal(st2["x"].view(np.uint64), ref["x"][-1].view(np.uint64))


def test_renumbered_chemistry_agrees_in_distribution():
    """The moving dam break's species (diffusion and its reaction): 64
    renumbered and 64 plain trajectories (one batch each), totals and the
    count in the left half of the fluid at every snapshot, compared with
    two-sample KS and mean tests at a family-wise 95% level."""
    sa = _dam(nt=101)
    ren = Engine(sa)
    plain = Engine(_with_flag(sa, _abi.FLAG_NO_RENUMBER))
    B = 64
    a = ren.run_batch(np.arange(B) + 1, fields=("x",))
    b = plain.run_batch(np.arange(B) + 1001, fields=("x",))
    assert np.abs(a["x"] - b["x"]).max() < 1e-8 * np.abs(b["x"]).max()
    xa, xb = a["xx"].astype(np.int64), b["xx"].astype(np.int64)
    K, S = xa.shape[1], xa.shape[3]
    x_last = b["x"]
    res = []
    for k in range(1, K):
        left = x_last[k][:, 0] < np.median(x_last[k][:, 0])
        for s in range(S):
            for name, fa, fb in (("total", xa[:, k, :, s].sum(1), xb[:, k, :, s].sum(1)),
                                 ("left", xa[:, k, left, s].sum(1), xb[:, k, left, s].sum(1))):
                if fa.std() == 0 and fb.std() == 0:
                    assert fa.mean() == fb.mean()
                    continue
                se = np.sqrt(fa.var(ddof=1) / B + fb.var(ddof=1) / B)
                res.append((k, s, name, "KS", stats.ks_2samp(fa, fb).pvalue))
                res.append((k, s, name, "mean", 2 * stats.norm.sf(abs(fa.mean() - fb.mean()) / se)))
    m = len(res)
    assert m > 0
    for r in res:
        assert r[-1] > 0.05 / m, (r, m)
