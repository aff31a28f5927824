"""GPU engine vs the FP64 oracle: bit-exact where the reference's arithmetic
is exact (neighbour sets, r^2 bits, list order, D_ij, rates, SSA events,
counts) and tolerance-checked for the SDPD forces."""
import numpy as np
import pytest

from spatialpy_amd import workloads
from spatialpy_amd.engine import Engine, fr_search, mt19937_64_stream

pytestmark = pytest.mark.gpu


def _lists_equal(a, b):
    rp1, i1, r1 = a
    rp2, i2, r2 = b
    return (np.array_equal(rp1, rp2) and np.array_equal(i1, i2)
            and np.array_equal(r1.view(np.uint64), r2.view(np.uint64)))


@pytest.mark.parametrize("case", ["c1_lattice", "c2_lattice", "jitter3d", "jitter2d", "mixed_time"])
def test_fr_search_matches_oracle_kdtree(case, oracle):
    rng = np.random.default_rng(3)
    if case == "c1_lattice":
        sa = workloads.birth_death_2d()
        pts, q, dim, h = sa.x, sa.x, 2, sa.h
    elif case == "c2_lattice":
        sa = workloads.static_rd_3d(n_side=14, nA=10, nB=10)
        pts, q, dim, h = sa.x, sa.x, 3, sa.h
    elif case == "jitter3d":
        pts, _ = workloads.jittered_lattice_3d(3000, seed=5)
        q, dim, h = pts, 3, 0.12
    elif case == "jitter2d":
        pts = rng.uniform(0, 1, size=(2000, 3))
        pts[:, 2] = 0
        q, dim, h = pts, 2, 0.05
    else:
        # moving-domain query: predicted positions against start-of-step tree
        sa = workloads.birth_death_2d()
        pts = sa.x
        q = pts + rng.normal(0, 1e-3, size=pts.shape) * np.array([1, 1, 0])
        dim, h = 2, sa.h
    gpu = fr_search(pts, q, dim, h)
    ref = oracle.fr_search(pts, q, dim, h)
    assert _lists_equal(gpu, ref)
    # the set and r^2 bits against brute force (independent of any tree)
    br = oracle.fr_brute(pts, q, dim, h)
    rp, idx, r2 = gpu
    for i in range(len(rp) - 1):
        s = slice(rp[i], rp[i + 1])
        a = sorted(zip(r2[s].tolist(), idx[s].tolist()))
        b = sorted(zip(br[2][br[0][i]:br[0][i + 1]].tolist(), br[1][br[0][i]:br[0][i + 1]].tolist()))
        assert a == b


def test_c1_replay_matches_oracle_event_by_event(oracle):
    sa = workloads.birth_death_2d(seed=1)
    words, expo = mt19937_64_stream(7, 1_200_000)
    eng = Engine(sa, trace_capacity=300_000)
    res = eng.run(words=words, expo=expo)
    st = eng.stats()
    ev, nev = eng.trace(300_000)
    o = oracle.Oracle(sa)
    o.set_trace(300_000)
    ref = o.run(seed=7)
    ost = o.stats()
    oev, onev = o.trace(300_000)
    assert (st.reactions, st.diffusions) == (ost.reactions, ost.diffusions)
    assert nev == onev
    assert np.array_equal(ev, oev)
    assert np.array_equal(res["xx"], ref["xx"])
    assert np.array_equal(res["C"], ref["C"])
    assert st.rng_words_used == ost.rng_words_used


@pytest.mark.parametrize("hfac", [4.7, 6.5])
def test_long_neighbour_rows_replay_matches_oracle(hfac, oracle):
    """Rows longer than one wavefront (about 65 and 130 neighbours): the
    diffusion walk runs 64 entries at a time; event-by-event parity."""
    sa = workloads.birth_death_2d(seed=3, t_end=2, num_points=3)
    sa.h = hfac * (sa.x[1, 1] - sa.x[0, 1] if sa.x[1, 1] != sa.x[0, 1] else 0.1)
    words, expo = mt19937_64_stream(11, 400_000)
    eng = Engine(sa, trace_capacity=100_000)
    res = eng.run(words=words, expo=expo)
    rp = eng.neighbors()[0]
    assert np.diff(rp).max() > 64
    ev, nev = eng.trace(100_000)
    o = oracle.Oracle(sa)
    o.set_trace(100_000)
    ref = o.run(seed=11)
    oev, onev = o.trace(100_000)
    assert nev == onev and nev > 500
    assert np.array_equal(ev, oev)
    assert np.array_equal(res["xx"], ref["xx"])


def test_c2_small_replay_matches_oracle(oracle):
    sa = workloads.static_rd_3d(n_side=12, nA=3000, nB=3000, t_end=3, num_points=4, seed=2)
    words, expo = mt19937_64_stream(11, 200_000)
    eng = Engine(sa)
    res = eng.run(words=words, expo=expo)
    o = oracle.Oracle(sa)
    ref = o.run(seed=11)
    assert (eng.stats().reactions, eng.stats().diffusions) == (o.stats().reactions, o.stats().diffusions)
    assert np.array_equal(res["xx"], ref["xx"])
    # neighbour lists, D_ij and the diffusion rates are bit-exact
    g = eng.neighbors()
    r = o.neighbors()
    assert np.array_equal(g[0], r[0]) and np.array_equal(g[1], r[1])
    for a, b in zip(g[2:], r[2:]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_moving_dam_break_small_matches_oracle(oracle):
    sa = workloads.dam_break_2d(nx=20, ny=20, spacing=0.025, n_mol=100, nt=45, out_every=15, dt=2e-4)
    sa.flags = 2   # ANN tie order on the moving domain as well
    words, expo = mt19937_64_stream(5, 200_000)
    eng = Engine(sa)
    res = eng.run(words=words, expo=expo)
    o = oracle.Oracle(sa, gauss_seidel=False)
    ref = o.run(seed=5)
    # positions/velocities: FP64 on both sides, pow(x,2) written as x*x here
    assert np.allclose(res["x"], ref["x"], rtol=0, atol=1e-12)
    assert np.allclose(res["v"], ref["v"], rtol=1e-9, atol=1e-12)
    assert np.allclose(res["rho"], ref["rho"], rtol=1e-12, atol=0)
    assert np.array_equal(res["xx"], ref["xx"])


@pytest.mark.parametrize("case,width", [("static18", 0), ("cylinder20k", 0), ("cylinder20k", 4)])
def test_large_event_tree_replay_matches_oracle(case, width, oracle, monkeypatch):
    """Domains with more than 64 level-1 nodes in the event tree, so the
    level held in registers is level 2 (as on the 1M cylinder and the 64k
    lattice): event-by-event parity with the oracle on a static 18^3
    lattice (5,832 voxels) and a moving 20k-particle cylinder (ANN tie
    order), the latter also at event-tree width 4."""
    if width:
        monkeypatch.setenv("SSASDPD_EVENT_TREE_PL", str(width))
    if case == "static18":
        sa = workloads.static_rd_3d(n_side=18, nA=1500, nB=1500, t_end=1, num_points=2, seed=4)
        o = oracle.Oracle(sa)
    else:
        sa = workloads.cylinder_3d(n_target=20_000, nt=6, seed=3)
        sa.flags = 2   # ANN tie order on the moving domain, as the oracle
        o = oracle.Oracle(sa, gauss_seidel=False)
    n_l1 = -(-sa.n // (64 * (width or 1)))
    assert n_l1 > 64
    words, expo = mt19937_64_stream(13, 2_000_000)
    eng = Engine(sa, trace_capacity=400_000)
    res = eng.run(words=words, expo=expo)
    ev, nev = eng.trace(400_000)
    o.set_trace(400_000)
    ref = o.run(seed=13)
    oev, onev = o.trace(400_000)
    assert nev == onev and nev > 200, (nev, onev)
    assert np.array_equal(ev, oev)
    assert np.array_equal(res["xx"], ref["xx"])
    assert eng.stats().rng_words_used == o.stats().rng_words_used


@pytest.mark.parametrize("case", ["dam_break", "cylinder20k"])
def test_advance_with_particle_lookahead_equals_run(case):
    """advance() starts the next step's particle phase on a second stream
    while the current step's event loop runs (moving domains); its state
    must equal run()'s, which takes the steps one after another, bit for
    bit.  Two advance() calls also cover the step where a call ends (no
    look-ahead there)."""
    if case == "dam_break":
        sa = workloads.dam_break_2d(nx=20, ny=20, spacing=0.025, n_mol=100, nt=45, out_every=45, dt=2e-4)
    else:
        sa = workloads.cylinder_3d(n_target=20_000, nt=6, seed=3)
    words, expo = mt19937_64_stream(17, 600_000)
    ref = Engine(sa).run(words=words, expo=expo, fields=("xx", "x", "v", "rho"))
    eng = Engine(sa)
    eng.reset(words=words, expo=expo)
    eng.advance(sa.nt // 2)
    eng.advance(sa.nt - sa.nt // 2)
    st = eng.state()
    assert np.array_equal(st["xx"], ref["xx"][-1])
    for f in ("x", "v", "rho"):
        assert np.array_equal(st[f].view(np.uint64), ref[f][-1].view(np.uint64)), f
