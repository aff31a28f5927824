"""CPU tests: pin the FP64 oracle against the reference engine's own outputs
(golden fixtures written by tests/golden/make_fixtures.py from SpatialPy's
C++ engine), against libstdc++'s generators, against brute force and
closed forms; check the lowering tables and the C-ABI library."""
import ctypes
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from spatialpy_amd import workloads, lowering

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
ROOT = os.path.dirname(HERE)


def load_case(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    sa = getattr(workloads, meta["gen"])(**meta["params"])
    # the generator must reproduce the fixture's inputs exactly
    assert np.array_equal(sa.x, z["x"]) and np.array_equal(sa.u0, z["u0"])
    return sa, meta, z


# ---------------------------------------------------------------- RNG pins
def test_mt19937_64_known_answer(oracle):
    """C++ standard [rand.predef]: the 10000th output of a default-constructed
    mt19937_64 (seed 5489) is 9981545732273789042."""
    w, _ = oracle.mt_stream(5489, 10000)
    assert int(w[-1]) == 9981545732273789042
    from spatialpy_amd.engine import mt19937_64_stream
    w2, e2 = mt19937_64_stream(5489, 10000)
    assert np.array_equal(w, w2)


def test_exponential_variates_match_libstdcxx(oracle, tmp_path):
    """The replay stream's exponential variates are those of
    std::exponential_distribution<double>(1) over std::mt19937_64 in this
    image's libstdc++/glibc."""
    src = tmp_path / "e.cpp"
    src.write_text(
        "#include <random>\n#include <cstdio>\nint main(){std::mt19937_64 g(77);"
        "std::exponential_distribution<> e(1.0);for(int i=0;i<20000;i++){double v=e(g);"
        "fwrite(&v,8,1,stdout);}}\n")
    exe = tmp_path / "e"
    subprocess.check_call(["g++", "-O2", str(src), "-o", str(exe)])
    ref = np.frombuffer(subprocess.check_output([str(exe)]), dtype=np.float64)
    _, expo = oracle.mt_stream(77, 20000)
    assert np.array_equal(ref.view(np.uint64), expo.view(np.uint64))
    from spatialpy_amd.engine import mt19937_64_stream
    _, e2 = mt19937_64_stream(77, 20000)
    assert np.array_equal(ref.view(np.uint64), e2.view(np.uint64))


# ---------------------------------------------------------------- kd-tree pins
def test_kdtree_matches_ann_test1_structure(oracle):
    """ANN's own golden file test1.save (20 points, bucket 1, sliding
    midpoint): leaves, splits and depth."""
    z = np.load(os.path.join(GOLDEN, "ann_test1.npz"))
    pts = np.zeros((z["data"].shape[0], 3))
    pts[:, :2] = z["data"]
    st = oracle.tree_stats(pts, 2)
    assert st["n_leaf"] == int(z["n_leaves"])
    assert st["n_split"] == int(z["n_splits"])
    assert st["n_empty"] == 0
    assert st["depth"] == int(z["depth"])


@pytest.mark.parametrize("case", ["lattice2d", "lattice3d", "jitter3d", "mixed"])
def test_fr_search_set_and_r2_equal_brute_force(oracle, case):
    rng = np.random.default_rng(0)
    if case == "lattice2d":
        sa = workloads.birth_death_2d()
        pts, q, dim, h = sa.x, sa.x, 2, sa.h
    elif case == "lattice3d":
        sa = workloads.static_rd_3d(n_side=9, nA=1, nB=1)
        pts, q, dim, h = sa.x, sa.x, 3, sa.h
    elif case == "jitter3d":
        pts, _ = workloads.jittered_lattice_3d(1500, seed=1)
        q, dim, h = pts, 3, 0.15
    else:
        pts = rng.uniform(-1, 1, (800, 3))
        q = pts + rng.normal(0, 1e-3, pts.shape)
        dim, h = 3, 0.2
    a = oracle.fr_search(pts, q, dim, h)
    b = oracle.fr_brute(pts, q, dim, h)
    assert np.array_equal(a[0], b[0])
    for i in range(len(a[0]) - 1):
        s = slice(a[0][i], a[0][i + 1])
        # ascending r^2 in the tree order too
        assert np.all(np.diff(a[2][s]) >= 0)
        assert sorted(zip(a[2][s].tolist(), a[1][s].tolist())) == list(zip(b[2][s].tolist(), b[1][s].tolist()))


# ------------------------------------------------- oracle vs reference engine
def _run_oracle_like_harness(oracle, sa, meta, gauss_seidel=True):
    o = oracle.Oracle(sa, gauss_seidel=gauss_seidel)
    res = o.run(seed=meta["seed"])
    return o, res


@pytest.mark.parametrize("name", ["c1_birth_death", "c2_static_rd_small"])
def test_oracle_static_bit_exact_with_reference(oracle, name):
    sa, meta, z = load_case(name)
    o, res = _run_oracle_like_harness(oracle, sa, meta)
    st = o.stats()
    assert [st.reactions, st.diffusions] == z["counters"].tolist()
    assert np.array_equal(res["xx"].ravel(), z["xx"])
    assert np.array_equal(res["C"].ravel().view(np.uint64), z["snap_C"].view(np.uint64))
    rp, idx, dist, dwdr, dij = o.neighbors()
    assert np.array_equal(rp, z["nbr_rowptr"]) and np.array_equal(idx, z["nbr_idx"])
    for mine, ref in ((dist, z["nbr_dist"]), (dwdr, z["nbr_dwdr"]), (dij, z["nbr_dij"])):
        assert np.array_equal(mine.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("name", ["c3_dam_break_small", "c5_cylinder_small"])
def test_oracle_moving_matches_reference(oracle, name):
    sa, meta, z = load_case(name)
    o, res = _run_oracle_like_harness(oracle, sa, meta)
    st = o.stats()
    assert [st.reactions, st.diffusions] == z["counters"].tolist()
    assert np.array_equal(res["xx"].ravel(), z["xx"])
    # FP64 on both sides with the same operation order; pow(x,2) may be
    # folded differently by the two compilers, hence 1e-12 relative
    for mine, ref in ((res["x"], z["snap_x"]), (res["v"], z["snap_v"]), (res["rho"], z["snap_rho"])):
        np.testing.assert_allclose(mine.ravel(), ref, rtol=1e-12, atol=1e-13)
    # neighbour lists and forces at the fixture's step
    o2 = oracle.Oracle(sa)
    o2.reset(seed=meta["seed"])
    o2.advance(meta["kw"]["force_step"] + 1)
    rp, idx, dist, dwdr, dij = o2.neighbors()
    assert np.array_equal(rp, z["nbr_rowptr"]) and np.array_equal(idx, z["nbr_idx"])
    assert np.array_equal(dist.view(np.uint64), z["nbr_dist"].view(np.uint64))
    s = o2.state()
    scale = np.abs(z["F"]).max()
    np.testing.assert_allclose(s["F"].ravel(), z["F"], rtol=0, atol=1e-12 * scale)


def test_oracle_gauss_seidel_and_jacobi_agree_on_small_dam_break(oracle):
    """The only order-dependent read (neighbour rho in the boundary volume
    fraction) does not change this run's outcome."""
    sa, meta, z = load_case("c3_dam_break_small")
    _, a = _run_oracle_like_harness(oracle, sa, meta, gauss_seidel=True)
    _, b = _run_oracle_like_harness(oracle, sa, meta, gauss_seidel=False)
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["xx"], b["xx"])


# ------------------------------------------------------------ SSA invariants
def test_diffusion_only_conserves_molecules(oracle):
    sa = workloads.static_rd_3d(n_side=7, nA=500, nB=400, kf=0.0, kr=0.0, t_end=5, num_points=6)
    o = oracle.Oracle(sa)
    res = o.run(seed=3)
    tot = res["xx"].astype(np.int64).sum(axis=1)
    assert np.all(tot == tot[0])
    assert o.stats().reactions == 0 and o.stats().diffusions > 0


def test_stoichiometric_balance(oracle):
    """A+B->C, C->A+B: A+C and B+C are conserved by every event."""
    sa = workloads.static_rd_3d(n_side=6, nA=300, nB=200, kf=5e-2, kr=0.5, t_end=4, num_points=5)
    res = oracle.Oracle(sa).run(seed=5)
    tot = res["xx"].astype(np.int64).sum(axis=1)
    assert np.all(tot[:, 0] + tot[:, 2] == 300) and np.all(tot[:, 1] + tot[:, 2] == 200)
    assert tot[-1, 2] > 0


def test_reference_selection_bias_signature(oracle):
    """The reference picks the reaction with rand1*srrate after testing
    rand1 <= srrate/totrate (simulate_rdme.cpp:255-261).  In the birth-death
    model every reaction event then becomes a birth (A->0 is unreachable
    while diffusion dominates totrate), so the total count obeys
    dN/dt = kb*V + kd*N: N(t) = (N0 + kb V/kd) e^{kd t} - kb V/kd.  With
    textbook selection it relaxes to the stationary mean kb V/kd instead."""
    sa = workloads.birth_death_2d(seed=1)
    V = sa.n * 0.01
    kb, kd, N0, t = 10.0, 0.1, 100, 10.0
    grow, tb = [], []
    for seed in range(8):
        grow.append(float(oracle.Oracle(sa).run(seed=seed)["xx"][-1].sum()))
        sb = workloads.birth_death_2d(seed=1)
        sb.flags = 1
        tb.append(float(oracle.Oracle(sb).run(seed=seed)["xx"][-1].sum()))
    expect_grow = (N0 + kb * V / kd) * np.exp(kd * t) - kb * V / kd
    stationary = kb * V / kd
    expect_tb = stationary + (N0 - stationary) * np.exp(-kd * t)
    assert abs(np.mean(grow) - expect_grow) < 0.08 * expect_grow
    assert abs(np.mean(tb) - expect_tb) < 4 * np.sqrt(expect_tb / 8) + 3


# ------------------------------------------------------------------ lowering
def test_output_schedule_matches_run_simulation():
    # output every step
    assert lowering.output_schedule(np.arange(11), 10).tolist() == list(range(11))
    # output every 10 steps: the reference also snapshots step 1 because the
    # first get_next_output() returns output_steps[0] == 0
    assert lowering.output_schedule(np.array([0, 10, 20]), 20).tolist() == [0, 1, 10, 20]


def test_timespan_output_steps_reference_unit_test():
    """test/unit_tests/test_timespan.py:39-46 restated: linspace(0,20,401) at
    dt=0.001 gives output steps every 50."""
    nt, steps = workloads.timespan_steps(20, 401, 0.001)
    assert nt == 20000 and steps[1] == 50 and steps[-1] == 20000 and steps.size == 401


def test_dependency_graph_and_stoichiometry_for_birth_death():
    sa = workloads.birth_death_2d()
    # birth: 0 -> A ; death: A -> 0
    assert sa.jcN.tolist() == [0, 1, 2] and sa.irN.tolist() == [0, 0] and sa.prN.tolist() == [1, -1]
    # species A feeds reaction 1 (death); both reactions change A so both
    # re-evaluate death
    assert sa.jcG.tolist() == [0, 1, 2, 3] and sa.irG.tolist() == [1, 1, 1]


def test_lattice_matches_reference_counts():
    # create_2D_domain(10x10) fills 11 x 11 points (numpy arange overshoot)
    assert workloads.birth_death_2d().n == 121
    assert workloads.static_rd_3d(n_side=40, nA=1, nB=1).n == 64000


# ------------------------------------------------------------ C-ABI library
def _declared_functions():
    text = open(os.path.join(ROOT, "include", "ssasdpd.h")).read()
    return sorted(set(re.findall(r"\b(ssasdpd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from spatialpy_amd import _abi
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 15
    for nm in names:
        assert hasattr(lib, nm), nm


def _grid_shape(lib, bbox, n, dim, h):
    bb = (ctypes.c_double * 6)(*bbox)
    dims = (ctypes.c_int32 * 3)()
    nc, cell = ctypes.c_int64(), ctypes.c_double()
    assert lib.ssasdpd_grid_shape(bb, n, dim, h, dims, ctypes.byref(nc), ctypes.byref(cell)) == 0
    return list(dims), nc.value, cell.value


def test_search_grid_stays_bounded_for_far_points():
    """The search grid for a bounding box: cells of side 1.01 h over an
    ordinary box, and a coarser grid of at most max(4n, 2^20) cells, never
    a wrapped count, when one point lies far from the rest (a box of 1e6 or
    1e8 units at h = 0.03 needs 1e22-1e27 cells of side h; the cell count
    wrapped to a negative number or to 0 in 64-bit integers)."""
    from spatialpy_amd import _abi
    lib = _abi.load_library()
    n, h = 1_005_649, 0.030223631139250296
    dims, nc, cell = _grid_shape(lib, [-1.0, 1.0, -0.5, 0.5, -0.25, 0.25], n, 3, h)
    assert cell == pytest.approx(1.01 * h, rel=1e-12)
    assert dims == [int(2.0 // cell) + 1, int(1.0 // cell) + 1, int(0.5 // cell) + 1]
    assert nc == dims[0] * dims[1] * dims[2]
    for far in (1e3, 1e6, 1e8, 1e12, 1e300):
        dims, nc, cell = _grid_shape(lib, [-far, 1.0, -far, 1.0, -far, 1.0], n, 3, h)
        assert 1 <= nc <= max(4 * n, 1 << 20), (far, dims, nc)
        assert nc == dims[0] * dims[1] * dims[2] and min(dims) >= 1, (far, dims, nc)
        # the cells cover the box
        assert all(d * cell >= far for d in dims), (far, dims, cell)
    # non-finite extents (NaN / Inf positions) get one cell on that side
    dims, nc, _ = _grid_shape(lib, [0.0, float("nan"), 0.0, float("inf"), 0.0, 1.0], 100, 3, 0.1)
    assert dims[:2] == [1, 1] and nc == dims[2]


def test_engine_fails_loudly_without_gpu():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    from spatialpy_amd.engine import Engine
    from spatialpy_amd import SimulationError
    with pytest.raises(SimulationError):
        Engine(workloads.birth_death_2d())


def test_bad_descriptor_is_rejected_before_gpu():
    from spatialpy_amd.engine import Engine
    from spatialpy_amd import SimulationError
    sa = workloads.birth_death_2d()
    sa.type[0] = 5      # out of 1..n_types
    with pytest.raises(SimulationError, match="BAD_DESC"):
        Engine(sa)


def test_mpn_oracle_backward_matches_finite_differences():
    """The surrogate's CPU oracle: its analytic gradient equals central
    differences of its FP64 loss on a random particle graph (three
    parameters from each block)."""
    from oracle import mpn_oracle as mo
    g = mo.random_graph(n=100, seed=2, h=0.3)
    sc = mo.default_scales(g)
    sel = np.nonzero(g["solid"] == 0)[0]
    y = np.random.default_rng(0).normal(0, 1, (len(sel), 3))
    w = mo.init_params(5)
    _, gr = mo.loss_grad(w, g, sc, sel, y)
    rng = np.random.default_rng(1)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mo._SHAPES])
    for a, b in zip(offs[:-1], offs[1:]):
        cand = [k for k in range(a, b) if abs(gr[k]) > 1e-6] or [a]
        for k in rng.choice(cand, size=min(3, len(cand)), replace=False):
            e = 1e-6 * max(1.0, abs(w[k]))
            wp, wm = w.copy(), w.copy()
            wp[k] += e
            wm[k] -= e
            fd = (mo.loss_grad(wp, g, sc, sel, y)[0] - mo.loss_grad(wm, g, sc, sel, y)[0]) / (2 * e)
            assert abs(fd - gr[k]) <= 1e-4 * max(abs(fd), 1e-3), (k, fd, gr[k])


def test_mpn_oracle_pair_terms_span_the_sdpd_force():
    """The network's output space holds the SDPD pair acceleration.  The
    pair loop's terms, restated as coefficients of dx, dv, v_i and v_j
    (mpn_oracle.sdpd_pair_terms), sum to the FP64 SDPD oracle's own F - g on
    a moving 2D dam break after 40 steps (densities off rho0, velocities,
    transport velocities, walls); and with those coefficients divided by the
    envelope as the network's outputs, the oracle's message assembly
    (features(): basis vectors dx/h, dv/v_s, v_i/v_s, v_j/v_s times the
    envelope) gives the same accelerations.  A wrong basis vector, sign or
    envelope breaks one of the two sums."""
    from oracle import mpn_oracle as mo
    from oracle import oracle_py
    from spatialpy_amd import workloads
    sa = workloads.dam_break_2d(nx=12, ny=12, spacing=0.025, n_mol=20, nt=60, out_every=60, dt=2e-4)
    o = oracle_py.Oracle(sa)
    o.reset(seed=3)
    o.advance(40)
    st = o.state()
    rp, idx, dist, dwdr, _ = o.neighbors()
    F, _, _ = o.eval_forces()
    gvec = np.asarray(sa.gravity)
    ref = F - gvec
    I, J, co = mo.sdpd_pair_terms(st["x"], st["v"], st["vt"], st["rho"], sa.mass, st["nu"], rp, idx, dist, dwdr,
                                  sa.h, sa.P0, sa.rho0, sa.dimension)
    x, v = st["x"], st["v"]
    terms = co[:, :1] * (x[I] - x[J]) + co[:, 1:2] * (v[I] - v[J]) + co[:, 2:3] * v[I] + co[:, 3:4] * v[J]
    terms[:, sa.dimension:] = 0.0
    acc = np.zeros_like(ref)
    np.add.at(acc, I, terms)
    scale = np.abs(ref[:, :sa.dimension]).max()
    assert scale > 0 and np.abs(ref[:, sa.dimension:]).max() == 0.0
    assert np.abs(acc - ref).max() <= 1e-10 * scale, np.abs(acc - ref).max() / scale
    # the network's assembly with the exact coefficients
    deg = np.diff(rp)
    W = int(deg.max())
    nidx = np.zeros((W, sa.n), np.int32)
    nr = np.zeros((W, sa.n))
    for k in range(W):
        s_ = np.nonzero(deg > k)[0]
        nidx[k, s_] = idx[rp[s_] + k]
        nr[k, s_] = dist[rp[s_] + k]
    g = {"h": sa.h, "x": x, "v": v, "vt": st["vt"], "rho": st["rho"], "mass": sa.mass, "nu": st["nu"],
         "solid": sa.fixed, "deg": deg, "nidx": nidx, "nr": nr}
    sc = np.array([sa.rho0, np.sqrt((v * v).sum(1).mean()), 0.01, np.abs(sa.nu).mean(), scale, 1e-3])
    P_, I2, J2, RD = mo.edges(g, np.arange(sa.n))
    X, B, u = mo.features(g, sc, I2, J2, RD)
    valid = u > 0
    assert np.array_equal(np.stack([I2[valid], J2[valid]]), np.stack([I, J]))
    o_ = np.zeros((len(I2), 4))
    o_[valid] = co * np.array([sa.h, sc[1], sc[1], sc[1]]) / (u[valid, None] * sc[4])
    pred = np.zeros((sa.n, 3))
    np.add.at(pred, P_, u[:, None] * np.einsum("ek,ekc->ec", o_, B))
    # basis and envelope are rounded to FP32 as the GPU holds them
    assert np.abs(pred * sc[4] - ref).max() <= 1e-5 * scale, np.abs(pred * sc[4] - ref).max() / scale


def test_mpn_parameter_count_matches_library():
    from oracle import mpn_oracle as mo
    from spatialpy_amd import _abi
    lib = _abi.load_library()
    assert int(lib.ssasdpd_mpn_num_params()) == mo.P == 6596


# ---------------------------------------------------------------------------
# VTK snapshots (spatialpy_amd/vtk.py) against files the reference engine's own
# writer produced (output.cpp:114-229, tests/golden/vtk_*/, made by
# tests/golden/make_fixtures.py).

def _parse_vtk(path):
    """Values of a snapshot file, by block name."""
    toks = open(path).read().split()
    out, i = {}, 0
    n = int(toks[toks.index("POINTS") + 1])
    i = toks.index("POINTS") + 3
    out["x"] = np.array(toks[i:i + 3 * n], dtype=np.float64).reshape(n, 3)
    i = toks.index("FIELD") + 3
    out["num_fields"] = int(toks[toks.index("FIELD") + 2])
    while i < len(toks):
        name, ncomp, cnt = toks[i], int(toks[i + 1]), int(toks[i + 2])
        vals = toks[i + 4:i + 4 + ncomp * cnt]
        out[name] = np.array(vals, dtype=np.float64).reshape(cnt, ncomp) if ncomp > 1 else np.array(vals, np.float64)
        i += 4 + ncomp * cnt
    return out


def _rewrite(ref_dir, tmp, k, species):
    from spatialpy_amd import vtk
    d = _parse_vtk(os.path.join(ref_dir, f"output{k}.vtk"))
    n = d["rho"].size
    C = np.column_stack([d[f"C[{s}]"] for s in species if f"C[{s}]" in d]) if any(
        f"C[{s}]" in d for s in species) else None
    xx = np.column_stack([d[f"D[{s}]"] for s in species if f"D[{s}]" in d]).astype(np.uint32) if any(
        f"D[{s}]" in d for s in species) else None
    S = 0 if xx is None else xx.shape[1]
    Sc = 0 if C is None else C.shape[1]
    initialized = d["num_fields"] == 7 + Sc + S
    vtk.write_snapshot(tmp, k, d["x"], d["v"], d["rho"], d["mass"], d["bvf_phi"], d["nu"],
                       d["type"].astype(np.int64), species, C=C, xx=xx, initialized=initialized,
                       ids=d["id"].astype(np.int64))


@pytest.mark.parametrize("case", ["vtk_c1_birth_death", "vtk_c3_dam_break_tiny"])
def test_vtk_writer_reproduces_reference_files_byte_for_byte(case, tmp_path):
    """The reference's own files, re-written from their parsed values, come out
    identical: formats, line breaks, trailing spaces, headers and the
    FieldData count quirk."""
    import json
    from spatialpy_amd import vtk
    ref = os.path.join(GOLDEN, case)
    meta = json.load(open(os.path.join(ref, "case.json")))
    species = ["s0", "s1", "s2"]
    for f in meta["files"]:
        if f == "output0_boundingBox.vtk":
            from spatialpy_amd import workloads
            sa = getattr(workloads, meta["gen"])(**meta["params"])
            vtk.write_bounding_box(str(tmp_path), sa.bbox)
        else:
            _rewrite(ref, str(tmp_path), int(f[6:-4]), species)
        assert open(os.path.join(ref, f)).read() == open(os.path.join(str(tmp_path), f)).read(), f


def test_vtk_from_oracle_run_matches_reference_birth_death(tmp_path, oracle):
    """Content and format together: the FP64 oracle's snapshots of the static
    birth-death case (bit-exact with the reference engine) written by this
    package's writer equal the reference engine's files byte for byte."""
    import json
    from spatialpy_amd import vtk, workloads
    from spatialpy_amd.lowering import output_schedule
    ref = os.path.join(GOLDEN, "vtk_c1_birth_death")
    meta = json.load(open(os.path.join(ref, "case.json")))
    sa = getattr(workloads, meta["gen"])(**meta["params"])
    o = oracle.Oracle(sa)
    res = o.run(seed=meta["seed"])
    steps = output_schedule(sa.output_steps, sa.nt)
    vtk.write_result(str(tmp_path), res, sa, ["s0"], steps)
    for f in meta["files"]:
        assert open(os.path.join(ref, f)).read() == open(os.path.join(str(tmp_path), f)).read(), f
