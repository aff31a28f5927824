"""Helpers for the fixtures written by SpatialPy's own generated program
(tests/golden/make_generated_fixtures.py): rebuild a case's model with this
package, lower it, and put engine / oracle results in the forms the
fixtures hold (event lines, VTK files)."""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)
import generated_models as GM  # noqa: E402

CASES = GM.CASES


def load(name):
    return np.load(os.path.join(GOLDEN, f"gen_{name}.npz"))


def lower(name):
    """The case's model built with spatialpy_amd's API (numpy.random seeded
    as the fixture script seeds it) and lowered by its Solver."""
    import spatialpy_amd as sp
    np.random.seed(GM.NUMPY_SEED[name])
    model, h = GM.build(sp, name)
    sol = sp.Solver(model)
    sol.h = h
    return sol.lower(), model


def event_lines(ev):
    """Events (EVENT_DTYPE records) as the canonical text of the fixture
    script: the time as printf's %e prints it, voxel, R/D, reaction index or
    destination (simulate_rdme.cpp:282-283, 405-406)."""
    out = []
    for e in ev:
        which = int(e["which"]) if int(e["kind"]) == 0 else int(e["dest"])
        out.append("%e %d %s %d\n" % (float(e["t"]), int(e["voxel"]), "R" if int(e["kind"]) == 0 else "D", which))
    return out


def check_events(name, fx, ev, n_events):
    """The event sequence against the program's: count, the first events
    field by field, and the SHA-256 of all of them."""
    lines = event_lines(ev[:n_events])
    k = min(fx["ev_voxel"].size, len(lines))
    for i in range(k):
        want = "%s %d %s %d\n" % (fx["ev_time_text"][i], fx["ev_voxel"][i], "R" if fx["ev_kind"][i] == 0 else "D",
                                  fx["ev_which"][i])
        assert lines[i] == want, (name, i, lines[i], want)
    assert n_events == int(fx["n_events"]), (name, n_events, int(fx["n_events"]))
    assert hashlib.sha256("".join(lines).encode()).hexdigest() == str(fx["events_sha256"]), name


def vtk_files(name):
    return json.loads(str(load(name)["vtk_files"]))


def write_vtk(folder, name, data, sa, model):
    from spatialpy_amd import vtk
    from spatialpy_amd.lowering import output_schedule
    steps = output_schedule(sa.output_steps, sa.nt)
    vtk.write_result(folder, data, sa, list(model.listOfSpecies), steps)


def _blocks(text):
    """A VTK file's value tokens by block (the header line's first word)."""
    out, cur = {}, None
    for line in text.split("\n"):
        if line[:1].isalpha() and not line.startswith(("nan", "inf")):
            cur = line.split()[0]
            out[cur] = []
        elif cur is not None:
            out[cur] += line.split()
    return out


def nd_overread(sa):
    """Whether the chem-RHS index R*rxn + s (model.cpp:181-189) reaches past
    the S x R array input_N_dense for some rxn < R, s < S.  The reference
    then reads whatever static array the compiler placed next (undefined
    behaviour); the oracle and the engine read 0 there."""
    S, R = sa.n_species, sa.n_reactions
    return S > 0 and R > 0 and R * (R - 1) + S - 1 >= S * R


def _close(p, q, key):
    """Two printed floats one unit apart in the last printed place (%lf:
    1e-6; %.10e: 1e-10 relative); bvf_phi within 1e-5 (the neighbour
    densities it weighs are read after the whole substep's update instead of
    in index order, which moves phi by O(dt * Frho / rho))."""
    a, b = float(p), float(q)
    if key == "bvf_phi":
        return abs(a - b) <= 1e-5
    return abs(a - b) <= 1.5e-6 + 2e-10 * abs(b)


def compare_vtk(folder, name, sa=None, printed_unit=False):
    """Every VTK file the fixture kept, byte for byte.  Where a file differs,
    the only differences allowed are the reference's undefined reads:
    * a value printed as 0 here and as a denormal (|x| < 1e-300) by the
      program: Particle::F, Fbp and vt start uninitialised (particle.cpp:
      70-83), so step 0's predictor adds dt times leftover memory, visible
      only on coordinates that are exactly 0;
    * the C[...] blocks of a model whose chem-RHS index reads past
      input_N_dense (nd_overread).
    printed_unit: floating-point values may also differ by one unit in their
    last printed place (the GPU on moving domains, whose positions and
    velocities agree with the FP64 reference arithmetic to ~1e-9 relative:
    it reads the neighbours' densities in the boundary volume fraction after
    every particle's update of the substep, where the one-thread reference
    reads them in index order).  Counts (D[...], id, type) stay exact.
    Returns the names of the files with any other difference."""
    bad, detail = [], []
    for f in vtk_files(name):
        with open(os.path.join(folder, f)) as a, open(os.path.join(GOLDEN, f"gen_{name}", f)) as b:
            mine, ref = a.read(), b.read()
        if mine == ref:
            continue
        ma, rb = _blocks(mine), _blocks(ref)
        ok = list(ma) == list(rb)
        for key in ma if ok else []:
            x, y = ma[key], rb[key]
            if len(x) != len(y):
                ok = False
                detail.append((f, key, len(x), len(y)))
                break
            if key.startswith("C[") and sa is not None and nd_overread(sa):
                continue
            exact = key.startswith("D[") or key in ("id", "type", "VERTICES")
            for p, q in zip(x, y):
                if p == q or (float(p) == 0.0 and abs(float(q)) < 1e-300):
                    continue
                if printed_unit and not exact and _close(p, q, key):
                    continue
                ok = False
                detail.append((f, key, p, q))
                break
            if not ok:
                break
        if not ok:
            bad.append(f)
    return bad + detail
