"""Golden vectors from the program SpatialPy itself generates.

Run only where the reference tree is (it imports spatialpy from there and
links the engine objects compiled from its sources):

    python tests/golden/make_generated_fixtures.py [case ...]

For each model of tests/golden/generated_models.py:
1. the model is built with spatialpy's own API, with numpy.random seeded
   first (ScatterInitialCondition draws from it);
2. SpatialPy's own generator, Solver._Solver__create_propensity_file
   (spatialpy/solvers/solver.py:100-158), writes the model program from
   propensity_file_template.cpp at debug level 1;
3. g++ -std=c++14 (no optimisation, as build/SConstruct does) links it with
   the engine and ANN objects that oracle/Makefile compiles from the
   reference sources (`make -C oracle ref`, oracle/_ref/O0);
4. it runs as `exe -t 1 -s SEED` in a scratch directory, as Solver.run
   starts it (solver.py:553-569).

Saved, as data only (no text of the reference or of the generated program):
* tests/golden/gen_<case>.npz: the arrays the generator baked into the
  program -- particles (x, type index, nu, mass, c, rho, fixed), input_u0,
  the stoichiometry and dependency CSC, input_N_dense, the diffusion matrix,
  the system scalars (dt, nt, h, rho0, c0, P0, bounding box, dimension,
  gravity, static_domain), the output steps, the type-name mapping and each
  boundary condition's bounds / type / target / values; the event lines the
  program printed (debug level 1, simulate_rdme.cpp:282-283, 405-413): the
  first MAX_EVENTS as (time, voxel, kind, reaction or destination), their
  total and the SHA-256 of all of them in a canonical form; and the final
  reaction / diffusion totals;
* tests/golden/gen_<case>/output*.vtk: the VTK files the program wrote
  (output.cpp), for the cases listed in VTK_FILES.

A second check runs here too: spatialpy_amd's Solver lowers the very same
spatialpy.Model object, and its arrays must equal the baked ones (the
script stops otherwise).
"""
import glob
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
ENG = os.path.join(REF, "spatialpy/solvers/c_base/ssa_sdpd-c-simulation-engine")
OBJ = os.path.join(ROOT, "oracle", "_ref", "O0")
MAX_EVENTS = 4000
# which VTK files to keep: all, or (moving domains, whose files are large)
# the bounding box, the first and the last snapshot
VTK_FILES = {"birth_death": "all", "static_rd": "all", "dam_break": "first_last", "cylinder": "first_last",
             "custom": "all", "custom_libm": "all"}

if os.environ.get("PYTHONHASHSEED") != "0":
    # the type-name mapping iterates a set of strings (core/domain.py:137-146):
    # a fixed hash seed makes it reproducible
    os.environ["PYTHONHASHSEED"] = "0"
    sys.exit(subprocess.call([sys.executable] + sys.argv))

sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
sys.path.insert(1, REF)
import generated_models as GM  # noqa: E402

EVENT_RE = re.compile(r"nsm: tt=(\S+) subvol=(-?\d+) type=(-?\d+) (?:Rxn (-?\d+)|Diff (-?\d+)->(-?\d+))")


def canonical_event(t, voxel, kind, which):
    """One event as text: time as printed (%e), voxel, R/D, reaction or
    destination.  Both sides hash this form."""
    return f"{t} {voxel} {'R' if kind == 0 else 'D'} {which}\n"


def parse_events(stdout):
    evs = []
    for m in EVENT_RE.finditer(stdout):
        t, vox = m.group(1), int(m.group(2))
        if m.group(4) is not None:
            evs.append((t, vox, 0, int(m.group(4))))
        else:
            evs.append((t, vox, 1, int(m.group(6))))
    return evs


def arr(text, name, dtype):
    m = re.search(r"\b" + re.escape(name) + r"\[\d*\]\s*=\s*\{([^}]*)\}", text)
    if m is None:
        return None
    body = m.group(1).strip()
    vals = [v for v in body.split(",") if v.strip()] if body else []
    if dtype is float:
        return np.array([float(v) for v in vals], dtype=np.float64)
    return np.array([int(v) for v in vals], dtype=np.int64)


def scalar(text, key):
    m = re.search(r"system->" + re.escape(key) + r"\s*=\s*([^;]+);", text)
    return None if m is None else m.group(1).strip()


def parse_program(text):
    """The arrays and scalars the generator wrote into the program."""
    out = {}
    types = dict((n, int(v)) for n, v in re.findall(r"const size_t (type_\w+) = (\d+);", text))
    out["type_names"] = json.dumps(types)
    parts = re.findall(r"init_create_particle\(sys,id\+\+,([^)]*)\);", text)
    rows = [p.split(",") for p in parts]
    out["x"] = np.array([[float(r[0]), float(r[1]), float(r[2])] for r in rows])
    out["type"] = np.array([types[r[3].strip()] for r in rows], dtype=np.int32)
    for k, name in ((4, "nu"), (5, "mass"), (6, "c"), (7, "rho")):
        out[name] = np.array([float(r[k]) for r in rows])
    out["fixed"] = np.array([int(r[8]) for r in rows], dtype=np.int32)
    out["u0"] = arr(text, "input_u0", int)
    for name in ("input_N_dense", "input_irN", "input_jcN", "input_prN", "input_irG", "input_jcG"):
        a = arr(text, name, int)
        out[name[6:]] = a if a is not None else np.zeros(0, np.int64)
    out["diffusion_matrix"] = arr(text, "input_subdomain_diffusion_matrix", float)
    df = arr(text, "input_data_fn", float)
    if df is not None:
        out["data_fn"] = df
    for key in ("dt", "nt", "h", "rho0", "c0", "P0", "xlo", "xhi", "ylo", "yhi", "zlo", "zhi", "dimension",
                "static_domain"):
        out[key] = float(scalar(text, key))
    out["gravity"] = np.array([float(scalar(text, f"gravity[{i}]") or 0.0) for i in range(3)])
    m = re.search(r"const std::vector<unsigned int> output_steps = \{([^}]*)\}", text)
    out["output_steps"] = np.array([int(v) for v in m.group(1).split(",")], dtype=np.int64)
    ps = re.search(r"new ParticleSystem\(([^)]*)\)", text).group(1)
    out["system_dims"] = np.array([int(v) for v in ps.split(",")], dtype=np.int64)
    # boundary conditions (core/boundarycondition.py:121-167)
    body = re.search(r"void applyBoundaryConditions\(Particle\* me, ParticleSystem\* system\)\{(.*?)\n\s*\}",
                     text, re.S).group(1)
    bcs = []
    for cond, act in re.findall(r"if\((.*?)\)\{(.*?)\}", body):
        b = {"lo": [None] * 3, "hi": [None] * 3, "type": None}
        for ax, op, v in re.findall(r"\(me->x\[(\d)\] (>=|<=) ([^)]+)\)", cond):
            (b["lo"] if op == ">=" else b["hi"])[int(ax)] = float(v)
        mt = re.search(r"\(me->type == (\w+)\)", cond)
        if mt:
            b["type"] = types[mt.group(1)]
        b["action"] = act
        bcs.append(b)
    out["bcs"] = json.dumps(bcs)
    return out


def vtk_files(d):
    return sorted(glob.glob(os.path.join(d, "output*.vtk")))


def make_case(name):
    import spatialpy
    from spatialpy_amd.solver import Solver as AmdSolver
    np.random.seed(GM.NUMPY_SEED[name])
    model, h = GM.build(spatialpy, name)
    stoich, dep = model.compile_prep()
    sol = spatialpy.Solver(model, debug_level=1)
    if h is not None:
        sol.h = h
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "model.cpp")
        sol._Solver__create_propensity_file(stoich, dep, file_name=src)
        text = open(src).read()
        exe = os.path.join(d, "ssa_sdpd.exe")
        objs = sorted(glob.glob(os.path.join(OBJ, "*.o")))
        subprocess.run(["g++", "-std=c++14", "-w", f"-I{ENG}/include", f"-I{ENG}/external/ANN/include", src]
                       + objs + ["-o", exe, "-pthread"], check=True)
        run_dir = os.path.join(d, "run")
        os.makedirs(run_dir)
        r = subprocess.run([exe, "-t", "1", "-s", str(GM.RUN_SEED[name])], cwd=run_dir, capture_output=True,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: generated program failed rc={r.returncode}: {r.stdout[-400:]}")
        out = parse_program(text)
        evs = parse_events(r.stdout)
        h_all = hashlib.sha256("".join(canonical_event(*e) for e in evs).encode()).hexdigest()
        tot = re.findall(r"NSM: total # reacton events (\d+)\nNSM: total # diffusion events (\d+)", r.stdout)
        first = evs[:MAX_EVENTS]
        out["ev_time"] = np.array([float(e[0]) for e in first])
        out["ev_time_text"] = np.array([e[0] for e in first])
        out["ev_voxel"] = np.array([e[1] for e in first], dtype=np.int32)
        out["ev_kind"] = np.array([e[2] for e in first], dtype=np.int32)
        out["ev_which"] = np.array([e[3] for e in first], dtype=np.int32)
        out["n_events"] = len(evs)
        out["events_sha256"] = h_all
        out["n_reactions"] = sum(1 for e in evs if e[2] == 0)
        out["n_diffusions"] = sum(1 for e in evs if e[2] == 1)
        out["meta"] = json.dumps(dict(case=name, numpy_seed=GM.NUMPY_SEED[name], run_seed=GM.RUN_SEED[name],
                                      h=h, printed_totals=tot[-1] if tot else None))
        files = vtk_files(run_dir)
        if VTK_FILES[name] == "first_last":
            snaps = sorted((f for f in files if "bounding" not in f),
                           key=lambda f: int(re.search(r"output(\d+)\.vtk", f).group(1)))
            files = sorted({f for f in files if "bounding" in f} | {snaps[0], snaps[-1]})
        dst = os.path.join(HERE, f"gen_{name}")
        os.makedirs(dst, exist_ok=True)
        for f in glob.glob(os.path.join(dst, "*.vtk")):
            os.remove(f)
        for f in files:
            with open(f) as a, open(os.path.join(dst, os.path.basename(f)), "w") as b:
                b.write(a.read())
        out["vtk_files"] = json.dumps([os.path.basename(f) for f in files])
    np.savez_compressed(os.path.join(HERE, f"gen_{name}.npz"), **out)

    # spatialpy_amd's Solver on the same spatialpy.Model object (its
    # compile_prep draws the scattered molecules again: same seed)
    np.random.seed(GM.NUMPY_SEED[name])
    amd = AmdSolver(model)
    amd.h = h
    sa = amd.lower()
    GM.check_lowering(name, sa, np.load(os.path.join(HERE, f"gen_{name}.npz")))
    print(f"gen_{name}: n={sa.n} events={len(evs)} (R {out['n_reactions']}, D {out['n_diffusions']}) "
          f"vtk={len(files)}")


if __name__ == "__main__":
    if not os.path.isdir(OBJ):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    only = sys.argv[1:]
    for case in GM.CASES:
        if not only or case in only:
            make_case(case)
