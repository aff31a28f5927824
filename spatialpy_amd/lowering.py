"""Model lowering to plain arrays (the C-ABI descriptor).

SpatialPy bakes these arrays into generated C++ text
(spatialpy/solvers/solver.py:208-299, 312-331, 375-419); here they stay numpy
arrays and are handed to libssasdpd.so through ``ssasdpd_desc``.
"""
from dataclasses import dataclass, field
import ctypes as C
import os

import numpy as np

from . import _abi

# propensity kinds (reaction.py:260-313); 4: expression-VM program
K_ZERO, K_FIRST, K_AB, K_2X, K_PROGRAM = 0, 1, 2, 3, 4


@dataclass
class SystemArrays:
    """Everything the engine needs for one model (one trajectory's inputs)."""
    x: np.ndarray                 # [n,3] float64
    type: np.ndarray              # [n] int32, 1..n_types
    mass: np.ndarray              # [n]
    rho: np.ndarray               # [n]
    nu: np.ndarray                # [n]
    fixed: np.ndarray             # [n] int32 (solidTag)
    n_types: int
    dimension: int
    h: float
    dt: float
    nt: int
    output_steps: np.ndarray      # [k] uint32
    static_domain: int = 1
    rho0: float = 1.0
    c0: float = 10.0
    P0: float = 100.0
    gravity: tuple = (0.0, 0.0, 0.0)
    bbox: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    v: np.ndarray = None          # [n,3] or None
    # stochastic chemistry
    u0: np.ndarray = None         # [n,S] uint32
    diffusion_matrix: np.ndarray = None   # [S,T] float64 (species-major)
    reactions: list = field(default_factory=list)   # (kind, a, b, k, type_mask[, reactants, products])
    jcN: np.ndarray = None
    irN: np.ndarray = None
    prN: np.ndarray = None
    jcG: np.ndarray = None
    irG: np.ndarray = None
    N_dense: np.ndarray = None    # [S,R] int32
    enable_rdme: bool = True
    enable_pde: bool = True
    C0: np.ndarray = None         # [n,Sc] float64 or None (= u0)
    bcs: list = field(default_factory=list)   # dicts: lo, hi, type, target, species, value
    flags: int = 0
    # custom propensities (kind-4 rows): VM code and constants; any custom
    # reaction makes the dependency graph dense (core/model.py:187-196)
    custom_propensities: bool = False
    vm_code: np.ndarray = None    # [n_code] uint32
    vm_consts: np.ndarray = None  # [n_consts] float64
    data_fn: np.ndarray = None    # [n, n_data_fn] float64
    # the C++ text of each reaction's functions and the parameter values
    # (kept for writing the reference harness's model)
    prop_cpp: list = None
    ode_cpp: list = None
    param_values: list = None

    @property
    def n(self):
        return int(self.x.shape[0])

    @property
    def n_species(self):
        return 0 if self.u0 is None else int(self.u0.shape[1])

    @property
    def n_reactions(self):
        return len(self.reactions)

    def normalized(self):
        """Return self with dtypes/contiguity fixed (in place)."""
        f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i4 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self.x = f8(self.x).reshape(-1, 3)
        self.type = i4(self.type)
        self.mass = f8(self.mass)
        self.rho = f8(self.rho)
        self.nu = f8(self.nu)
        self.fixed = i4(self.fixed)
        if self.v is not None:
            self.v = f8(self.v).reshape(-1, 3)
        self.output_steps = np.ascontiguousarray(self.output_steps, dtype=np.uint32)
        if self.u0 is not None:
            self.u0 = np.ascontiguousarray(self.u0, dtype=np.uint32)
            S = self.u0.shape[1]
            self.diffusion_matrix = f8(self.diffusion_matrix).reshape(S, self.n_types)
            if self.jcG is None:
                self.build_sparse()
            for k in ("jcN", "irN", "prN", "jcG", "irG", "N_dense"):
                setattr(self, k, i4(getattr(self, k)))
        if self.C0 is not None:
            self.C0 = f8(self.C0)
        if self.vm_code is not None:
            self.vm_code = np.ascontiguousarray(self.vm_code, dtype=np.uint32)
            self.vm_consts = f8(self.vm_consts if self.vm_consts is not None else np.zeros(0))
        if self.data_fn is not None:
            self.data_fn = f8(self.data_fn).reshape(self.n, -1)
        return self

    def build_sparse(self):
        """Stoichiometry CSC and dependency graph CSC exactly as
        core/model.py:187-268 builds them (all ones when any reaction has a
        custom propensity)."""
        S, R = self.n_species, self.n_reactions
        reactants = [rx[5] if len(rx) > 5 else _default_reactants(rx) for rx in self.reactions]
        products = [rx[6] if len(rx) > 6 else {} for rx in self.reactions]
        N = np.zeros((S, R), dtype=np.int64)
        for r in range(R):
            for s, c in reactants[r].items():
                N[s, r] -= c
            for s, c in products[r].items():
                N[s, r] += c
        jc, ir, pr = [0], [], []
        for r in range(R):
            for s in range(S):
                if N[s, r] != 0:
                    ir.append(s)
                    pr.append(int(N[s, r]))
            jc.append(len(ir))
        self.jcN, self.irN, self.prN = np.array(jc), np.array(ir), np.array(pr)
        self.N_dense = N.astype(np.int32)
        # dependency graph (model.py:187-250)
        raw = np.zeros((R, R + S))
        if self.custom_propensities:
            raw[:] = 1
        spec_to_rx = [[r for r in range(R) if s in reactants[r]] for s in range(S)]
        for s in range(S):
            for r in spec_to_rx[s]:
                raw[r, s] = 1
        for i in range(R if not self.custom_propensities else 0):
            deps = set()
            for s in list(reactants[i].keys()) + list(products[i].keys()):
                deps.update(spec_to_rx[s])
            for r in deps:
                raw[r, S + i] = 1
        jc, ir = [0], []
        for col in range(R + S):
            for row in range(R):
                if raw[row, col] != 0:
                    ir.append(row)
            jc.append(len(ir))
        self.jcG, self.irG = np.array(jc), np.array(ir)


def _default_reactants(rx):
    kind, a, b = rx[0], rx[1], rx[2]
    if kind == K_ZERO:
        return {}
    if kind == K_FIRST:
        return {a: 1}
    if kind == K_AB:
        return {a: 1, b: 1}
    return {a: 2}


class DescHandle:
    """A ctypes descriptor plus the arrays it points into (kept alive)."""

    def __init__(self, sa: SystemArrays):
        sa.normalized()
        self.sa = sa
        d = _abi.Desc()
        d.abi_version = _abi.ABI_VERSION
        d.n = sa.n
        d.dimension = sa.dimension
        d.n_types = sa.n_types
        S = sa.n_species if sa.enable_rdme else 0
        R = sa.n_reactions if sa.enable_rdme else 0
        Sc = sa.n_species if sa.enable_pde else 0
        Rc = sa.n_reactions if sa.enable_pde else 0
        d.n_species, d.n_reactions = S, R
        d.n_chem_species, d.n_chem_reactions = Sc, Rc
        d.static_domain = int(sa.static_domain)
        d.flags = int(sa.flags)
        d.nt = int(sa.nt)
        d.dt, d.h, d.rho0, d.c0, d.P0 = sa.dt, sa.h, sa.rho0, sa.c0, sa.P0
        for k in range(6):
            d.bbox[k] = sa.bbox[k]
        for k in range(3):
            d.gravity[k] = sa.gravity[k]
        self.keep = []
        d.x = _abi.ptr(sa.x, C.c_double)
        d.v = _abi.ptr(sa.v, C.c_double)
        d.type = _abi.ptr(sa.type, C.c_int32)
        d.nu = _abi.ptr(sa.nu, C.c_double)
        d.mass = _abi.ptr(sa.mass, C.c_double)
        d.rho = _abi.ptr(sa.rho, C.c_double)
        d.fixed = _abi.ptr(sa.fixed, C.c_int32)
        if sa.u0 is not None and (S > 0 or Sc > 0):
            d.u0 = _abi.ptr(sa.u0, C.c_uint32)
            d.diffusion_matrix = _abi.ptr(sa.diffusion_matrix, C.c_double)
            d.jcN = _abi.ptr(sa.jcN, C.c_int32)
            d.irN = _abi.ptr(sa.irN, C.c_int32) if sa.irN.size else None
            d.prN = _abi.ptr(sa.prN, C.c_int32) if sa.prN.size else None
            d.jcG = _abi.ptr(sa.jcG, C.c_int32)
            d.irG = _abi.ptr(sa.irG, C.c_int32) if sa.irG.size else None
            d.N_dense = _abi.ptr(sa.N_dense, C.c_int32) if sa.N_dense.size else None
            rx = (_abi.Reaction * max(1, sa.n_reactions))()
            for i, r in enumerate(sa.reactions):
                rx[i].kind, rx[i].a, rx[i].b = int(r[0]), int(r[1]), int(r[2])
                rx[i].k = float(r[3])
                rx[i].type_mask = int(r[4])
            self.keep.append(rx)
            d.reactions = C.cast(rx, C.POINTER(_abi.Reaction))
        if sa.C0 is not None:
            d.C0 = _abi.ptr(sa.C0, C.c_double)
        nb = len(sa.bcs)
        bcs = (_abi.BC * max(1, nb))()
        for i, b in enumerate(sa.bcs):
            for k in range(3):
                bcs[i].lo[k] = b["lo"][k]
                bcs[i].hi[k] = b["hi"][k]
                bcs[i].value[k] = b["value"][k] if k < len(b["value"]) else 0.0
            bcs[i].type = int(b.get("type", -1))
            bcs[i].target = int(b["target"])
            bcs[i].species = int(b.get("species", 0))
        self.keep.append(bcs)
        d.n_bcs = nb
        d.bcs = C.cast(bcs, C.POINTER(_abi.BC))
        d.n_output_steps = int(sa.output_steps.size)
        d.output_steps = _abi.ptr(sa.output_steps, C.c_uint32)
        if sa.vm_code is not None and sa.vm_code.size:
            d.n_code, d.n_consts = int(sa.vm_code.size), int(sa.vm_consts.size)
            d.code = _abi.ptr(sa.vm_code, C.c_uint32)
            d.consts = _abi.ptr(sa.vm_consts, C.c_double) if sa.vm_consts.size else None
        if sa.data_fn is not None and sa.data_fn.shape[1] > 0:
            d.n_data_fn = int(sa.data_fn.shape[1])
            d.data_fn = _abi.ptr(sa.data_fn, C.c_double)
        self.desc = d

    def byref(self):
        return C.byref(self.desc)


def output_schedule(output_steps, nt):
    """Steps at which run_simulation takes snapshots (simulate_threads.cpp:
    231-289 with the generated get_next_output, solver.py:290-299)."""
    steps = []
    nxt, idx = 0, 0
    for step in range(nt):
        if step >= nxt:
            steps.append(step)
            nxt = int(output_steps[idx]) if idx < len(output_steps) else 0xFFFFFFFF
            idx += 1
    steps.append(nt)
    return np.array(steps, dtype=np.uint32)


def write_harness_dir(sa: SystemArrays, path, seed, run_steps=None, debug_flag=0,
                      force_step=0, nbr_step=0, threads=0, write_vtk=0, state=None, start_step=0):
    """Write a model in the raw format read by oracle/ref_harness.cpp.

    state / start_step: continue a run mid-way.  state holds the particle
    state after step start_step - 1 (x, v, rho, F, Fbp, Frho, C, Q and the
    counts xx, as Engine.state() returns them); the harness then runs steps
    start_step .. start_step + run_steps - 1."""
    sa.normalized()
    os.makedirs(path, exist_ok=True)
    S = sa.n_species if sa.enable_rdme else 0
    R = sa.n_reactions if sa.enable_rdme else 0
    Sc = sa.n_species if sa.enable_pde else 0
    Rc = sa.n_reactions if sa.enable_pde else 0
    meta = {
        "n": sa.n, "dimension": sa.dimension, "n_types": sa.n_types, "n_species": S,
        "n_reactions": R, "n_chem_species": Sc, "n_chem_reactions": Rc, "nt": sa.nt,
        "seed": seed, "debug_flag": debug_flag, "static_domain": int(sa.static_domain),
        "dt": repr(float(sa.dt)), "h": repr(float(sa.h)), "rho0": repr(float(sa.rho0)),
        "c0": repr(float(sa.c0)), "P0": repr(float(sa.P0)),
        "xlo": repr(float(sa.bbox[0])), "xhi": repr(float(sa.bbox[1])),
        "ylo": repr(float(sa.bbox[2])), "yhi": repr(float(sa.bbox[3])),
        "zlo": repr(float(sa.bbox[4])), "zhi": repr(float(sa.bbox[5])),
        "g0": repr(float(sa.gravity[0])), "g1": repr(float(sa.gravity[1])),
        "g2": repr(float(sa.gravity[2])), "n_bcs": len(sa.bcs),
        "n_output_steps": int(sa.output_steps.size), "force_step": force_step,
        "nbr_step": nbr_step, "threads": threads, "write_vtk": int(write_vtk),
    }
    if run_steps is not None:
        meta["run_steps"] = run_steps
    if state is not None:
        meta["start_step"] = int(start_step)
    with open(os.path.join(path, "meta.txt"), "w") as f:
        for k, v in meta.items():
            f.write(f"{k} {v}\n")
    w = lambda name, a, dt: np.ascontiguousarray(a, dtype=dt).tofile(os.path.join(path, name))
    w("x.f64", sa.x if state is None else state["x"], "<f8")
    w("v.f64", (sa.v if sa.v is not None else np.zeros((sa.n, 3))) if state is None else state["v"], "<f8")
    if state is not None:
        w("F.f64", state["F"], "<f8")
        w("Fbp.f64", state["Fbp"], "<f8")
        w("Frho.f64", state["Frho"], "<f8")
        w("Q.f64", state["Q"], "<f8")
        w("C.f64", state["C"], "<f8")
    w("type.i32", sa.type, "<i4")
    w("nu.f64", sa.nu, "<f8")
    w("mass.f64", sa.mass, "<f8")
    w("rho.f64", sa.rho if state is None else state["rho"], "<f8")
    w("fixed.i32", sa.fixed, "<i4")
    nsp = sa.n_species
    u0 = sa.u0 if sa.u0 is not None else np.zeros((sa.n, 0), np.uint32)
    if state is not None and nsp:
        u0 = state["xx"]
    w("u0.u32", u0, "<u4")
    dm = sa.diffusion_matrix if sa.diffusion_matrix is not None else np.zeros((0,))
    w("dmat.f64", dm, "<f8")
    ri = np.array([[r[0], r[1], r[2]] for r in sa.reactions], dtype=np.int32).reshape(-1, 3)
    w("rx_int.i32", ri, "<i4")
    w("rx_k.f64", np.array([r[3] for r in sa.reactions], dtype=np.float64), "<f8")
    w("rx_mask.u64", np.array([r[4] for r in sa.reactions], dtype=np.uint64), "<u8")
    if nsp:
        w("jcN.i32", sa.jcN, "<i4")
        w("irN.i32", sa.irN, "<i4")
        w("prN.i32", sa.prN, "<i4")
        w("jcG.i32", sa.jcG, "<i4")
        w("irG.i32", sa.irG, "<i4")
        w("ndense.i32", sa.N_dense, "<i4")
    bd = np.array([list(b["lo"]) + list(b["hi"]) + list(b["value"]) + [0.0] * (3 - len(b["value"]))
                   for b in sa.bcs], dtype=np.float64).reshape(-1, 9)
    bi = np.array([[b.get("type", -1), b["target"], b.get("species", 0)] for b in sa.bcs],
                  dtype=np.int32).reshape(-1, 3)
    w("bc.f64", bd, "<f8")
    w("bc.i32", bi, "<i4")
    w("out_steps.u32", sa.output_steps, "<u4")
    if sa.data_fn is not None and sa.data_fn.shape[1] > 0:
        # the generated input_data_fn layout: [function][particle]
        w("data_fn.f64", sa.data_fn.T, "<f8")
        with open(os.path.join(path, "meta.txt"), "a") as f:
            f.write(f"n_data_fn {sa.data_fn.shape[1]}\n")


def harness_propensity_source(sa: SystemArrays):
    """C++ source of a model's reaction functions in the form the
    reference generator writes them (solver.py:160-190 det<name>, 300-310
    parameter definitions, 344-373 propensities), one extern "C" symbol per
    reaction (prop_<r>, det_<r>), for oracle/ref_harness.cpp."""
    out = ["#include <math.h>", "#include <stddef.h>"]
    for i, v in enumerate(sa.param_values or []):
        out.append(f"static const double P{i} = {float(v)!r};")

    def body(expr, mask):
        if not mask:
            return f"return {expr};"
        conds = "||".join(f"sd == {t}" for t in range(64) if (mask >> t) & 1)
        return f"if({conds}){{\nreturn {expr};\n}}else{{\n\treturn 0.0;}}"
    for r, rx in enumerate(sa.reactions):
        mask = int(rx[4])
        out.append(f'extern "C" double prop_{r}(const int *x, double t, const double vol, const double *data_fn, '
                   f'int sd)\n{{\n{body(sa.prop_cpp[r], mask)}\n}}')
        out.append(f'extern "C" double det_{r}(const double *x, double t, const double vol, const double *data_fn, '
                   f'int sd)\n{{\n{body(sa.ode_cpp[r], mask)}\n}}')
    return "\n".join(out) + "\n"
