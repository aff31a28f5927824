"""Solver drop-in: the same constructor / compile / run interface as
spatialpy/solvers/solver.py:46-607, running the HIP engine in-process
through the C ABI instead of generating, compiling and spawning a C++
program."""
import math
import os

import numpy as np

from . import _abi, expression
from .core.errors import ModelError, SimulationError
from .core.model import Parameter, Species
from .lowering import SystemArrays, output_schedule


class Result:
    """In-memory result with spatialpy.core.result.Result's getters
    (result.py:270-402, 601-660).  Snapshot k is what the reference writes
    to output{k}.vtk."""

    def __init__(self, model, data, steps, sa):
        self.model = model
        self.data = data
        self.steps = steps
        self.sa = sa
        self.success = True
        self.timeout = False
        self.trajectories = [self]

    def get_timespan(self):
        return self.model.tspan.items

    def _tindex(self, timepoints):
        idx = np.arange(len(self.get_timespan()))
        if timepoints is not None:
            if isinstance(timepoints, float):
                raise SimulationError("timepoints argument must be an integer, the index of time timespan")
            idx = np.atleast_1d(idx[timepoints])
        return idx

    def get_species(self, species, timepoints=None, concentration=False, deterministic=False, debug=False):
        name = species.name if isinstance(species, Species) else species
        if name not in self.model.listOfSpecies:
            raise SimulationError(f"Species '{name}' not found")
        s = list(self.model.listOfSpecies).index(name)
        idx = self._tindex(timepoints)
        if deterministic:
            out = self.data["C"][idx, :, s].astype(float)
        else:
            out = self.data["xx"][idx, :, s].astype(float)
            if concentration:
                out = out / (self.sa.mass / self.data["rho"][idx])
        return out[0] if out.shape[0] == 1 else out

    def get_property(self, property_name, timepoints=None):
        idx = self._tindex(timepoints)
        if property_name == "v":
            out = self.data["v"][idx]
        elif property_name in ("rho", "bvf_phi", "nu"):
            out = self.data[property_name][idx]
        elif property_name == "mass":
            out = np.broadcast_to(self.sa.mass, (len(idx), self.sa.n))
        elif property_name == "id":
            out = np.broadcast_to(self.sa.type, (len(idx), self.sa.n))
        elif property_name == "type":
            names = {v: k for k, v in self.model.domain.typeNdxMapping.items()}
            out = np.broadcast_to(np.array([names[t] for t in self.sa.type], dtype=object), (len(idx), self.sa.n))
        else:
            raise SimulationError(f"unknown property {property_name}")
        return out[0] if out.shape[0] == 1 else out

    def export_vtk(self, folder):
        """Write this trajectory's snapshots as the reference engine does:
        output{k}.vtk and output0_boundingBox.vtk in folder (the files the
        reference's Result reads, core/result.py:173-183)."""
        from .lowering import output_schedule
        from .vtk import write_result
        names = list(self.model.listOfSpecies)
        steps = output_schedule(self.sa.output_steps, self.sa.nt)
        write_result(folder, self.data, self.sa, names, steps)
        return folder

    def __eq__(self, other):
        if not isinstance(other, Result):
            return False
        return all(np.array_equal(self.data[k], other.data[k]) for k in self.data)

    def __ne__(self, other):
        return not self.__eq__(other)

    def append(self, other):
        self.trajectories.append(other)

    def __len__(self):
        return len(self.trajectories)

    def __getitem__(self, i):
        return self.trajectories[i]


class Solver:
    """solvers/solver.py:46-607.  `exact_replay=True` feeds the engine the
    std::mt19937_64(seed) stream (with host-libm exponentials), which makes a
    run event-for-event identical to the reference engine run with -s seed;
    otherwise the GPU's own counter-based stream is used."""

    def __init__(self, model, debug_level=0, exact_replay=False, flags=0):
        if model is None:
            raise SimulationError("A model is required to run the simulation.")
        self.model = model
        self.debug_level = debug_level
        self.is_compiled = False
        self.h = None
        self.exact_replay = exact_replay
        self.flags = flags
        self.sa = None
        self.engine = None

    # -- lowering (solver.py:100-419, core/model.py:187-268)
    def compile(self, debug=False, profile=False):
        """Lower the model and create the GPU engine (the reference's
        build step, solver.py:443-507)."""
        self.sa = self.lower()
        from .engine import Engine
        self.engine = Engine(self.sa)
        self.is_compiled = True

    def lower(self):
        """The model as engine arrays (no GPU needed)."""
        m = self.model
        prep = m.compile_prep()
        if isinstance(prep, tuple):
            # a spatialpy.Model: compile_prep returns the stoichiometry and
            # dependency graph and resolves each Parameter's value in place
            # (core/model.py:982-1019)
            params = {name: float(p.value) for name, p in m.listOfParameters.items()}
        else:
            params = prep
        # parameter names as the generated code spells them (P0, P1, ...;
        # core/model.py:703-715): compile_prep may leave a species' diffusion
        # coefficient as one of these
        sanitized = {f"P{i}": name for i, name in enumerate(m.listOfParameters)}
        d = m.domain
        types = d.typeNdxMapping
        T = len(types) - 1
        tp = np.array([types[t] for t in d.type_id], dtype=np.int32)
        species = list(m.listOfSpecies.values())
        S = len(species)
        sidx = {s.name: i for i, s in enumerate(species)}

        def sindex(x):
            return sidx[x if isinstance(x, str) else x.name]

        # diffusion matrix [S][T] (solver.py:265-286)
        D = np.zeros((S, T))
        for i, sp in enumerate(species):
            coef = sp.diffusion_coefficient
            if isinstance(coef, str):
                coef = params[sanitized.get(coef, coef)]
            elif isinstance(coef, Parameter) or hasattr(coef, "value"):
                coef = coef.value
            for name, j in types.items():
                if j == 0:
                    continue
                if sp.restrict_to is None or name in sp.restrict_to:
                    D[i, j - 1] = float(coef)
        # reactions: the propensity and deterministic-rate strings become C++
        # text (solver.py:160-190, 344-373 via BuildExpression.getexpr_cpp),
        # read with C types; the four mass-action shapes run from the table,
        # anything else as VM programs
        dfs = list(m.listOfDataFunctions.values())
        ns = expression.model_namespace(list(m.listOfSpecies), list(m.listOfParameters), [f.name for f in dfs])
        cparams = {f"P{i}": float(params[name]) for i, name in enumerate(m.listOfParameters)}
        prog = expression.Program()
        rx, prop_cpp, ode_cpp = [], [], []
        for r in m.listOfReactions.values():
            try:
                pc = expression.cpp_text(r.propensity_function, ns)
                oc = expression.cpp_text(r.ode_propensity_function, ns)
                tree_p = expression.parse_c(pc, cparams, S, len(dfs), expression.I)
                tree_o = expression.parse_c(oc, cparams, S, len(dfs), expression.D)
            except (SyntaxError, expression.ExpressionError) as e:
                raise SimulationError(f"Reaction {r.name}: {e}") from e
            prop_cpp.append(pc)
            ode_cpp.append(oc)
            shape = expression.table_shape(tree_p, tree_o)
            if shape is not None:
                kind, a, b, k = shape
            else:
                try:
                    kind, a, b, k = expression.K_PROGRAM, prog.add(tree_p), prog.add(tree_o), 0.0
                except expression.ExpressionError as e:
                    raise SimulationError(f"Reaction {r.name}: {e}") from e
            reac = {sindex(sp): c for sp, c in r.reactants.items()}
            prod = {sindex(sp): c for sp, c in r.products.items()}
            mask = 0
            if r.restrict_to:
                for t in r.restrict_to:
                    mask |= 1 << types[t]
            rx.append((kind, a, b, k, mask, reac, prod))
        x = d.vertices
        if not np.count_nonzero(x[:, 1]):
            dim = 1
        elif not np.count_nonzero(x[:, 2]):
            dim = 2
        else:
            dim = 3
        d.dimensions = dim
        if self.h is None:
            self.h = d.find_h()
        if self.h == 0.0:
            raise ModelError("h (basis function width) can not be zero.")
        bcs = []
        for bc in m.listOfBoundaryConditions:
            bounds = getattr(bc, "bounds", None)
            if bounds is None:      # a spatialpy.BoundaryCondition
                bounds = (bc.xmin, bc.xmax, bc.ymin, bc.ymax, bc.zmin, bc.zmax)
            lo = [-math.inf if v is None else float(v) for v in bounds[0::2]]
            hi = [math.inf if v is None else float(v) for v in bounds[1::2]]
            tt = -1 if bc.type_id is None else types[bc.type_id]
            target = bc.target if isinstance(bc.target, str) else bc.target.name
            if target in m.listOfSpecies:
                if not bc.deterministic:
                    raise SimulationError("BoundaryConditions don't work for stochastic species yet.")
                bcs.append(dict(lo=lo, hi=hi, type=tt, target=_abi.BC_CHEM, species=sidx[target],
                                value=[float(bc.value)]))
            elif target == "v":
                bcs.append(dict(lo=lo, hi=hi, type=tt, target=_abi.BC_V, value=[float(v) for v in bc.value]))
            elif target in ("nu", "rho"):
                bcs.append(dict(lo=lo, hi=hi, type=tt, target=_abi.BC_NU if target == "nu" else _abi.BC_RHO,
                                value=[float(bc.value)]))
            else:
                raise SimulationError(f"Unable handle boundary condition for property '{target}'")
        ts = m.tspan
        g = tuple(d.gravity) if d.gravity is not None else (0.0, 0.0, 0.0)
        sa = SystemArrays(
            x=x, type=tp, mass=d.mass, rho=d.rho, nu=d.nu, fixed=d.fixed.astype(np.int32), n_types=T,
            dimension=dim, h=float(self.h), dt=float(ts.timestep_size), nt=int(ts.num_timesteps),
            output_steps=ts.output_steps, static_domain=int(bool(m.staticDomain)), rho0=float(d.rho0),
            c0=float(d.c0), P0=float(d.P0), gravity=g,
            bbox=(d.xlim[0], d.xlim[1], d.ylim[0], d.ylim[1], d.zlim[0], d.zlim[1]),
            u0=np.ascontiguousarray(m.u0.T).astype(np.uint32) if S else None,
            diffusion_matrix=D if S else None, reactions=rx, enable_rdme=m.enable_rdme, enable_pde=m.enable_pde,
            bcs=bcs, flags=self.flags,
            custom_propensities=any(not r.massaction for r in m.listOfReactions.values()))
        if prog.code:
            sa.vm_code, sa.vm_consts = prog.arrays()
        if dfs and S:
            # solver.py:192-198, 239-249: each function's value at the initial
            # positions (one column per function)
            sa.data_fn = np.array([[float(f.map([p[0], p[1], p[2]])) for f in dfs] for p in x],
                                  dtype=np.float64).reshape(len(x), len(dfs))
        sa.prop_cpp, sa.ode_cpp, sa.param_values = prop_cpp, ode_cpp, list(cparams.values())
        if S:
            sa.build_sparse()
        return sa.normalized()

    def run(self, number_of_trajectories=1, seed=None, timeout=None, number_of_threads=None, debug=False,
            profile=False, verbose=True):
        """solver.py:510-607: trajectory i uses seed+i.  timeout applies to
        each trajectory as the reference's Popen.wait(timeout) does
        (solver.py:574-586): a trajectory that runs past it stops at the next
        step boundary, keeps the snapshots taken so far and is marked
        result.timeout; the remaining trajectories still run.  A batched run
        (several seeded trajectories at once) shares one limit, since its
        trajectories advance together."""
        if not self.is_compiled:
            self.compile(debug=debug, profile=profile)
        self.engine.set_timeout(timeout)
        fields = ("xx", "C", "x", "v", "rho", "bvf_phi", "nu")
        first = None
        if number_of_trajectories > 1 and not self.exact_replay and self.sa.n_species > 0:
            # one batched GPU run: the particle state is shared, every
            # trajectory has its own counts, queue and stream (seed + i)
            base = seed if seed is not None else int.from_bytes(os.urandom(8), "little")
            seeds = [(base + i) & 0xFFFFFFFFFFFFFFFF for i in range(number_of_trajectories)]
            data = self.engine.run_batch(seeds, fields=fields, partial_ok=True)
            for i in range(number_of_trajectories):
                d = {f: data[f] for f in fields if f != "xx" and f in data}
                d["xx"] = data["xx"][i]
                d["steps"] = data["steps"]
                res = Result(self.model, d, d["steps"], self.sa)
                res.timeout = bool(data["timeout"])
                if first is None:
                    first = res
                else:
                    first.append(res)
            return first
        for i in range(number_of_trajectories):
            s = (seed + i) if seed is not None else int.from_bytes(os.urandom(8), "little")
            if self.exact_replay:
                data = self._run_replay(s)
            else:
                data = self.engine.run(seed=s, fields=fields, partial_ok=True)
            res = Result(self.model, data, data["steps"], self.sa)
            res.timeout = bool(data.pop("timeout", False))
            if first is None:
                first = res
            else:
                first.append(res)
        return first

    def _run_replay(self, seed):
        from .engine import mt19937_64_stream
        n = 1 << 20
        while True:
            words, expo = mt19937_64_stream(seed & 0xFFFFFFFFFFFFFFFF, n)
            try:
                return self.engine.run(words=words, expo=expo, fields=("xx", "C", "x", "v", "rho", "bvf_phi", "nu"),
                                       partial_ok=True)
            except SimulationError as e:
                if e.code != 9 or n >= (1 << 31):
                    raise
                n *= 4
