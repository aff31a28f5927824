"""Model description API mirroring spatialpy.core (Model, Species, Parameter,
Reaction, TimeSpan, initial and boundary conditions).  Only what the engine's
lowering needs is restated; the semantics follow the reference files cited
in each docstring."""
import math
from collections import OrderedDict

import numpy as np

from .errors import ModelError


class Species:
    """spatialpy/core/species.py: name, diffusion coefficient, restrict_to."""

    def __init__(self, name=None, diffusion_coefficient=None, restrict_to=None):
        if name is None or not isinstance(name, str):
            raise ModelError("Species name must be a string")
        if diffusion_coefficient is None:
            raise ModelError("Species diffusion_coefficient is required")
        self.name = name
        self.diffusion_coefficient = diffusion_coefficient
        if restrict_to is not None and not isinstance(restrict_to, list):
            restrict_to = [restrict_to]
        self.restrict_to = None if restrict_to is None else [f"type_{t}" for t in restrict_to]

    def __repr__(self):
        return f"Species({self.name}, D={self.diffusion_coefficient})"


class Parameter:
    """spatialpy/core/parameter.py: a named value given as a number or an
    expression over earlier parameters."""

    def __init__(self, name=None, expression=None, value=None):
        if name is None:
            raise ModelError("Parameter name is required")
        self.name = name
        self.expression = expression if expression is not None else value
        self.value = None if value is None else float(value)

    def _evaluate(self, namespace):
        if self.value is None or self.expression is not None:
            expr = self.expression
            if isinstance(expr, (int, float)):
                self.value = float(expr)
            else:
                self.value = float(eval(str(expr), {"__builtins__": {}, "math": math},
                                        dict(namespace)))
        return self.value


class Reaction:
    """spatialpy/core/reaction.py:73-140.  A mass-action reaction (rate: a
    Parameter, its name or a number) or a custom one (propensity_function
    and/or ode_propensity_function, each defaulting to the other).  Both
    strings are kept as the reference keeps them, after
    Reaction._create_custom_propensity (reaction.py:260-330); the solver
    turns them into C++ text and VM programs (spatialpy_amd/expression.py)."""

    def __init__(self, name=None, reactants=None, products=None, propensity_function=None,
                 ode_propensity_function=None, rate=None, annotation=None, restrict_to=None):
        from ..expression import ExpressionError, mass_action_texts, reaction_text
        self.name = name
        self.reactants = OrderedDict()
        for s, c in (reactants or {}).items():
            key = s.name if isinstance(s, Species) else s
            self.reactants[key] = self.reactants.get(key, 0) + c
        self.products = OrderedDict()
        for s, c in (products or {}).items():
            key = s.name if isinstance(s, Species) else s
            self.products[key] = self.products.get(key, 0) + c
        if isinstance(rate, Parameter):
            rate = rate.name
        elif isinstance(rate, (int, float)) and not isinstance(rate, bool):
            rate = str(rate)
        self.marate = rate
        self.annotation = annotation
        if restrict_to is not None and not isinstance(restrict_to, list):
            restrict_to = [restrict_to]
        # stored as type names, as reaction.py:136-139 stores them
        self.restrict_to = None if restrict_to is None else [f"type_{t}" for t in restrict_to]
        if rate is not None:
            self.massaction = True
            self.type = "mass-action"
            try:
                prop, ode = mass_action_texts(rate, self.reactants)
            except ExpressionError as e:
                raise ModelError(str(e)) from e
        else:
            self.massaction = False
            self.type = "customized"
            if propensity_function is None and ode_propensity_function is None:
                raise ModelError(f"Reaction {name}: give either rate (mass action) or propensity_function")
            prop = propensity_function if propensity_function is not None else ode_propensity_function
            ode = ode_propensity_function if ode_propensity_function is not None else propensity_function
        try:
            self.propensity_function = reaction_text(prop)
            self.ode_propensity_function = reaction_text(ode)
        except (SyntaxError, ExpressionError) as e:
            raise ModelError(f"Reaction {name}: invalid propensity function: {e}") from e

    def propensity_expression(self):
        """The propensity string as the reference keeps it (stage A of the
        expression pipeline, core/reaction.py:325-330)."""
        return self.propensity_function


class DataFunction:
    """spatialpy/core/datafunction.py: a named function of the initial
    particle position, subclassed with map(point) -> float.  Its values
    are taken once at t = 0 and read by propensities as the data function's
    name."""

    def __init__(self, name=None):
        if name is not None:
            self.name = name
        if getattr(self, "name", None) is None:
            raise ModelError("DataFunction must have a 'name'")

    def map(self, point):
        raise ModelError(f"{self.name}: DataFunction.map() must be implemented.")


class TimeSpan:
    """spatialpy/core/timespan.py: output times and the output steps
    (TimeSpan._set_timesteps, timespan.py:72-85)."""

    def __init__(self, items, timestep_size=None):
        items = np.asarray(items, dtype=float)
        if items.size < 2:
            raise ModelError("TimeSpan needs at least two points")
        diff = np.diff(items)
        if not np.allclose(diff, diff[0]):
            raise ModelError("TimeSpan items must be uniformly spaced")
        self.timestep_size = timestep_size
        self._set_timesteps(diff[0], len(diff))

    def _set_timesteps(self, output_interval, num_steps):
        if self.timestep_size is None:
            self.timestep_size = output_interval
        self.output_freq = output_interval / self.timestep_size
        self.num_timesteps = math.ceil(num_steps * self.output_freq)
        steps = np.arange(0, self.num_timesteps + self.timestep_size, self.output_freq)
        self.output_steps = np.unique(np.round(steps).astype(int))
        self.items = self.output_steps * self.timestep_size

    @classmethod
    def linspace(cls, t=20, num_points=None, timestep_size=None):
        if num_points is None:
            num_points = int(t / 0.05) + 1
        return cls(np.linspace(0, t, num_points), timestep_size=timestep_size)

    @classmethod
    def arange(cls, increment, t=20, timestep_size=None):
        return cls(np.arange(0, t + increment, increment), timestep_size=timestep_size)

    def __len__(self):
        return len(self.items)

    def __iter__(self):
        return iter(self.items)

    def __getitem__(self, k):
        return self.items[k]


class PlaceInitialCondition:
    """initialcondition.py:35-75: count molecules at the vertex that
    Domain.closest_vertex returns for location (core/domain.py:460-477):
    the first vertex of smallest positive distance -- a vertex at distance
    exactly 0 is skipped, so a location on a vertex places the molecules on
    its nearest other vertex."""

    def __init__(self, species, count, location=None):
        self.species, self.count, self.location = species, count, location if location is not None else [0, 0, 0]

    def apply(self, model):
        s = model._species_index(self.species)
        vtx = closest_vertex(model.domain.vertices, self.location)
        model.u0[s, vtx] += self.count


def closest_vertex(vertices, point):
    """core/domain.py:460-477: the first index i with the smallest positive
    numpy.linalg.norm(vertices[i] - point).  Candidates are found with
    vectorised squared distances; the reference's own norm decides among
    those within rounding of the minimum, in index order."""
    v = np.asarray(vertices, dtype=float)
    p = np.asarray(point, dtype=float)
    d2 = np.einsum("ij,ij->i", v - p, v - p)
    pos = d2 > 0
    if not pos.any():
        return None
    m = d2[pos].min()
    best, best_d = None, None
    for i in np.nonzero(pos & (d2 <= m * (1 + 1e-9)))[0]:
        d = np.linalg.norm(v[i, :] - p)
        if d > 0 and (best_d is None or d < best_d):
            best, best_d = int(i), d
    return best


class UniformInitialCondition:
    """initialcondition.py:77-130: count molecules in every voxel (of types)."""

    def __init__(self, species, count, types=None):
        self.species, self.count = species, count
        self.types = None if types is None else [f"type_{t}" for t in types]

    def apply(self, model):
        s = model._species_index(self.species)
        for i, t in enumerate(model.domain.type_id):
            if self.types is None or t in self.types:
                model.u0[s, i] += self.count


class ScatterInitialCondition:
    """initialcondition.py:133-200: count molecules on uniformly random voxels
    (of types).  The reference draws from the global numpy.random; a seed
    here (or a Generator) makes it reproducible."""

    def __init__(self, species, count, types=None, seed=None):
        self.species, self.count = species, count
        self.types = None if types is None else [f"type_{t}" for t in types]
        self.seed = seed

    def apply(self, model):
        s = model._species_index(self.species)
        if self.types is None:
            allowed = np.arange(model.domain.get_num_voxels())
        else:
            allowed = np.array([i for i, t in enumerate(model.domain.type_id) if t in self.types])
            if allowed.size == 0:
                raise ModelError("ScatterInitialCondition has no voxel of the given types")
        rng = np.random.default_rng(self.seed) if self.seed is not None else np.random
        idx = allowed[rng.integers(0, allowed.size, size=self.count)] if self.seed is not None else \
            allowed[np.random.randint(0, allowed.size, size=self.count)]
        np.add.at(model.u0[s], idx, 1)


class BoundaryCondition:
    """boundarycondition.py: a box (any bound may be None), an optional type,
    and a target (deterministic species, 'v', 'nu' or 'rho') set to value."""

    def __init__(self, xmin=None, xmax=None, ymin=None, ymax=None, zmin=None, zmax=None,
                 type_id=None, deterministic=True, target=None, value=None, model=None):
        if target is None:
            raise ModelError("BoundaryCondition needs a target")
        self.bounds = (xmin, xmax, ymin, ymax, zmin, zmax)
        self.type_id = None if type_id is None else f"type_{type_id}"
        self.deterministic = deterministic
        self.target = target.name if isinstance(target, Species) else target
        if isinstance(value, int):
            value = float(value)
        self.value = value
        if all(b is None for b in self.bounds) and self.type_id is None:
            raise ModelError("need at least one condition on the BoundaryCondition")


class Model:
    """spatialpy/core/model.py:65-1056 (the parts the engine consumes)."""

    def __init__(self, name="spatialpy"):
        self.name = name
        self.listOfParameters = OrderedDict()
        self.listOfSpecies = OrderedDict()
        self.listOfReactions = OrderedDict()
        self.listOfInitialConditions = []
        self.listOfBoundaryConditions = []
        self.listOfDataFunctions = OrderedDict()
        self.domain = None
        self.staticDomain = True
        self.enable_rdme = True
        self.enable_pde = True
        self.tspan = None
        self.u0 = None

    # -- building
    def add_species(self, species):
        for s in species if isinstance(species, list) else [species]:
            self.listOfSpecies[s.name] = s
        return species

    def add_parameter(self, params):
        for p in params if isinstance(params, list) else [params]:
            self.listOfParameters[p.name] = p
        return params

    def add_reaction(self, reactions):
        for r in reactions if isinstance(reactions, list) else [reactions]:
            self.listOfReactions[r.name] = r
        return reactions

    def add_domain(self, domain, allow_all_types=False):
        self.domain = domain

    def add_initial_condition(self, ic):
        self.listOfInitialConditions.append(ic)

    def add_boundary_condition(self, bc):
        self.listOfBoundaryConditions.append(bc)

    def add_data_function(self, data_function):
        """core/model.py add_data_function: one DataFunction or a list."""
        for df in data_function if isinstance(data_function, list) else [data_function]:
            if not isinstance(df, DataFunction) and type(df).__name__ != "DataFunction" and \
                    not hasattr(df, "map"):
                raise ModelError("data_function must be a DataFunction")
            self.listOfDataFunctions[df.name] = df
        return data_function

    def timespan(self, time_span, timestep_size=None):
        if isinstance(time_span, TimeSpan):
            self.tspan = time_span
        else:
            self.tspan = TimeSpan(time_span, timestep_size=timestep_size)

    def set_timesteps(self, output_interval, num_steps, timestep_size=None):
        self.tspan = TimeSpan(np.arange(0, (num_steps + 1) * output_interval, output_interval)[: num_steps + 1],
                              timestep_size=timestep_size)

    def get_num_species(self):
        return len(self.listOfSpecies)

    def get_num_reactions(self):
        return len(self.listOfReactions)

    def _species_index(self, sp):
        name = sp.name if isinstance(sp, Species) else sp
        if name not in self.listOfSpecies:
            raise ModelError(f"Species {name} does not exist in the model.")
        return list(self.listOfSpecies).index(name)

    # -- lowering helpers
    def resolve_parameters(self):
        ns = OrderedDict()
        for name, p in self.listOfParameters.items():
            ns[name] = p._evaluate(ns)
        return ns

    def compile_prep(self):
        """Validate and fill u0[species, voxel] (model.py:178-185, 982-1019)."""
        if self.tspan is None:
            raise ModelError("The model's timespan is not set.")
        if self.domain is None:
            raise ModelError("The model's domain is not set.  Use 'add_domain()'.")
        self.domain.compile_prep()
        self.u0 = np.zeros((self.get_num_species(), self.domain.get_num_voxels()))
        for ic in self.listOfInitialConditions:
            ic.apply(self)
        return self.resolve_parameters()

    def run(self, number_of_trajectories=1, seed=None, timeout=None, number_of_threads=None,
            debug_level=0, debug=False, profile=False, solver=None):
        """model.py:1021-1056 plus `solver=`: a Solver class or instance
        (defaults to the GPU Solver)."""
        from ..solver import Solver
        if solver is None:
            sol = Solver(self, debug_level=debug_level)
        elif isinstance(solver, type):
            sol = solver(self, debug_level=debug_level)
        else:
            sol = solver
        return sol.run(number_of_trajectories=number_of_trajectories, seed=seed, timeout=timeout,
                       number_of_threads=number_of_threads, debug=debug, profile=profile)
