"""spatialpy_amd -- SpatialPy's sSSA-SDPD engine rebuilt for AMD MI355X (gfx950).

The package keeps SpatialPy's Model / Solver / Result interface and runs the
engine as hand-written HIP kernels in libssasdpd.so, called through a C ABI
(include/ssasdpd.h).
"""
from .core.errors import ModelError, DomainError, SimulationError, SimulationTimeout
from .core.model import (Model, Species, Parameter, Reaction, TimeSpan, PlaceInitialCondition,
                         UniformInitialCondition, ScatterInitialCondition, BoundaryCondition,
                         DataFunction)
from .core.domain import Domain
from .lowering import SystemArrays, output_schedule
from .solver import Solver, Result

__version__ = "0.1.0"
