"""Host-side model API mirroring spatialpy.core."""
from .errors import ModelError, DomainError, SimulationError, SimulationTimeout
