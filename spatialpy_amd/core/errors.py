"""Exception types mirroring spatialpy/core/spatialpyerror.py:18-118."""


class ModelError(Exception):
    """Model construction failed."""


class DomainError(ModelError):
    """Domain construction failed."""


class SimulationError(Exception):
    """The engine reported a failure (the reference: non-zero exit code,
    spatialpy/solvers/solver.py:595-597)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class SimulationTimeout(SimulationError):
    """Run exceeded its timeout (solver.py:583-586)."""
