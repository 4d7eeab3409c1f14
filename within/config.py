"""Advanced configuration objects of ``within`` (python/within/config.py:33-51)."""
from within._within import (
    AdditiveSchwarz,
    ApproxCholConfig,
    ApproxSchurConfig,
    LocalSolverConfig,
    ReductionStrategy,
    ScalingConfig,
    Schur,
)

__all__ = [
    "AdditiveSchwarz",
    "ApproxCholConfig",
    "ApproxSchurConfig",
    "LocalSolverConfig",
    "ReductionStrategy",
    "ScalingConfig",
    "Schur",
]
