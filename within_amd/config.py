"""Advanced configuration objects (mirrors python/within/config.py:33-51)."""
from ._within import (
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
