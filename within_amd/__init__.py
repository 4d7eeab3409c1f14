"""within_amd -- MI355X-native fixed-effects projection (drop-in for ``within``).

Solves D'W D x = D'W z with modified LSMR and an additive-Schwarz preconditioner on one
AMD Instinct MI355X, returning x and the demeaned output z - D x.  The public names mirror
the reference package ``within`` (python/within/__init__.py:50-78).
"""
from ._within import (
    BatchSolveResult,
    CoefficientLayout,
    Effect,
    LsmrOptions,
    Preconditioner,
    PreconditionerConfig,
    SolveResult,
    Solver,
    UnidentifiedDirection,
    fwl,
    gram,
    last_many_shape,
    set_many_shape,
    solve,
    solve_batch,
    solve_many,
)
from . import config  # noqa: F401

__all__ = [
    "BatchSolveResult",
    "CoefficientLayout",
    "Effect",
    "LsmrOptions",
    "Preconditioner",
    "PreconditionerConfig",
    "SolveResult",
    "Solver",
    "UnidentifiedDirection",
    "solve",
    "solve_batch",
    "solve_many",
    "gram",
    "fwl",
    "last_many_shape",
    "set_many_shape",
    "config",
]
