"""``within`` -- the drop-in package name of within_amd.

Code written against the reference package (``import within``, ``from within import solve,
Solver``, ``from within.config import AdditiveSchwarz``) runs unchanged: this package
re-exports the names of python/within/__init__.py:50-78 from within_amd, whose solves run on
the MI355X through libwithin_amd.so.  The MI355X additions (``solve_many``, ``gram``,
``fwl``) stay in ``within_amd``.
"""
from within._within import (
    BatchSolveResult,
    CoefficientLayout,
    Effect,
    LsmrOptions,
    Preconditioner,
    PreconditionerConfig,
    SolveResult,
    Solver,
    UnidentifiedDirection,
    solve,
    solve_batch,
)
from within import config  # noqa: F401 -- expose the submodule as `within.config`

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
    "config",
]
