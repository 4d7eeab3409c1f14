"""``within._within`` -- the module the reference's PyO3 extension registers
(crates/within-py/src/lib.rs:26-47), here the ctypes binding of libwithin_amd.so."""
from within_amd._within import (  # noqa: F401
    AdditiveSchwarz,
    ApproxCholConfig,
    ApproxSchurConfig,
    BatchSolveResult,
    CoefficientLayout,
    Effect,
    LocalSolverConfig,
    LsmrOptions,
    Preconditioner,
    PreconditionerConfig,
    ReductionStrategy,
    ScalingConfig,
    Schur,
    SolveResult,
    Solver,
    UnidentifiedDirection,
    solve,
    solve_batch,
)
