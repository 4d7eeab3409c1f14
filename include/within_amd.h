/*
 * within_amd.h -- C ABI of the MI355X fixed-effects projection library (libwithin_amd.so).
 *
 * The library solves the weighted fixed-effects normal equations D'W D x = D'W z with
 * (modified) LSMR and an additive-Schwarz / diagonal / no preconditioner, on one AMD
 * Instinct MI355X (gfx950), and returns x and the demeaned output z - D x.
 *
 * Every entry point replaces one item of the reference's public surface (a Rust crate
 * plus a PyO3 module); the comment above each declaration names the reference
 * interface (file:line, relative to the reference repository root).
 *
 * Conventions
 *   - Plain pointers and sizes only.  Host pointers unless the name ends in `_device`.
 *   - Inputs are borrowed for the duration of the call and never modified.
 *   - Outputs are written into caller-provided buffers.  Matrices are column-major
 *     (Fortran order) with an explicit leading dimension.
 *   - Handles are immutable after construction and reference counted; every handle
 *     returned through an out-pointer must be released with its `*_release`.
 *   - All functions return a wx_status; on failure `wx_last_error()` returns a
 *     thread-local message naming the offending input.
 *   - Calls on one handle from several threads are safe: each solve owns its
 *     workspace and HIP stream.
 */
#ifndef WITHIN_AMD_H
#define WITHIN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes mirror the reference error taxonomy:
 *   crates/within/src/error.rs:12-146 (BuildError, WithinError) and
 *   crates/schwarz-precond/src/error.rs:22-90 (SolveError, LocalSolveError). */
typedef enum wx_status {
    WX_OK = 0,
    WX_ERR_INVALID_INPUT = 1,  /* SolveError::InvalidInput -> Python ValueError            */
    WX_ERR_BUILD = 2,          /* BuildError::*            -> Python ValueError            */
    WX_ERR_RUNTIME = 3,        /* other SolveError         -> Python RuntimeError          */
    WX_ERR_DEVICE = 4,         /* HIP runtime failure      -> Python RuntimeError          */
    WX_ERR_NOT_POSITIVE_DEF = 5, /* SolveError::InvalidInput "preconditioner not positive
                                    definite" (lsmr/bidiag.rs:99-111) -> ValueError        */
    WX_ERR_INDEX = 6           /* out-of-range layout lookup -> Python IndexError          */
} wx_status;

/* LSMR stop reasons, crates/schwarz-precond/src/lsmr.rs:54-72 (declaration order kept). */
typedef enum wx_stop_reason {
    WX_STOP_ZERO_RHS = 0,
    WX_STOP_INITIAL_NORMAL_EQ_ZERO = 1,
    WX_STOP_RESIDUAL_TOLERANCE = 2,
    WX_STOP_NORMAL_EQ_TOLERANCE = 3,
    WX_STOP_WARM_START_EXACT = 4,
    WX_STOP_MAX_ITERATIONS = 5
} wx_stop_reason;

/* Arithmetic of the preconditioner apply.  The Krylov vectors, recurrences and the
 * final residual check are FP64 in both modes. */
typedef enum wx_precision {
    WX_PRECISION_FP64 = 0,   /* FP64 preconditioner coefficients and arithmetic            */
    WX_PRECISION_MIXED = 1   /* FP32 preconditioner coefficients (factors, C, 1/D), FP64
                                vector arithmetic                                           */
} wx_precision;

/* Preconditioner kind, crates/within/src/config.rs:133-158 (PreconditionerConfig). */
typedef enum wx_precond_kind {
    WX_PRECOND_DEFAULT = 0,   /* Additive with LocalSolverConfig::default()                */
    WX_PRECOND_OFF = 1,
    WX_PRECOND_DIAGONAL = 2,
    WX_PRECOND_ADDITIVE = 3
} wx_precond_kind;

/* Schur mode, crates/within/src/config.rs:47-59. */
typedef enum wx_schur_mode { WX_SCHUR_APPROXIMATE = 0, WX_SCHUR_EXACT = 1 } wx_schur_mode;

/* Which kept blocks get the dense exact factor.  The reference rule (block_elim/solver.rs:311-347)
 * always applies: dense iff n_kept <= dense_threshold.  WX_DENSE_DEVICE (the default) also
 * factors a kept block densely where the device cost model finds the dense apply cheaper (up
 * to 4,096 kept levels, or up to 16,384 when the reduced graph is at least 0.5 % dense); the
 * choice depends on the config and the design only, never on whether a GPU is visible. */
typedef enum wx_dense_policy { WX_DENSE_REFERENCE = 0, WX_DENSE_DEVICE = 1 } wx_dense_policy;

/* LocalSolverConfig / ApproxCholConfig / ApproxSchurConfig / ScalingConfig,
 * crates/within/src/config.rs:20-131 (defaults: dense_threshold 24, dense_policy
 * WX_DENSE_DEVICE).  dense_threshold < 0 disables the reference rule. */
typedef struct wx_precond_config {
    int32_t kind;              /* wx_precond_kind */
    int32_t schur;             /* wx_schur_mode */
    uint64_t approx_chol_seed;
    int32_t split_merge;       /* <= 0: None */
    uint64_t schur_seed;
    int32_t schur_split;       /* >= 1 */
    int64_t dense_threshold;
    double scaling_tolerance;
    int64_t scaling_max_sweeps;
    int32_t scaling_error_on_failure; /* 0 Warn, 1 Error */
    int32_t reduction;         /* ReductionStrategy; accepted, the device combine is always
                                  the deterministic gather */
    int32_t dense_policy;      /* wx_dense_policy; no reference counterpart */
} wx_precond_config;

/* LsmrOptions, crates/within/src/config.rs:160-178, plus the device precision. */
typedef struct wx_lsmr_options {
    double tol;
    int64_t maxiter;
    int64_t local_size;        /* <= 0: None */
    int32_t precision;         /* wx_precision */
    int32_t check_every;       /* host convergence poll interval in iterations; <= 0: auto */
    int32_t timing;            /* != 0: time the solve's kernels with HIP events (wx_last_timing) */
    int32_t true_residual;     /* != 0 (default): compute wx_rhs_stats.true_residual, one more D' pass
                                  over the demeaned output; 0: skip it (reported as NaN), leaving the
                                  reference's work: x, demeaned and the LSMR estimate `residual` */
} wx_lsmr_options;

/* Per-RHS result, the per-column fields of BatchSolveResult (crates/within/src/solver.rs:248-293)
 * plus the FP64 true normal-equation residual ||D'W(z - Dx)|| / ||D'Wz||. */
typedef struct wx_rhs_stats {
    int32_t converged;
    int32_t stop_reason;       /* wx_stop_reason */
    int64_t iterations;
    double residual;           /* LSMR estimate |zeta_k| / |zeta_0| (M-metric when preconditioned) */
    double residual_norm;      /* |phi_bar| */
    double true_residual;      /* ||D'W(z - D x)|| / ||D'W z|| computed in FP64 on the device; with
                                  slope covariates D holds the per-level reparametrised slope
                                  columns the solver works in (the residual z - D x is the same) */
    double time_solve;         /* seconds */
} wx_rhs_stats;

typedef struct wx_design wx_design;
typedef struct wx_solver wx_solver;
typedef struct wx_precond wx_precond;

/* Per-phase device times of this thread's last solve run with `timing` set, measured with
 * HIP events on the solve stream: out[2p] = total milliseconds, out[2p+1] = launches, for
 * p = 0 D v gather (fused u update), 1 D' reduction (all terms), 2 preconditioner apply,
 * 3 whole iterations.  Returns the number of phases written (<= n / 2).  No reference
 * counterpart: the measurement hook of the benchmark. */
uint64_t wx_last_timing(double *out, uint64_t n);

/* Thread-local message of the last failed call on this thread ("" if none). */
const char *wx_last_error(void);

/* Name of the reference error variant the last failed call on this thread stands for, as
 * "BuildError::InvalidWeight", "SolveError::InvalidInput", ... (crates/within/src/error.rs:12-146,
 * crates/schwarz-precond/src/error.rs:22-90), so that a binding can rebuild the exact variant
 * that a wx_status collapses; "" when the failure has no reference counterpart (a HIP runtime
 * failure, WX_ERR_DEVICE) or no call failed. */
const char *wx_last_error_variant(void);

/* Library / device information: fills `name` (NUL terminated) and returns the
 * device count seen by the HIP runtime (0 when no GPU is visible). */
int32_t wx_device_info(char *name, uint64_t name_len);

/* Default configs (config.rs:74-131, 160-178). */
void wx_precond_config_default(wx_precond_config *cfg);
void wx_lsmr_options_default(wx_lsmr_options *opts);

/* ---- design ---------------------------------------------------------------- */

/* Replaces `IntoDesign for ArrayView2<u32>` (crates/within/src/solver.rs:34-48) +
 * `Design::from_frame` (crates/within/src/domain.rs:174-178): intercept-only factors,
 * element (i, f) at codes[i*row_stride + f*col_stride]. */
wx_status wx_design_from_codes(const uint32_t *codes, uint64_t n_obs, uint64_t n_factors,
                               int64_t row_stride, int64_t col_stride, wx_design **out);

/* Replaces `Design::new(effects)` (crates/within/src/domain.rs:156-171) with
 * `Effect::new(levels, intercept, slopes)` (crates/within/src/domain/effect.rs:13-42).
 * levels[e] has n_obs codes; effect e has n_slopes[e] slope columns taken in order from
 * the flat list `slopes` (each n_obs long). */
wx_status wx_design_from_effects(uint64_t n_obs, uint64_t n_effects, const uint32_t *const *levels,
                                 const uint8_t *intercept, const uint64_t *n_slopes,
                                 const double *const *slopes, wx_design **out);

void wx_design_release(wx_design *d);

/* ---- solver ------------------------------------------------------------------ */

/* Replaces `Solver::new(design, weights, preconditioner)` (crates/within/src/solver.rs:361-411).
 * Exactly one of `cfg` / `prebuilt` may be non-NULL; both NULL = PreconditionerInput::Default. */
wx_status wx_solver_new(const wx_design *design, const double *weights, const wx_precond_config *cfg,
                        const wx_precond *prebuilt, wx_solver **out);

void wx_solver_release(wx_solver *s);

/* The Frisch-Waugh-Lovell step after demeaning (the reference's README.md:55-66, where numpy's
 * lstsq does it on the host): G = A'A (m x m, column-major) for the column-major n x m matrix
 * A = [y~, X~] of demeaned columns, so that X~'X~ and X~'y~ are G's blocks.  One pass over A,
 * deterministic (partial sums combined in a fixed order); 1 <= m <= 64, lda >= n.
 * wx_gram_device: A and G in device memory, enqueued on `stream` (a hipStream_t, NULL: the
 * default stream) and synchronised before returning.  wx_gram: A and G in host memory. */
wx_status wx_gram_device(const double *A, uint64_t n, uint64_t m, uint64_t lda, double *G, void *stream);
wx_status wx_gram(const double *A, uint64_t n, uint64_t m, uint64_t lda, double *G);

/* wx_gram_device without any device allocation, for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments, kernels and launch grid as wx_gram_device, so G is bitwise the
 * same; 1 <= m <= 32.  The partial sums live in a static device array of 8.6 MB per GPU, held
 * until the process ends and shared by every call on that GPU: calls that overlap on one GPU
 * (from several threads or streams) race and must be serialised by the caller. */
wx_status wx_gram_device_noalloc(const double *A, uint64_t n, uint64_t m, uint64_t lda, double *G, void *stream);

/* A Solver for the same design with new observation weights (NULL: unweighted), as
 * wx_solver_new(design, weights, cfg, prebuilt) would build it, reusing `base`'s design and
 * device index structures (copied on the device) so that only the weights and the
 * preconditioner are formed anew; with `prebuilt` (e.g. the base's preconditioner, still an SPD
 * preconditioner of the reweighted system) nothing is built on the host.  No counterpart in the
 * reference, which builds every Solver from its design; for callers that re-solve one design
 * under changing weights (the differentiable projection's training loops).  Designs with slope
 * columns are rejected: their whitened columns depend on the weights. */
wx_status wx_solver_reweight(const wx_solver *base, const double *weights, const wx_precond_config *cfg,
                             const wx_precond *prebuilt, wx_solver **out);

/* Replaces `Solver::n_obs / n_dofs` (crates/within/src/solver.rs:586-596). */
uint64_t wx_solver_n_obs(const wx_solver *s);
uint64_t wx_solver_n_dofs(const wx_solver *s);

/* Seconds spent in wx_solver_new (design + reparametrisation + preconditioner build). */
double wx_solver_time_setup(const wx_solver *s);

/* Replaces `Solver::solve_batch` (crates/within/src/solver.rs:530-577) and `Solver::solve`
 * (solver.rs:503-527, k = 1).  Y is n_obs x k column-major with leading dimension ldy;
 * X (n_dofs x k, ld n_dofs) and demeaned (n_obs x k, ld n_obs) are column-major and may be
 * NULL when not wanted.  `opts` NULL = LsmrOptions::default(). */
wx_status wx_solver_solve_batch(const wx_solver *s, const double *Y, uint64_t ldy, uint64_t k,
                                const wx_lsmr_options *opts, double *X, double *demeaned,
                                wx_rhs_stats *stats);

/* Same as wx_solver_solve_batch with every array in device memory (HBM) and the work
 * enqueued on `stream` (a hipStream_t, NULL = a library-owned stream).  Returns after the
 * solve has completed.  No reference counterpart: the reference has no device boundary;
 * this is the entry point for callers whose data already lives on the GPU. */
wx_status wx_solver_solve_batch_device(const wx_solver *s, const double *Y_device, uint64_t ldy, uint64_t k,
                                       const wx_lsmr_options *opts, double *X_device, double *demeaned_device,
                                       wx_rhs_stats *stats, void *stream);

/* Many independent problems in one call: problem p solves Y[p] (n_obs(p) x k[p] column-major,
 * leading dimension ldy[p]) with solvers[p] and writes X[p], demeaned[p] (either table or
 * entry may be NULL) and stats[p][0..k[p]).  Replaces a loop of `Solver::solve_batch` calls
 * (crates/within/src/solver.rs:530-577) over independent solvers.  Problems of one shape
 * (same terms, weighting and preconditioner kind, no slopes, width class) are packed into one
 * block-diagonal run whose every iteration kernel covers all of them, with per-problem LSMR
 * state and convergence; the runs iterate concurrently on several streams as captured
 * iteration graphs.  The launch shape (packing, streams, iterations per graph, graphs) is
 * chosen by an autotuner on the first call for a problem-count class.  _device: every array in device memory, work ordered after `stream`;
 * the host form copies. */
wx_status wx_solve_many_device(const wx_solver *const *solvers, uint64_t n_problems, const double *const *Y,
                               const uint64_t *ldy, const uint64_t *k, const wx_lsmr_options *opts, double *const *X,
                               double *const *demeaned, wx_rhs_stats *const *stats, void *stream);
wx_status wx_solve_many(const wx_solver *const *solvers, uint64_t n_problems, const double *const *Y,
                        const uint64_t *ldy, const uint64_t *k, const wx_lsmr_options *opts, double *const *X,
                        double *const *demeaned, wx_rhs_stats *const *stats);

/* The shape the autotuner chose for the last many-problem call: streams, iterations per
 * captured graph, whether graphs were used (1) or iterations launched one by one (0), and
 * whether problems of one shape were packed (1) or run one by one (0). */
void wx_last_many_shape(int32_t *streams, int32_t *iterations_per_graph, int32_t *graph, int32_t *pack);

/* Fixes the shape of later many-problem calls instead of the autotuner's choice (streams 1-64,
 * iterations per graph 1-64, graph 0/1, pack 0/1); streams = 0 returns the choice to the
 * autotuner. */
void wx_set_many_shape(int32_t streams, int32_t iterations_per_graph, int32_t graph, int32_t pack);

/* Work-item tile of the packed many-problem runs: the most rows of one D' or fused-gather work
 * item (a larger level is split into pieces of that many rows whose partial sums are combined in
 * order).  wx_set_many_tile fixes it for later calls (16-4096; 0 returns it to the autotuner,
 * which tries 32, 64, 128 and 256 for the packed shape it picked); wx_last_many_tile reports
 * the last call's.  No reference counterpart: the reference runs problems one after another
 * (crates/within/src/solver.rs:530-577). */
void wx_set_many_tile(int32_t rows);
int32_t wx_last_many_tile(void);

/* Adjoint of the weighted projection demeaned = z - D x(z, w) for reverse-mode
 * differentiation.  No counterpart in the reference (its bindings are not differentiable);
 * it serves torch.autograd callers of within_amd.autograd.  Given the upstream gradient
 * G = dL/d(demeaned) and the forward's output R (both n_obs x k column-major, device, caller
 * row order), one adjoint solve D'WD lambda = D'G reuses the solver's preconditioner and a
 * fused gather forms grad_z = G - w .* (D lambda) (n_obs x k) and, when grad_w is not NULL,
 * grad_w = -sum_j R_j .* (D lambda)_j (n_obs).  Needs strictly positive weights. */
wx_status wx_solver_projection_grad_device(const wx_solver *s, const double *G, uint64_t ldg, const double *R,
                                           uint64_t ldr, uint64_t k, const wx_lsmr_options *opts, double *grad_z,
                                           double *grad_w, wx_rhs_stats *stats, void *stream);

/* Replaces `Operator::apply` / `Operator::apply_adjoint` of the weighted design operator
 * A = W^{1/2} D (crates/schwarz-precond/src/lib.rs:48-57 implemented by
 * crates/within/src/operator/design.rs:92-117), in the caller's row order:
 * adjoint = 0: Y (n_obs x k) = A X (n_dofs x k);  adjoint = 1: Y (n_dofs x k) = A' X (n_obs x k).
 * Column-major host arrays with leading dimensions n_obs / n_dofs. */
wx_status wx_solver_apply_operator(const wx_solver *s, int32_t adjoint, const double *X, uint64_t k, double *Y);

/* Replaces `Solver::preconditioner()` (crates/within/src/solver.rs:580-583): *out is a new
 * reference (NULL when the solver runs unpreconditioned). */
wx_status wx_solver_preconditioner(const wx_solver *s, wx_precond **out);

/* Replaces `CoefficientLayout` (crates/within/src/solver.rs:125-200): per-term
 * (offset, n_levels, n_columns).  Arrays must hold wx_solver_n_terms() entries. */
uint64_t wx_solver_n_terms(const wx_solver *s);
void wx_solver_layout(const wx_solver *s, uint64_t *offset, uint64_t *n_levels, uint64_t *n_columns);

/* Replaces `SolveResult::unidentified` (crates/within/src/solver.rs:221-222,
 * solver/reparam.rs): triples (term, level, column), `count` in/out. */
wx_status wx_solver_unidentified(const wx_solver *s, uint64_t *triples, uint64_t *count);

/* Replaces `Solver::warnings` (crates/within/src/solver.rs:415-419, BuildWarning
 * error.rs:107-134): number of warnings and the text of warning i. */
uint64_t wx_solver_n_warnings(const wx_solver *s);
const char *wx_solver_warning(const wx_solver *s, uint64_t i);

/* ---- preconditioner --------------------------------------------------------- */

/* Replaces `build_preconditioner(design, weights, config)` (crates/within/src/operator/schwarz.rs:232-271):
 * builds a preconditioner on the host only (no GPU needed), for offline construction and
 * persistence; *out is NULL when the config resolves to no preconditioner (Off, or an
 * Additive build without factor pairs). */
wx_status wx_precond_build(const wx_design *design, const double *weights, const wx_precond_config *cfg,
                           wx_precond **out);

void wx_precond_release(wx_precond *p);

/* Replaces `Preconditioner::nrows/ncols/variant_name` (crates/within/src/operator/schwarz.rs:129-147). */
uint64_t wx_precond_n(const wx_precond *p);
const char *wx_precond_variant(const wx_precond *p);

/* Replaces `Preconditioner::apply` (crates/within/src/operator/schwarz.rs:150-152):
 * y = M^{-1} x for n x k column-major host arrays (FP64 coefficients). */
wx_status wx_precond_apply(const wx_precond *p, const double *x, uint64_t k, double *y);

/* Replaces the postcard persistence of `Preconditioner` (crates/within-py/src/config.rs:367-388):
 * a versioned little-endian binary format of this library.  The buffer is allocated by
 * the library and freed with wx_buffer_free. */
wx_status wx_precond_serialize(const wx_precond *p, uint8_t **buf, uint64_t *len);
wx_status wx_precond_deserialize(const uint8_t *buf, uint64_t len, wx_precond **out);
void wx_buffer_free(uint8_t *buf);

#ifdef __cplusplus
}
#endif
#endif /* WITHIN_AMD_H */
