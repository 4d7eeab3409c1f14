/*
 * within_amd_testing.h -- test hooks of libwithin_amd.so.
 *
 * Not part of the drop-in surface: these entry points expose the host preconditioner
 * builder's component pipeline on one bipartite block given directly, so that the reference's
 * unit tests of its local solver, Schur sampler and approximate Cholesky, which build such
 * blocks by hand, can be restated against this library's builder.  Host only; no GPU needed.
 */
#ifndef WITHIN_AMD_TESTING_H
#define WITHIN_AMD_TESTING_H

#include "within_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A preconditioner of dimension n_rows + n_cols whose single subdomain is the symmetric block
 * A = [[diag(d_r), B], [B', diag(d_c)]] (B: n_rows x n_cols, row-major, any signs; diag holds
 * d_r then d_c), oriented, converted to SDDM form and factored as the builder does for a slope
 * pair's component, with unit partition weights: its apply is the block-elimination local solve
 * of A (crates/within/src/domain/factor_pairs/sddm.rs:86-463, block_elim/solver.rs:311-466).
 * Stands in for the reference's `LocalComponent::general_for_test` + `BlockElimSolver::build`
 * (domain/factor_pairs/sddm/tests.rs:16-27). */
wx_status wx_testing_block_precond(uint32_t n_rows, uint32_t n_cols, const double *B, const double *diag,
                                   const wx_precond_config *cfg, wx_precond **out);

/* The reduced Laplacian that block's kept side is factored from: the exact or sampled Schur
 * complement (config's schur mode, seed and split) plus the explicit ground vertex of a grounded
 * block, as a dense row-major n x n matrix (n = kept levels, + 1 when grounded; of the Gremban
 * cover's kept levels for a frustrated block).  `out` holds `cap` doubles; *n_out receives n.
 * Stands in for the reference's `schur::sampled` / `schur::exact_for_factor`
 * (block_elim/schur.rs:29-79, 307-403). */
wx_status wx_testing_reduced_laplacian(uint32_t n_rows, uint32_t n_cols, const double *B, const double *diag,
                                       const wx_precond_config *cfg, double *out, uint64_t cap, uint32_t *n_out);

/* The index structures of `design` as the solver builds them (crates/within/src/domain.rs:80-247):
 * per term t < n_terms: n_levels[t], offset[t] (term-major coefficient layout) and sorted[t] (the
 * column is non-decreasing in internal order); *dominant = the term the locality sort uses (last
 * argmax of n_dofs); *permuted = 1 when the rows were sorted, and then obs_perm (n_obs entries,
 * may be NULL) = the internal row order (internal row i is caller row obs_perm[i]).
 * Stands in for the reference's crate-internal Design fields (`obs_perm`, `terms[t].sorted`,
 * `TermMeta`), which its own unit tests read (domain.rs:355-518). */
wx_status wx_testing_design_index(const wx_design *design, uint64_t n_terms, uint64_t *n_levels, uint64_t *offset,
                                  uint8_t *sorted, int64_t *dominant, int32_t *permuted, uint32_t *obs_perm);

/* The slope reparametrisation of a design (crates/within/src/solver/reparam.rs:78-276), built by
 * the host (on_device = 0) or on the device (on_device = 1, needs a GPU), as a Solver would build
 * it.  slopes_out (may be NULL): the whitened loading columns, n_obs x n_slope_columns, F-order,
 * caller row order.  transform_out / mean_out (may be NULL): every slope term's per-level V x V
 * whitening rows and V centres, terms in order.  unidentified_out (may be NULL): the
 * (term, level, column) triples the reparametrisation could not identify, ascending; their count
 * goes to *n_unidentified.  Stands in for the reference's crate-internal SlopeReparam fields, which
 * its own tests read through solve results (tests/slopes.rs:19-195). */
wx_status wx_testing_reparam(const wx_design *design, const double *weights, int32_t on_device, double *slopes_out,
                             double *transform_out, uint64_t transform_cap, double *mean_out, uint64_t mean_cap,
                             uint64_t *unidentified_out, uint64_t unidentified_cap, uint64_t *n_unidentified);

/* wx_precond_apply with the arithmetic of a solve of the given wx_precision: WX_PRECISION_MIXED
 * applies M^{-1} as a mixed-precision solve does (FP32 coefficients and FP32 slot-space vectors,
 * FP64 accumulation), so tests can hold that operator to symmetry and definiteness bounds.
 * Needs a GPU. */
wx_status wx_testing_precond_apply(const wx_precond *p, const double *x, uint64_t k, double *y, int32_t precision);

/* Process-wide switch for later solves with an additive Schwarz preconditioner: on (the default)
 * the LSMR steps on either side of the apply -- p~ <- A'u / beta - (theta / rho) p~ before it and
 * the partial sums of <v,p~>, <v,v>, <p~,p~> after it (lsmr/bidiag.rs:99-111, 475-485) -- run
 * inside the apply's first and last launch; off they run as launches of their own.  Both forms
 * compute the same values in the same order, which tests hold them to.  Returns the previous
 * setting (1 on, 0 off).  Stands in for nothing in the reference, whose steps are separate loops. */
int32_t wx_testing_set_fused_steps(int32_t on);

#ifdef __cplusplus
}
#endif
#endif /* WITHIN_AMD_TESTING_H */
