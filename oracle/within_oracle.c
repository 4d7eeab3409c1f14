/*
 * within_oracle.c -- CPU restatement of the reference's fixed-effects projection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product in
 * `within_amd/`: only tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  Nothing in the product links or calls it.
 *
 * Every function restates one piece of the reference (a Rust crate) in plain C,
 * citing the reference file:line it follows.  Summation orders follow the
 * reference's sequential (single-thread) branches, so small cases reproduce the
 * reference's own known-answer vectors exactly.
 *
 * Pinning (what checks each function against something other than itself):
 *   wo_stable_argsort     reference known answers (domain.rs:356-392, 433-466) + numpy argsort(kind="stable")
 *   wo_design_layout      reference known answers (domain.rs:482-518 offsets 0/6/9, n_dofs 11) + last-argmax rule
 *   wo_gather / wo_scatter reference exact vectors (operator/design/tests.rs:6-57), scipy.sparse D and the
 *                         adjoint identity <Dx,u> = <x,D'u> (weighted and unweighted)
 *   wo_build_diagonal     closed form 1/diag(D'WD) via numpy bincount
 *   wo_lsmr               reference LSMR fixtures (lsmr/tests/solve.rs, breakdown.rs: x=[1,2,3], stop reasons,
 *                         iteration counts), scipy.sparse.linalg.lsmr and the exact least-squares solution
 *   wo_schwarz_apply      exact local solves compared with a dense numpy pseudo-inverse of each subdomain
 *                         Gramian; the additive sum compared with sum_i R_i^T W_i A_i^+ W_i R_i built densely
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Design layout                                                              */
/* ------------------------------------------------------------------------- */

/* Stable ascending argsort of `key` (values < n_levels).
 * Follows crates/within/src/domain.rs:110-140: counting sort when n >= L, else a
 * comparison sort; both give the same (stable) permutation. */
static const uint32_t *g_sort_key;
static int cmp_by_key(const void *a, const void *b) {
    uint32_t ia = *(const uint32_t *)a, ib = *(const uint32_t *)b;
    uint32_t ka = g_sort_key[ia], kb = g_sort_key[ib];
    if (ka != kb) return ka < kb ? -1 : 1;
    return ia < ib ? -1 : (ia > ib); /* index tiebreak == stability */
}

void wo_stable_argsort(const uint32_t *key, uint64_t n, uint64_t n_levels, uint32_t *perm) {
    if (n < n_levels) {
        for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
        g_sort_key = key;
        qsort(perm, n, sizeof(uint32_t), cmp_by_key);
        return;
    }
    uint64_t *cursor = (uint64_t *)calloc(n_levels + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) cursor[key[i] + 1]++;
    for (uint64_t l = 1; l <= n_levels; ++l) cursor[l] += cursor[l - 1];
    for (uint64_t i = 0; i < n; ++i) perm[cursor[key[i]]++] = (uint32_t)i;
    free(cursor);
}

/* Term layout of a design.  Follows crates/within/src/domain.rs:187-247:
 * n_levels = max + 1, term-major offsets (coefficient c of level l at offset + c*L + l),
 * a per-term `sorted` flag, and the dominant term = LAST argmax of n_dofs
 * (Rust max_by_key, domain.rs:226).  `codes` is F-order (n x T).
 * Returns n_dofs, or -1 on empty observations. */
int64_t wo_design_layout(const uint32_t *codes, uint64_t n, uint64_t T, const int32_t *ncols,
                         uint64_t *n_levels, uint64_t *offset, uint8_t *sorted, int64_t *dominant) {
    if (n == 0) return -1;
    uint64_t off = 0;
    int64_t dom = -1;
    uint64_t best = 0;
    for (uint64_t t = 0; t < T; ++t) {
        const uint32_t *col = codes + t * n;
        uint32_t mx = 0, prev = 0;
        uint8_t s = 1;
        for (uint64_t i = 0; i < n; ++i) {
            uint32_t v = col[i];
            if (v > mx) mx = v;
            s &= (uint8_t)(v >= prev);
            prev = v;
        }
        n_levels[t] = (uint64_t)mx + 1;
        offset[t] = off;
        sorted[t] = s;
        uint64_t nd = (uint64_t)ncols[t] * n_levels[t];
        off += nd;
        if (dom < 0 || nd >= best) { /* >= keeps the last maximum */
            dom = (int64_t)t;
            best = nd;
        }
    }
    *dominant = dom;
    return (int64_t)off;
}

/* ------------------------------------------------------------------------- */
/* Design operator D (gather) and D' (scatter)                                */
/* ------------------------------------------------------------------------- */

/* Column loadings: for term t, column c, colload[colbase[t] + c] is -1 for the
 * intercept (loading 1) or the index of a slope column in `slopes` (F-order n x S). */
typedef struct {
    uint64_t n, T;
    const uint32_t *codes;  /* n x T, F-order */
    const uint64_t *n_levels, *offset;
    const int32_t *ncols;
    const int64_t *colload; /* sum(ncols) entries */
    const double *slopes;   /* n x S, F-order (may be NULL) */
} wo_design;

static inline double loading(const wo_design *d, int64_t li, uint64_t i) {
    return li < 0 ? 1.0 : d->slopes[(uint64_t)li * d->n + i];
}

/* dst[i] = (sum_t fold_{-0.0} over c of x[off_t + c*L_t + code_t(i)] * load_c(i)) * scale[i].
 * Follows crates/within/src/operator/design/gather.rs:10-78 and the per-term fold at
 * gather.rs:95-109 (fold from -0.0; the scale multiply comes after all terms). */
void wo_gather(const wo_design *d, const double *x, const double *scale, double *dst) {
    for (uint64_t i = 0; i < d->n; ++i) dst[i] = 0.0;
    uint64_t cb = 0;
    for (uint64_t t = 0; t < d->T; ++t) {
        const uint32_t *lev = d->codes + t * d->n;
        const uint64_t L = d->n_levels[t], off = d->offset[t];
        const int32_t nc = d->ncols[t];
        for (uint64_t i = 0; i < d->n; ++i) {
            double acc = -0.0;
            for (int32_t c = 0; c < nc; ++c)
                acc = acc + x[off + (uint64_t)c * L + lev[i]] * loading(d, d->colload[cb + c], i);
            dst[i] += acc;
        }
        cb += (uint64_t)nc;
    }
    if (scale)
        for (uint64_t i = 0; i < d->n; ++i) dst[i] *= scale[i];
}

/* dst = D' (sw .* u): dst zeroed, then block[c*L + code(i)] += load_c(i) * sw[i]*u[i] in row order.
 * Follows crates/within/src/operator/design.rs:104-116 (zero then scatter of sqrt(w_i) x_i)
 * and the sequential strategy crates/within/src/operator/design/scatter.rs:124-138. */
void wo_scatter(const wo_design *d, const double *u, const double *sw, double *dst) {
    uint64_t ndofs = 0;
    for (uint64_t t = 0; t < d->T; ++t) ndofs += (uint64_t)d->ncols[t] * d->n_levels[t];
    for (uint64_t j = 0; j < ndofs; ++j) dst[j] = 0.0;
    uint64_t cb = 0;
    for (uint64_t t = 0; t < d->T; ++t) {
        const uint32_t *lev = d->codes + t * d->n;
        const uint64_t L = d->n_levels[t], off = d->offset[t];
        const int32_t nc = d->ncols[t];
        for (uint64_t i = 0; i < d->n; ++i) {
            double b = sw ? sw[i] * u[i] : u[i];
            for (int32_t c = 0; c < nc; ++c) {
                int64_t li = d->colload[cb + c];
                double v = li < 0 ? b : d->slopes[(uint64_t)li * d->n + i] * b;
                dst[off + (uint64_t)c * L + lev[i]] += v;
            }
        }
        cb += (uint64_t)nc;
    }
}

/* Diagonal (Jacobi) preconditioner: inv_diag = pinv(diag(D'WD)).
 * Follows crates/within/src/operator/schwarz.rs:185-229 (w*z*z left to right; zero
 * diagonal keeps 0; a non-finite reciprocal is an error).  Returns -1 - index on error. */
int64_t wo_build_diagonal(const wo_design *d, const double *w, double *inv_diag) {
    uint64_t ndofs = 0;
    for (uint64_t t = 0; t < d->T; ++t) ndofs += (uint64_t)d->ncols[t] * d->n_levels[t];
    for (uint64_t j = 0; j < ndofs; ++j) inv_diag[j] = 0.0;
    uint64_t cb = 0;
    for (uint64_t t = 0; t < d->T; ++t) {
        const uint32_t *lev = d->codes + t * d->n;
        const uint64_t L = d->n_levels[t], off = d->offset[t];
        for (int32_t c = 0; c < d->ncols[t]; ++c) {
            int64_t li = d->colload[cb + c];
            double *slice = inv_diag + off + (uint64_t)c * L;
            for (uint64_t i = 0; i < d->n; ++i) {
                double wi = w ? w[i] : 1.0;
                if (li < 0) slice[lev[i]] += wi;
                else {
                    double z = d->slopes[(uint64_t)li * d->n + i];
                    slice[lev[i]] += wi * z * z;
                }
            }
        }
        cb += (uint64_t)d->ncols[t];
    }
    for (uint64_t j = 0; j < ndofs; ++j) {
        if (inv_diag[j] == 0.0) continue;
        double inv = 1.0 / inv_diag[j];
        if (!isfinite(inv)) return -1 - (int64_t)j;
        inv_diag[j] = inv;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* LSMR / modified LSMR                                                       */
/* ------------------------------------------------------------------------- */

/* Max-scaled, NaN-propagating 2-norm.  Follows crates/schwarz-precond/src/lsmr.rs:24-36. */
double wo_vec_norm(const double *v, uint64_t n) {
    double scale = 0.0;
    int any_nan = 0;
    for (uint64_t i = 0; i < n; ++i) {
        double a = fabs(v[i]);
        if (isnan(v[i])) any_nan = 1;
        else if (a > scale) scale = a; /* f64::max ignores NaN */
    }
    if (scale == 0.0) return any_nan ? NAN : 0.0;
    double s = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        double q = v[i] / scale;
        s += q * q;
    }
    return scale * sqrt(s);
}

static double dotp(const double *a, const double *b, uint64_t n) {
    double s = 0.0;
    for (uint64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Abstract operator + preconditioner used by the LSMR restatement. */
typedef struct wo_op {
    uint64_t m, n;
    void (*apply)(const struct wo_op *, const double *x, double *y);         /* y = A x  (m) */
    void (*apply_adjoint)(const struct wo_op *, const double *x, double *y); /* y = A'x  (n) */
    const wo_design *design;
    const double *sw;
} wo_op;

typedef struct wo_prec {
    uint64_t n;
    /* returns 0 on success */
    int (*apply)(const struct wo_prec *, const double *x, double *y);
    const double *inv_diag; /* kind diag */
    const double *dense;    /* kind dense (n x n row-major) */
    const void *ctx;        /* kind schwarz */
} wo_prec;

static void design_apply(const wo_op *op, const double *x, double *y) { wo_gather(op->design, x, op->sw, y); }
static void design_adjoint(const wo_op *op, const double *x, double *y) { wo_scatter(op->design, x, op->sw, y); }

static int prec_diag(const wo_prec *p, const double *x, double *y) {
    for (uint64_t i = 0; i < p->n; ++i) y[i] = p->inv_diag[i] * x[i];
    return 0;
}
static int prec_dense(const wo_prec *p, const double *x, double *y) {
    for (uint64_t i = 0; i < p->n; ++i) y[i] = dotp(p->dense + i * p->n, x, p->n);
    return 0;
}

enum {
    WO_STOP_ZERO_RHS = 0,
    WO_STOP_INIT_NORMAL_ZERO = 1,
    WO_STOP_RESIDUAL_TOL = 2,
    WO_STOP_NORMAL_EQ_TOL = 3,
    WO_STOP_MAX_ITER = 5,
};

typedef struct {
    int64_t iterations;
    int32_t stop_reason;
    int32_t converged;
    double residual_norm;
    double normal_eq_residual;
} wo_lsmr_result;

/* Givens rotation zeroing b against a.  recurrence.rs:20-27. */
static void givens(double a, double b, double *c, double *s, double *r) {
    double rr = hypot(a, b);
    *r = rr;
    if (rr > 0.0) { *c = a / rr; *s = b / rr; }
    else { *c = 1.0; *s = 0.0; }
}

/* Modified / standard LSMR.  `prec == NULL` runs the standard Golub-Kahan stream
 * (lsmr.rs:177-203 + bidiag.rs:235-273, 347-384); otherwise the modified stream with
 * one M^{-1} per step (lsmr.rs:206-295 + bidiag.rs:278-309, 428-512).  The recurrence
 * and stopping tests follow lsmr.rs:298-371 and recurrence.rs:58-256.
 * Returns 0, or -1 for invalid input (non-finite rhs/tol), -2 for an indefinite
 * preconditioner (bidiag.rs:99-111), -3 for a preconditioner failure. */
int wo_lsmr_run(const wo_op *A, const wo_prec *M, const double *b, double tol, int64_t maxiter,
                int64_t local_size, double *x, wo_lsmr_result *res) {
    const uint64_t m = A->m, n = A->n;
    if (!isfinite(tol) || tol < 0.0) return -1;
    for (uint64_t i = 0; i < m; ++i)
        if (!isfinite(b[i])) return -1;
    for (uint64_t j = 0; j < n; ++j) x[j] = 0.0;

    const double b_norm = wo_vec_norm(b, m);
    if (b_norm == 0.0) {
        res->iterations = 0; res->stop_reason = WO_STOP_ZERO_RHS; res->converged = 1;
        res->residual_norm = 0.0; res->normal_eq_residual = 0.0;
        return 0;
    }
    double *u = calloc(m, sizeof(double)), *av = calloc(m, sizeof(double));
    double *v = calloc(n, sizeof(double)), *atu = calloc(n, sizeof(double));
    double *pt = M ? calloc(n, sizeof(double)) : NULL;
    double *h = calloc(n, sizeof(double)), *hb = calloc(n, sizeof(double));
    /* windowed MGS ring (bidiag.rs:114-217): lane 0 = v, lane 1 = p_tilde/alpha */
    uint64_t cap = 0;
    if (local_size > 0) {
        cap = (uint64_t)local_size;
        if (cap > m) cap = m;
        if (cap > n) cap = n;
    }
    double *ring_v = cap ? calloc(cap * n, sizeof(double)) : NULL;
    double *ring_p = (cap && M) ? calloc(cap * n, sizeof(double)) : NULL;
    uint64_t ring_next = 0, ring_count = 0;
    int rc = 0;

    /* --- init (bidiag.rs:347-384 / 428-472) --- */
    double beta = b_norm; /* vec_norm(b) */
    {
        double inv = 1.0 / beta;
        for (uint64_t i = 0; i < m; ++i) u[i] = b[i] * inv;
    }
    double alpha;
    if (!M) {
        A->apply_adjoint(A, u, v);
        alpha = sqrt(dotp(v, v, n));
        if (alpha > 0.0) { double s = 1.0 / alpha; for (uint64_t j = 0; j < n; ++j) v[j] *= s; }
        if (cap) { uint64_t slot = ring_next; ring_next = (ring_next + 1) % cap; if (ring_count < cap) ring_count++;
                   memcpy(ring_v + slot * n, v, n * sizeof(double)); }
    } else {
        A->apply_adjoint(A, u, pt);
        if (M->apply(M, pt, v)) { rc = -3; goto done; }
        double vp = dotp(v, pt, n);
        if (vp < 0.0) {
            double bound = sqrt(DBL_EPSILON) * sqrt(dotp(v, v, n) * dotp(pt, pt, n));
            if (vp < -bound) { rc = -2; goto done; }
        }
        alpha = sqrt(vp > 0.0 ? vp : 0.0);
        if (alpha > 0.0) { double s = 1.0 / alpha; for (uint64_t j = 0; j < n; ++j) v[j] *= s; }
        if (cap) { uint64_t slot = ring_next; ring_next = (ring_next + 1) % cap; if (ring_count < cap) ring_count++;
                   double ia = alpha > 0.0 ? 1.0 / alpha : 0.0;
                   memcpy(ring_v + slot * n, v, n * sizeof(double));
                   for (uint64_t j = 0; j < n; ++j) ring_p[slot * n + j] = pt[j] * ia; }
    }
    double beta_prev_inv = 1.0; /* u was normalized by init */

    if (alpha == 0.0) {
        res->iterations = 0; res->stop_reason = WO_STOP_INIT_NORMAL_ZERO; res->converged = 1;
        res->residual_norm = b_norm; res->normal_eq_residual = 0.0;
        goto done;
    }

    /* recurrence state (recurrence.rs:68-79) */
    double alpha_bar = alpha, phi_bar = beta, c_bar = 1.0, s_bar = 0.0;
    double zeta_bar = alpha * beta;
    double zeta0 = fabs(zeta_bar) > DBL_MIN ? fabs(zeta_bar) : DBL_MIN;
    double a_norm_sq = alpha * alpha;
    const double abs_tol = tol * b_norm, rel_tol = tol;
    double prev_rho = 1.0, prev_rho_bar = 1.0; /* RotationStep::initial */
    memcpy(h, v, n * sizeof(double));

    for (int64_t itn = 1; itn <= maxiter; ++itn) {
        double alpha_new, beta_new;
        /* --- bidiagonal step --- */
        A->apply(A, v, av);
        double scale = M ? -(alpha * beta_prev_inv) : -alpha;
        double bsq = 0.0;
        for (uint64_t i = 0; i < m; ++i) { double t = av[i] + scale * u[i]; u[i] = t; bsq += t * t; }
        beta_new = sqrt(bsq);
        if (beta_new == 0.0) {
            for (uint64_t j = 0; j < n; ++j) v[j] = 0.0;
            if (M) for (uint64_t j = 0; j < n; ++j) pt[j] = 0.0;
            alpha_new = 0.0;
            beta_prev_inv = 0.0;
        } else if (!M) {
            double s = 1.0 / beta_new;
            for (uint64_t i = 0; i < m; ++i) u[i] *= s;
            A->apply_adjoint(A, u, atu);
            double asq = 0.0;
            for (uint64_t j = 0; j < n; ++j) { double t = atu[j] + (-beta_new) * v[j]; v[j] = t; asq += t * t; }
            if (cap) {
                for (uint64_t q = 0; q < ring_count; ++q) {
                    uint64_t start = ring_count < cap ? 0 : ring_next;
                    uint64_t slot = (start + q) % cap;
                    double c = dotp(v, ring_v + slot * n, n);
                    for (uint64_t j = 0; j < n; ++j) v[j] = -c * ring_v[slot * n + j] + 1.0 * v[j];
                }
                asq = dotp(v, v, n);
            }
            alpha_new = sqrt(asq);
            if (alpha_new > 0.0) { double s2 = 1.0 / alpha_new; for (uint64_t j = 0; j < n; ++j) v[j] *= s2; }
            if (cap) { uint64_t slot = ring_next; ring_next = (ring_next + 1) % cap; if (ring_count < cap) ring_count++;
                       memcpy(ring_v + slot * n, v, n * sizeof(double)); }
        } else {
            double beta_inv = 1.0 / beta_new;
            /* update_p_tilde: bidiag.rs:475-485 */
            A->apply_adjoint(A, u, atu);
            double p_coeff = beta_new / alpha;
            for (uint64_t j = 0; j < n; ++j) pt[j] = beta_inv * atu[j] + (-p_coeff) * pt[j];
            /* reorthonormalize_v: bidiag.rs:488-512 */
            if (M->apply(M, pt, v)) { rc = -3; goto done; }
            if (cap) {
                uint64_t start = ring_count < cap ? 0 : ring_next;
                for (uint64_t q = 0; q < ring_count; ++q) {
                    uint64_t slot = (start + q) % cap;
                    double c = dotp(v, ring_p + slot * n, n);
                    for (uint64_t j = 0; j < n; ++j) v[j] = -c * ring_v[slot * n + j] + 1.0 * v[j];
                    for (uint64_t j = 0; j < n; ++j) pt[j] = -c * ring_p[slot * n + j] + 1.0 * pt[j];
                }
            }
            double vp = dotp(v, pt, n);
            if (vp < 0.0) {
                double bound = sqrt(DBL_EPSILON) * sqrt(dotp(v, v, n) * dotp(pt, pt, n));
                if (vp < -bound) { rc = -2; goto done; }
            }
            alpha_new = sqrt(vp > 0.0 ? vp : 0.0);
            if (alpha_new > 0.0) { double s2 = 1.0 / alpha_new; for (uint64_t j = 0; j < n; ++j) v[j] *= s2; }
            if (cap) { uint64_t slot = ring_next; ring_next = (ring_next + 1) % cap; if (ring_count < cap) ring_count++;
                       double ia = alpha_new > 0.0 ? 1.0 / alpha_new : 0.0;
                       memcpy(ring_v + slot * n, v, n * sizeof(double));
                       for (uint64_t j = 0; j < n; ++j) ring_p[slot * n + j] = pt[j] * ia; }
            beta_prev_inv = beta_inv;
        }
        alpha = alpha_new;

        /* convergence.observe (recurrence.rs:239-241) */
        a_norm_sq += alpha_new * alpha_new + beta_new * beta_new;

        /* recurrence.step (recurrence.rs:81-107) */
        double ph_c, ph_s, ph_r;
        givens(alpha_bar, beta_new, &ph_c, &ph_s, &ph_r);
        double theta_new = ph_s * alpha_new;
        double alpha_bar_new = -ph_c * alpha_new;
        double phi_bar_new = ph_s * phi_bar;
        double theta_bar = s_bar * ph_r;
        double pb_c, pb_s, pb_r;
        givens(c_bar * ph_r, theta_new, &pb_c, &pb_s, &pb_r);
        double zeta = pb_c * zeta_bar;
        double zeta_bar_new = -pb_s * zeta_bar;
        alpha_bar = alpha_bar_new;
        phi_bar = phi_bar_new;
        c_bar = pb_c;
        s_bar = pb_s;
        zeta_bar = zeta_bar_new;

        /* solution.update (recurrence.rs:143-191) */
        double t_x_denom = ph_r * pb_r;
        double t_x = fabs(t_x_denom) > DBL_EPSILON ? zeta / t_x_denom : 0.0;
        double t_hbar_denom = prev_rho * prev_rho_bar;
        double t_hbar = fabs(t_hbar_denom) > DBL_EPSILON ? theta_bar * ph_r / t_hbar_denom : 0.0;
        double t_h = fabs(ph_r) > DBL_EPSILON ? theta_new / ph_r : 0.0;
        for (uint64_t j = 0; j < n; ++j) {
            double h_old = h[j];
            double hbj = h_old - t_hbar * hb[j];
            hb[j] = hbj;
            x[j] += t_x * hbj;
            h[j] = v[j] - t_h * h_old;
        }

        /* convergence.check (recurrence.rs:244-255) */
        double residual = fabs(phi_bar);
        int stop = -1;
        if (residual <= abs_tol) stop = WO_STOP_RESIDUAL_TOL;
        else {
            double a_norm = sqrt(a_norm_sq);
            if (a_norm < DBL_MIN) a_norm = DBL_MIN;
            double normar = fabs(zeta_bar);
            double rr = residual > DBL_MIN ? residual : DBL_MIN;
            if (normar / (a_norm * rr) <= rel_tol) stop = WO_STOP_NORMAL_EQ_TOL;
        }
        if (stop >= 0) {
            res->iterations = itn; res->stop_reason = stop; res->converged = 1;
            res->residual_norm = fabs(phi_bar); res->normal_eq_residual = fabs(zeta_bar) / zeta0;
            goto done;
        }
        prev_rho = ph_r;
        prev_rho_bar = pb_r;
        if (itn == maxiter) {
            res->iterations = maxiter; res->stop_reason = WO_STOP_MAX_ITER; res->converged = 0;
            res->residual_norm = fabs(phi_bar); res->normal_eq_residual = fabs(zeta_bar) / zeta0;
        }
    }
    if (maxiter <= 0) {
        res->iterations = 0; res->stop_reason = WO_STOP_MAX_ITER; res->converged = 0;
        res->residual_norm = fabs(phi_bar); res->normal_eq_residual = fabs(zeta_bar) / zeta0;
    }
done:
    free(u); free(av); free(v); free(atu); free(pt); free(h); free(hb); free(ring_v); free(ring_p);
    return rc;
}

/* Dense-operator LSMR entry (for the reference's fake-operator fixtures,
 * lsmr/fixtures.rs): A is m x n row-major; prec_kind 0 = none (standard lsmr),
 * 1 = diag (inv_diag), 2 = dense M^{-1} (n x n row-major), 3 = identity through
 * the modified stream (mlsmr with M = I). */
typedef struct { const double *a; } dense_ctx;
static const double *g_dense_a;
static void dense_apply(const wo_op *op, const double *x, double *y) {
    for (uint64_t i = 0; i < op->m; ++i) y[i] = dotp(g_dense_a + i * op->n, x, op->n);
}
static void dense_adjoint(const wo_op *op, const double *x, double *y) {
    for (uint64_t j = 0; j < op->n; ++j) y[j] = 0.0;
    for (uint64_t i = 0; i < op->m; ++i)
        for (uint64_t j = 0; j < op->n; ++j) y[j] += g_dense_a[i * op->n + j] * x[i];
}
static int prec_identity(const wo_prec *p, const double *x, double *y) {
    memcpy(y, x, p->n * sizeof(double));
    return 0;
}

int wo_lsmr_dense(const double *a, uint64_t m, uint64_t n, const double *b, int32_t prec_kind,
                  const double *prec_data, double tol, int64_t maxiter, int64_t local_size, double *x,
                  wo_lsmr_result *res) {
    g_dense_a = a;
    wo_op op = {m, n, dense_apply, dense_adjoint, NULL, NULL};
    wo_prec p = {n, NULL, NULL, NULL, NULL};
    const wo_prec *pp = NULL;
    if (prec_kind == 1) { p.apply = prec_diag; p.inv_diag = prec_data; pp = &p; }
    else if (prec_kind == 2) { p.apply = prec_dense; p.dense = prec_data; pp = &p; }
    else if (prec_kind == 3) { p.apply = prec_identity; pp = &p; }
    return wo_lsmr_run(&op, pp, b, tol, maxiter, local_size, x, res);
}

/* ------------------------------------------------------------------------- */
/* Additive Schwarz apply (restated over an exported preconditioner)          */
/* ------------------------------------------------------------------------- */

/* Flattened additive-Schwarz preconditioner over plain (Laplacian, floating,
 * canonical-fold) factor-pair components, the layout the product's host builder
 * exports.  Per component c:
 *   slots [slot_start[c], slot_start[c] + ne[c] + nk[c]) -- eliminated block first
 *   glob[slot], pw[slot]         restriction indices and partition weights
 *   inv_de[elim slot]            1 / D_elim
 *   C   (ne x nk) CSR: c_ptr[elim slot .. +1] into c_idx (local kept index) / c_val
 *   reduced factor of the kept-block Laplacian S (nk x nk), either
 *     dense  : fac_kind 0, lower-triangular L (nk x nk row-major at fac_off), unit
 *              diagonal, with pivots d[] (the last pivot of a floating Laplacian is 0
 *              and is pseudo-inverted to 0), S = L diag(d) L'
 *     sparse : fac_kind 1, L strictly-lower in CSR by rows (row i lists j < i), pivots d[]
 *     inverse: fac_kind 2, an explicit symmetric nk x nk operator (the pseudo-inverse S^+
 *              of a floating Laplacian, S^-1 of a definite block) at fac_off, applied as a
 *              matrix-vector product.
 *     cover  : fac_kind 3, a sparse factor (as kind 1) of the Gremban cover of a signed kept
 *              block: 2 nk rows in elimination order, row t holding cover vertex
 *              cover_perm[cover_off + t] (vertex v < nk = kept level v, else level v - nk on
 *              the second sheet); solved on the antisymmetric embed [b, -b] and read back as
 *              (y+ - y-)/2 (block_elim/factor.rs:93-118).
 *   fold[slot]   coordinate map into SDDM form (sddm.rs:53-82): +1 / -1 on the two blocks of
 *                a plain pair, sign * dominance scale of a general one; applied on restriction
 *                and again on prolongation.
 *   gauge[comp]  0 floating: the component mean is removed before and after the solve and the
 *                reduced rhs mean before the reduced solve; 1 none (grounded or cover).
 * The local solve follows crates/within/src/block_elim/solver.rs:350-466 (fold, mean,
 * eliminate, reduced solve, back-substitution, mean, unfold) and the subdomain wrapper
 * crates/schwarz-precond/src/local_solve.rs:106-147 with domain.rs:135-188 (weighted
 * restrict / prolongate).  Mean subtraction is the compensated mean of block_elim.rs:16-29. */
typedef struct {
    uint64_t n_dofs, n_comp;
    const uint64_t *slot_start; /* n_comp + 1 */
    const uint64_t *ne, *nk;
    const uint32_t *glob;
    const double *pw;
    const double *inv_de;       /* indexed by slot (kept slots unused) */
    const uint64_t *c_ptr;      /* per slot: row pointer for elim slots (n_slots + 1) */
    const uint32_t *c_idx;
    const double *c_val;
    const int32_t *fac_kind;    /* per component */
    const uint64_t *fac_off;    /* per component: into fac_val (dense) / fac_ptr base (sparse) */
    const double *fac_val;      /* dense L entries or sparse L values */
    const uint64_t *fac_ptr;    /* sparse row pointers (global) */
    const uint32_t *fac_idx;    /* sparse column indices (local) */
    const uint64_t *piv_off;    /* per component: into piv */
    const double *piv;          /* pivots d */
    const double *fold;         /* per slot */
    const int32_t *gauge;       /* per component */
    const uint64_t *cover_off;  /* per component: into cover_perm */
    const uint32_t *cover_perm;
} wo_schwarz;

static double compensated_sum(const double *v, uint64_t n) {
    double sum = 0.0, comp = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        double next = sum + v[i];
        if (fabs(sum) >= fabs(v[i])) comp += (sum - next) + v[i];
        else comp += (v[i] - next) + sum;
        sum = next;
    }
    return sum + comp;
}
static void subtract_mean(double *v, uint64_t n) {
    if (!n) return;
    double mean = compensated_sum(v, n) / (double)n;
    for (uint64_t i = 0; i < n; ++i) v[i] -= mean;
}

/* Sparse unit-lower LDL' solve in place over n rows (a zero pivot, the floating gauge, gives 0). */
static void sparse_ldl_solve(const wo_schwarz *s, uint64_t c, double *x, uint64_t n) {
    const double *d = s->piv + s->piv_off[c];
    const uint64_t *ptr = s->fac_ptr + s->fac_off[c];
    for (uint64_t i = 0; i < n; ++i) {
        double acc = x[i];
        for (uint64_t p = ptr[i]; p < ptr[i + 1]; ++p) acc -= s->fac_val[p] * x[s->fac_idx[p]];
        x[i] = acc;
    }
    for (uint64_t i = 0; i < n; ++i) x[i] = d[i] != 0.0 ? x[i] / d[i] : 0.0;
    for (uint64_t ii = n; ii-- > 0;) {
        double xi = x[ii];
        for (uint64_t p = ptr[ii]; p < ptr[ii + 1]; ++p) x[s->fac_idx[p]] -= s->fac_val[p] * xi;
    }
}

/* LDL' solve of the reduced kept-block system, x overwritten in place. */
static void reduced_solve(const wo_schwarz *s, uint64_t c, double *x, uint64_t nk) {
    const double *d = s->piv + s->piv_off[c];
    if (s->fac_kind[c] == 3) { /* Gremban cover: embed [x, -x] in elimination order, solve, read back */
        const uint32_t *perm = s->cover_perm + s->cover_off[c];
        double *y = malloc((2 * nk ? 2 * nk : 1) * sizeof(double));
        for (uint64_t t = 0; t < 2 * nk; ++t) y[t] = perm[t] < nk ? x[perm[t]] : -x[perm[t] - nk];
        sparse_ldl_solve(s, c, y, 2 * nk);
        for (uint64_t t = 0; t < 2 * nk; ++t)
            if (perm[t] < nk) x[perm[t]] = 0.5 * y[t];
        for (uint64_t t = 0; t < 2 * nk; ++t)
            if (perm[t] >= nk) x[perm[t] - nk] -= 0.5 * y[t];
        free(y);
        return;
    }
    if (s->fac_kind[c] == 2) { /* explicit symmetric operator S^+ (nk x nk row-major) */
        const double *P = s->fac_val + s->fac_off[c];
        double *t = malloc((nk ? nk : 1) * sizeof(double));
        for (uint64_t i = 0; i < nk; ++i) t[i] = dotp(P + i * nk, x, nk);
        memcpy(x, t, nk * sizeof(double));
        free(t);
        return;
    }
    if (s->fac_kind[c] == 0) {
        const double *L = s->fac_val + s->fac_off[c];
        for (uint64_t i = 0; i < nk; ++i) { /* forward: L y = b (unit diagonal) */
            double acc = x[i];
            for (uint64_t j = 0; j < i; ++j) acc -= L[i * nk + j] * x[j];
            x[i] = acc;
        }
        for (uint64_t i = 0; i < nk; ++i) x[i] = d[i] != 0.0 ? x[i] / d[i] : 0.0;
        for (uint64_t ii = nk; ii-- > 0;) { /* backward: L' x = y */
            double acc = x[ii];
            for (uint64_t j = ii + 1; j < nk; ++j) acc -= L[j * nk + ii] * x[j];
            x[ii] = acc;
        }
    } else {
        sparse_ldl_solve(s, c, x, nk);
    }
}

/* z = sum_c R_c' W_c A_c^+ W_c R_c r.  Sequential over components (executor order). */
int wo_schwarz_apply(const wo_schwarz *s, const double *r, double *z) {
    for (uint64_t j = 0; j < s->n_dofs; ++j) z[j] = 0.0;
    uint64_t maxn = 0;
    for (uint64_t c = 0; c < s->n_comp; ++c) {
        uint64_t nl = s->ne[c] + s->nk[c];
        if (nl > maxn) maxn = nl;
    }
    double *rl = malloc((maxn + 1) * sizeof(double)), *sl = malloc((maxn + 1) * sizeof(double));
    for (uint64_t c = 0; c < s->n_comp; ++c) {
        const uint64_t base = s->slot_start[c], ne = s->ne[c], nk = s->nk[c], nl = ne + nk;
        const int floating = s->gauge[c] == 0;
        /* restrict_weighted (domain.rs:135-152) */
        for (uint64_t i = 0; i < nl; ++i) rl[i] = r[s->glob[base + i]] * s->pw[base + i];
        /* fold into SDDM coordinates (sddm.rs:62-82) */
        for (uint64_t i = 0; i < nl; ++i) rl[i] *= s->fold[base + i];
        if (floating) subtract_mean(rl, nl);
        /* eliminate_and_recover (block_elim/solver.rs:350-423) */
        for (uint64_t i = 0; i < ne; ++i) rl[i] *= s->inv_de[base + i];
        double *red = sl + ne; /* reduced rhs built in place in the kept half of sol */
        for (uint64_t k = 0; k < nk; ++k) red[k] = rl[ne + k];
        for (uint64_t i = 0; i < ne; ++i) { /* Ct spmv: red[k] += C[i,k] * b_e[i] */
            for (uint64_t p = s->c_ptr[base + i]; p < s->c_ptr[base + i + 1]; ++p)
                red[s->c_idx[p]] += s->c_val[p] * rl[i];
        }
        if (floating) subtract_mean(red, nk);
        reduced_solve(s, c, red, nk);
        for (uint64_t i = 0; i < ne; ++i) { /* x_e = b_e + D_e^{-1} (C x_k) */
            double acc = 0.0;
            for (uint64_t p = s->c_ptr[base + i]; p < s->c_ptr[base + i + 1]; ++p)
                acc += s->c_val[p] * red[s->c_idx[p]];
            sl[i] = rl[i] + s->inv_de[base + i] * acc;
        }
        if (floating) subtract_mean(sl, nl);
        for (uint64_t i = 0; i < nl; ++i) sl[i] *= s->fold[base + i]; /* unfold */
        for (uint64_t i = 0; i < nl; ++i) z[s->glob[base + i]] += sl[i] * s->pw[base + i];
    }
    free(rl);
    free(sl);
    return 0;
}

static int prec_schwarz(const wo_prec *p, const double *x, double *y) {
    return wo_schwarz_apply((const wo_schwarz *)p->ctx, x, y);
}

/* ------------------------------------------------------------------------- */
/* Full solve: Solver::solve_rhs restated (solver.rs:423-491)                 */
/* ------------------------------------------------------------------------- */

/* One RHS through the design operator.  `y` and `w` are in the design's internal
 * (already permuted) row order; prec_kind 0 = Off, 1 = Diagonal (inv_diag),
 * 4 = Schwarz (schwarz).  Writes x (n_dofs) and demeaned = y - D x (unweighted). */
int wo_solve_rhs(const wo_design *d, const double *sw, const double *y, int32_t prec_kind,
                 const double *inv_diag, const wo_schwarz *schwarz, double tol, int64_t maxiter,
                 int64_t local_size, double *x, double *demeaned, wo_lsmr_result *res) {
    uint64_t ndofs = 0;
    for (uint64_t t = 0; t < d->T; ++t) ndofs += (uint64_t)d->ncols[t] * d->n_levels[t];
    for (uint64_t i = 0; i < d->n; ++i)
        if (!isfinite(y[i])) return -1;
    wo_op op = {d->n, ndofs, design_apply, design_adjoint, d, sw};
    double *b = malloc(d->n * sizeof(double));
    for (uint64_t i = 0; i < d->n; ++i) b[i] = sw ? sw[i] * y[i] : y[i];
    wo_prec p = {ndofs, NULL, NULL, NULL, NULL};
    const wo_prec *pp = NULL;
    if (prec_kind == 1) { p.apply = prec_diag; p.inv_diag = inv_diag; pp = &p; }
    else if (prec_kind == 4) { p.apply = prec_schwarz; p.ctx = schwarz; pp = &p; }
    int rc = wo_lsmr_run(&op, pp, b, tol, maxiter, local_size, x, res);
    free(b);
    if (rc) return rc;
    wo_gather(d, x, NULL, demeaned);
    for (uint64_t i = 0; i < d->n; ++i) demeaned[i] = y[i] - demeaned[i];
    return 0;
}

/* RHS-parallel batch (solver.rs:530-577: par_iter over RHS).  Y and outputs are
 * column-major (n x k, N x k).  `threads` <= 0 uses the OpenMP default. */
int wo_solve_batch(const wo_design *d, const double *sw, const double *Y, uint64_t k, int32_t prec_kind,
                   const double *inv_diag, const wo_schwarz *schwarz, double tol, int64_t maxiter,
                   int64_t local_size, int32_t threads, double *X, double *Dm, wo_lsmr_result *res) {
    uint64_t ndofs = 0;
    for (uint64_t t = 0; t < d->T; ++t) ndofs += (uint64_t)d->ncols[t] * d->n_levels[t];
    int err = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(| : err)
#endif
    for (int64_t j = 0; j < (int64_t)k; ++j) {
        int rc = wo_solve_rhs(d, sw, Y + (uint64_t)j * d->n, prec_kind, inv_diag, schwarz, tol, maxiter,
                              local_size, X + (uint64_t)j * ndofs, Dm + (uint64_t)j * d->n, res + j);
        if (rc) err |= 1;
    }
    (void)threads;
    return err ? -1 : 0;
}

int wo_omp_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
