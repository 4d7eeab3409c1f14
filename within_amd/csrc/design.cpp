// design.cpp -- design construction, locality sort and slope reparametrisation (host).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>

#include <omp.h>

#include "core.hpp"

namespace wx {

std::vector<uint32_t> stable_argsort(const uint32_t *key, uint64_t n, uint64_t n_levels) {
    // crates/within/src/domain.rs:110-140: counting sort when n >= L, else a stable
    // comparison sort on the key; both produce the same permutation.
    std::vector<uint32_t> perm(n);
    if (n < n_levels) {
        std::iota(perm.begin(), perm.end(), 0u);
        std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
        return perm;
    }
    // The counting sort over contiguous row blocks, one per thread: block b's rows of level l
    // go after level l's rows of blocks < b, so the order is the serial one.
    const int nb = n >= (1u << 20) ? std::max(1, std::min(omp_get_max_threads(), 32)) : 1;
    std::vector<std::vector<uint64_t>> cnt(nb, std::vector<uint64_t>(n_levels, 0));
    auto lo = [&](int b) { return n * (uint64_t)b / (uint64_t)nb; };
#pragma omp parallel for schedule(static, 1) num_threads(nb)
    for (int b = 0; b < nb; ++b) {
        uint64_t *c = cnt[b].data();
        for (uint64_t i = lo(b); i < lo(b + 1); ++i) c[key[i]]++;
    }
    uint64_t run = 0;
    for (uint64_t l = 0; l < n_levels; ++l)
        for (int b = 0; b < nb; ++b) {
            const uint64_t c = cnt[b][l];
            cnt[b][l] = run;
            run += c;
        }
#pragma omp parallel for schedule(static, 1) num_threads(nb)
    for (int b = 0; b < nb; ++b) {
        uint64_t *c = cnt[b].data();
        for (uint64_t i = lo(b); i < lo(b + 1); ++i) perm[c[key[i]]++] = (uint32_t)i;
    }
    return perm;
}


Design build_design(uint64_t n_obs, const std::vector<std::vector<uint32_t>> &columns,
                    const std::vector<std::vector<int64_t>> &structure,
                    const std::vector<std::vector<double>> &slopes, bool locality_sort) {
    // crates/within/src/domain.rs:187-247.
    if (n_obs == 0) fail(WX_ERR_BUILD, "no observations provided", "BuildError::EmptyObservations");
    const uint64_t T = columns.size();
    for (uint64_t t = 0; t < T; ++t)
        if (columns[t].size() != n_obs)
            fail(WX_ERR_BUILD, "column " + std::to_string(t) + " has " + std::to_string(columns[t].size()) +
                                   " observations, expected " + std::to_string(n_obs),
                 "BuildError::ObservationCountMismatch");
    for (uint64_t s = 0; s < slopes.size(); ++s)
        if (slopes[s].size() != n_obs)
            fail(WX_ERR_BUILD, "column " + std::to_string(T + s) + " has " + std::to_string(slopes[s].size()) +
                                   " observations, expected " + std::to_string(n_obs),
                 "BuildError::SlopeLengthMismatch");
    Design d;
    d.n_obs = n_obs;
    d.n_slopes = slopes.size();
    d.terms.resize(T);
    uint64_t offset = 0;
    for (uint64_t t = 0; t < T; ++t) {
        const uint32_t *col = columns[t].data();
        uint32_t mx = 0;
        int unsorted = 0;
#pragma omp parallel for schedule(static) reduction(max : mx) reduction(| : unsorted) if (n_obs >= (1u << 20))
        for (int64_t i = 0; i < (int64_t)n_obs; ++i) {
            mx = std::max(mx, col[i]);
            unsorted |= i > 0 && col[i] < col[i - 1];
        }
        const bool sorted = !unsorted;
        Term &tm = d.terms[t];
        tm.n_levels = (uint64_t)mx + 1;
        tm.offset = offset;
        tm.sorted = sorted;
        tm.loads = structure[t];
        offset += tm.n_dofs();
    }
    if (offset > 0xFFFFFFFFull)
        fail(WX_ERR_BUILD, "DOF space of " + std::to_string(offset) + " coefficients exceeds u32 indexing",
             "BuildError::DofSpaceExceedsU32");
    d.n_dofs = offset;

    // Dominant term = last argmax of n_dofs (Rust max_by_key, domain.rs:226).
    int64_t dom = -1;
    uint64_t best = 0;
    for (uint64_t t = 0; t < T; ++t)
        if (dom < 0 || d.terms[t].n_dofs() >= best) {
            dom = (int64_t)t;
            best = d.terms[t].n_dofs();
        }
    bool sort = locality_sort && dom >= 0 && !d.terms[dom].sorted && n_obs <= 0xFFFFFFFFull;
    d.codes.resize(T * n_obs);
    d.slopes.resize(d.n_slopes * n_obs);
    if (sort) {
        // the reference's locality sort (domain.rs:226-234): a stable argsort by the dominant
        // term, so obs_perm, the permuted columns and the sorted flags are the reference's
        d.obs_perm = stable_argsort(columns[dom].data(), n_obs, d.terms[dom].n_levels);
        const uint32_t *perm = d.obs_perm.data();
        const bool par = n_obs >= (1u << 20);
        for (uint64_t t = 0; t < T; ++t) {
            uint32_t *dst = d.codes.data() + t * n_obs;
            const uint32_t *src = columns[t].data();
            int unsorted = 0;
#pragma omp parallel for schedule(static) if (par)
            for (int64_t i = 0; i < (int64_t)n_obs; ++i) dst[i] = src[perm[i]];
#pragma omp parallel for schedule(static) reduction(| : unsorted) if (par)
            for (int64_t i = 1; i < (int64_t)n_obs; ++i) unsorted |= dst[i] < dst[i - 1];
            d.terms[t].sorted = !unsorted; // nested factors come out sorted (domain.rs:231-234)
        }
        for (uint64_t s = 0; s < d.n_slopes; ++s) {
            double *dst = d.slopes.data() + s * n_obs;
            const double *src = slopes[s].data();
#pragma omp parallel for schedule(static) if (par)
            for (int64_t i = 0; i < (int64_t)n_obs; ++i) dst[i] = src[perm[i]];
        }
    } else {
        for (uint64_t t = 0; t < T; ++t) std::memcpy(d.codes.data() + t * n_obs, columns[t].data(), n_obs * 4);
        for (uint64_t s = 0; s < d.n_slopes; ++s)
            std::memcpy(d.slopes.data() + s * n_obs, slopes[s].data(), n_obs * 8);
    }
    return d;
}

void permute_in(const Design &d, const double *src, double *dst) {
    if (d.obs_perm.empty()) {
        std::memcpy(dst, src, d.n_obs * sizeof(double));
        return;
    }
#pragma omp parallel for schedule(static) if (d.n_obs >= (1u << 20))
    for (int64_t k = 0; k < (int64_t)d.n_obs; ++k) dst[k] = src[d.obs_perm[k]];
}

void validate_weights(const Design &d, const double *w, uint64_t len) {
    // crates/within/src/domain.rs:261-279.
    if (len != d.n_obs)
        fail(WX_ERR_BUILD, "weights has length " + std::to_string(len) + ", expected " + std::to_string(d.n_obs),
             "BuildError::WeightCountMismatch");
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad) if (len >= (1u << 20))
    for (int64_t i = 0; i < (int64_t)len; ++i) bad |= !(w[i] >= 0.0 && std::isfinite(w[i]));
    if (!bad) return;
    for (uint64_t i = 0; i < len; ++i) // the first offending weight
        if (!(w[i] >= 0.0 && std::isfinite(w[i])))
            fail(WX_ERR_BUILD, "weight at index " + std::to_string(i) + " must be finite and non-negative, got " +
                                   std::to_string(w[i]),
                 "BuildError::InvalidWeight");
}

std::vector<std::vector<uint8_t>> find_active_levels(const Design &d) {
    std::vector<std::vector<uint8_t>> active(d.terms.size());
    for (uint64_t t = 0; t < d.terms.size(); ++t) {
        active[t].assign(d.terms[t].n_levels, 0);
        const uint32_t *col = d.level_column(t);
        uint8_t *a = active[t].data();
#pragma omp parallel for schedule(static) if (d.n_obs >= (1u << 20))
        for (int64_t i = 0; i < (int64_t)d.n_obs; ++i) // read first: stores only the first time
            if (!__atomic_load_n(&a[col[i]], __ATOMIC_RELAXED)) __atomic_store_n(&a[col[i]], (uint8_t)1, __ATOMIC_RELAXED);
    }
    return active;
}

// ---------------------------------------------------------------------------
// Slope reparametrisation (crates/within/src/solver/reparam.rs)
// ---------------------------------------------------------------------------

namespace {
constexpr double RANK_TOL = 1e-10; // reparam.rs:13

// Pivoted Gram-Schmidt in coordinates of the level Gram g (v x v row-major).
// Returns rank x v orthonormal rows w (w g w' = I) and the kept mask (reparam.rs:284-316).
// Pivot = largest residual variance, ties to the last index (Rust max_by).
void pivoted_gram_schmidt(const double *g, uint64_t v, double tol, std::vector<double> &basis,
                          std::vector<uint8_t> &kept) {
    std::vector<double> residual(v);
    for (uint64_t j = 0; j < v; ++j) residual[j] = g[j * v + j];
    kept.assign(v, 0);
    basis.clear();
    std::vector<double> q(v);
    for (;;) {
        int64_t p = -1;
        for (uint64_t j = 0; j < v; ++j) {
            if (kept[j] || !std::isfinite(residual[j]) || !(residual[j] > tol * g[j * v + j])) continue;
            if (p < 0 || residual[j] >= residual[p]) p = (int64_t)j;
        }
        if (p < 0) break;
        kept[p] = 1;
        std::fill(q.begin(), q.end(), 0.0);
        q[p] = 1.0;
        for (uint64_t t = 0; t * v < basis.size(); ++t) {
            const double *qt = basis.data() + t * v;
            double c = 0.0;
            for (uint64_t j = 0; j < v; ++j) c += g[p * v + j] * qt[j];
            for (uint64_t j = 0; j < v; ++j) q[j] -= c * qt[j];
        }
        double norm = std::sqrt(residual[p]);
        for (auto &qj : q) qj /= norm;
        for (uint64_t j = 0; j < v; ++j) {
            double c = 0.0;
            for (uint64_t k = 0; k < v; ++k) c += g[j * v + k] * q[k];
            residual[j] -= c * c;
        }
        basis.insert(basis.end(), q.begin(), q.end());
    }
}
} // namespace

std::unique_ptr<SlopeReparam> build_reparam(Design &d, const double *weights) {
    auto rp = std::make_unique<SlopeReparam>();
    for (uint64_t term = 0; term < d.terms.size(); ++term) {
        const Term &tm = d.terms[term];
        std::vector<uint64_t> slope_cols, zcols;
        int64_t icol = -1;
        for (uint64_t c = 0; c < tm.loads.size(); ++c) {
            if (tm.loads[c] >= 0) {
                slope_cols.push_back(c);
                zcols.push_back((uint64_t)tm.loads[c]);
            } else icol = (int64_t)c;
        }
        if (zcols.empty()) continue;
        const uint64_t v = zcols.size(), L = tm.n_levels, n = d.n_obs;
        const bool intercept = icol >= 0;
        const uint32_t *lev = d.level_column(term);
        // weighted Welford moments per level (reparam.rs:221-252)
        const uint64_t tl = v * (v + 1) / 2;
        std::vector<double> wsum(L, 0.0), mean(L * v, 0.0), com(L * tl, 0.0), delta(v), z(v);
        for (uint64_t i = 0; i < n; ++i) {
            double w = weights ? weights[i] : 1.0;
            if (w <= 0.0) continue;
            uint64_t l = lev[i];
            for (uint64_t j = 0; j < v; ++j) z[j] = d.slope_column(zcols[j])[i];
            wsum[l] += w;
            double ratio = w / wsum[l];
            double *m = mean.data() + l * v;
            for (uint64_t j = 0; j < v; ++j) {
                delta[j] = z[j] - m[j];
                m[j] += ratio * delta[j];
            }
            double *cm = com.data() + l * tl;
            for (uint64_t j = 0; j < v; ++j) {
                double dev = w * (z[j] - m[j]);
                for (uint64_t k = 0; k <= j; ++k) cm[j * (j + 1) / 2 + k] += delta[k] * dev;
            }
        }
        SlopeReparam::TermBlock blk;
        blk.term = term;
        blk.n_slopes = v;
        blk.intercept = intercept;
        blk.transform.assign(L * v * v, 0.0); // rows = output basis index, cols = original slope j
        blk.mean.assign(L * v, 0.0);
        std::vector<double> g(v * v), basis;
        std::vector<uint8_t> kept;
        std::vector<uint64_t> rank(L, 0);
        for (uint64_t l = 0; l < L; ++l) {
            const double *cm = com.data() + l * tl;
            const double *m = mean.data() + l * v;
            for (uint64_t j = 0; j < v; ++j)
                for (uint64_t k = 0; k <= j; ++k) {
                    double gv = cm[j * (j + 1) / 2 + k];
                    if (!intercept) gv += wsum[l] * m[j] * m[k];
                    g[j * v + k] = gv;
                    g[k * v + j] = gv;
                }
            pivoted_gram_schmidt(g.data(), v, RANK_TOL, basis, kept);
            if (intercept && wsum[l] == 0.0) {
                rp->unidentified.insert(rp->unidentified.end(), {term, l, (uint64_t)icol});
            }
            for (uint64_t j = 0; j < v; ++j)
                if (!kept[j]) rp->unidentified.insert(rp->unidentified.end(), {term, l, slope_cols[j]});
            rank[l] = basis.size() / v;
            std::copy(basis.begin(), basis.end(), blk.transform.begin() + l * v * v);
            if (intercept) std::copy(m, m + v, blk.mean.begin() + l * v);
        }
        // rewrite the loading columns: u = W (z - center) (reparam.rs:151-163)
        std::vector<std::vector<double>> ucols(v, std::vector<double>(n, 0.0));
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t l = lev[i];
            const double *W = blk.transform.data() + l * v * v;
            const double *c = blk.mean.data() + l * v;
            for (uint64_t j = 0; j < v; ++j) z[j] = d.slope_column(zcols[j])[i] - c[j];
            for (uint64_t r = 0; r < rank[l]; ++r) {
                double acc = 0.0;
                for (uint64_t j = 0; j < v; ++j) acc += W[r * v + j] * z[j];
                ucols[r][i] = acc;
            }
        }
        for (uint64_t j = 0; j < v; ++j)
            std::memcpy(d.slopes.data() + zcols[j] * n, ucols[j].data(), n * sizeof(double));
        blk.first_col = slope_cols.front();
        // rows >= rank[l] of a level's transform stay zero; an all-zero transform marks a
        // fully dropped level, which back_transform skips (reparam.rs:177-179)
        rp->blocks.push_back(std::move(blk));
    }
    if (rp->blocks.empty()) return nullptr;
    // unidentified ascending in (term, level, column)
    std::vector<std::array<uint64_t, 3>> tr;
    for (uint64_t i = 0; i < rp->unidentified.size(); i += 3)
        tr.push_back({rp->unidentified[i], rp->unidentified[i + 1], rp->unidentified[i + 2]});
    std::sort(tr.begin(), tr.end());
    rp->unidentified.clear();
    for (auto &t : tr) rp->unidentified.insert(rp->unidentified.end(), t.begin(), t.end());
    return rp;
}

void SlopeReparam::back_transform(const Design &d, double *x) const {
    // reparam.rs:173-194: b = sum_k W[k,:] * x[slot(k)]; x[slot(j)] = b_j; intercept -= b . center.
    for (const auto &blk : blocks) {
        const Term &tm = d.terms[blk.term];
        const uint64_t v = blk.n_slopes, L = tm.n_levels;
        std::vector<uint64_t> slope_cols;
        int64_t icol = -1;
        for (uint64_t c = 0; c < tm.loads.size(); ++c) {
            if (tm.loads[c] >= 0) slope_cols.push_back(c);
            else icol = (int64_t)c;
        }
        std::vector<double> b(v);
        for (uint64_t l = 0; l < L; ++l) {
            const double *W = blk.transform.data() + l * v * v;
            bool empty = true;
            for (uint64_t q = 0; q < v * v; ++q) empty &= W[q] == 0.0;
            if (empty) continue;
            std::fill(b.begin(), b.end(), 0.0);
            for (uint64_t k = 0; k < v; ++k) {
                double bk = x[tm.offset + slope_cols[k] * L + l];
                for (uint64_t j = 0; j < v; ++j) b[j] += W[k * v + j] * bk;
            }
            for (uint64_t j = 0; j < v; ++j) x[tm.offset + slope_cols[j] * L + l] = b[j];
            if (icol >= 0) {
                double acc = 0.0;
                for (uint64_t j = 0; j < v; ++j) acc += b[j] * blk.mean[l * v + j];
                x[tm.offset + (uint64_t)icol * L + l] -= acc;
            }
        }
    }
}

} // namespace wx
