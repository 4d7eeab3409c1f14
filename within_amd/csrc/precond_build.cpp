// precond_build.cpp -- host construction of the diagonal and additive Schwarz preconditioners.
//
// The additive Schwarz build restates crates/within/src/domain/factor_pairs.rs:36-194
// (channel pairs, components, partition of unity), domain/cross_tab.rs (cross-tabulation
// and bipartite components), domain/factor_pairs/sddm.rs:86-197 (orientation, Laplacian
// validation) and block_elim/{solver,schur,factor}.rs (block elimination of the larger
// side and a factorisation of the kept block).  The kept block is factored either
//   * densely and exactly (exact Schur complement, then its pseudo-inverse S^+), the
//     reference's `ExactBelow` backend, chosen by a device cost model, or
//   * by approximate Cholesky with clique-tree sampling (Gao, Kyng & Spielman 2023),
//     this library's own implementation of the published algorithm, as a unit-lower LDL'
//     factor in elimination order, ready for level-scheduled triangular solves.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <functional>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <queue>
#include <string>
#include <tuple>
#include <chrono>

#include <omp.h>

#include "schwarz.hpp"

namespace wx {

// dense_gpu.hip

std::shared_ptr<Precond> build_diagonal(const Design &d, const std::vector<double> &w) {
    // crates/within/src/operator/schwarz.rs:185-229
    auto p = std::make_shared<Precond>();
    p->kind = Precond::DIAGONAL;
    p->n = d.n_dofs;
    p->inv_diag.assign(d.n_dofs, 0.0);
    for (uint64_t t = 0; t < d.terms.size(); ++t) {
        const Term &tm = d.terms[t];
        const uint32_t *lev = d.level_column(t);
        for (uint64_t c = 0; c < tm.ncols(); ++c) {
            double *slice = p->inv_diag.data() + tm.offset + c * tm.n_levels;
            const int64_t li = tm.loads[c];
            for (uint64_t i = 0; i < d.n_obs; ++i) {
                const double wi = w.empty() ? 1.0 : w[i];
                if (li < 0) slice[lev[i]] += wi;
                else {
                    const double z = d.slope_column((uint64_t)li)[i];
                    slice[lev[i]] += wi * z * z;
                }
            }
        }
    }
    for (uint64_t j = 0; j < d.n_dofs; ++j) {
        double &v = p->inv_diag[j];
        if (v == 0.0) continue;
        const double inv = 1.0 / v;
        if (!std::isfinite(inv))
            fail(WX_ERR_BUILD, "zero diagonal in preconditioner at index " + std::to_string(j),
                 "BuildError::SingularDiagonal");
        v = inv;
    }
    return p;
}

namespace {

// splitmix64 stream; one stream per (seed, star) so sampling is independent of scheduling.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

uint64_t mix_seed(uint64_t seed, uint64_t stream, uint64_t salt) {
    Rng r(seed ^ (stream * 0xD1B54A32D192ED03ull) ^ (salt * 0x8CB92BA72F3D8DD7ull));
    r.next();
    return r.next();
}

struct Edge {
    uint32_t lo, hi;
    double w;
};

// Clique-tree sampling of the clique sum_{i<j} w_i w_j / W L_ij of one eliminated star
// (Gao, Kyng & Spielman 2023): neighbours in ascending weight order; neighbour i joins a
// later neighbour j drawn with probability w_j / S_i (S_i = sum of later weights) at weight
// w_i S_i / W.  The expectation is the clique exactly.  `split` independent trees at weight
// 1/split reduce the variance.  `nb` must be merged (one entry per neighbour).
void sample_clique_tree(std::vector<std::pair<uint32_t, double>> &nb, double W, int split, Rng &rng,
                        std::vector<Edge> &out) {
    const size_t d = nb.size();
    if (d < 2 || W <= 0.0) return;
    std::sort(nb.begin(), nb.end(), [](const auto &a, const auto &b) {
        return a.second < b.second || (a.second == b.second && a.first < b.first);
    });
    thread_local std::vector<double> suffix; // reused across stars: no allocation per star
    if (suffix.size() < d + 1) suffix.resize(2 * d + 2);
    suffix[d] = 0.0;
    for (size_t i = d; i-- > 0;) suffix[i] = suffix[i + 1] + nb[i].second;
    const double inv_split = 1.0 / split;
    for (int rep = 0; rep < split; ++rep) {
        for (size_t i = 0; i + 1 < d; ++i) {
            const double rest = suffix[i + 1];
            if (rest <= 0.0) continue;
            // draw j in (i, d) with P(j) = w_j / rest: find first j with suffix[i+1] - suffix[j+1] > r
            const double r = rng.uniform() * rest;
            size_t lo = i + 1, hi = d - 1;
            while (lo < hi) {
                size_t mid = (lo + hi) / 2;
                if (suffix[i + 1] - suffix[mid + 1] > r) hi = mid;
                else lo = mid + 1;
            }
            const double w = nb[i].second * rest / W * inv_split;
            uint32_t a = nb[i].first, b = nb[lo].first;
            if (a == b) continue;
            out.push_back({std::min(a, b), std::max(a, b), w});
        }
    }
}

// Merges edges by (lo, hi), summing weights in a fixed (lo, hi, w) order.
void sort_and_merge(std::vector<Edge> &e) {
    auto less = [](const Edge &a, const Edge &b) {
        if (a.lo != b.lo) return a.lo < b.lo;
        if (a.hi != b.hi) return a.hi < b.hi;
        return a.w < b.w;
    };
    std::sort(e.begin(), e.end(), less);
    size_t wr = 0;
    for (size_t r = 0; r < e.size(); ++r) {
        if (wr > 0 && e[wr - 1].lo == e[r].lo && e[wr - 1].hi == e[r].hi) e[wr - 1].w += e[r].w;
        else e[wr++] = e[r];
    }
    e.resize(wr);
}

// Approximate Cholesky of a connected Laplacian given by its off-diagonal edges, as
// L D L' with L unit-lower in elimination order.  Vertices are eliminated in approximate
// minimum-degree order (the count of live edge instances, a seeded hash breaking ties), as
// approximate Cholesky is usually run: low-degree vertices first keeps the sampled fill and
// the triangular solves' level depth small on hub-heavy graphs.  Each star's clique is
// replaced by `split` sampled clique trees.  The last vertex has an empty star and pivot 0
// (the floating gauge, pseudo-inverted to 0).  `last` >= 0 holds that vertex back to the
// end: with the ground vertex of a grounded system there, dropping it leaves a factor of
// the grounded (positive definite) kept block.
struct AcFactor {
    std::vector<uint32_t> order;    // order[t] = vertex eliminated t-th
    std::vector<uint64_t> rptr;     // L by rows (relabelled), strictly lower
    std::vector<uint32_t> ridx;
    std::vector<double> rval;
    std::vector<double> d;          // pivots (relabelled)
};

AcFactor approx_cholesky(uint32_t n, const std::vector<Edge> &edges, uint64_t seed, int split, int64_t last = -1) {
    AcFactor f;
    std::vector<std::vector<std::pair<uint32_t, double>>> adj(n);
    for (const Edge &e : edges) {
        adj[e.lo].push_back({e.hi, e.w});
        adj[e.hi].push_back({e.lo, e.w});
    }
    std::vector<uint64_t> cnt(n), tie(n);
    Rng tr(mix_seed(seed, n, 0xAC0));
    for (uint32_t v = 0; v < n; ++v) {
        cnt[v] = adj[v].size();
        tie[v] = tr.next();
    }
    using Key = std::tuple<uint64_t, uint64_t, uint32_t>; // (degree, tie, vertex), smallest first
    std::priority_queue<Key, std::vector<Key>, std::greater<Key>> pq;
    for (uint32_t v = 0; v < n; ++v)
        if ((int64_t)v != last) pq.push({cnt[v], tie[v], v});
    std::vector<uint8_t> gone(n, 0);
    // columns of L in elimination order: (row vertex, value)
    std::vector<std::vector<std::pair<uint32_t, double>>> cols;
    cols.reserve(n);
    f.order.reserve(n);
    f.d.reserve(n);
    std::vector<std::pair<uint32_t, double>> nb;
    std::vector<Edge> fill;
    auto eliminate = [&](uint32_t v) {
        nb.clear();
        for (auto &e : adj[v])
            if (!gone[e.first]) {
                nb.push_back(e);
                --cnt[e.first];
                pq.push({cnt[e.first], tie[e.first], e.first});
            }
        std::vector<std::pair<uint32_t, double>>().swap(adj[v]);
        gone[v] = 1;
        f.order.push_back(v);
        std::sort(nb.begin(), nb.end());
        size_t wr = 0;
        for (size_t r = 0; r < nb.size(); ++r) {
            if (wr > 0 && nb[wr - 1].first == nb[r].first) nb[wr - 1].second += nb[r].second;
            else nb[wr++] = nb[r];
        }
        nb.resize(wr);
        double W = 0.0;
        for (auto &e : nb) W += e.second;
        f.d.push_back(W);
        cols.emplace_back();
        if (W <= 0.0) return;
        for (auto &e : nb) cols.back().push_back({e.first, -e.second / W});
        fill.clear();
        Rng rng(mix_seed(seed, v, 0x57A));
        sample_clique_tree(nb, W, split, rng, fill);
        for (const Edge &e : fill) {
            adj[e.lo].push_back({e.hi, e.w});
            adj[e.hi].push_back({e.lo, e.w});
            for (uint32_t u : {e.lo, e.hi}) {
                ++cnt[u];
                if ((int64_t)u != last) pq.push({cnt[u], tie[u], u});
            }
        }
    };
    // every degree change pushes a fresh entry, so an entry whose degree disagrees is stale
    while (!pq.empty()) {
        const auto [c, tb, v] = pq.top();
        pq.pop();
        if (gone[v] || c != cnt[v] || (int64_t)v == last) continue;
        eliminate(v);
    }
    if (last >= 0 && n > 0) eliminate((uint32_t)last);
    std::vector<uint32_t> pos(n);
    for (uint32_t t = 0; t < n; ++t) pos[f.order[t]] = t;
    // transpose columns into rows (row = later vertex, col = earlier vertex)
    f.rptr.assign(n + 1, 0);
    for (uint32_t t = 0; t < n; ++t)
        for (auto &e : cols[t]) f.rptr[pos[e.first] + 1]++;
    for (uint32_t i = 0; i < n; ++i) f.rptr[i + 1] += f.rptr[i];
    f.ridx.resize(f.rptr[n]);
    f.rval.resize(f.rptr[n]);
    std::vector<uint64_t> cur(f.rptr.begin(), f.rptr.end() - 1);
    for (uint32_t t = 0; t < n; ++t)
        for (auto &e : cols[t]) {
            const uint32_t row = pos[e.first];
            f.ridx[cur[row]] = t;
            f.rval[cur[row]] = e.second;
            cur[row]++;
        }
    return f;
}

// Dense inverse of a kept block S (n x n row-major, in place).  Floating (a connected
// Laplacian): ground the last vertex, invert the SPD minor through Cholesky, pad, and project
// with P = I - 11'/n on both sides, so the result is the Moore-Penrose inverse S^+.
// Otherwise S is positive definite (grounded or signed) and is inverted whole.
bool dense_laplacian_pinv(std::vector<double> &S, uint32_t n, bool floating = true) {
    if (floating && n <= 1) {
        std::fill(S.begin(), S.end(), 0.0);
        return true;
    }
    const uint32_t m = floating ? n - 1 : n;
    std::vector<double> A((size_t)m * m);
    for (uint32_t i = 0; i < m; ++i) std::memcpy(&A[(size_t)i * m], &S[(size_t)i * n], m * sizeof(double));
    // Cholesky A = L L' (lower, in place), right-looking with the trailing update in parallel
    for (uint32_t k = 0; k < m; ++k) {
        double akk = A[(size_t)k * m + k];
        if (!(akk > 0.0)) return false;
        akk = std::sqrt(akk);
        A[(size_t)k * m + k] = akk;
        const double inv = 1.0 / akk;
        for (uint32_t i = k + 1; i < m; ++i) A[(size_t)i * m + k] *= inv;
#pragma omp parallel for schedule(static) if (m - k > 256)
        for (int64_t i = k + 1; i < (int64_t)m; ++i) {
            const double lik = A[(size_t)i * m + k];
            double *row = &A[(size_t)i * m];
            const double *colk = &A[0];
            for (uint32_t j = k + 1; j <= (uint32_t)i; ++j) row[j] -= lik * colk[(size_t)j * m + k];
        }
    }
    // Inverse of L (lower) column by column: Linv e_c
    std::vector<double> Li((size_t)m * m, 0.0);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t c = 0; c < (int64_t)m; ++c) {
        for (uint32_t i = (uint32_t)c; i < m; ++i) {
            double acc = (i == (uint32_t)c) ? 1.0 : 0.0;
            const double *row = &A[(size_t)i * m];
            for (uint32_t j = (uint32_t)c; j < i; ++j) acc -= row[j] * Li[(size_t)j * m + c];
            Li[(size_t)i * m + c] = acc / row[i];
        }
    }
    // G = Linv' Linv
    std::vector<double> G((size_t)m * m);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)m; ++i)
        for (uint32_t j = 0; j <= (uint32_t)i; ++j) {
            double acc = 0.0;
            for (uint32_t r = (uint32_t)i; r < m; ++r) acc += Li[(size_t)r * m + i] * Li[(size_t)r * m + j];
            G[(size_t)i * m + j] = acc;
            G[(size_t)j * m + i] = acc;
        }
    if (!floating) {
        std::copy(G.begin(), G.end(), S.begin());
        return true;
    }
    // pad and project
    std::vector<double> rowmean(n, 0.0);
    double total = 0.0;
    for (uint32_t i = 0; i < m; ++i) {
        double s = 0.0;
        for (uint32_t j = 0; j < m; ++j) s += G[(size_t)i * m + j];
        rowmean[i] = s / n;
        total += s;
    }
    total /= (double)n * n;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < n; ++j) {
            const double g = (i < m && j < m) ? G[(size_t)i * m + j] : 0.0;
            S[(size_t)i * n + j] = g - rowmean[i] - rowmean[j] + total;
        }
    return true;
}

struct Channel {
    uint32_t term, column;
};

// Stable CSR transpose: output row c lists the input rows holding column c in ascending order
// (with their values when `val` is given) -- the serial counting transpose's result, formed over
// contiguous blocks of input rows, one per thread, each placing its entries after the earlier
// blocks' entries of the same column.
void par_transpose(uint64_t nrows, uint64_t ncols, const uint64_t *ptr, const uint32_t *idx, const double *val,
                   std::vector<uint64_t> &tptr, std::vector<uint32_t> &tidx, std::vector<double> *tval) {
    const uint64_t nnz = ptr[nrows];
    tptr.assign(ncols + 1, 0);
    tidx.resize(nnz);
    if (tval) tval->resize(nnz);
    int nb = nnz >= (1u << 18) ? std::max(1, std::min(omp_get_max_threads(), 32)) : 1;
    if ((uint64_t)nb * ncols > 8 * nnz + 4096) nb = 1; // counters would outweigh the entries
    std::vector<uint64_t> rb(nb + 1, nrows); // row blocks of about equal entry counts
    rb[0] = 0;
    for (int b = 1; b < nb; ++b)
        rb[b] = (uint64_t)(std::upper_bound(ptr, ptr + nrows + 1, nnz * (uint64_t)b / nb) - ptr) - 1;
    for (int b = 1; b <= nb; ++b) rb[b] = std::max(rb[b], rb[b - 1]);
    std::vector<std::vector<uint64_t>> cnt(nb, std::vector<uint64_t>(ncols, 0));
#pragma omp parallel for schedule(static, 1) num_threads(nb)
    for (int b = 0; b < nb; ++b)
        for (uint64_t p = ptr[rb[b]]; p < ptr[rb[b + 1]]; ++p) cnt[b][idx[p]]++;
    uint64_t run = 0;
    for (uint64_t c = 0; c < ncols; ++c) {
        tptr[c] = run;
        for (int b = 0; b < nb; ++b) {
            const uint64_t k = cnt[b][c];
            cnt[b][c] = run;
            run += k;
        }
    }
    tptr[ncols] = run;
#pragma omp parallel for schedule(static, 1) num_threads(nb)
    for (int b = 0; b < nb; ++b) {
        uint64_t *cur = cnt[b].data();
        for (uint64_t r = rb[b]; r < rb[b + 1]; ++r)
            for (uint64_t p = ptr[r]; p < ptr[r + 1]; ++p) {
                const uint64_t q = cur[idx[p]]++;
                tidx[q] = (uint32_t)r;
                if (tval) (*tval)[q] = val[p];
            }
    }
}

struct Component {
    // eliminated-major after orientation, in SDDM coordinates
    std::vector<uint32_t> glob;  // ne + nk
    std::vector<double> diag;    // ne + nk
    uint32_t ne = 0, nk = 0;
    // C: ne x nk CSR (local kept index); the matrix is [[D_e, -C], [-C', D_k]]
    std::vector<uint64_t> cptr;
    std::vector<uint32_t> cidx;
    std::vector<double> cval;
    std::vector<double> fold;    // coordinate map: original = fold .* SDDM coordinates
    std::vector<double> surplus; // ground-edge weight, diag - sum |C| (zero when floating)
    bool grounded = false;       // surplus retained: no gauge, the kept block is definite
    bool signed_form = false;    // frustrated: C keeps both signs (solved through a cover)
};

// Magnitude-weighted neighbours of a component node in the bipartite graph: node i < ne is
// an eliminated level (row i of C), node ne + k the kept level k (column k, rows ascending).
struct Adjacency {
    std::vector<uint64_t> tp;
    std::vector<uint32_t> te;
    std::vector<double> tv;
    explicit Adjacency(const Component &cm) {
        par_transpose(cm.ne, cm.nk, cm.cptr.data(), cm.cidx.data(), cm.cval.data(), tp, te, &tv);
    }
    template <typename F> void each(const Component &cm, uint32_t i, F &&fn) const {
        if (i < cm.ne) {
            for (uint64_t p = cm.cptr[i]; p < cm.cptr[i + 1]; ++p) fn(cm.ne + cm.cidx[p], cm.cval[p]);
        } else {
            const uint32_t k = i - cm.ne;
            for (uint64_t q = tp[k]; q < tp[k + 1]; ++q) fn(te[q], tv[q]);
        }
    }
};

// Conversion of a general (slope-channel) component to SDDM form, following
// crates/within/src/domain/factor_pairs/sddm.rs:199-463: a folding signature when the signed
// graph is balanced (no cycle with an odd number of negative entries), else the block sign
// flip of the frustrated form; the dominance scaling of the diagonal (sweeps that raise mu_i
// to its neighbours' scaled weight, renormalised by the peak); the folded block
// -f_i f_j C_ij; and the surplus diag - sum |C| kept as ground-edge weight (grounded), or
// dropped when it is round-off (floating).  Returns false when no SDDM form exists.
bool convert_general(Component &cm, const LocalSolverSettings &cfg, int64_t &sweeps_out, double &violation_out) {
    const uint32_t ne = cm.ne, n = cm.ne + cm.nk;
    const double tol = cfg.scaling_tolerance;
    Adjacency adj(cm);
    // folding signature (sddm.rs:355-380): neighbours take the opposite sign of a positive entry
    std::vector<double> sg(n, 0.0);
    bool balanced = true;
    std::vector<uint32_t> stack;
    for (uint32_t root = 0; root < n && balanced; ++root) {
        if (sg[root] != 0.0) continue;
        sg[root] = 1.0;
        stack.assign(1, root);
        while (!stack.empty() && balanced) {
            const uint32_t i = stack.back();
            stack.pop_back();
            adj.each(cm, i, [&](uint32_t j, double v) {
                if (!balanced) return;
                const double expected = -sg[i] * (std::signbit(v) ? -1.0 : 1.0);
                if (sg[j] == 0.0) {
                    sg[j] = expected;
                    stack.push_back(j);
                } else if (sg[j] != expected) {
                    balanced = false;
                }
            });
        }
    }
    // dominance scaling (sddm.rs:389-463)
    for (double dv : cm.diag)
        if (!std::isfinite(dv) || !(dv > 0.0)) return false;
    std::vector<double> scales(n, 1.0);
    double violation = 0.0;
    int64_t sweeps = 0;
    bool dominant = true;
    for (uint32_t i = 0; i < n && dominant; ++i) {
        double rs = 0.0;
        adj.each(cm, i, [&](uint32_t, double v) { rs += std::fabs(v); });
        dominant = rs <= cm.diag[i] * (1.0 + tol);
    }
    if (!dominant) {
        std::vector<double> isq(n), mu(n, 1.0);
        for (uint32_t i = 0; i < n; ++i) isq[i] = 1.0 / std::sqrt(cm.diag[i]);
        auto target = [&](uint32_t i) {
            double t = 0.0;
            adj.each(cm, i, [&](uint32_t j, double v) { t += std::fabs(v) * isq[i] * isq[j] * mu[j]; });
            return t;
        };
        auto violation_of = [&]() {
            double m = 0.0;
            for (uint32_t i = 0; i < n; ++i) m = std::fmax(m, target(i) / mu[i] - 1.0);
            return m;
        };
        violation = violation_of();
        while (violation > tol && sweeps < cfg.scaling_max_sweeps) {
            for (uint32_t i = 0; i < n; ++i) {
                const double t = target(i);
                if (t > mu[i]) mu[i] = t;
            }
            double peak = 0.0;
            for (double v : mu) peak = std::fmax(peak, v);
            if (!std::isfinite(peak)) return false;
            for (double &v : mu) v /= peak;
            violation = violation_of();
            ++sweeps;
        }
        for (uint32_t i = 0; i < n; ++i) scales[i] = mu[i] * isq[i];
        for (double v : scales)
            if (!std::isfinite(v) || !(v > 0.0)) return false;
        violation = std::fmax(violation, 0.0);
    }
    if (violation > tol && cfg.scaling_error) return false;
    std::vector<double> f(n);
    for (uint32_t i = 0; i < n; ++i) f[i] = balanced ? sg[i] * scales[i] : (i < ne ? scales[i] : -scales[i]);
    // folded block; a balanced component must fold to a Z-matrix (sddm.rs:230-278)
    for (uint32_t e = 0; e < ne; ++e)
        for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) {
            const double folded = -f[e] * f[ne + cm.cidx[p]] * cm.cval[p];
            if (!std::isfinite(folded) || (balanced && folded < 0.0)) return false;
            cm.cval[p] = folded;
        }
    for (uint32_t i = 0; i < n; ++i) cm.diag[i] = cm.diag[i] * f[i] * f[i];
    // surplus and gauge (sddm.rs:281-336)
    const Adjacency folded_adj(cm);
    std::vector<double> rsum(n, 0.0), sur(n, 0.0);
    for (uint32_t i = 0; i < n; ++i) folded_adj.each(cm, i, [&](uint32_t, double v) { rsum[i] += std::fabs(v); });
    double total_d = 0.0, total_s = 0.0, clamped = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double &dg = cm.diag[i];
        if (!std::isfinite(dg) || !(dg > 0.0)) return false;
        const double deficit = (rsum[i] - dg) / dg;
        if (deficit > tol) {
            if (cfg.scaling_error) return false;
            clamped = std::fmax(clamped, deficit);
        }
        dg = std::fmax(dg, rsum[i]);
        sur[i] = std::fmax(dg - rsum[i], 0.0);
        total_d += dg;
        total_s += sur[i];
    }
    const bool floats = total_s <= 4.0 * 2.220446049250313e-16 * std::sqrt((double)std::max<uint32_t>(n, 1)) * total_d;
    if (floats) {
        cm.diag = rsum;
        std::fill(sur.begin(), sur.end(), 0.0);
    }
    cm.fold = std::move(f);
    cm.surplus = std::move(sur);
    cm.grounded = !floats;
    cm.signed_form = !balanced;
    sweeps_out = sweeps;
    violation_out = std::fmax(violation, clamped);
    return true;
}

std::string pair_name(Channel a, Channel b) {
    return "term " + std::to_string(a.term) + " column " + std::to_string(a.column) + " and term " +
           std::to_string(b.term) + " column " + std::to_string(b.column);
}

// Connected components of a bipartite graph (rows 0..nr-1, columns nr..nr+nc-1, edges from a
// CSR) by a lock-free union-find: every edge hooks the larger of its two roots under the smaller,
// so each component's root is its smallest node.  label[v] = that root.
std::vector<uint32_t> bipartite_roots(uint32_t nr, uint32_t nc, const std::vector<uint64_t> &cptr,
                                      const std::vector<uint32_t> &cidx) {
    const uint32_t nl = nr + nc;
    std::vector<std::atomic<uint32_t>> parent(nl);
#pragma omp parallel for schedule(static) if (nl >= (1u << 16))
    for (int64_t v = 0; v < (int64_t)nl; ++v) parent[v].store((uint32_t)v, std::memory_order_relaxed);
    auto find = [&](uint32_t x) {
        for (;;) {
            const uint32_t p = parent[x].load(std::memory_order_relaxed);
            if (p == x) return x;
            const uint32_t g = parent[p].load(std::memory_order_relaxed);
            if (g != p) { // path halving: point x at its grandparent (still an ancestor)
                uint32_t seen = p;
                parent[x].compare_exchange_weak(seen, g, std::memory_order_relaxed);
            }
            x = g;
        }
    };
#pragma omp parallel for schedule(dynamic, 1024) if (cidx.size() >= (1u << 16))
    for (int64_t r = 0; r < (int64_t)nr; ++r)
        for (uint64_t q = cptr[r]; q < cptr[r + 1]; ++q) {
            uint32_t a = find((uint32_t)r), b = find(nr + cidx[q]);
            while (a != b) {
                if (a > b) std::swap(a, b); // hook b (the larger root) under a
                uint32_t expect = b;
                if (parent[b].compare_exchange_strong(expect, a, std::memory_order_relaxed)) break;
                a = find(a);
                b = find(b);
            }
        }
    std::vector<uint32_t> label(nl);
#pragma omp parallel for schedule(static) if (nl >= (1u << 16))
    for (int64_t v = 0; v < (int64_t)nl; ++v) label[v] = find((uint32_t)v);
    return label;
}

// A plain pair's components from its device cross-tab cells: the same components, in the same
// order and with the same bits as the serial path below (components ordered by their smallest
// node, each row's entries in column order, every kept level's sum over rows in row order), with
// the cross block, its transpose, the components and each component's extraction in parallel.
void plain_pair_components_from_cells(const Design &d, Channel rc, Channel cc,
                                      const std::vector<std::vector<uint8_t>> &active, const CellList &cells,
                                      std::vector<Component> &out, std::string &error) {
    const Term &tr = d.terms[rc.term], &tc = d.terms[cc.term];
    std::vector<uint32_t> rmap(tr.n_levels, UINT32_MAX), cmap(tc.n_levels, UINT32_MAX);
    uint32_t nr = 0, nc = 0;
    for (uint64_t l = 0; l < tr.n_levels; ++l)
        if (active[rc.term][l]) rmap[l] = nr++;
    for (uint64_t l = 0; l < tc.n_levels; ++l)
        if (active[cc.term][l]) cmap[l] = nc++;
    if (nr == 0 || nc == 0) return;
    std::vector<uint32_t> l2g;
    l2g.reserve(nr + nc);
    const uint64_t base_r = tr.offset + rc.column * tr.n_levels, base_c = tc.offset + cc.column * tc.n_levels;
    for (uint64_t l = 0; l < tr.n_levels; ++l)
        if (active[rc.term][l]) l2g.push_back((uint32_t)(base_r + l));
    for (uint64_t l = 0; l < tc.n_levels; ++l)
        if (active[cc.term][l]) l2g.push_back((uint32_t)(base_c + l));
    // the cross block C (rows = compact row levels, nonzero cells only) from the key-sorted cells
    const uint64_t ncell = cells.key.size(), lcn = tc.n_levels;
    const bool par = ncell >= (1u << 16);
    const int nt = par ? std::max(1, std::min(omp_get_max_threads(), 32)) : 1;
    std::vector<uint64_t> cptr(nr + 1, 0), chunk_nz(nt + 1, 0);
    auto clo = [&](int b) { return ncell * (uint64_t)b / (uint64_t)nt; };
#pragma omp parallel for schedule(static, 1) num_threads(nt)
    for (int b = 0; b < nt; ++b) {
        uint64_t z = 0;
        for (uint64_t q = clo(b); q < clo(b + 1); ++q) z += cells.val[q] != 0.0;
        chunk_nz[b + 1] = z;
    }
    for (int b = 0; b < nt; ++b) chunk_nz[b + 1] += chunk_nz[b];
    std::vector<uint32_t> cidx(chunk_nz[nt]), crow(chunk_nz[nt]);
    std::vector<double> cval(chunk_nz[nt]);
#pragma omp parallel for schedule(static, 1) num_threads(nt)
    for (int b = 0; b < nt; ++b) {
        uint64_t o = chunk_nz[b];
        for (uint64_t q = clo(b); q < clo(b + 1); ++q) {
            const double v = cells.val[q];
            if (v == 0.0) continue;
            crow[o] = rmap[cells.key[q] / lcn];
            cidx[o] = cmap[cells.key[q] % lcn];
            cval[o] = v;
            ++o;
        }
    }
    for (uint64_t q = 0; q < crow.size(); ++q) // rows ascending: count, then prefix
        cptr[crow[q] + 1]++;
    std::vector<uint32_t>().swap(crow);
    for (uint32_t r = 0; r < nr; ++r) cptr[r + 1] += cptr[r];
    // weight sums per level (zero-valued cells add nothing), for the dead-component test
    std::vector<double> rdiag(nr, 0.0), cdiag(nc, 0.0);
    std::vector<uint64_t> tptr;
    std::vector<uint32_t> tidx;
    std::vector<double> tval;
    par_transpose(nr, nc, cptr.data(), cidx.data(), cval.data(), tptr, tidx, &tval);
#pragma omp parallel for schedule(static) if (par)
    for (int64_t r = 0; r < (int64_t)nr; ++r) {
        double a = 0.0;
        for (uint64_t q = cptr[r]; q < cptr[r + 1]; ++q) a += cval[q];
        rdiag[r] = a;
    }
#pragma omp parallel for schedule(static) if (par)
    for (int64_t c = 0; c < (int64_t)nc; ++c) {
        double a = 0.0;
        for (uint64_t q = tptr[c]; q < tptr[c + 1]; ++q) a += tval[q];
        cdiag[c] = a;
    }
    // components, ordered by their smallest node; members ascending
    const std::vector<uint32_t> root = bipartite_roots(nr, nc, cptr, cidx);
    const uint32_t nl = nr + nc;
    std::vector<uint32_t> comp_of(nl, UINT32_MAX), first;
    for (uint32_t v = 0; v < nl; ++v)
        if (root[v] == v) {
            comp_of[v] = (uint32_t)first.size();
            first.push_back(v);
        }
    const uint32_t ncomp = (uint32_t)first.size();
    std::vector<uint64_t> mptr(ncomp + 1, 0);
    for (uint32_t v = 0; v < nl; ++v) mptr[comp_of[root[v]] + 1]++;
    for (uint32_t c = 0; c < ncomp; ++c) mptr[c + 1] += mptr[c];
    std::vector<uint32_t> members(nl);
    {
        std::vector<uint64_t> cur(mptr.begin(), mptr.end() - 1);
        for (uint32_t v = 0; v < nl; ++v) members[cur[comp_of[root[v]]]++] = v;
    }
    std::vector<uint32_t> kmap(std::max(nr, nc), UINT32_MAX); // kept level -> local index (shared, reset per component)
    for (uint32_t ci = 0; ci < ncomp; ++ci) {
        const uint32_t *mb = members.data() + mptr[ci], *me = members.data() + mptr[ci + 1];
        const uint32_t *split = std::lower_bound(mb, me, nr); // rows, then columns (as nr + c)
        const uint64_t n_rows = split - mb, n_cols = me - split;
        bool dead = true;
        for (const uint32_t *v = mb; v < split; ++v) dead &= rdiag[*v] == 0.0;
        for (const uint32_t *v = split; v < me; ++v) dead &= cdiag[*v - nr] == 0.0;
        if (dead) continue;
        // orientation: eliminate the larger block; swap iff n_cols > n_rows (sddm.rs:86-106)
        const bool swap = n_cols > n_rows;
        std::vector<uint32_t> E, K;
        for (const uint32_t *v = mb; v < split; ++v) (swap ? K : E).push_back(*v);
        for (const uint32_t *v = split; v < me; ++v) (swap ? E : K).push_back(*v - nr);
        Component comp;
        comp.ne = (uint32_t)E.size();
        comp.nk = (uint32_t)K.size();
        comp.glob.reserve(E.size() + K.size());
        for (uint32_t e : E) comp.glob.push_back(swap ? l2g[nr + e] : l2g[e]);
        for (uint32_t k : K) comp.glob.push_back(swap ? l2g[k] : l2g[nr + k]);
        for (uint32_t i = 0; i < (uint32_t)K.size(); ++i) kmap[K[i]] = i;
        // an eliminated node's entries: its C row (no swap) or its C' row (swap), kept-ascending
        const uint64_t *eptr = swap ? tptr.data() : cptr.data();
        const uint32_t *eidx = swap ? tidx.data() : cidx.data();
        const double *evalp = swap ? tval.data() : cval.data();
        const uint64_t *kptr = swap ? cptr.data() : tptr.data();
        const double *kval = swap ? cval.data() : tval.data();
        comp.cptr.assign(E.size() + 1, 0);
        for (size_t ei = 0; ei < E.size(); ++ei) comp.cptr[ei + 1] = comp.cptr[ei] + (eptr[E[ei] + 1] - eptr[E[ei]]);
        const uint64_t nnz = comp.cptr[E.size()];
        comp.cidx.resize(nnz);
        comp.cval.resize(nnz);
        std::vector<double> rowsum_e(E.size(), 0.0), rowsum_k(K.size(), 0.0);
        const bool cpar = nnz >= (1u << 16);
        (void)cpar;
#pragma omp parallel for schedule(dynamic, 256) if (cpar)
        for (int64_t ei = 0; ei < (int64_t)E.size(); ++ei) {
            uint64_t o = comp.cptr[ei];
            double a = 0.0;
            for (uint64_t q = eptr[E[ei]]; q < eptr[E[ei] + 1]; ++q, ++o) {
                comp.cidx[o] = kmap[eidx[q]];
                comp.cval[o] = evalp[q];
                a += evalp[q];
            }
            rowsum_e[ei] = a;
        }
        // a kept level's sum over the eliminated levels in ascending order
#pragma omp parallel for schedule(dynamic, 64) if (cpar)
        for (int64_t ki = 0; ki < (int64_t)K.size(); ++ki) {
            double a = 0.0;
            for (uint64_t q = kptr[K[ki]]; q < kptr[K[ki] + 1]; ++q) a += kval[q];
            rowsum_k[ki] = a;
        }
        for (uint32_t k : K) kmap[k] = UINT32_MAX;
        // validate the Laplacian claim and adopt the row sums (sddm.rs:173-197)
        double total_diag = 0.0, mismatch = 0.0;
        for (size_t ei = 0; ei < E.size(); ++ei) {
            const double dg = swap ? cdiag[E[ei]] : rdiag[E[ei]];
            total_diag += dg;
            mismatch += std::fabs(dg - rowsum_e[ei]);
        }
        for (size_t ki = 0; ki < K.size(); ++ki) {
            const double dg = swap ? rdiag[K[ki]] : cdiag[K[ki]];
            total_diag += dg;
            mismatch += std::fabs(dg - rowsum_k[ki]);
        }
        const double tol =
            64.0 * 2.220446049250313e-16 * std::sqrt((double)std::max<size_t>(1, E.size() + K.size())) * total_diag;
        if (mismatch > tol) {
            error = "signed component between " + pair_name(rc, cc) + " is not diagonally scalable to SDDM form";
            return;
        }
        comp.diag = std::move(rowsum_e);
        comp.diag.insert(comp.diag.end(), rowsum_k.begin(), rowsum_k.end());
        comp.fold.assign(E.size() + K.size(), 1.0);
        std::fill(comp.fold.begin() + E.size(), comp.fold.end(), -1.0);
        comp.surplus.assign(E.size() + K.size(), 0.0);
        out.push_back(std::move(comp));
    }
}

// Cross-tab of one channel pair and its components (cross_tab.rs, accumulate.rs, factor_pairs.rs:95-152),
// each converted to SDDM form (sddm.rs).  Runs inside a parallel loop, so a failure is returned
// as a message in `error` rather than thrown.
// `cells`, when given (a plain pair), holds the cross-tab computed on the device, bitwise equal to
// the host accumulation below.
void build_pair_components(const Design &d, const std::vector<double> &w, Channel rc, Channel cc,
                           const std::vector<std::vector<uint8_t>> &active, const LocalSolverSettings &cfg,
                           std::vector<Component> &out, std::vector<std::string> &warnings, std::string &error,
                           const CellList *cells = nullptr) {
    const Term &tr = d.terms[rc.term], &tc = d.terms[cc.term];
    const int64_t lr = tr.loads[rc.column], lc = tc.loads[cc.column];
    const bool general = lr >= 0 || lc >= 0;
    if (cells && !general) {
        plain_pair_components_from_cells(d, rc, cc, active, *cells, out, error);
        return;
    }
    // compact active levels
    std::vector<uint32_t> rmap(tr.n_levels, UINT32_MAX), cmap(tc.n_levels, UINT32_MAX);
    uint32_t nr = 0, nc = 0;
    for (uint64_t l = 0; l < tr.n_levels; ++l)
        if (active[rc.term][l]) rmap[l] = nr++;
    for (uint64_t l = 0; l < tc.n_levels; ++l)
        if (active[cc.term][l]) cmap[l] = nc++;
    if (nr == 0 || nc == 0) return;
    std::vector<uint32_t> l2g;
    l2g.reserve(nr + nc);
    const uint64_t base_r = tr.offset + rc.column * tr.n_levels, base_c = tc.offset + cc.column * tc.n_levels;
    for (uint64_t l = 0; l < tr.n_levels; ++l)
        if (active[rc.term][l]) l2g.push_back((uint32_t)(base_r + l));
    for (uint64_t l = 0; l < tc.n_levels; ++l)
        if (active[cc.term][l]) l2g.push_back((uint32_t)(base_c + l));
    std::vector<double> rdiag(nr, 0.0), cdiag(nc, 0.0);
    std::vector<uint64_t> cptr(nr + 1, 0);
    std::vector<uint32_t> cidx;
    std::vector<double> cval;
    if (cells && !general) {
        // the device's cells (row-major keys level_r * L_c + level_c); zero-weight cells dropped as below
        const uint64_t lcn = tc.n_levels;
        cidx.reserve(cells->key.size());
        cval.reserve(cells->key.size());
        for (size_t q = 0; q < cells->key.size(); ++q) {
            const uint32_t cj = rmap[cells->key[q] / lcn], ck = cmap[cells->key[q] % lcn];
            const double v = cells->val[q];
            rdiag[cj] += v;
            cdiag[ck] += v;
            if (v != 0.0) {
                cidx.push_back(ck);
                cval.push_back(v);
                cptr[cj + 1]++;
            }
        }
        for (uint32_t r = 0; r < nr; ++r) cptr[r + 1] += cptr[r];
    } else {
        // accumulate: bucket by row (observation order), dedup per row with sorted columns
        const uint32_t *rlev = d.level_column(rc.term), *clev = d.level_column(cc.term);
        const double *zr = lr >= 0 ? d.slope_column((uint64_t)lr) : nullptr;
        const double *zc = lc >= 0 ? d.slope_column((uint64_t)lc) : nullptr;
        std::vector<uint64_t> rcount(nr + 1, 0);
        for (uint64_t i = 0; i < d.n_obs; ++i) {
            const uint32_t cj = rmap[rlev[i]], ck = cmap[clev[i]];
            const double wi = w.empty() ? 1.0 : w[i];
            const double a = zr ? zr[i] : 1.0, b = zc ? zc[i] : 1.0;
            rdiag[cj] += wi * a * a;
            cdiag[ck] += wi * b * b;
            rcount[cj + 1]++;
        }
        for (uint32_t r = 0; r < nr; ++r) rcount[r + 1] += rcount[r];
        std::vector<uint32_t> bcol(d.n_obs);
        std::vector<double> bval(d.n_obs);
        {
            std::vector<uint64_t> cur(rcount.begin(), rcount.end() - 1);
            for (uint64_t i = 0; i < d.n_obs; ++i) {
                const uint32_t cj = rmap[rlev[i]];
                const double wi = w.empty() ? 1.0 : w[i];
                const double a = zr ? zr[i] : 1.0, b = zc ? zc[i] : 1.0;
                bcol[cur[cj]] = cmap[clev[i]];
                bval[cur[cj]] = wi * a * b;
                cur[cj]++;
            }
        }
        cidx.reserve(d.n_obs);
        cval.reserve(d.n_obs);
        {
            std::vector<double> work(nc, 0.0);
            std::vector<uint32_t> touched;
            std::vector<uint8_t> seen(nc, 0);
            for (uint32_t r = 0; r < nr; ++r) {
                for (uint64_t p = rcount[r]; p < rcount[r + 1]; ++p) {
                    const uint32_t c = bcol[p];
                    if (!seen[c]) {
                        seen[c] = 1;
                        touched.push_back(c);
                    }
                    work[c] += bval[p];
                }
                std::sort(touched.begin(), touched.end());
                for (uint32_t c : touched) {
                    if (work[c] != 0.0) {
                        cidx.push_back(c);
                        cval.push_back(work[c]);
                    }
                    work[c] = 0.0;
                    seen[c] = 0;
                }
                touched.clear();
                cptr[r + 1] = cidx.size();
            }
        }
        std::vector<uint32_t>().swap(bcol);
        std::vector<double>().swap(bval);
    }
    // transpose
    std::vector<uint64_t> tptr(nc + 1, 0);
    for (uint32_t c : cidx) tptr[c + 1]++;
    for (uint32_t c = 0; c < nc; ++c) tptr[c + 1] += tptr[c];
    std::vector<uint32_t> tidx(cidx.size());
    {
        std::vector<uint64_t> cur(tptr.begin(), tptr.end() - 1);
        for (uint32_t r = 0; r < nr; ++r)
            for (uint64_t p = cptr[r]; p < cptr[r + 1]; ++p) tidx[cur[cidx[p]]++] = r;
    }
    // components by DFS from nodes 0..n_local (cross_tab.rs:175-210)
    const uint32_t nl = nr + nc;
    std::vector<uint8_t> visited(nl, 0);
    std::vector<uint32_t> stack, rows, cols;
    std::vector<uint32_t> rremap(nr, UINT32_MAX), cremap(nc, UINT32_MAX);
    for (uint32_t s0 = 0; s0 < nl; ++s0) {
        if (visited[s0]) continue;
        visited[s0] = 1;
        stack.push_back(s0);
        rows.clear();
        cols.clear();
        while (!stack.empty()) {
            uint32_t node = stack.back();
            stack.pop_back();
            if (node < nr) {
                rows.push_back(node);
                for (uint64_t p = cptr[node]; p < cptr[node + 1]; ++p) {
                    uint32_t j = nr + cidx[p];
                    if (!visited[j]) {
                        visited[j] = 1;
                        stack.push_back(j);
                    }
                }
            } else {
                uint32_t c = node - nr;
                cols.push_back(c);
                for (uint64_t p = tptr[c]; p < tptr[c + 1]; ++p) {
                    uint32_t j = tidx[p];
                    if (!visited[j]) {
                        visited[j] = 1;
                        stack.push_back(j);
                    }
                }
            }
        }
        std::sort(rows.begin(), rows.end());
        std::sort(cols.begin(), cols.end());
        // dead component: all-zero diagonal (factor_pairs.rs:124-126)
        bool dead = true;
        for (uint32_t r : rows) dead &= rdiag[r] == 0.0;
        for (uint32_t c : cols) dead &= cdiag[c] == 0.0;
        if (dead) continue;
        Component comp;
        // orientation: eliminate the larger block; swap iff n_cols > n_rows (sddm.rs:86-106)
        const bool swap = cols.size() > rows.size();
        const std::vector<uint32_t> &E = swap ? cols : rows, &K = swap ? rows : cols;
        comp.ne = (uint32_t)E.size();
        comp.nk = (uint32_t)K.size();
        for (uint32_t e : E) comp.glob.push_back(swap ? l2g[nr + e] : l2g[e]);
        for (uint32_t k : K) comp.glob.push_back(swap ? l2g[k] : l2g[nr + k]);
        for (uint32_t i = 0; i < (uint32_t)K.size(); ++i) (swap ? rremap[K[i]] : cremap[K[i]]) = i;
        comp.cptr.push_back(0);
        std::vector<double> rowsum_e(E.size(), 0.0), rowsum_k(K.size(), 0.0);
        for (size_t ei = 0; ei < E.size(); ++ei) {
            const uint32_t e = E[ei];
            const uint64_t b = swap ? tptr[e] : cptr[e], en = swap ? tptr[e + 1] : cptr[e + 1];
            // build the row; for the swapped orientation the row comes from the transpose,
            // whose column indices are rows of C (sorted ascending)
            for (uint64_t p = b; p < en; ++p) {
                uint32_t other;
                double val;
                if (swap) {
                    other = tidx[p];
                    // value of C[other, e]: search in row `other`
                    auto it = std::lower_bound(cidx.begin() + cptr[other], cidx.begin() + cptr[other + 1], e);
                    val = cval[it - cidx.begin()];
                } else {
                    other = cidx[p];
                    val = cval[p];
                }
                const uint32_t kl = swap ? rremap[other] : cremap[other];
                comp.cidx.push_back(kl);
                comp.cval.push_back(val);
                rowsum_e[ei] += val;
                rowsum_k[kl] += val;
            }
            comp.cptr.push_back(comp.cidx.size());
        }
        // diagonal: plain components adopt the adjacency row sums (sddm.rs:173-197)
        double total_diag = 0.0, mismatch = 0.0;
        for (size_t ei = 0; ei < E.size(); ++ei) {
            const double dg = swap ? cdiag[E[ei]] : rdiag[E[ei]];
            total_diag += dg;
            mismatch += std::fabs(dg - rowsum_e[ei]);
        }
        for (size_t ki = 0; ki < K.size(); ++ki) {
            const double dg = swap ? rdiag[K[ki]] : cdiag[K[ki]];
            total_diag += dg;
            mismatch += std::fabs(dg - rowsum_k[ki]);
        }
        for (uint32_t k : K) (swap ? rremap[k] : cremap[k]) = UINT32_MAX;
        if (!general) {
            // plain pair: validate the Laplacian claim and adopt the row sums (sddm.rs:173-197);
            // the coordinate map negates the kept block
            const double tol = 64.0 * 2.220446049250313e-16 * std::sqrt((double)std::max<size_t>(1, E.size() + K.size())) * total_diag;
            if (mismatch > tol) {
                error = "signed component between " + pair_name(rc, cc) + " is not diagonally scalable to SDDM form";
                return;
            }
            comp.diag = rowsum_e;
            comp.diag.insert(comp.diag.end(), rowsum_k.begin(), rowsum_k.end());
            comp.fold.assign(E.size() + K.size(), 1.0);
            std::fill(comp.fold.begin() + E.size(), comp.fold.end(), -1.0);
            comp.surplus.assign(E.size() + K.size(), 0.0);
        } else {
            for (size_t ei = 0; ei < E.size(); ++ei) comp.diag.push_back(swap ? cdiag[E[ei]] : rdiag[E[ei]]);
            for (size_t ki = 0; ki < K.size(); ++ki) comp.diag.push_back(swap ? rdiag[K[ki]] : cdiag[K[ki]]);
            int64_t sweeps = 0;
            double violation = 0.0;
            if (!convert_general(comp, cfg, sweeps, violation)) {
                error = "signed component between " + pair_name(rc, cc) + " is not diagonally scalable to SDDM form";
                return;
            }
            if (violation > cfg.scaling_tolerance) {
                char buf[64];
                std::snprintf(buf, sizeof buf, "%.2e", violation);
                warnings.push_back("signed component between " + pair_name(rc, cc) +
                                   ": dominance scaling uncertified after " + std::to_string(sweeps) +
                                   " sweeps (max relative violation " + buf +
                                   "); deficits clamped, preconditioner quality may degrade");
            }
        }
        out.push_back(std::move(comp));
    }
}

// Dense reduced kept block of a component (row-major nk x nk, rows in parallel): the exact
// Schur complement's off-diagonals -sum_e (C_ei C_ej) / D_e (schur.rs:96-113) and its diagonal
// rebuilt from the adjacency plus the reduced surplus (schur.rs:197-268), so a floating
// Laplacian's rows sum to zero exactly.  For a signed component this is the single-sized
// operator whose action on antisymmetric vectors equals the Gremban cover's complement; its
// diagonal is the cover row's adjacency, sum_e |C_ei| (sum_j |C_ej| - |C_ei|) / D_e.
void dense_schur(const Component &cm, std::vector<double> &Sm, bool parallel) {
    const uint32_t ne = cm.ne, nk = cm.nk;
    const Adjacency adj(cm);
    std::vector<double> rowabs(ne, 0.0);
    for (uint32_t e = 0; e < ne; ++e)
        for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) rowabs[e] += std::fabs(cm.cval[p]);
    Sm.assign((size_t)nk * nk, 0.0);
#pragma omp parallel for schedule(dynamic, 8) if (parallel)
    for (int64_t i = 0; i < (int64_t)nk; ++i) {
        double *row = &Sm[(size_t)i * nk];
        double surplus = cm.grounded ? cm.surplus[ne + i] : 0.0, mag = 0.0;
        for (uint64_t q = adj.tp[i]; q < adj.tp[i + 1]; ++q) {
            const uint32_t e = adj.te[q];
            const double wi = adj.tv[q], inv = 1.0 / cm.diag[e];
            for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) row[cm.cidx[p]] -= (wi * cm.cval[p]) * inv;
            if (cm.grounded) surplus += std::fabs(wi) * (inv * cm.surplus[e]);
            if (cm.signed_form) mag += std::fabs(wi) * (rowabs[e] - std::fabs(wi)) * inv;
        }
        double adjacency = 0.0;
        if (cm.signed_form) adjacency = mag;
        else
            for (uint32_t j = 0; j < nk; ++j)
                if (j != (uint32_t)i) adjacency -= row[j];
        row[i] = adjacency + surplus;
    }
}

// Schur-complement edges of a Laplacian-form component over its kept vertices 0..nk-1 plus,
// when grounded, the ground vertex nk (schur.rs:64-68, 337-372): per eliminated star the
// sampled clique trees (or the exact clique), a grounded star's surplus joining it as one
// more neighbour (its capacity is then its diagonal), and the kept surplus as ground edges.
void schur_edges(const Component &cm, const LocalSolverSettings &cfg, uint64_t stream, std::vector<Edge> &edges) {
    const uint32_t ne = cm.ne, nk = cm.nk, ground = nk;
    std::vector<std::pair<uint32_t, double>> nb;
    for (uint32_t e = 0; e < ne; ++e) {
        nb.clear();
        for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) nb.push_back({cm.cidx[p], cm.cval[p]});
        if (cm.grounded && cm.surplus[e] > 0.0) nb.push_back({ground, cm.surplus[e]});
        if (nb.size() < 2) continue;
        const double De = cm.diag[e];
        if (cfg.schur_exact) {
            for (size_t p = 0; p < nb.size(); ++p)
                for (size_t q = p + 1; q < nb.size(); ++q)
                    edges.push_back({std::min(nb[p].first, nb[q].first), std::max(nb[p].first, nb[q].first),
                                     (nb[p].second * nb[q].second) / De});
        } else {
            Rng rng(mix_seed(cfg.schur_seed, e + (stream << 32), 0x5C4));
            sample_clique_tree(nb, De, std::max(1, cfg.schur_split), rng, edges);
        }
    }
    if (cm.grounded)
        for (uint32_t k = 0; k < nk; ++k)
            if (cm.surplus[ne + k] > 0.0) edges.push_back({k, ground, cm.surplus[ne + k]});
    sort_and_merge(edges);
}

// Gremban cover of a signed component (block_elim/solver.rs:221-277): two sheets of every
// vertex; a positive entry joins equal sheets and a negative one opposite sheets, each with
// its magnitude.  The cover is a Laplacian-form SDDM of twice the size (surplus and gauge
// carried over) acting on antisymmetric vectors [z, -z] as the signed matrix does.
Component make_cover(const Component &cm) {
    Component cv;
    const uint32_t ne = cm.ne, nk = cm.nk;
    cv.ne = 2 * ne;
    cv.nk = 2 * nk;
    cv.grounded = cm.grounded;
    cv.cptr.push_back(0);
    for (int copy = 0; copy < 2; ++copy)
        for (uint32_t e = 0; e < ne; ++e) {
            for (int shifted = 0; shifted < 2; ++shifted) {
                const bool negative = shifted != copy;
                for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) {
                    if ((cm.cval[p] < 0.0) != negative) continue;
                    cv.cidx.push_back(cm.cidx[p] + (shifted ? nk : 0));
                    cv.cval.push_back(std::fabs(cm.cval[p]));
                }
            }
            cv.cptr.push_back(cv.cidx.size());
        }
    auto doubled = [&](const std::vector<double> &v) {
        std::vector<double> o;
        o.insert(o.end(), v.begin(), v.begin() + ne);
        o.insert(o.end(), v.begin(), v.begin() + ne);
        o.insert(o.end(), v.begin() + ne, v.end());
        o.insert(o.end(), v.begin() + ne, v.end());
        return o;
    };
    cv.diag = doubled(cm.diag);
    cv.surplus = doubled(cm.surplus);
    return cv;
}

} // namespace

namespace {
std::shared_ptr<Precond> assemble_additive(std::vector<Component> &comps, uint64_t n_dofs, bool any_slope,
                                           const LocalSolverSettings &cfg,
                                           const std::function<void(const char *)> &phase);
} // namespace

std::shared_ptr<Precond> build_additive(const Design &d, const std::vector<double> &w, const LocalSolverSettings &cfg,
                                        std::vector<std::string> &warnings) {
    // WITHIN_AMD_VERBOSE: per-phase wall times of the build on stderr
    const bool verbose = std::getenv("WITHIN_AMD_VERBOSE") != nullptr;
    auto t_last = std::chrono::steady_clock::now();
    auto phase = [&](const char *name) {
        if (!verbose) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[within_amd] preconditioner build: %s %.1f ms\n", name,
                     std::chrono::duration<double, std::milli>(now - t_last).count());
        t_last = now;
    };
    // channels and cross-term pairs (factor_pairs.rs:36-56)
    std::vector<Channel> channels;
    for (uint32_t t = 0; t < d.terms.size(); ++t)
        for (uint32_t c = 0; c < d.terms[t].ncols(); ++c) channels.push_back({t, c});
    std::vector<std::pair<Channel, Channel>> pairs;
    for (size_t i = 0; i < channels.size(); ++i)
        for (size_t j = i + 1; j < channels.size(); ++j)
            if (channels[j].term != channels[i].term) pairs.push_back({channels[i], channels[j]});
    bool any_slope = false;
    for (auto &c : channels) any_slope |= d.terms[c.term].loads[c.column] >= 0;
    const auto active = find_active_levels(d);
    // the cross-tabs of the plain (intercept) pairs on the device when a GPU is visible (the same
    // cells as the host accumulation, bit for bit)
    std::vector<std::pair<uint32_t, uint32_t>> plain;
    std::vector<int64_t> plain_of(pairs.size(), -1);
    for (size_t p = 0; p < pairs.size(); ++p)
        if (d.terms[pairs[p].first.term].loads[pairs[p].first.column] < 0 &&
            d.terms[pairs[p].second.term].loads[pairs[p].second.column] < 0) {
            plain_of[p] = (int64_t)plain.size();
            plain.push_back({pairs[p].first.term, pairs[p].second.term});
        }
    std::vector<CellList> cells;
    const bool on_device = d.n_obs <= 0x7FFFFFFFull && device_cross_tabs(d, w, plain, cells);
    if (on_device) phase("cross-tabs on the device");
    std::vector<std::vector<Component>> per_pair(pairs.size());
    std::vector<std::vector<std::string>> pair_warn(pairs.size());
    std::vector<std::string> pair_err(pairs.size());
    // pairs accumulated on the host run side by side; pairs with device cells run one after
    // another, each using every thread inside
    auto from_cells = [&](int64_t p) { return on_device && plain_of[p] >= 0; };
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t p = 0; p < (int64_t)pairs.size(); ++p)
        if (!from_cells(p))
            build_pair_components(d, w, pairs[p].first, pairs[p].second, active, cfg, per_pair[p], pair_warn[p],
                                  pair_err[p], nullptr);
    for (int64_t p = 0; p < (int64_t)pairs.size(); ++p)
        if (from_cells(p))
            build_pair_components(d, w, pairs[p].first, pairs[p].second, active, cfg, per_pair[p], pair_warn[p],
                                  pair_err[p], &cells[plain_of[p]]);
    phase("cross-tabs and components");
    for (size_t p = 0; p < pairs.size(); ++p) {
        if (!pair_err[p].empty()) fail(WX_ERR_BUILD, pair_err[p], "BuildError::UnscalableComponent");
        warnings.insert(warnings.end(), pair_warn[p].begin(), pair_warn[p].end());
    }
    std::vector<Component> comps;
    for (auto &v : per_pair)
        for (auto &c : v) comps.push_back(std::move(c));
    if (comps.empty()) return nullptr;
    return assemble_additive(comps, d.n_dofs, any_slope, cfg, phase);
}

namespace {

// Factorisation of every component's kept block and assembly of the slot arrays
// (operator/schwarz.rs:90-113, block_elim/solver.rs:311-347).
std::shared_ptr<Precond> assemble_additive(std::vector<Component> &comps, uint64_t n_dofs, bool any_slope,
                                           const LocalSolverSettings &cfg,
                                           const std::function<void(const char *)> &phase) {
    auto pre = std::make_shared<Precond>();
    pre->kind = Precond::ADDITIVE;
    pre->n = n_dofs;
    auto S = std::make_shared<Precond::Schwarz>();
    S->n_dofs = n_dofs;
    // partition of unity 1/sqrt(count), only without slope channels (factor_pairs.rs:84-89,155-194)
    std::vector<uint32_t> counts(n_dofs, 0);
    for (auto &c : comps)
        for (uint32_t g : c.glob) counts[g]++;
    const uint64_t nc = comps.size();
    S->slot_start.assign(nc + 1, 0);
    S->ne.resize(nc);
    S->nk.resize(nc);
    S->fac_kind.resize(nc);
    S->gauge.resize(nc);
    S->dense_off.assign(nc, 0);
    S->l_row_off.assign(nc, 0);
    S->piv_off.assign(nc, 0);
    S->cover_off.assign(nc, 0);
    for (uint64_t c = 0; c < nc; ++c) {
        S->ne[c] = comps[c].ne;
        S->nk[c] = comps[c].nk;
        S->slot_start[c + 1] = S->slot_start[c] + comps[c].ne + comps[c].nk;
        S->gauge[c] = (comps[c].grounded || comps[c].signed_form) ? 1 : 0;
    }
    const uint64_t ns = S->slot_start[nc];
    S->glob.resize(ns);
    S->pw.resize(ns);
    S->fold.resize(ns);
    S->inv_de.assign(ns, 0.0);
    S->c_ptr.assign(ns + 1, 0);
    // factor choice per component.  The reference rule (block_elim/solver.rs:311-347): the
    // exact dense factor iff n_kept <= dense_threshold.  With the device policy a kept block is
    // also factored densely where the device cost model prefers it: small kept blocks and
    // dense-ish reduced graphs get the exact dense inverse -- one parallel symmetric product per
    // apply -- while large sparse reduced graphs keep approximate Cholesky (of the Gremban cover
    // for a signed block), whose triangular solves run level by level.  The density estimate is
    // the sampled Schur complement's edge count sum_e (deg(e) - 1) over nk^2 / 2.  The choice
    // depends on the config and the component only; a visible GPU merely computes the large
    // dense blocks (`big`) instead of the host.
    // WITHIN_AMD_HOST_SETUP=1 builds the large blocks on the host too (the device-independent
    // reference the device build is tested against)
    const bool have_gpu = gpu_available() && std::getenv("WITHIN_AMD_HOST_SETUP") == nullptr;
    std::vector<uint8_t> dense(nc), big(nc, 0);
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t nk = comps[c].nk;
        bool d = cfg.dense_threshold >= 0 && nk <= (uint64_t)cfg.dense_threshold;
        if (!d && cfg.dense_device) {
            if (nk <= 4096) d = true;
            else if (nk <= 16384) {
                double m = 0.0;
                for (uint32_t e = 0; e < comps[c].ne; ++e) {
                    const uint64_t deg = comps[c].cptr[e + 1] - comps[c].cptr[e];
                    if (deg > 1) m += (double)(deg - 1);
                }
                d = 2.0 * m / ((double)nk * nk) >= 0.005;
            }
        }
        dense[c] = d;
        big[c] = dense[c] && nk > 1024 && have_gpu;
        S->fac_kind[c] = dense[c] ? 0 : (comps[c].signed_form ? 2 : 1);
    }
    // dense blocks' place in the concatenated storage; the GPU path writes its blocks there
    uint64_t dense_total = 0;
    for (uint64_t c = 0; c < nc; ++c)
        if (dense[c]) {
            S->dense_off[c] = dense_total;
            dense_total += (uint64_t)comps[c].nk * comps[c].nk;
        }
    S->dense.resize(dense_total); // uninitialised
    // the host-built blocks' pages first touched by all threads; the GPU-built blocks stay on the
    // device and their host pages untouched until the host needs them
    {
        std::vector<std::pair<uint64_t, uint64_t>> host_ranges;
        for (uint64_t c = 0; c < nc; ++c)
            if (dense[c] && !big[c]) host_ranges.push_back({S->dense_off[c], (uint64_t)comps[c].nk * comps[c].nk});
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t r = 0; r < (int64_t)host_ranges.size(); ++r)
            std::fill(S->dense.begin() + host_ranges[r].first,
                      S->dense.begin() + host_ranges[r].first + host_ranges[r].second, 0.0);
    }
    // per-component factorisation (parallel over components).  A dense block whose reduced
    // block costs many adjacency products (a small kept side next to a large eliminated one) is
    // built first, its rows in parallel; each row sums its terms in adjacency order, so the
    // result does not depend on the split.
    std::vector<std::vector<double>> dense_mats(nc);
    std::vector<AcFactor> acs(nc);
    std::vector<int> failed(nc, 0);
    std::vector<uint8_t> heavy(nc, 0);
    for (uint64_t c = 0; c < nc; ++c) {
        if (!dense[c] || big[c]) continue;
        double work = 0.0;
        for (uint32_t e = 0; e < comps[c].ne; ++e) {
            const double deg = (double)(comps[c].cptr[e + 1] - comps[c].cptr[e]);
            work += deg * deg;
        }
        heavy[c] = work > 4e6;
    }
    for (uint64_t c = 0; c < nc; ++c) {
        if (!heavy[c]) continue;
        const Component &cm = comps[c];
        for (uint32_t e = 0; e < cm.ne; ++e)
            if (!(cm.diag[e] > 0.0)) failed[c] = 1;
        if (failed[c]) continue;
        std::vector<double> Sm;
        dense_schur(cm, Sm, true);
        if (!dense_laplacian_pinv(Sm, cm.nk, S->gauge[c] == 0)) failed[c] = 2;
        dense_mats[c] = std::move(Sm);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < (int64_t)nc; ++c) {
        const Component &cm = comps[c];
        if (heavy[c]) continue;
        for (uint32_t e = 0; e < cm.ne; ++e)
            if (!(cm.diag[e] > 0.0)) failed[c] = 1;
        if (failed[c] || big[c]) continue;
        const bool floating = S->gauge[c] == 0;
        if (dense[c]) {
            std::vector<double> Sm;
            dense_schur(cm, Sm, false);
            if (!dense_laplacian_pinv(Sm, cm.nk, floating)) failed[c] = 2;
            dense_mats[c] = std::move(Sm);
        } else {
            // Laplacian-form system to factor: the component itself, or its Gremban cover
            const Component cover = cm.signed_form ? make_cover(cm) : Component{};
            const Component &lf = cm.signed_form ? cover : cm;
            std::vector<Edge> edges;
            const auto t0 = std::chrono::steady_clock::now();
            schur_edges(lf, cfg, (uint64_t)c, edges);
            const auto t1 = std::chrono::steady_clock::now();
            const uint32_t n_ac = lf.nk + (lf.grounded ? 1 : 0);
            acs[c] = approx_cholesky(n_ac, edges, mix_seed(cfg.approx_chol_seed, c, 0xC401),
                                     cfg.split_merge > 0 ? cfg.split_merge : 1, lf.grounded ? (int64_t)lf.nk : -1);
            static const bool verbose_factor = std::getenv("WITHIN_AMD_VERBOSE") != nullptr;
            if (verbose_factor && lf.ne + lf.nk > 10000)
                std::fprintf(stderr, "[within_amd] component %lld: %u eliminated, %u kept, %zu Schur edges: "
                             "sampling %.1f ms, approximate Cholesky %.1f ms\n", (long long)c, lf.ne, lf.nk,
                             edges.size(), std::chrono::duration<double, std::milli>(t1 - t0).count(),
                             std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count());
            if (lf.grounded) { // drop the ground vertex, eliminated last
                AcFactor &f = acs[c];
                f.order.resize(lf.nk);
                f.d.resize(lf.nk);
                f.rptr.resize(lf.nk + 1);
                f.ridx.resize(f.rptr[lf.nk]);
                f.rval.resize(f.rptr[lf.nk]);
            }
        }
    }
    phase("local factorisations");
    // large dense kept blocks: the reduced block and its pseudo-inverse formed on the GPU
    for (uint64_t c = 0; c < nc; ++c) {
        if (!big[c] || failed[c]) continue;
        const Component &cm = comps[c];
        const Adjacency adj(cm);
        std::vector<double> rowabs;
        if (cm.signed_form) {
            rowabs.assign(cm.ne, 0.0);
            for (uint32_t e = 0; e < cm.ne; ++e)
                for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) rowabs[e] += std::fabs(cm.cval[p]);
        }
        const DenseSchurArgs a{cm.ne,          cm.nk,          cm.cptr.data(),
                               cm.cidx.data(), cm.cval.data(), (uint64_t)cm.cidx.size(),
                               cm.diag.data(), cm.grounded ? cm.surplus.data() : nullptr,
                               cm.signed_form ? rowabs.data() : nullptr,
                               adj.tp.data(),  adj.te.data(),  adj.tv.data(),
                               cm.signed_form};
        std::shared_ptr<double> dev;
        if (!dense_schur_pinv_gpu(a, S->gauge[c] == 0, dev)) {
            failed[c] = 2;
            continue;
        }
        S->dense_dev.push_back({S->dense_off[c], (uint64_t)cm.nk * cm.nk, std::move(dev)});
        S->dense_on_host = false;
    }
    phase("large dense blocks");
    for (uint64_t c = 0; c < nc; ++c) {
        if (failed[c] == 1)
            fail(WX_ERR_BUILD, "zero diagonal in preconditioner at index " + std::to_string(c),
                 "BuildError::SingularDiagonal");
        if (failed[c] == 2) fail(WX_ERR_BUILD, "local solver build failed: Schur complement is not positive definite");
    }
    // assemble slots (kept block of sparse components relabelled into elimination order;
    // cover components keep their kept slots and carry the cover's elimination order)
    uint64_t l_total = 0, piv_total = 0, rows_total = 0, cover_total = 0;
    for (uint64_t c = 0; c < nc; ++c) {
        if (!dense[c]) {
            const uint64_t nf = acs[c].d.size();
            l_total += acs[c].ridx.size();
            piv_total += nf;
            rows_total += nf + 1;
            if (S->fac_kind[c] == 2) cover_total += nf;
        }
    }
    S->l_idx.resize(l_total);
    S->l_val.resize(l_total);
    S->piv.resize(piv_total);
    S->l_ptr.resize(rows_total);
    S->cover_perm.resize(cover_total);
    uint64_t cnnz = 0;
    for (auto &cm : comps) cnnz += cm.cidx.size();
    S->c_idx.resize(cnnz);
    S->c_val.resize(cnnz);
    // every component's place in the concatenated arrays, then the components one after another,
    // each filling its own ranges with its rows in parallel
    std::vector<uint64_t> c_at(nc + 1, 0), l_at(nc, 0), p_at(nc, 0), r_at(nc, 0), v_at(nc, 0);
    {
        uint64_t loff = 0, poff = 0, rptr_off = 0, coff = 0;
        for (uint64_t c = 0; c < nc; ++c) {
            c_at[c + 1] = c_at[c] + comps[c].cidx.size();
            if (dense[c]) continue;
            l_at[c] = loff;
            p_at[c] = poff;
            r_at[c] = rptr_off;
            v_at[c] = coff;
            const uint64_t nf = acs[c].d.size();
            loff += acs[c].ridx.size();
            rptr_off += nf + 1;
            poff += nf;
            if (S->fac_kind[c] == 2) coff += nf;
        }
    }
    for (uint64_t c = 0; c < nc; ++c) {
        const Component &cm = comps[c];
        const uint64_t base = S->slot_start[c];
        const uint32_t ne = cm.ne, nk = cm.nk;
        const bool par = cm.cidx.size() + ne + nk >= (1u << 16);
        (void)par;
        const bool relabelled = S->fac_kind[c] == 1;
        std::vector<uint32_t> relabel(nk); // old kept index -> new
        if (relabelled)
            for (uint32_t t = 0; t < nk; ++t) relabel[acs[c].order[t]] = t;
        else std::iota(relabel.begin(), relabel.end(), 0u);
#pragma omp parallel for schedule(static) if (par)
        for (int64_t e = 0; e < (int64_t)ne; ++e) {
            S->glob[base + e] = cm.glob[e];
            S->fold[base + e] = cm.fold[e];
            S->inv_de[base + e] = 1.0 / cm.diag[e];
        }
        for (uint32_t k = 0; k < nk; ++k) {
            S->glob[base + ne + relabel[k]] = cm.glob[ne + k];
            S->fold[base + ne + relabel[k]] = cm.fold[ne + k];
        }
#pragma omp parallel for schedule(static) if (par)
        for (int64_t s = 0; s < (int64_t)(ne + nk); ++s) {
            const uint32_t g = S->glob[base + s];
            S->pw[base + s] = (!any_slope && counts[g] > 1) ? 1.0 / std::sqrt((double)counts[g]) : 1.0;
        }
        // C rows, each sorted by its (relabelled) kept index; rows of a component come out of
        // the cross-tab kept-ascending, so only a relabelled block needs the sort
#pragma omp parallel for schedule(dynamic, 512) if (par)
        for (int64_t e = 0; e < (int64_t)ne; ++e) {
            const uint64_t at = c_at[c] + cm.cptr[e];
            S->c_ptr[base + e] = at;
            const uint64_t len = cm.cptr[e + 1] - cm.cptr[e];
            if (!relabelled) {
                std::copy(cm.cidx.begin() + cm.cptr[e], cm.cidx.begin() + cm.cptr[e + 1], S->c_idx.begin() + at);
                std::copy(cm.cval.begin() + cm.cptr[e], cm.cval.begin() + cm.cptr[e + 1], S->c_val.begin() + at);
                continue;
            }
            std::vector<std::pair<uint32_t, double>> row(len);
            for (uint64_t q = 0; q < len; ++q)
                row[q] = {relabel[cm.cidx[cm.cptr[e] + q]], cm.cval[cm.cptr[e] + q]};
            std::sort(row.begin(), row.end(), [](auto &a, auto &b) { return a.first < b.first; });
            for (uint64_t q = 0; q < len; ++q) {
                S->c_idx[at + q] = row[q].first;
                S->c_val[at + q] = row[q].second;
            }
        }
        for (uint32_t k = 0; k < nk; ++k) S->c_ptr[base + ne + k] = c_at[c + 1];
        if (dense[c]) { // the large blocks are on the device already
            if (!big[c]) {
                const uint64_t m = dense_mats[c].size();
#pragma omp parallel for schedule(static) if (m >= (1u << 20))
                for (int64_t q = 0; q < (int64_t)m; ++q) S->dense[S->dense_off[c] + q] = dense_mats[c][q];
            }
            std::vector<double>().swap(dense_mats[c]);
        } else {
            const AcFactor &f = acs[c];
            const uint64_t nf = f.d.size();
            S->l_row_off[c] = r_at[c];
            S->piv_off[c] = p_at[c];
            for (uint64_t q = 0; q <= nf; ++q) S->l_ptr[r_at[c] + q] = l_at[c] + f.rptr[q];
            std::copy(f.ridx.begin(), f.ridx.end(), S->l_idx.begin() + l_at[c]);
            std::copy(f.rval.begin(), f.rval.end(), S->l_val.begin() + l_at[c]);
            std::copy(f.d.begin(), f.d.end(), S->piv.begin() + p_at[c]);
            if (S->fac_kind[c] == 2) {
                S->cover_off[c] = v_at[c];
                std::copy(f.order.begin(), f.order.end(), S->cover_perm.begin() + v_at[c]);
            }
        }
    }
    S->c_ptr[ns] = c_at[nc];
    pre->schwarz = S;
    phase("assembly");
    return pre;
}

// One general bipartite block [[diag_r, B], [B', diag_c]] (B: n_rows x n_cols row-major, any
// signs) as an oriented, SDDM-converted component with global DOFs 0..n_rows + n_cols - 1, as
// build_pair_components forms a slope pair's component (sddm.rs:86-106, 199-463).
Component component_from_block(uint32_t nr, uint32_t nc, const double *B, const double *diag,
                               const LocalSolverSettings &cfg, std::vector<std::string> &warnings) {
    Component comp;
    const bool swap = nc > nr;
    const uint32_t ne = swap ? nc : nr, nk = swap ? nr : nc;
    comp.ne = ne;
    comp.nk = nk;
    for (uint32_t e = 0; e < ne; ++e) comp.glob.push_back(swap ? nr + e : e);
    for (uint32_t k = 0; k < nk; ++k) comp.glob.push_back(swap ? k : nr + k);
    comp.cptr.push_back(0);
    for (uint32_t e = 0; e < ne; ++e) {
        for (uint32_t k = 0; k < nk; ++k) {
            const double v = swap ? B[(uint64_t)k * nc + e] : B[(uint64_t)e * nc + k];
            if (v != 0.0) {
                comp.cidx.push_back(k);
                comp.cval.push_back(v);
            }
        }
        comp.cptr.push_back(comp.cidx.size());
    }
    for (uint32_t i = 0; i < ne + nk; ++i) comp.diag.push_back(diag[comp.glob[i]]);
    int64_t sweeps = 0;
    double violation = 0.0;
    if (!convert_general(comp, cfg, sweeps, violation))
        fail(WX_ERR_BUILD, "signed component is not diagonally scalable to SDDM form", "BuildError::UnscalableComponent");
    if (violation > cfg.scaling_tolerance) warnings.push_back("dominance scaling uncertified");
    return comp;
}

} // namespace

std::shared_ptr<Precond> build_block_precond(uint32_t n_rows, uint32_t n_cols, const double *B, const double *diag,
                                             const LocalSolverSettings &cfg) {
    std::vector<std::string> warnings;
    std::vector<Component> comps;
    comps.push_back(component_from_block(n_rows, n_cols, B, diag, cfg, warnings));
    return assemble_additive(comps, (uint64_t)n_rows + n_cols, true, cfg, [](const char *) {});
}

std::vector<double> block_reduced_laplacian(uint32_t n_rows, uint32_t n_cols, const double *B, const double *diag,
                                            const LocalSolverSettings &cfg, uint32_t &n_out) {
    std::vector<std::string> warnings;
    Component cm = component_from_block(n_rows, n_cols, B, diag, cfg, warnings);
    const Component cover = cm.signed_form ? make_cover(cm) : Component{};
    const Component &lf = cm.signed_form ? cover : cm;
    std::vector<Edge> edges;
    schur_edges(lf, cfg, 0, edges);
    const uint32_t n = lf.nk + (lf.grounded ? 1 : 0);
    n_out = n;
    std::vector<double> L((size_t)n * n, 0.0);
    for (const Edge &e : edges) {
        L[(size_t)e.lo * n + e.hi] -= e.w;
        L[(size_t)e.hi * n + e.lo] -= e.w;
        L[(size_t)e.lo * n + e.lo] += e.w;
        L[(size_t)e.hi * n + e.hi] += e.w;
    }
    return L;
}

} // namespace wx
