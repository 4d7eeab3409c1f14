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
#include <cmath>
#include <cstring>
#include <numeric>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "schwarz.hpp"

namespace wx {

// dense_gpu.hip
bool gpu_available();
bool dense_laplacian_pinv_gpu(std::vector<double> &S, uint32_t n);

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
        if (!std::isfinite(inv)) fail(WX_ERR_BUILD, "zero diagonal in preconditioner at index " + std::to_string(j));
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
    std::vector<double> suffix(d + 1, 0.0);
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
    std::sort(e.begin(), e.end(), [](const Edge &a, const Edge &b) {
        if (a.lo != b.lo) return a.lo < b.lo;
        if (a.hi != b.hi) return a.hi < b.hi;
        return a.w < b.w;
    });
    size_t wr = 0;
    for (size_t r = 0; r < e.size(); ++r) {
        if (wr > 0 && e[wr - 1].lo == e[r].lo && e[wr - 1].hi == e[r].hi) e[wr - 1].w += e[r].w;
        else e[wr++] = e[r];
    }
    e.resize(wr);
}

// Approximate Cholesky of a connected Laplacian given by its off-diagonal edges, as
// L D L' with L unit-lower in elimination order.  Vertices are eliminated in a seeded random
// order; each star's clique is replaced by `split` sampled clique trees.  The last vertex
// has an empty star and pivot 0 (the floating gauge, pseudo-inverted to 0).
struct AcFactor {
    std::vector<uint32_t> order;    // order[t] = vertex eliminated t-th
    std::vector<uint64_t> rptr;     // L by rows (relabelled), strictly lower
    std::vector<uint32_t> ridx;
    std::vector<double> rval;
    std::vector<double> d;          // pivots (relabelled)
};

AcFactor approx_cholesky(uint32_t n, const std::vector<Edge> &edges, uint64_t seed, int split) {
    AcFactor f;
    std::vector<std::vector<std::pair<uint32_t, double>>> adj(n);
    for (const Edge &e : edges) {
        adj[e.lo].push_back({e.hi, e.w});
        adj[e.hi].push_back({e.lo, e.w});
    }
    f.order.resize(n);
    std::iota(f.order.begin(), f.order.end(), 0u);
    Rng ord(mix_seed(seed, n, 0xAC0));
    for (uint32_t i = n; i > 1; --i) std::swap(f.order[i - 1], f.order[ord.next() % i]);
    std::vector<uint32_t> pos(n);
    for (uint32_t t = 0; t < n; ++t) pos[f.order[t]] = t;
    std::vector<uint8_t> gone(n, 0);
    // columns of L in elimination order: (row position, value)
    std::vector<std::vector<std::pair<uint32_t, double>>> cols(n);
    f.d.assign(n, 0.0);
    std::vector<std::pair<uint32_t, double>> nb;
    std::vector<Edge> fill;
    for (uint32_t t = 0; t < n; ++t) {
        const uint32_t v = f.order[t];
        nb.clear();
        for (auto &e : adj[v])
            if (!gone[e.first]) nb.push_back(e);
        std::vector<std::pair<uint32_t, double>>().swap(adj[v]);
        gone[v] = 1;
        std::sort(nb.begin(), nb.end());
        size_t wr = 0;
        for (size_t r = 0; r < nb.size(); ++r) {
            if (wr > 0 && nb[wr - 1].first == nb[r].first) nb[wr - 1].second += nb[r].second;
            else nb[wr++] = nb[r];
        }
        nb.resize(wr);
        double W = 0.0;
        for (auto &e : nb) W += e.second;
        f.d[t] = W;
        if (W <= 0.0) continue;
        for (auto &e : nb) cols[t].push_back({pos[e.first], -e.second / W});
        fill.clear();
        Rng rng(mix_seed(seed, v, 0x57A));
        sample_clique_tree(nb, W, split, rng, fill);
        for (const Edge &e : fill) {
            adj[e.lo].push_back({e.hi, e.w});
            adj[e.hi].push_back({e.lo, e.w});
        }
    }
    // transpose columns into rows (row = later vertex, col = earlier vertex)
    f.rptr.assign(n + 1, 0);
    for (uint32_t t = 0; t < n; ++t)
        for (auto &e : cols[t]) f.rptr[e.first + 1]++;
    for (uint32_t i = 0; i < n; ++i) f.rptr[i + 1] += f.rptr[i];
    f.ridx.resize(f.rptr[n]);
    f.rval.resize(f.rptr[n]);
    std::vector<uint64_t> cur(f.rptr.begin(), f.rptr.end() - 1);
    for (uint32_t t = 0; t < n; ++t)
        for (auto &e : cols[t]) {
            f.ridx[cur[e.first]] = t;
            f.rval[cur[e.first]] = e.second;
            cur[e.first]++;
        }
    return f;
}

// Dense pseudo-inverse of a connected Laplacian S (n x n row-major, in place):
// ground the last vertex, invert the SPD minor through Cholesky, pad, and project with
// P = I - 11'/n on both sides, so the result is the Moore-Penrose inverse S^+.
bool dense_laplacian_pinv(std::vector<double> &S, uint32_t n) {
    if (n <= 1) {
        std::fill(S.begin(), S.end(), 0.0);
        return true;
    }
    const uint32_t m = n - 1;
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

struct Component {
    // eliminated-major after orientation
    std::vector<uint32_t> glob;  // ne + nk
    std::vector<double> diag;    // ne + nk (row sums adopted)
    uint32_t ne = 0, nk = 0;
    // C: ne x nk CSR (local kept index)
    std::vector<uint64_t> cptr;
    std::vector<uint32_t> cidx;
    std::vector<double> cval;
};

// Cross-tab of one channel pair and its components (cross_tab.rs, accumulate.rs, factor_pairs.rs:95-152).
void build_pair_components(const Design &d, const std::vector<double> &w, Channel rc, Channel cc,
                           const std::vector<std::vector<uint8_t>> &active, std::vector<Component> &out,
                           bool &general) {
    const Term &tr = d.terms[rc.term], &tc = d.terms[cc.term];
    const int64_t lr = tr.loads[rc.column], lc = tc.loads[cc.column];
    general = lr >= 0 || lc >= 0;
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
    // accumulate: bucket by row (observation order), dedup per row with sorted columns
    const uint32_t *rlev = d.level_column(rc.term), *clev = d.level_column(cc.term);
    const double *zr = lr >= 0 ? d.slope_column((uint64_t)lr) : nullptr;
    const double *zc = lc >= 0 ? d.slope_column((uint64_t)lc) : nullptr;
    std::vector<double> rdiag(nr, 0.0), cdiag(nc, 0.0);
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
    std::vector<uint64_t> cptr(nr + 1, 0);
    std::vector<uint32_t> cidx;
    std::vector<double> cval;
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
        if (!general) {
            const double tol = 64.0 * 2.220446049250313e-16 * std::sqrt((double)std::max<size_t>(1, E.size() + K.size())) * total_diag;
            if (mismatch > tol) fail(WX_ERR_BUILD, "factor-pair component failed Laplacian validation");
            comp.diag = rowsum_e;
            comp.diag.insert(comp.diag.end(), rowsum_k.begin(), rowsum_k.end());
        } else {
            for (size_t ei = 0; ei < E.size(); ++ei) comp.diag.push_back(swap ? cdiag[E[ei]] : rdiag[E[ei]]);
            for (size_t ki = 0; ki < K.size(); ++ki) comp.diag.push_back(swap ? rdiag[K[ki]] : cdiag[K[ki]]);
        }
        for (uint32_t k : K) (swap ? rremap[k] : cremap[k]) = UINT32_MAX;
        out.push_back(std::move(comp));
    }
}

} // namespace

std::shared_ptr<Precond> build_additive(const Design &d, const std::vector<double> &w, const LocalSolverSettings &cfg,
                                        std::vector<std::string> &warnings) {
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
    std::vector<std::vector<Component>> per_pair(pairs.size());
    std::vector<uint8_t> pair_general(pairs.size(), 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t p = 0; p < (int64_t)pairs.size(); ++p) {
        bool general = false;
        build_pair_components(d, w, pairs[p].first, pairs[p].second, active, per_pair[p], general);
        pair_general[p] = general;
    }
    std::vector<Component> comps;
    for (auto &v : per_pair)
        for (auto &c : v) comps.push_back(std::move(c));
    for (size_t p = 0; p < pairs.size(); ++p)
        if (pair_general[p] && !per_pair[p].empty())
            fail(WX_ERR_BUILD, "additive Schwarz over slope (signed) channel pairs is not supported by this build; "
                               "use PreconditionerConfig.Diagonal for varying-slope designs");
    if (comps.empty()) return nullptr;

    auto pre = std::make_shared<Precond>();
    pre->kind = Precond::ADDITIVE;
    pre->n = d.n_dofs;
    auto S = std::make_shared<Precond::Schwarz>();
    S->n_dofs = d.n_dofs;
    // partition of unity 1/sqrt(count), only without slope channels (factor_pairs.rs:84-89,155-194)
    std::vector<uint32_t> counts(d.n_dofs, 0);
    for (auto &c : comps)
        for (uint32_t g : c.glob) counts[g]++;
    const uint64_t nc = comps.size();
    S->slot_start.assign(nc + 1, 0);
    S->ne.resize(nc);
    S->nk.resize(nc);
    S->fac_kind.resize(nc);
    S->dense_off.assign(nc, 0);
    S->l_row_off.assign(nc, 0);
    S->piv_off.assign(nc, 0);
    for (uint64_t c = 0; c < nc; ++c) {
        S->ne[c] = comps[c].ne;
        S->nk[c] = comps[c].nk;
        S->slot_start[c + 1] = S->slot_start[c] + comps[c].ne + comps[c].nk;
    }
    const uint64_t ns = S->slot_start[nc];
    S->glob.resize(ns);
    S->pw.resize(ns);
    S->inv_de.assign(ns, 0.0);
    S->c_ptr.assign(ns + 1, 0);
    // factor choice per component.  Device cost model (dense_threshold < 0): small kept blocks
    // and dense-ish reduced graphs get the exact dense pseudo-inverse -- one parallel
    // symmetric product per apply -- while large sparse reduced graphs get approximate
    // Cholesky, whose triangular solves run level by level.  The density estimate is the
    // sampled Schur complement's edge count sum_e (deg(e) - 1) over nk^2 / 2.
    const bool have_gpu = gpu_available();
    std::vector<uint8_t> dense(nc), big(nc, 0);
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t nk = comps[c].nk;
        if (cfg.dense_threshold >= 0) dense[c] = nk <= (uint64_t)cfg.dense_threshold;
        else if (nk <= 4096) dense[c] = 1;
        else if (nk <= 16384 && have_gpu) {
            double m = 0.0;
            for (uint32_t e = 0; e < comps[c].ne; ++e) {
                const uint64_t deg = comps[c].cptr[e + 1] - comps[c].cptr[e];
                if (deg > 1) m += (double)(deg - 1);
            }
            dense[c] = 2.0 * m / ((double)nk * nk) >= 0.005;
        } else dense[c] = 0;
        big[c] = dense[c] && nk > 1024 && have_gpu;
        S->fac_kind[c] = dense[c] ? 0 : 1;
    }
    // per-component factorisation (parallel)
    std::vector<std::vector<double>> dense_mats(nc);
    std::vector<AcFactor> acs(nc);
    std::vector<int> failed(nc, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < (int64_t)nc; ++c) {
        const Component &cm = comps[c];
        const uint32_t ne = cm.ne, nk = cm.nk;
        for (uint32_t e = 0; e < ne; ++e)
            if (!(cm.diag[e] > 0.0)) failed[c] = 1;
        if (failed[c] || big[c]) continue;
        if (dense[c]) {
            // exact Schur complement S = D_k - C' D_e^{-1} C (schur.rs:29-61)
            std::vector<double> Sm((size_t)nk * nk, 0.0);
            for (uint32_t k = 0; k < nk; ++k) Sm[(size_t)k * nk + k] = cm.diag[ne + k];
            for (uint32_t e = 0; e < ne; ++e) {
                const double inv = 1.0 / cm.diag[e];
                for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) {
                    const uint32_t i = cm.cidx[p];
                    const double wi = cm.cval[p];
                    for (uint64_t q = cm.cptr[e]; q < cm.cptr[e + 1]; ++q)
                        Sm[(size_t)i * nk + cm.cidx[q]] -= (wi * cm.cval[q]) * inv;
                }
            }
            if (!dense_laplacian_pinv(Sm, nk)) failed[c] = 2;
            dense_mats[c] = std::move(Sm);
        } else {
            // Schur complement edges: sampled clique trees per eliminated star, or exact
            std::vector<Edge> edges;
            for (uint32_t e = 0; e < ne; ++e) {
                const uint64_t b = cm.cptr[e], en = cm.cptr[e + 1];
                if (en - b < 2) continue;
                const double De = cm.diag[e];
                if (cfg.schur_exact) {
                    for (uint64_t p = b; p < en; ++p)
                        for (uint64_t q = p + 1; q < en; ++q)
                            edges.push_back({std::min(cm.cidx[p], cm.cidx[q]), std::max(cm.cidx[p], cm.cidx[q]),
                                             (cm.cval[p] * cm.cval[q]) / De});
                } else {
                    std::vector<std::pair<uint32_t, double>> nb;
                    for (uint64_t p = b; p < en; ++p) nb.push_back({cm.cidx[p], cm.cval[p]});
                    Rng rng(mix_seed(cfg.schur_seed, e + ((uint64_t)c << 32), 0x5C4));
                    sample_clique_tree(nb, De, std::max(1, cfg.schur_split), rng, edges);
                }
            }
            sort_and_merge(edges);
            acs[c] = approx_cholesky(nk, edges, mix_seed(cfg.approx_chol_seed, c, 0xC401),
                                     cfg.split_merge > 0 ? cfg.split_merge : 1);
        }
    }
    // large dense kept blocks: exact Schur complement accumulated row-parallel on the host
    // (each thread owns kept rows, entries ordered (w*v)*inv as in schur.rs:112-113), then the
    // pseudo-inverse on the GPU
    for (uint64_t c = 0; c < nc; ++c) {
        if (!big[c] || failed[c]) continue;
        const Component &cm = comps[c];
        const uint32_t ne = cm.ne, nk = cm.nk;
        std::vector<uint64_t> tp(nk + 1, 0);
        for (uint32_t k : cm.cidx) tp[k + 1]++;
        for (uint32_t k = 0; k < nk; ++k) tp[k + 1] += tp[k];
        std::vector<uint32_t> te(cm.cidx.size());
        std::vector<double> tv(cm.cidx.size());
        {
            std::vector<uint64_t> cur(tp.begin(), tp.end() - 1);
            for (uint32_t e = 0; e < ne; ++e)
                for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) {
                    te[cur[cm.cidx[p]]] = e;
                    tv[cur[cm.cidx[p]]] = cm.cval[p];
                    cur[cm.cidx[p]]++;
                }
        }
        std::vector<double> Sm((size_t)nk * nk, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
        for (int64_t i = 0; i < (int64_t)nk; ++i) {
            double *row = &Sm[(size_t)i * nk];
            row[i] = cm.diag[ne + i];
            for (uint64_t q = tp[i]; q < tp[i + 1]; ++q) {
                const uint32_t e = te[q];
                const double wi = tv[q], inv = 1.0 / cm.diag[e];
                for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) row[cm.cidx[p]] -= (wi * cm.cval[p]) * inv;
            }
        }
        if (!dense_laplacian_pinv_gpu(Sm, nk)) failed[c] = 2;
        dense_mats[c] = std::move(Sm);
    }
    for (uint64_t c = 0; c < nc; ++c) {
        if (failed[c] == 1) fail(WX_ERR_BUILD, "zero diagonal in preconditioner at index " + std::to_string(c));
        if (failed[c] == 2) fail(WX_ERR_BUILD, "local solver build failed: Schur complement is not positive definite");
    }
    // assemble slots (kept block of sparse components relabelled into elimination order)
    uint64_t dense_total = 0, l_total = 0, piv_total = 0;
    for (uint64_t c = 0; c < nc; ++c) {
        if (dense[c]) dense_total += (uint64_t)comps[c].nk * comps[c].nk;
        else {
            l_total += acs[c].ridx.size();
            piv_total += comps[c].nk;
        }
    }
    S->dense.resize(dense_total);
    S->l_idx.resize(l_total);
    S->l_val.resize(l_total);
    S->piv.resize(piv_total);
    uint64_t doff = 0, loff = 0, poff = 0, rows_total = 0;
    for (uint64_t c = 0; c < nc; ++c)
        if (!dense[c]) rows_total += comps[c].nk + 1;
    S->l_ptr.resize(rows_total);
    uint64_t rptr_off = 0;
    uint64_t cnnz = 0;
    for (auto &cm : comps) cnnz += cm.cidx.size();
    S->c_idx.resize(cnnz);
    S->c_val.resize(cnnz);
    uint64_t cpos = 0;
    for (uint64_t c = 0; c < nc; ++c) {
        const Component &cm = comps[c];
        const uint64_t base = S->slot_start[c];
        const uint32_t ne = cm.ne, nk = cm.nk;
        std::vector<uint32_t> relabel(nk); // old kept index -> new
        if (dense[c]) std::iota(relabel.begin(), relabel.end(), 0u);
        else
            for (uint32_t t = 0; t < nk; ++t) relabel[acs[c].order[t]] = t;
        for (uint32_t e = 0; e < ne; ++e) {
            S->glob[base + e] = cm.glob[e];
            S->inv_de[base + e] = 1.0 / cm.diag[e];
        }
        for (uint32_t k = 0; k < nk; ++k) S->glob[base + ne + relabel[k]] = cm.glob[ne + k];
        for (uint32_t s = 0; s < ne + nk; ++s) {
            const uint32_t g = S->glob[base + s];
            S->pw[base + s] = (!any_slope && counts[g] > 1) ? 1.0 / std::sqrt((double)counts[g]) : 1.0;
        }
        for (uint32_t e = 0; e < ne; ++e) {
            S->c_ptr[base + e] = cpos;
            for (uint64_t p = cm.cptr[e]; p < cm.cptr[e + 1]; ++p) {
                S->c_idx[cpos] = relabel[cm.cidx[p]];
                S->c_val[cpos] = cm.cval[p];
                ++cpos;
            }
            // keep each C row sorted by (relabelled) kept index
            std::vector<std::pair<uint32_t, double>> row;
            for (uint64_t p = S->c_ptr[base + e]; p < cpos; ++p) row.push_back({S->c_idx[p], S->c_val[p]});
            std::sort(row.begin(), row.end(), [](auto &a, auto &b) { return a.first < b.first; });
            for (size_t q = 0; q < row.size(); ++q) {
                S->c_idx[S->c_ptr[base + e] + q] = row[q].first;
                S->c_val[S->c_ptr[base + e] + q] = row[q].second;
            }
        }
        for (uint32_t k = 0; k < nk; ++k) S->c_ptr[base + ne + k] = cpos;
        if (dense[c]) {
            S->dense_off[c] = doff;
            std::copy(dense_mats[c].begin(), dense_mats[c].end(), S->dense.begin() + doff);
            doff += (uint64_t)nk * nk;
            std::vector<double>().swap(dense_mats[c]);
        } else {
            const AcFactor &f = acs[c];
            S->l_row_off[c] = rptr_off;
            S->piv_off[c] = poff;
            for (uint32_t i = 0; i <= nk; ++i) S->l_ptr[rptr_off + i] = loff + f.rptr[i];
            std::copy(f.ridx.begin(), f.ridx.end(), S->l_idx.begin() + loff);
            std::copy(f.rval.begin(), f.rval.end(), S->l_val.begin() + loff);
            std::copy(f.d.begin(), f.d.end(), S->piv.begin() + poff);
            loff += f.ridx.size();
            rptr_off += nk + 1;
            poff += nk;
        }
    }
    S->c_ptr[ns] = cpos;
    pre->schwarz = S;
    (void)warnings;
    return pre;
}

} // namespace wx
