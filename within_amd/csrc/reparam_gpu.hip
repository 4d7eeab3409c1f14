// reparam_gpu.hip -- slope reparametrisation on the device
// (crates/within/src/solver/reparam.rs:78-276).
//
// Per level of a term with slope columns, the weighted moments of its loadings give a V x V Gram
// matrix; a pivoted Gram-Schmidt in that Gram's coordinates yields the whitening rows W (W G W' = I)
// and the kept slopes, and every loading is rewritten as W (z - mean).  On the device the term's
// rows are sorted by level (stable radix sort), one block per level forms the moments as fixed-order
// reductions, its first thread runs the Gram-Schmidt, and one thread per row rewrites the loadings.
//
// The moments are two passes shifted by the level's first positive-weight loading z0:
//   wsum = sum w,  mean = z0 + sum w (z - z0) / wsum,  C = sum w (z - mean)(z - mean)'
// where the reference runs Welford's one-pass update per row (reparam.rs:221-252).  The two are
// equal in exact arithmetic and agree to rounding; the shift keeps a loading that is constant within
// a level exactly constant (z - z0 = 0), so such a slope is dropped exactly as by the one-pass update.
// Rows of zero weight are skipped, as there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "solver.hpp"

namespace wx {

namespace {
constexpr int RB = 256;                // threads per level block
constexpr double RANK_TOL_DEV = 1e-10; // reparam.rs:13
constexpr uint32_t MAX_V = 3;          // slopes per term on the device path (MAX_COLS - 1 for the intercept)

template <int V> struct Cols {
    const double *z[V];
};
template <int V> struct OutCols {
    double *u[V];
};

__global__ __launch_bounds__(RB) void rp_fill(uint64_t n, uint32_t *__restrict__ a, uint32_t value) {
    const uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x;
    if (i < n) a[i] = value;
}

// [start, end) of every level in the level-sorted order (empty levels keep start = end = 0)
__global__ __launch_bounds__(RB) void rp_bounds(uint64_t n, const uint32_t *__restrict__ slev,
                                                uint32_t *__restrict__ start, uint32_t *__restrict__ end) {
    const uint64_t p = (uint64_t)blockIdx.x * RB + threadIdx.x;
    if (p >= n) return;
    const uint32_t l = slev[p];
    if (p == 0 || slev[p - 1] != l) start[l] = (uint32_t)p;
    if (p + 1 == n || slev[p + 1] != l) end[l] = (uint32_t)(p + 1);
}

// Sum of M values over the block in a fixed order: a butterfly inside each wavefront, then the
// wavefronts' sums added in order.  Every thread gets the totals.
template <int M> __device__ void block_sum(double (&a)[M], double *lds) {
    for (int j = 0; j < M; ++j)
        for (int off = 32; off > 0; off >>= 1) a[j] += __shfl_xor(a[j], off, 64);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0)
        for (int j = 0; j < M; ++j) lds[wv * M + j] = a[j];
    __syncthreads();
    for (int j = 0; j < M; ++j) {
        double s = lds[j];
        for (int q = 1; q < RB / 64; ++q) s += lds[q * M + j];
        a[j] = s;
    }
}

// Pivoted Gram-Schmidt in the coordinates of g (V x V): rows q of `basis` with q g q' = I, kept
// slopes in `kept` (reparam.rs:284-316).  The pivot is the largest residual variance, ties going to
// the last index; a slope whose residual is not above RANK_TOL times its variance is dropped.
template <int V>
__device__ int gram_schmidt(const double (&g)[V][V], double (&basis)[V][V], bool (&kept)[V]) {
#pragma clang fp contract(off)
    double residual[V];
    for (int j = 0; j < V; ++j) {
        residual[j] = g[j][j];
        kept[j] = false;
        for (int k = 0; k < V; ++k) basis[j][k] = 0.0;
    }
    int rank = 0;
    for (int it = 0; it < V; ++it) {
        int p = -1;
        for (int j = 0; j < V; ++j) {
            if (kept[j] || !isfinite(residual[j]) || !(residual[j] > RANK_TOL_DEV * g[j][j])) continue;
            if (p < 0 || residual[j] >= residual[p]) p = j;
        }
        if (p < 0) break;
        kept[p] = true;
        double q[V];
        for (int j = 0; j < V; ++j) q[j] = 0.0;
        q[p] = 1.0;
        for (int t = 0; t < rank; ++t) {
            double c = 0.0;
            for (int j = 0; j < V; ++j) c += g[p][j] * basis[t][j];
            for (int j = 0; j < V; ++j) q[j] -= c * basis[t][j];
        }
        const double norm = sqrt(residual[p]);
        for (int j = 0; j < V; ++j) q[j] /= norm;
        for (int j = 0; j < V; ++j) {
            double c = 0.0;
            for (int k = 0; k < V; ++k) c += g[j][k] * q[k];
            residual[j] -= c * c;
        }
        for (int j = 0; j < V; ++j) basis[rank][j] = q[j];
        ++rank;
    }
    return rank;
}

// One block per level: moments, Gram, Gram-Schmidt.  transform: L x V x V (rows = whitened
// coordinates), mean: L x V (zero without an intercept), info: per level the rank in bits 0-3,
// the kept slopes in bits 4-6 and bit 7 set when the level's weight sum is zero.
template <int V>
__global__ __launch_bounds__(RB) void rp_moments(const uint32_t *__restrict__ start, const uint32_t *__restrict__ end,
                                                 const uint32_t *__restrict__ perm, const double *__restrict__ w,
                                                 Cols<V> z, int intercept, double *__restrict__ transform,
                                                 double *__restrict__ mean, uint8_t *__restrict__ info) {
#pragma clang fp contract(off)
    constexpr int TL = V * (V + 1) / 2;
    __shared__ double lds[(RB / 64) * (TL > V + 1 ? TL : V + 1)];
    __shared__ uint32_t first;
    const uint32_t l = blockIdx.x;
    const uint32_t b = start[l], e = end[l];
    if (threadIdx.x == 0) first = e;
    __syncthreads();
    // the level's first positive-weight row (in observation order)
    for (uint32_t p = b + threadIdx.x; p < e; p += RB) {
        const double wi = w ? w[perm[p]] : 1.0;
        if (wi > 0.0) {
            atomicMin(&first, p);
            break;
        }
    }
    __syncthreads();
    const uint32_t f = first;
    double z0[V];
    for (int j = 0; j < V; ++j) z0[j] = f < e ? z.z[j][perm[f]] : 0.0;
    // pass 1: weight sum and shifted first moments
    double s1[V + 1];
    for (int j = 0; j <= V; ++j) s1[j] = 0.0;
    for (uint32_t p = f + threadIdx.x; p < e; p += RB) {
        const uint32_t i = perm[p];
        const double wi = w ? w[i] : 1.0;
        if (!(wi > 0.0)) continue;
        s1[0] += wi;
        for (int j = 0; j < V; ++j) s1[j + 1] += wi * (z.z[j][i] - z0[j]);
    }
    block_sum<V + 1>(s1, lds);
    const double wsum = s1[0];
    double m[V];
    for (int j = 0; j < V; ++j) m[j] = wsum > 0.0 ? z0[j] + s1[j + 1] / wsum : 0.0;
    // pass 2: centred second moments (lower triangle)
    double s2[TL];
    for (int q = 0; q < TL; ++q) s2[q] = 0.0;
    for (uint32_t p = f + threadIdx.x; p < e; p += RB) {
        const uint32_t i = perm[p];
        const double wi = w ? w[i] : 1.0;
        if (!(wi > 0.0)) continue;
        double dz[V];
        for (int j = 0; j < V; ++j) dz[j] = z.z[j][i] - m[j];
        for (int j = 0; j < V; ++j) {
            const double dev = wi * dz[j];
            for (int k = 0; k <= j; ++k) s2[j * (j + 1) / 2 + k] += dz[k] * dev;
        }
    }
    block_sum<TL>(s2, lds);
    if (threadIdx.x != 0) return;
    double g[V][V];
    for (int j = 0; j < V; ++j)
        for (int k = 0; k <= j; ++k) {
            double gv = s2[j * (j + 1) / 2 + k];
            if (!intercept) gv += wsum * m[j] * m[k];
            g[j][k] = gv;
            g[k][j] = gv;
        }
    double basis[V][V];
    bool kept[V];
    const int rank = gram_schmidt<V>(g, basis, kept);
    for (int r = 0; r < V; ++r)
        for (int j = 0; j < V; ++j) transform[((uint64_t)l * V + r) * V + j] = basis[r][j];
    for (int j = 0; j < V; ++j) mean[(uint64_t)l * V + j] = intercept ? m[j] : 0.0;
    uint32_t bits = (uint32_t)rank;
    for (int j = 0; j < V; ++j) bits |= kept[j] ? (16u << j) : 0u;
    if (wsum == 0.0) bits |= 128u;
    info[l] = (uint8_t)bits;
}

// u = W (z - mean) per row, rows of W beyond the level's rank giving zero (reparam.rs:151-163)
template <int V>
__global__ __launch_bounds__(RB) void rp_rewrite(uint64_t n, const uint32_t *__restrict__ lev, Cols<V> z,
                                                 const double *__restrict__ transform, const double *__restrict__ mean,
                                                 const uint8_t *__restrict__ info, OutCols<V> u) {
#pragma clang fp contract(off)
    const uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = lev[i];
    const uint32_t rank = info[l] & 15u;
    double d[V];
    for (int j = 0; j < V; ++j) d[j] = z.z[j][i] - mean[(uint64_t)l * V + j];
    for (int r = 0; r < V; ++r) {
        double acc = 0.0;
        if ((uint32_t)r < rank)
            for (int j = 0; j < V; ++j) acc += transform[((uint64_t)l * V + r) * V + j] * d[j];
        u.u[r][i] = acc;
    }
}

template <int V>
void run_term(const Design &d, uint64_t term, const std::vector<uint64_t> &zcols, bool intercept, DevBuf &lev,
              DevBuf &wdev, std::vector<double> &transform, std::vector<double> &mean, std::vector<uint8_t> &info,
              std::vector<std::vector<double>> &ucols) {
    const uint64_t n = d.n_obs, L = d.terms[term].n_levels;
    hipStream_t st = nullptr;
    const uint32_t gn = (uint32_t)((n + RB - 1) / RB);
    DevBuf zb, ub, perm, slev, start, end, tr, mn, inf;
    zb.alloc(n * V * 8);
    for (int j = 0; j < V; ++j)
        WX_HIP(hipMemcpyAsync(zb.as<double>() + j * n, d.slope_column(zcols[j]), n * 8, hipMemcpyHostToDevice, st));
    perm.alloc(n * 4);
    slev.alloc(n * 4);
    dev_stable_argsort(lev.as<uint32_t>(), n, L, perm.as<uint32_t>(), st);
    dev_gather_u32(lev.as<uint32_t>(), perm.as<uint32_t>(), n, slev.as<uint32_t>(), st);
    start.alloc(L * 4);
    end.alloc(L * 4);
    const uint32_t gl = (uint32_t)((L + RB - 1) / RB);
    hipLaunchKernelGGL(rp_fill, dim3(gl), dim3(RB), 0, st, L, start.as<uint32_t>(), 0u);
    hipLaunchKernelGGL(rp_fill, dim3(gl), dim3(RB), 0, st, L, end.as<uint32_t>(), 0u);
    hipLaunchKernelGGL(rp_bounds, dim3(gn), dim3(RB), 0, st, n, slev.as<uint32_t>(), start.as<uint32_t>(),
                       end.as<uint32_t>());
    tr.alloc(L * V * V * 8);
    mn.alloc(L * V * 8);
    inf.alloc(L);
    Cols<V> zc;
    for (int j = 0; j < V; ++j) zc.z[j] = zb.as<double>() + j * n;
    hipLaunchKernelGGL((rp_moments<V>), dim3((uint32_t)L), dim3(RB), 0, st, start.as<uint32_t>(),
                       end.as<uint32_t>(), perm.as<uint32_t>(), wdev.p ? wdev.as<double>() : nullptr, zc,
                       intercept ? 1 : 0, tr.as<double>(), mn.as<double>(), inf.as<uint8_t>());
    ub.alloc(n * V * 8);
    OutCols<V> uc;
    for (int j = 0; j < V; ++j) uc.u[j] = ub.as<double>() + j * n;
    hipLaunchKernelGGL((rp_rewrite<V>), dim3(gn), dim3(RB), 0, st, n, lev.as<uint32_t>(), zc, tr.as<double>(),
                       mn.as<double>(), inf.as<uint8_t>(), uc);
    WX_HIP(hipGetLastError());
    transform.resize(L * V * V);
    mean.resize(L * V);
    info.resize(L);
    WX_HIP(hipMemcpyAsync(transform.data(), tr.p, L * V * V * 8, hipMemcpyDeviceToHost, st));
    WX_HIP(hipMemcpyAsync(mean.data(), mn.p, L * V * 8, hipMemcpyDeviceToHost, st));
    WX_HIP(hipMemcpyAsync(info.data(), inf.p, L, hipMemcpyDeviceToHost, st));
    ucols.assign(V, std::vector<double>());
    for (int j = 0; j < V; ++j) {
        ucols[j].resize(n);
        WX_HIP(hipMemcpyAsync(ucols[j].data(), ub.as<double>() + j * n, n * 8, hipMemcpyDeviceToHost, st));
    }
    WX_HIP(hipStreamSynchronize(st));
}
} // namespace

std::unique_ptr<SlopeReparam> build_reparam_device(Design &d, const double *weights) {
    auto rp = std::make_unique<SlopeReparam>();
    const uint64_t n = d.n_obs;
    if (n >= 0xFFFFFFFFull) fail(WX_ERR_INVALID_INPUT, "the device reparametrisation supports fewer than 2^32 rows");
    DevBuf wdev;
    if (weights) wdev.upload(weights, n);
    std::vector<std::array<uint64_t, 3>> unid;
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
        const uint64_t v = zcols.size(), L = tm.n_levels;
        if (v > MAX_V)
            fail(WX_ERR_INVALID_INPUT, "the device reparametrisation supports at most " + std::to_string(MAX_V) +
                                           " slopes per effect");
        const bool intercept = icol >= 0;
        DevBuf lev;
        lev.upload(d.level_column(term), n);
        std::vector<double> transform, mean;
        std::vector<uint8_t> info;
        std::vector<std::vector<double>> ucols;
        if (v == 1) run_term<1>(d, term, zcols, intercept, lev, wdev, transform, mean, info, ucols);
        else if (v == 2) run_term<2>(d, term, zcols, intercept, lev, wdev, transform, mean, info, ucols);
        else run_term<3>(d, term, zcols, intercept, lev, wdev, transform, mean, info, ucols);
        SlopeReparam::TermBlock blk;
        blk.term = term;
        blk.n_slopes = v;
        blk.intercept = intercept;
        blk.transform = std::move(transform);
        blk.mean = std::move(mean);
        blk.first_col = slope_cols.front();
        for (uint64_t l = 0; l < L; ++l) {
            if (intercept && (info[l] & 128u)) unid.push_back({term, l, (uint64_t)icol});
            for (uint64_t j = 0; j < v; ++j)
                if (!(info[l] & (16u << j))) unid.push_back({term, l, slope_cols[j]});
        }
        for (uint64_t j = 0; j < v; ++j)
            std::memcpy(d.slopes.data() + zcols[j] * n, ucols[j].data(), n * sizeof(double));
        rp->blocks.push_back(std::move(blk));
    }
    if (rp->blocks.empty()) return nullptr;
    std::sort(unid.begin(), unid.end());
    for (auto &t : unid) rp->unidentified.insert(rp->unidentified.end(), t.begin(), t.end());
    return rp;
}

} // namespace wx
