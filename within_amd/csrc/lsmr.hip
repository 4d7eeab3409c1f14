// lsmr.hip -- device D / D' operators and the batched (modified) LSMR loop for gfx950.
//
// Vectors of a batch of k right-hand sides are interleaved: observation-space arrays are
// n x k row-major (row i holds its k columns contiguously), DOF-space arrays N x k.  One
// code load then serves all k columns and every access of a k-wide row is one coalesced
// transaction.  Scalars of the LSMR recurrence live on the device, one lane per column,
// so an iteration never waits for the host; the host polls the per-column "active" flags
// every few iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>

#include "schwarz.hpp"
#include "solver.hpp"

namespace wx {

// Whether runs fold the LSMR steps around an additive Schwarz apply into it (set_fused_steps).
std::atomic<bool> g_fused_steps{true};

bool set_fused_steps(bool on) { return g_fused_steps.exchange(on); }

namespace {

constexpr int BLK = (int)KERNEL_BLOCK;
constexpr uint32_t SEG_CHUNK = 64;  // rows per D' work item before a level is split
constexpr uint32_t FIX_CHUNK = 64;  // partials per combine item
// rows per work item of a fused pair's primary term (64, 128, 256, 512 and 1,024 measured the
// same or slower at config 2: fewer, longer items keep fewer loads in flight)
constexpr uint32_t PAIR_CHUNK = 64;
constexpr uint32_t MAX_FUSED_BLOCKS = 32768; // grid cap of the fused gather (one item per row group)
constexpr int GATHER_WAVE_STEPS = 2;             // row steps per pass of k_gather_wave
constexpr uint32_t GATHER_WAVE_BLOCKS = 8192;    // its grid cap (4 waves, each looping over items)

// Per-column LSMR state (structure of arrays, each of length MAX_BATCH).
struct ColState {
    double *alpha, *beta, *beta_prev_inv, *beta_inv, *p_coeff, *gscale;
    double *alpha_bar, *phi_bar, *c_bar, *s_bar, *zeta_bar, *zeta0, *a_norm_sq, *abs_tol;
    double *prev_rho, *prev_rho_bar, *t_x, *t_hbar, *t_h, *inv_alpha, *bnorm_scale;
    double *den_norm; // ||D'W y|| = beta_1 ||p~_1||, the denominator of the true residual
    double *rel_tol;  // the solve's tol, written with beta_1: a replayed iteration graph reads
                      // the current call's tolerance, never one fixed at capture
    int32_t *active, *upd, *brk, *stop, *iters, *err;
};
constexpr int N_DOUBLE_FIELDS = 23;
constexpr int N_INT_FIELDS = 6;

// Partial blocks of one packed member: the ranges lh[2r], lh[2r+1] for r in
// [ptr[m], ptr[m+1]).  A null ptr means one problem and the blocks [0, nb).
struct Ranges {
    const uint32_t *ptr;
    const uint32_t *lh;
};

// The beta reduction (k_beta) of ncol columns, run by ncol workgroups appended to the grid of
// a launch that neither writes the column state nor reads the fields k_beta writes; part
// holds nb norm partials per column.
struct BetaSide {
    const double *part;
    uint32_t nb, kp, ncol, k;
    Ranges rg;
};

__device__ inline double term_load(const DesignArgs &d, int32_t li, uint64_t i) {
    return li < 0 ? 1.0 : d.slopes[(uint64_t)li * d.n + i];
}

// Sum of the k-lane block reduction: threads sharing (tid & (kp-1)) hold the same column.
__device__ inline double block_reduce_cols(double v, uint32_t kp, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t s = BLK / 2; s >= kp && s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[threadIdx.x];
}

// (D x)[row] for R rows of one column j, as in crates/within/src/operator/design/gather.rs:10-78:
// terms summed in order from 0.0, each term folded from -0.0 over its coefficient columns.
// TT > 0 fixes the number of terms at compile time; PLAIN says every term is one intercept
// column, so the fold reduces to the value itself (-0.0 + x*1.0 == x bitwise) and all R*TT
// code loads are issued before any x load.  Every row must be valid: callers clamp rows past
// their range to a valid one and discard those results, so no load sits behind a branch.
template <int R, int TT, bool PLAIN>
__device__ inline void gather_rows(const DesignArgs &d, const double *__restrict__ x, const uint64_t (&rows)[R],
                                   uint32_t k, uint32_t j, double (&acc)[R]) {
    if constexpr (TT > 0 && PLAIN) {
        uint32_t code[TT][R];
#pragma unroll
        for (int t = 0; t < TT; ++t)
#pragma unroll
            for (int q = 0; q < R; ++q) code[t][q] = d.codes[(uint64_t)t * d.n + rows[q]];
        double val[TT][R];
#pragma unroll
        for (int t = 0; t < TT; ++t)
#pragma unroll
            for (int q = 0; q < R; ++q) val[t][q] = x[(d.terms[t].offset + code[t][q]) * k + j];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            double a = 0.0;
#pragma unroll
            for (int t = 0; t < TT; ++t) a += val[t][q];
            acc[q] = a;
        }
    } else {
#pragma unroll
        for (int q = 0; q < R; ++q) acc[q] = 0.0;
        for (uint32_t t = 0; t < d.T; ++t) {
            const TermInfo &ti = d.terms[t];
            uint64_t code[R];
#pragma unroll
            for (int q = 0; q < R; ++q) code[q] = d.codes[(uint64_t)t * d.n + rows[q]];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                double tv = -0.0;
                for (uint32_t c = 0; c < ti.ncols; ++c)
                    tv = tv + x[(ti.offset + c * ti.n_levels + code[q]) * k + j] * term_load(d, ti.load[c], rows[q]);
                acc[q] += tv;
            }
        }
    }
}


// ---------------------------------------------------------------------------
// D v gather fused with the u update:  u <- sqrt(w) .* (D v) + gscale_j * u_old,  ||u||^2
// partials (axpy_with_sq_norm, crates/schwarz-precond/src/lsmr/bidiag.rs:24-43).  u is kept
// unnormalised (u_k = beta_k * u_k-hat) as in ModifiedGolubKahan::step (bidiag.rs:275-309).
// FIRST: u_old is the right-hand side b = sqrt(w) .* y, formed from y_int on the fly.
// ---------------------------------------------------------------------------
template <bool WEIGHTED, bool FIRST, int TT, bool PLAIN>
__global__ __launch_bounds__(BLK) void k_gather_axpy(DesignArgs d, const double *__restrict__ sw,
                                                     const double *__restrict__ v, const double *uin, double *u,
                                                     ColState cs, double *__restrict__ part, uint32_t k,
                                                     uint32_t kshift) {
    __shared__ double sh[BLK];
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    double sq = 0.0;
    const bool act = j < k && cs.active[j];
    if (act) {
        const double scale = cs.gscale[j];
        // R rows per thread per pass, their loads issued together for memory-level parallelism
        constexpr int R = 4;
        for (uint64_t i0 = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; i0 < d.n; i0 += R * rstride) {
            uint64_t rows[R];
            bool ok[R];
            double uo[R], wq[R], acc[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                ok[q] = i0 + q * rstride < d.n;
                rows[q] = ok[q] ? i0 + q * rstride : i0; // clamped: loads stay unconditional
                uo[q] = uin[rows[q] * k + j];
                wq[q] = WEIGHTED ? sw[rows[q]] : 1.0;
            }
            gather_rows<R, TT, PLAIN>(d, v, rows, k, j, acc);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (!ok[q]) continue;
                double a = acc[q];
                if (WEIGHTED) a *= wq[q];
                const double b = (FIRST && WEIGHTED) ? wq[q] * uo[q] : uo[q];
                const double nu = a + scale * b;
                u[rows[q] * k + j] = nu;
                sq += nu * nu;
            }
        }
    }
    double tot = block_reduce_cols(sq, kp, sh);
    if (threadIdx.x < kp) part[(uint64_t)blockIdx.x * kp + threadIdx.x] = tot;
}

// The same update walked along the D' work items of a sorted single-column term (the term
// the observations are ordered by), so the term's reduction D_t' sqrt(w) u comes out of the
// same pass: a row group owns one item (<= 256 consecutive rows of one level), forms u for
// its rows four at a time, and sums sqrt(w_i) u_i in row order -- the order k_segreduce uses,
// so the result is bitwise the same as the separate reduction.
template <bool WEIGHTED, bool FIRST, int TT, bool PLAIN, bool NT>
__global__ __launch_bounds__(BLK) void k_gather_seg(DesignArgs d, TermInfo ts, const uint32_t *__restrict__ fcodes,
                                                    const SegItem *__restrict__ items,
                                                    uint32_t n_items, const double *__restrict__ sw,
                                                    const double *__restrict__ v, const double *uin, double *u,
                                                    double *__restrict__ atu, double *__restrict__ partials,
                                                    ColState cs, double *__restrict__ part, uint32_t k,
                                                    uint32_t kshift, const uint32_t *__restrict__ blk_col) {
    __shared__ double sh[BLK];
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    const uint32_t c = (blk_col ? blk_col[blockIdx.x] : 0u) + j; // column-state index
    double sq = 0.0;
    const bool act = j < k && cs.active[c];
    if (act) {
        const double scale = cs.gscale[c];
        constexpr int R = 4;
        const uint32_t groups = (gridDim.x * BLK) >> kshift;
        for (uint32_t it = (blockIdx.x * BLK + threadIdx.x) >> kshift; it < n_items; it += groups) {
            const SegItem item = items[it];
            // an item of several complete levels writes each level's sum when the level
            // changes, and zero for the levels in its range that have no rows
            uint32_t cur = item.level;
            double adj = 0.0;
            for (uint32_t p0 = item.begin; p0 < item.end; p0 += R) {
                uint64_t rows[R];
                bool ok[R];
                double uo[R], wq[R], acc[R];
                uint32_t lv[R];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    ok[q] = p0 + q < item.end;
                    rows[q] = ok[q] ? p0 + q : p0; // clamped: loads stay unconditional
                    uo[q] = NT ? __builtin_nontemporal_load(&uin[rows[q] * k + j]) : uin[rows[q] * k + j];
                    wq[q] = WEIGHTED ? sw[rows[q]] : 1.0;
                    lv[q] = fcodes[rows[q]]; // equals cur throughout a single-level item
                }
                gather_rows<R, TT, PLAIN>(d, v, rows, k, j, acc);
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    if (!ok[q]) continue;
                    double a = acc[q];
                    if (WEIGHTED) a *= wq[q];
                    const double b = (FIRST && WEIGHTED) ? wq[q] * uo[q] : uo[q];
                    const double nu = a + scale * b;
                    // u streams through once per iteration: non-temporal (wide rows), so the
                    // gathered coefficient rows keep their place in L2
                    if (NT) __builtin_nontemporal_store(nu, &u[rows[q] * k + j]);
                    else u[rows[q] * k + j] = nu;
                    sq += nu * nu;
                    if (lv[q] != cur) {
                        atu[(ts.offset + cur) * k + j] = adj;
                        for (uint32_t l = cur + 1; l < lv[q]; ++l) atu[(ts.offset + l) * k + j] = 0.0;
                        adj = 0.0;
                        cur = lv[q];
                    }
                    adj += WEIGHTED ? wq[q] * nu : nu;
                }
            }
            if (item.dest >= 0) {
                atu[(ts.offset + cur) * k + j] = adj;
                for (uint32_t l = cur + 1; l <= (uint32_t)item.dest; ++l) atu[(ts.offset + l) * k + j] = 0.0;
            } else if (item.dest != SEG_SKIP) {
                partials[(uint64_t)(-item.dest - 1) * k + j] = adj;
            }
        }
    }
    double tot = block_reduce_cols(sq, kp, sh);
    if (threadIdx.x < kp) part[(uint64_t)blockIdx.x * kp + threadIdx.x] = tot;
}

typedef double dbl2 __attribute__((ext_vector_type(2)));

// Orders a wavefront's LDS accesses: what one lane wrote before is visible to the others after.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum of a lane's value over the wave's row groups (lanes jv, jv + kv, ...), in a fixed tree
// order; every lane of the column pair gets the total.
__device__ inline dbl2 sum_over_groups(dbl2 a, uint32_t kv) {
    for (uint32_t off = kv; off < 64; off <<= 1) {
        a.x += __shfl_xor(a.x, (int)off);
        a.y += __shfl_xor(a.y, (int)off);
    }
    return a;
}

// The fused gather pass (k_gather_seg's arithmetic) for plain designs and an even number of
// columns, two columns per lane (16-byte loads and stores), one wavefront per work item: the
// wave's row groups take
// consecutive rows (row group g of step s owns row base + s*G + g, G = 64 / lanes per row), so
// each load instruction covers G adjacent rows.  Their code records, sqrt-weights and u rows are
// contiguous, and rows that share a level of the sorted term -- or of a small factor, which the
// observation order keeps together within a level of the sorted term -- share coefficient lines,
// so an instruction touches few distinct lines instead of one per row group.
// The sorted term's level sums: while a pass stays in the current level (most rows: big levels
// are split into chunks of one level), each lane adds its own rows and the row groups' sums are
// combined when the level ends; a pass that crosses levels goes through LDS, where the first row
// group adds its rows in order.  The order is fixed by the work items, so the sums are
// deterministic (they differ from k_segreduce's row-order sums by rounding only).
template <bool WEIGHTED, bool FIRST, int TT, int S>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_gather_wave(
    DesignArgs d, TermInfo ts, const uint4 *__restrict__ codes4, uint32_t fterm, const SegItem *__restrict__ items,
    uint32_t n_items, const double *__restrict__ sw, const double *__restrict__ v, const double *uin, double *u,
    double *__restrict__ atu, double *__restrict__ partials, ColState cs, double *__restrict__ part, uint32_t k,
    uint32_t kshift) {
    constexpr uint32_t WAVE = 64, NW = BLK / WAVE;
    __shared__ double sh[BLK];
    __shared__ dbl2 cbuf[NW][S * WAVE];   // per wave: row contributions of a pass, slot = step * 64 + lane
    __shared__ uint32_t lbuf[NW][S * 16]; // sorted-term level per row of the pass (G <= 16)
    const uint32_t kp = 1u << kshift, kv = kp >> 1; // lanes per row
    const uint32_t lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const uint32_t G = WAVE / kv;                   // rows per step
    const uint32_t jv = lane & (kv - 1), g = lane / kv, j0 = 2 * jv;
    double sq0 = 0.0, sq1 = 0.0;
    const bool act = j0 < k && (cs.active[j0] || cs.active[j0 + 1]);
    if (act) {
        const double sc0 = cs.gscale[j0], sc1 = cs.gscale[j0 + 1];
        const uint32_t waves = gridDim.x * NW;
        for (uint32_t it = blockIdx.x * NW + wv; it < n_items; it += waves) {
            const SegItem item = items[it];
            uint32_t cur = item.level;
            dbl2 adj = {0.0, 0.0}; // this lane's part of level cur's sum
            for (uint32_t base = item.begin; base < item.end; base += S * G) {
                uint32_t rows[S];
                bool ok[S];
                dbl2 uo[S];
                double wq[S];
                uint32_t lv[S], code[TT][S];
#pragma unroll
                for (int q = 0; q < S; ++q) {
                    const uint32_t r = base + q * G + g;
                    ok[q] = r < item.end;
                    rows[q] = ok[q] ? r : item.begin; // clamped: loads stay unconditional
                    uo[q] = __builtin_nontemporal_load((const dbl2 *)&uin[(uint64_t)rows[q] * k + j0]);
                    wq[q] = WEIGHTED ? sw[rows[q]] : 1.0;
                    const uint4 c4 = codes4[rows[q]];
                    const uint32_t cc[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
                    for (int t = 0; t < TT; ++t) code[t][q] = cc[t];
                    lv[q] = fterm == 0 ? c4.x : fterm == 1 ? c4.y : fterm == 2 ? c4.z : c4.w;
                }
                dbl2 val[TT][S];
#pragma unroll
                for (int t = 0; t < TT; ++t)
#pragma unroll
                    for (int q = 0; q < S; ++q)
                        val[t][q] = *(const dbl2 *)&v[(d.terms[t].offset + code[t][q]) * k + j0];
                dbl2 cq[S];
                bool cross = false;
#pragma unroll
                for (int q = 0; q < S; ++q) {
                    double a0 = 0.0, a1 = 0.0;
#pragma unroll
                    for (int t = 0; t < TT; ++t) {
                        a0 += val[t][q].x;
                        a1 += val[t][q].y;
                    }
                    if (WEIGHTED) {
                        a0 *= wq[q];
                        a1 *= wq[q];
                    }
                    const double b0 = (FIRST && WEIGHTED) ? wq[q] * uo[q].x : uo[q].x;
                    const double b1 = (FIRST && WEIGHTED) ? wq[q] * uo[q].y : uo[q].y;
                    dbl2 nu;
                    nu.x = a0 + sc0 * b0;
                    nu.y = a1 + sc1 * b1;
                    cq[q].x = ok[q] ? (WEIGHTED ? wq[q] * nu.x : nu.x) : 0.0;
                    cq[q].y = ok[q] ? (WEIGHTED ? wq[q] * nu.y : nu.y) : 0.0;
                    if (ok[q]) {
                        __builtin_nontemporal_store(nu, (dbl2 *)&u[(uint64_t)rows[q] * k + j0]);
                        sq0 += nu.x * nu.x;
                        sq1 += nu.y * nu.y;
                        cross |= lv[q] != cur;
                    }
                }
                if (!__any(cross)) {
#pragma unroll
                    for (int q = 0; q < S; ++q) {
                        adj.x += cq[q].x;
                        adj.y += cq[q].y;
                    }
                    continue;
                }
                // the pass crosses levels: the first row group carries level cur's sum so far
                // and adds the pass's rows in order, writing each level that ends
                adj = sum_over_groups(adj, kv);
                // every lane of a row group writes the row's level (the same value): a column
                // pair that has stopped skips this loop, so no one pair can be relied on for it
#pragma unroll
                for (int q = 0; q < S; ++q) {
                    cbuf[wv][q * WAVE + lane] = cq[q];
                    lbuf[wv][q * G + g] = lv[q];
                }
                wave_sync();
                if (g == 0) {
                    const uint32_t nrow = min(S * G, item.end - base); // row i: slot i * kv + jv
                    for (uint32_t i = 0; i < nrow; ++i) {
                        const uint32_t l = lbuf[wv][i];
                        const dbl2 c = cbuf[wv][i * kv + jv];
                        if (l != cur) {
                            *(dbl2 *)&atu[(ts.offset + cur) * k + j0] = adj;
                            const dbl2 zero = {0.0, 0.0};
                            for (uint32_t m = cur + 1; m < l; ++m) *(dbl2 *)&atu[(ts.offset + m) * k + j0] = zero;
                            adj = zero;
                            cur = l;
                        }
                        adj.x += c.x;
                        adj.y += c.y;
                    }
                } else {
                    adj.x = adj.y = 0.0;
                }
                cur = __shfl(cur, (int)jv); // the first row group's level, for every group
                wave_sync(); // the next pass overwrites the buffers
            }
            adj = sum_over_groups(adj, kv);
            if (g == 0) {
                if (item.dest >= 0) {
                    *(dbl2 *)&atu[(ts.offset + cur) * k + j0] = adj;
                    const dbl2 zero = {0.0, 0.0};
                    for (uint32_t m = cur + 1; m <= (uint32_t)item.dest; ++m)
                        *(dbl2 *)&atu[(ts.offset + m) * k + j0] = zero;
                } else if (item.dest != SEG_SKIP) {
                    *(dbl2 *)&partials[(uint64_t)(-item.dest - 1) * k + j0] = adj;
                }
            }
        }
    }
    const double t0 = block_reduce_cols(sq0, kv, sh);
    __syncthreads();
    if (threadIdx.x < kv) part[(uint64_t)blockIdx.x * kp + 2 * threadIdx.x] = t0;
    __syncthreads();
    const double t1 = block_reduce_cols(sq1, kv, sh);
    if (threadIdx.x < kv) part[(uint64_t)blockIdx.x * kp + 2 * threadIdx.x + 1] = t1;
}

// ---------------------------------------------------------------------------
// D' reduction: out[off + c*L + l] = sum over rows of level l (code-sorted order) of
// load_c(row) * scale(row) * src[row].  Deterministic: each item sums its rows in order,
// split levels add their partials in slot order.  Replaces the strategies of
// crates/within/src/operator/design/scatter.rs:14-242 (sequential / fold / CAS atomics /
// run-coalesced atomics) with one atomic-free code-sorted segmented sum.
// ---------------------------------------------------------------------------
// SCALE: multiply each row by a per-row factor (sqrt-weights or weights).  The factor array is
// indexed by position in the term's sorted order (a copy permuted at upload), so it is read
// sequentially; only the k-wide rows of src are gathered through the permutation.
template <int NC, bool SCALE>
__global__ __launch_bounds__(BLK) void k_segreduce(DesignArgs d, TermInfo ti, const uint32_t *__restrict__ perm,
                                                   const SegItem *__restrict__ items, uint32_t n_items,
                                                   const double *__restrict__ scale_pos, const double *__restrict__ src,
                                                   double *__restrict__ out, double *__restrict__ partials,
                                                   const int32_t *__restrict__ active, uint32_t k, uint32_t kshift) {
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    if (j >= k || (active && !active[j])) return;
    const uint32_t groups = (gridDim.x * BLK) >> kshift;
    for (uint32_t it = (blockIdx.x * BLK + threadIdx.x) >> kshift; it < n_items; it += groups) {
        const SegItem item = items[it];
        double acc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = 0.0;
        // rows are read eight at a time (independent loads in flight) but summed strictly in order
        constexpr uint32_t U = 8;
        uint32_t p = item.begin;
        for (; p + U <= item.end; p += U) {
            uint64_t row[U];
            double b[U], sc[U];
#pragma unroll
            for (int q = 0; q < (int)U; ++q) {
                row[q] = perm ? perm[p + q] : p + q;
                sc[q] = SCALE ? scale_pos[p + q] : 1.0;
            }
#pragma unroll
            for (int q = 0; q < (int)U; ++q) {
                b[q] = src[row[q] * k + j];
                if (SCALE) b[q] = sc[q] * b[q];
            }
#pragma unroll
            for (int q = 0; q < (int)U; ++q)
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    acc[c] += (ti.load[c] < 0 ? b[q] : term_load(d, ti.load[c], row[q]) * b[q]);
        }
        for (; p < item.end; ++p) {
            const uint64_t row = perm ? perm[p] : p;
            double b = src[row * k + j];
            if (SCALE) b = scale_pos[p] * b;
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[c] += (ti.load[c] < 0 ? b : term_load(d, ti.load[c], row) * b);
        }
        if (item.dest >= 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c) out[(ti.offset + c * ti.n_levels + item.level) * k + j] = acc[c];
        } else if (item.dest != SEG_SKIP) {
            const uint64_t slot = (uint64_t)(-item.dest - 1);
#pragma unroll
            for (int c = 0; c < NC; ++c) partials[(slot * NC + c) * k + j] = acc[c];
        }
    }
}

// The same reduction for a single intercept column with two columns per lane (16-byte row
// loads, half the lanes per row) for an even number of columns; per column the sums are
// exactly those of k_segreduce<1, SCALE>.
template <bool SCALE>
__global__ __launch_bounds__(BLK) void k_segreduce2(TermInfo ti, const uint32_t *__restrict__ perm,
                                                    const SegItem *__restrict__ items, uint32_t n_items,
                                                    const double *__restrict__ scale_pos,
                                                    const double *__restrict__ src, double *__restrict__ out,
                                                    double *__restrict__ partials, const int32_t *__restrict__ active,
                                                    uint32_t k, uint32_t kshift) {
    const uint32_t kv = (1u << kshift) >> 1;
    const uint32_t j0 = 2 * (threadIdx.x & (kv - 1));
    if (j0 >= k || (active && !active[j0] && !active[j0 + 1])) return;
    const uint32_t groups = (gridDim.x * BLK) / kv;
    for (uint32_t it = (blockIdx.x * BLK + threadIdx.x) / kv; it < n_items; it += groups) {
        const SegItem item = items[it];
        double a0 = 0.0, a1 = 0.0;
        constexpr uint32_t U = 8;
        uint32_t p = item.begin;
        for (; p + U <= item.end; p += U) {
            uint32_t row[U];
            double sc[U];
            dbl2 b[U];
#pragma unroll
            for (int q = 0; q < (int)U; ++q) {
                row[q] = perm ? perm[p + q] : p + q;
                sc[q] = SCALE ? scale_pos[p + q] : 1.0;
            }
#pragma unroll
            for (int q = 0; q < (int)U; ++q) b[q] = *(const dbl2 *)&src[(uint64_t)row[q] * k + j0];
#pragma unroll
            for (int q = 0; q < (int)U; ++q) {
                a0 += SCALE ? sc[q] * b[q].x : b[q].x;
                a1 += SCALE ? sc[q] * b[q].y : b[q].y;
            }
        }
        if (p < item.end) { // the last rows: their loads in flight together, added in order
            const uint32_t rem = item.end - p;
            uint32_t row[U - 1];
            double sc[U - 1];
            dbl2 b[U - 1];
#pragma unroll
            for (int q = 0; q < (int)U - 1; ++q)
                if ((uint32_t)q < rem) {
                    row[q] = perm ? perm[p + q] : p + q;
                    sc[q] = SCALE ? scale_pos[p + q] : 1.0;
                }
#pragma unroll
            for (int q = 0; q < (int)U - 1; ++q)
                if ((uint32_t)q < rem) b[q] = *(const dbl2 *)&src[(uint64_t)row[q] * k + j0];
#pragma unroll
            for (int q = 0; q < (int)U - 1; ++q)
                if ((uint32_t)q < rem) {
                    a0 += SCALE ? sc[q] * b[q].x : b[q].x;
                    a1 += SCALE ? sc[q] * b[q].y : b[q].y;
                }
        }
        dbl2 av;
        av.x = a0;
        av.y = a1;
        if (item.dest >= 0) *(dbl2 *)&out[(ti.offset + item.level) * k + j0] = av;
        else if (item.dest != SEG_SKIP) *(dbl2 *)&partials[(uint64_t)(-item.dest - 1) * k + j0] = av;
    }
}

// The primary term's reduction of a fused pair (PairPlan): k_segreduce2 over P's rows in (P, Q)
// order, plus the sum of every run of equal Q codes inside the item (a piece), written to
// pieces[piece] in run order from the item's first piece.  Each row's (scaled) u is read once
// and added to both sums; P's sums are exactly k_segreduce2's for the same row order.
__device__ inline void beta_column(const double *part, uint32_t nb, uint32_t kp, const ColState &cs, uint32_t k,
                                   Ranges rg, uint32_t c, double *sh);

template <bool SCALE>
__global__ __launch_bounds__(BLK) void k_segreduce2_pair(TermInfo ti, const uint32_t *__restrict__ perm,
                                                         const SegItem *__restrict__ items, uint32_t n_items,
                                                         const uint32_t *__restrict__ piece0,
                                                         const uint32_t *__restrict__ piece_slot,
                                                         const uint32_t *__restrict__ qcode,
                                                         const double *__restrict__ scale_pos,
                                                         const double *__restrict__ src, double *__restrict__ out,
                                                         double *__restrict__ partials, double *__restrict__ pieces,
                                                         const int32_t *__restrict__ active, uint32_t k,
                                                         uint32_t kshift, ColState cs, BetaSide bs) {
    const uint32_t nmain = gridDim.x - bs.ncol;
    if (blockIdx.x >= nmain) { // the appended beta workgroups
        __shared__ double sh[BLK];
        beta_column(bs.part, bs.nb, bs.kp, cs, bs.k, bs.rg, blockIdx.x - nmain, sh);
        return;
    }
    const uint32_t kv = (1u << kshift) >> 1;
    const uint32_t j0 = 2 * (threadIdx.x & (kv - 1));
    if (j0 >= k || (active && !active[j0] && !active[j0 + 1])) return;
    const uint32_t groups = (nmain * BLK) / kv;
    for (uint32_t it = (blockIdx.x * BLK + threadIdx.x) / kv; it < n_items; it += groups) {
        const SegItem item = items[it];
        const bool any = item.begin < item.end; // an unobserved level still writes its zero sum
        double a0 = 0.0, a1 = 0.0, r0 = 0.0, r1 = 0.0;
        uint32_t piece = piece0[it], cq = any ? qcode[item.begin] : 0u;
        constexpr uint32_t U = 8;
        uint32_t p = item.begin;
        for (; p + U <= item.end; p += U) {
            uint32_t row[U], qc[U];
            double sc[U];
            dbl2 b[U];
#pragma unroll
            for (int q = 0; q < (int)U; ++q) {
                row[q] = perm[p + q];
                qc[q] = qcode[p + q];
                sc[q] = SCALE ? scale_pos[p + q] : 1.0;
            }
#pragma unroll
            for (int q = 0; q < (int)U; ++q) b[q] = *(const dbl2 *)&src[(uint64_t)row[q] * k + j0];
#pragma unroll
            for (int q = 0; q < (int)U; ++q) {
                const double x0 = SCALE ? sc[q] * b[q].x : b[q].x;
                const double x1 = SCALE ? sc[q] * b[q].y : b[q].y;
                a0 += x0;
                a1 += x1;
                if (qc[q] != cq) {
                    dbl2 rv;
                    rv.x = r0;
                    rv.y = r1;
                    *(dbl2 *)&pieces[(uint64_t)piece_slot[piece] * k + j0] = rv;
                    ++piece;
                    r0 = 0.0;
                    r1 = 0.0;
                    cq = qc[q];
                }
                r0 += x0;
                r1 += x1;
            }
        }
        for (; p < item.end; ++p) {
            const uint64_t row = perm[p];
            const uint32_t qc = qcode[p];
            const dbl2 b = *(const dbl2 *)&src[row * k + j0];
            const double x0 = SCALE ? scale_pos[p] * b.x : b.x;
            const double x1 = SCALE ? scale_pos[p] * b.y : b.y;
            a0 += x0;
            a1 += x1;
            if (qc != cq) {
                dbl2 rv;
                rv.x = r0;
                rv.y = r1;
                *(dbl2 *)&pieces[(uint64_t)piece_slot[piece] * k + j0] = rv;
                ++piece;
                r0 = 0.0;
                r1 = 0.0;
                cq = qc;
            }
            r0 += x0;
            r1 += x1;
        }
        if (any) {
            dbl2 rv;
            rv.x = r0;
            rv.y = r1;
            *(dbl2 *)&pieces[(uint64_t)piece_slot[piece] * k + j0] = rv;
        }
        dbl2 av;
        av.x = a0;
        av.y = a1;
        if (item.dest >= 0) *(dbl2 *)&out[(ti.offset + item.level) * k + j0] = av;
        else if (item.dest != SEG_SKIP) *(dbl2 *)&partials[(uint64_t)(-item.dest - 1) * k + j0] = av;
    }
}

template <int NC>
__global__ void k_segfix(TermInfo ti, const SegFix *__restrict__ fixes, uint32_t n_fixes,
                         const double *__restrict__ partials, double *__restrict__ out,
                         const int32_t *__restrict__ active, uint32_t k, uint32_t kshift) {
    // one workgroup per split level: row groups stride over the partials, then a fixed-order
    // LDS tree across the groups
    __shared__ double sh[BLK];
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1), grp = threadIdx.x >> kshift;
    const bool on = j < k && !(active && !active[j]);
    const SegFix f = fixes[blockIdx.x];
    const uint32_t step = BLK >> kshift;
    for (int c = 0; c < NC; ++c) {
        double acc = 0.0;
        if (on) { // eight loads in flight, added in slot order
            uint32_t q = grp;
            for (; q + 7 * step < f.count; q += 8 * step) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = partials[((uint64_t)(f.first + q + u * step) * NC + c) * k + j];
#pragma unroll
                for (int u = 0; u < 8; ++u) acc += v[u];
            }
            for (; q < f.count; q += step) acc += partials[((uint64_t)(f.first + q) * NC + c) * k + j];
        }
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (uint32_t s = BLK / 2; s >= kp && s > 0; s >>= 1) {
            if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x < kp && on) out[(ti.offset + c * ti.n_levels + f.level) * k + j] = sh[threadIdx.x];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Right-hand side loading and vec_norm (crates/schwarz-precond/src/lsmr.rs:24-36)
// ---------------------------------------------------------------------------
// Caller-order column-major Y -> internal-order interleaved y_int through 64-row LDS tiles
// (k+1 doubles per tile row): Y is read along rows (coalesced) and each internal row's k
// values are written as one contiguous segment.  Non-finite entries are counted, and every
// block leaves per-column partials of sum(b^2) and max|b| for b = sqrt(w) .* y, so beta_1
// needs no further pass over the observations.
template <bool WEIGHTED, uint32_t QM>
__global__ __launch_bounds__(BLK) void k_load_rhs(uint64_t n, const uint32_t *__restrict__ inv_perm,
                                                  const double *__restrict__ sw_caller, const double *__restrict__ Y,
                                                  uint64_t ldy, double *__restrict__ y_int, double *__restrict__ part,
                                                  int32_t *__restrict__ bad, uint32_t k, uint32_t kshift, uint32_t kv) {
    // kv <= k: Y holds kv columns; columns kv .. k of y_int are zero (a packed member
    // narrower than its pack)
    extern __shared__ double tile[]; // 64 x (k + 1)
    __shared__ double sh[BLK];
    __shared__ uint64_t t_int[64]; // the tile rows' internal rows
    __shared__ double t_sw[64];    // and their sqrt-weights
    const uint32_t ld = k + 1, nq = 64 * k;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1), grp = threadIdx.x >> kshift;
    const uint32_t ngrp = BLK >> kshift;
    const uint64_t n_tiles = (n + 63) / 64;
    double mx = 0.0, sq = 0.0;
    int32_t nbad = 0;
    // a tile's loads into registers: element q = threadIdx.x + m * BLK of the 64 x k tile (column
    // q / 64, row q % 64; QM * BLK >= 64 k), and for threads < 64 their row's internal row and
    // sqrt-weight; the next tile's are issued before the current tile's stores
    double yr[QM];
    uint64_t ir = 0;
    double wr = 1.0;
    auto fetch = [&](uint64_t tb) {
        const uint64_t r0 = tb * 64;
#pragma unroll
        for (uint32_t m = 0; m < QM; ++m) {
            const uint32_t q = threadIdx.x + m * BLK, c = q >> 6;
            const uint64_t r = r0 + (q & 63);
            yr[m] = q < nq && r < n && c < kv ? Y[r + c * ldy] : 0.0;
        }
        if (threadIdx.x < 64) {
            const uint64_t r = r0 + threadIdx.x;
            if (r < n) {
                ir = inv_perm ? (uint64_t)inv_perm[r] : r;
                if (WEIGHTED) wr = sw_caller[r];
            }
        }
    };
    if (blockIdx.x < n_tiles) fetch(blockIdx.x);
    for (uint64_t tb = blockIdx.x; tb < n_tiles; tb += gridDim.x) {
        const uint64_t r0 = tb * 64;
#pragma unroll
        for (uint32_t m = 0; m < QM; ++m) {
            const uint32_t q = threadIdx.x + m * BLK;
            if (q < nq && r0 + (q & 63) < n) {
                nbad += !isfinite(yr[m]);
                tile[(q & 63) * ld + (q >> 6)] = yr[m];
            }
        }
        if (threadIdx.x < 64) {
            t_int[threadIdx.x] = ir;
            t_sw[threadIdx.x] = wr;
        }
        __syncthreads();
        if (tb + gridDim.x < n_tiles) fetch(tb + gridDim.x);
        if (j < k)
            for (uint32_t rr = grp; rr < 64; rr += ngrp) {
                if (r0 + rr >= n) break;
                const uint64_t i = t_int[rr];
                const double y = tile[rr * ld + j];
                y_int[i * k + j] = y;
                const double b = WEIGHTED ? t_sw[rr] * y : y;
                mx = fmax(mx, fabs(b));
                sq += b * b;
            }
        __syncthreads();
    }
    if (nbad) atomicAdd(bad, nbad);
    const double tsq = block_reduce_cols(sq, kp, sh);
    __syncthreads();
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (uint32_t s = BLK / 2; s >= kp && s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < kp) {
        part[((uint64_t)blockIdx.x * kp + threadIdx.x) * 2 + 0] = tsq;
        part[((uint64_t)blockIdx.x * kp + threadIdx.x) * 2 + 1] = sh[threadIdx.x];
    }
}

// Per-column reduction of block partials (one block per column, fixed tree order).  The
// reduction kernels run one workgroup per column-state index c = m * k + j: column j of
// packed member m (m = 0 without packing).
__device__ inline double reduce_partials(const double *part, uint32_t nb, uint32_t kp, uint32_t stride_fields,
                                         uint32_t field, uint32_t j, double *sh, bool is_max, Ranges rg = {},
                                         uint32_t m = 0) {
    // each thread keeps four accumulators (four loads in flight), combined in a fixed order;
    // then a butterfly inside each wavefront and the wavefronts' results in order
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const uint32_t r0 = rg.ptr ? rg.ptr[m] : 0u, r1 = rg.ptr ? rg.ptr[m + 1] : 1u;
    auto at = [&](uint32_t b) { return part[((uint64_t)b * kp + j) * stride_fields + field]; };
    for (uint32_t r = r0; r < r1; ++r) {
        const uint32_t lo = rg.ptr ? rg.lh[2 * r] : 0u, hi = rg.ptr ? rg.lh[2 * r + 1] : nb;
        uint32_t b = lo + threadIdx.x;
        for (; b + 3 * BLK < hi; b += 4 * BLK) {
            const double v0 = at(b), v1 = at(b + BLK), v2 = at(b + 2 * BLK), v3 = at(b + 3 * BLK);
            if (is_max) {
                a0 = fmax(a0, v0);
                a1 = fmax(a1, v1);
                a2 = fmax(a2, v2);
                a3 = fmax(a3, v3);
            } else {
                a0 += v0;
                a1 += v1;
                a2 += v2;
                a3 += v3;
            }
        }
        for (; b < hi; b += BLK) a0 = is_max ? fmax(a0, at(b)) : a0 + at(b);
    }
    double acc = is_max ? fmax(fmax(a0, a1), fmax(a2, a3)) : (a0 + a1) + (a2 + a3);
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(acc, off, 64);
        acc = is_max ? fmax(acc, o) : acc + o;
    }
    constexpr uint32_t NWAVE = BLK / 64;
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    double r = sh[0];
    for (uint32_t w = 1; w < NWAVE; ++w) r = is_max ? fmax(r, sh[w]) : r + sh[w];
    __syncthreads();
    return r;
}

// beta_1 of a column and the stop test for a zero right-hand side (lsmr.rs:266-282).
__device__ inline void set_beta1(ColState &cs, uint32_t j, double beta, double tol) {
    cs.beta[j] = beta;
    cs.abs_tol[j] = tol * beta;
    cs.rel_tol[j] = tol;
    cs.iters[j] = 0;
    cs.err[j] = 0;
    cs.brk[j] = 0;
    cs.upd[j] = 0;
    cs.den_norm[j] = 0.0;
    if (beta == 0.0) {
        cs.active[j] = 0;
        cs.stop[j] = WX_STOP_ZERO_RHS;
        cs.beta_inv[j] = 0.0;
    } else {
        cs.active[j] = 1;
        cs.stop[j] = -1;
        cs.beta_inv[j] = 1.0 / beta;
    }
}

// vec_norm (crates/schwarz-precond/src/lsmr.rs:24-36) scales by max|b| to avoid overflow and
// underflow of the squares.  When max|b| lies in [1e-140, 1e140] no square can overflow and
// the largest is normal, so sqrt(sum b^2) from the load partials is used directly; other
// columns raise `need_scaled` and get the scaled two-pass form (k_scaled_sumsq).
constexpr double NORM_SAFE_LO = 1e-140, NORM_SAFE_HI = 1e140;

// nonfinite: the load's count of non-finite entries; when it is not 0 every column stops at once
// (the solve then reports the error).  host_bad (pinned, or null): receives that count.
__global__ __launch_bounds__(BLK) void k_beta1(const double *part, uint32_t nb, uint32_t kp, ColState cs, double tol,
                                               int32_t *need_scaled, uint32_t k, Ranges rg,
                                               const int32_t *nonfinite, int32_t *host_bad) {
    __shared__ double sh[BLK];
    const uint32_t c = blockIdx.x, j = c % k;
    const double s = reduce_partials(part, nb, kp, 2, 0, j, sh, false, rg, c / k);
    const double m = reduce_partials(part, nb, kp, 2, 1, j, sh, true, rg, c / k);
    if (threadIdx.x == 0) {
        cs.bnorm_scale[c] = m;
        double beta = sqrt(s);
        if (m != 0.0 && !(m >= NORM_SAFE_LO && m <= NORM_SAFE_HI)) {
            beta = m; // provisional; replaced by the scaled pass
            need_scaled[0] = 1;
        }
        set_beta1(cs, c, m == 0.0 ? 0.0 : beta, tol);
        const int32_t bad = nonfinite ? *nonfinite : 0;
        if (bad) cs.active[c] = 0;
        if (host_bad && c == 0) host_bad[0] = bad;
    }
}

// sum((b/scale)^2) partials, b = sqrt(w) .* y_int, for the columns outside the safe range
template <bool WEIGHTED>
// (need: when set, the launch returns at once unless *need, k_beta1's flag, is set)
__global__ __launch_bounds__(BLK) void k_scaled_sumsq(uint64_t n, const double *__restrict__ sw,
                                                      const double *__restrict__ y_int, ColState cs,
                                                      double *__restrict__ part, uint32_t k, uint32_t kshift,
                                                      const int32_t *need) {
    if (need && *need == 0) return;
    __shared__ double sh[BLK];
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    double sq = 0.0;
    if (j < k && cs.bnorm_scale[j] > 0.0) {
        const double sc = cs.bnorm_scale[j];
        for (uint64_t i = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; i < n; i += rstride) {
            const double b = WEIGHTED ? sw[i] * y_int[i * k + j] : y_int[i * k + j];
            const double q = b / sc;
            sq += q * q;
        }
    }
    double tot = block_reduce_cols(sq, kp, sh);
    if (threadIdx.x < kp) part[(uint64_t)blockIdx.x * kp + threadIdx.x] = tot;
}

__global__ __launch_bounds__(BLK) void k_beta1_scaled(const double *part, uint32_t nb, uint32_t kp, ColState cs,
                                                      double tol, const int32_t *need, const int32_t *nonfinite) {
    if (need && *need == 0) return;
    __shared__ double sh[BLK];
    const uint32_t j = blockIdx.x;
    const double s = reduce_partials(part, nb, kp, 1, 0, j, sh, false);
    if (threadIdx.x == 0) {
        const double m = cs.bnorm_scale[j];
        if (m != 0.0 && !(m >= NORM_SAFE_LO && m <= NORM_SAFE_HI)) set_beta1(cs, j, m * sqrt(s), tol);
        if (nonfinite && *nonfinite) cs.active[j] = 0;
    }
}

// ---------------------------------------------------------------------------
// DOF-space kernels
// ---------------------------------------------------------------------------
// PREC: 0 identity, 1 diagonal, 2 external (only p~ is written).  INIT: p~ holds A'b and
// becomes A'b / beta_1 = A'u_1.
// p~ <- A'u * (1/beta) - (beta/alpha) p~   (bidiag.rs:475-485), v = M^{-1} p~, and partial
// sums of <v,p~>, <v,v>, <p~,p~> for alpha_from_vp (bidiag.rs:99-111).
template <int PREC, bool INIT>
__global__ __launch_bounds__(BLK) void k_ptilde(uint64_t N, const double *__restrict__ atu, double *__restrict__ pt,
                                                double *__restrict__ v, const double *__restrict__ inv_diag,
                                                ColState cs, double *__restrict__ part, uint32_t k, uint32_t kshift,
                                                const uint32_t *__restrict__ blk_col) {
    __shared__ double sh[BLK];
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    const uint32_t c = (blk_col ? blk_col[blockIdx.x] : 0u) + j;
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    double vp = 0.0, vv = 0.0, pp = 0.0;
    if (j < k && cs.active[c]) {
        const bool brk = !INIT && cs.brk[c];
        const double bi = cs.beta_inv[c];
        const double pc = INIT ? 0.0 : cs.p_coeff[c];
        for (uint64_t dof = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; dof < N; dof += rstride) {
            const uint64_t q = dof * k + j;
            double p;
            if (brk) p = 0.0;
            else if (INIT) p = bi * pt[q];
            else p = bi * atu[q] + (-pc) * pt[q];
            pt[q] = p;
            if (PREC != 2) {
                const double z = brk ? 0.0 : (PREC == 1 ? inv_diag[dof] * p : p);
                v[q] = z;
                vp += z * p;
                vv += z * z;
                pp += p * p;
            } else if (brk) {
                v[q] = 0.0;
            }
        }
    }
    double a = block_reduce_cols(vp, kp, sh);
    __syncthreads();
    double b = block_reduce_cols(vv, kp, sh);
    __syncthreads();
    double cc = block_reduce_cols(pp, kp, sh);
    if (threadIdx.x < kp) {
        double *o = part + ((uint64_t)blockIdx.x * kp + threadIdx.x) * 3;
        o[0] = a;
        o[1] = b;
        o[2] = cc;
    }
}

// <v,p~>, <v,v>, <p~,p~> partials after an external preconditioner apply.
__global__ __launch_bounds__(BLK) void k_dots(uint64_t N, const double *__restrict__ pt, const double *__restrict__ v,
                                              ColState cs, double *__restrict__ part, uint32_t k, uint32_t kshift,
                                              const uint32_t *__restrict__ blk_col) {
    __shared__ double sh[BLK];
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    const uint32_t c = (blk_col ? blk_col[blockIdx.x] : 0u) + j;
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    double vp = 0.0, vv = 0.0, pp = 0.0;
    if (j < k && cs.active[c]) {
        for (uint64_t dof = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; dof < N; dof += rstride) {
            const double z = v[dof * k + j], p = pt[dof * k + j];
            vp += z * p;
            vv += z * z;
            pp += p * p;
        }
    }
    double a = block_reduce_cols(vp, kp, sh);
    __syncthreads();
    double b = block_reduce_cols(vv, kp, sh);
    __syncthreads();
    double cc = block_reduce_cols(pp, kp, sh);
    if (threadIdx.x < kp) {
        double *o = part + ((uint64_t)blockIdx.x * kp + threadIdx.x) * 3;
        o[0] = a;
        o[1] = b;
        o[2] = cc;
    }
}

// Windowed reorthogonalisation (WindowRing<2>, crates/schwarz-precond/src/lsmr/bidiag.rs:114-217):
// the ring holds the last `cap` normalised v and M v = p~ / alpha of every column.  One MGS
// sweep over the stored slots, oldest first: c = <v, P_slot>, v -= c V_slot, p~ -= c P_slot.
__global__ __launch_bounds__(BLK) void k_ring_dot(uint64_t N, const double *__restrict__ v,
                                                  const double *__restrict__ P, ColState cs,
                                                  double *__restrict__ part, uint32_t k, uint32_t kshift) {
    __shared__ double sh[BLK];
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    double acc = 0.0;
    if (j < k && cs.active[j])
        for (uint64_t dof = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; dof < N; dof += rstride)
            acc += v[dof * k + j] * P[dof * k + j];
    const double tot = block_reduce_cols(acc, kp, sh);
    if (threadIdx.x < kp) part[(uint64_t)blockIdx.x * kp + threadIdx.x] = tot;
}

__global__ __launch_bounds__(BLK) void k_ring_axpy(uint64_t N, double *__restrict__ v, double *__restrict__ pt,
                                                   const double *__restrict__ V, const double *__restrict__ P,
                                                   const double *__restrict__ coef, ColState cs, uint32_t k,
                                                   uint32_t kshift) {
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    if (j >= k || !cs.active[j]) return;
    const double c = coef[j];
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t dof = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; dof < N; dof += rstride) {
        const uint64_t q = dof * k + j;
        v[q] = -c * V[q] + 1.0 * v[q];
        pt[q] = -c * P[q] + 1.0 * pt[q];
    }
}

__device__ inline void givens(double a, double b, double &c, double &s, double &r) {
    // crates/schwarz-precond/src/lsmr/recurrence.rs:20-27
    r = hypot(a, b);
    if (r > 0.0) {
        c = a / r;
        s = b / r;
    } else {
        c = 1.0;
        s = 0.0;
    }
}

__device__ inline double alpha_from_parts(double vp, double vv, double pp, bool &indefinite) {
    // crates/schwarz-precond/src/lsmr/bidiag.rs:99-111
    indefinite = false;
    if (vp < 0.0) {
        double bound = sqrt(2.220446049250313e-16) * sqrt(vv * pp);
        if (vp < -bound) indefinite = true;
    }
    return sqrt(vp > 0.0 ? vp : 0.0);
}

// beta = ||u|| after the gather; breakdown flags and the p~ update coefficients (column c,
// one workgroup).
__device__ inline void beta_column(const double *part, uint32_t nb, uint32_t kp, const ColState &cs, uint32_t k,
                                   Ranges rg, uint32_t c, double *sh) {
    const uint32_t j = c % k;
    double s = reduce_partials(part, nb, kp, 1, 0, j, sh, false, rg, c / k);
    if (threadIdx.x == 0 && cs.active[c]) {
        double beta = sqrt(s);
        cs.beta[c] = beta;
        if (beta == 0.0) { // lucky breakdown (bidiag.rs:282-291)
            cs.brk[c] = 1;
            cs.beta_inv[c] = 0.0;
            cs.p_coeff[c] = 0.0;
        } else {
            cs.brk[c] = 0;
            cs.beta_inv[c] = 1.0 / beta;
            cs.p_coeff[c] = beta / cs.alpha[c];
        }
    }
}

__global__ __launch_bounds__(BLK) void k_beta(const double *part, uint32_t nb, uint32_t kp, ColState cs, uint32_t k,
                                              Ranges rg) {
    __shared__ double sh[BLK];
    beta_column(part, nb, kp, cs, k, rg, blockIdx.x, sh);
}

// First alpha (ModifiedGolubKahan::init, bidiag.rs:428-472) and the recurrence seed
// (LsmrRecurrenceState::init, recurrence.rs:68-79; lsmr_from_bidiag early exit lsmr.rs:306-316).
__global__ __launch_bounds__(BLK) void k_alpha1(const double *part, uint32_t nb, uint32_t kp, ColState cs, uint32_t k, Ranges rg) {
    __shared__ double sh[BLK];
    const uint32_t c = blockIdx.x, j = c % k;
    double vp = reduce_partials(part, nb, kp, 3, 0, j, sh, false, rg, c / k);
    double vv = reduce_partials(part, nb, kp, 3, 1, j, sh, false, rg, c / k);
    double pp = reduce_partials(part, nb, kp, 3, 2, j, sh, false, rg, c / k);
    if (threadIdx.x == 0 && cs.active[c]) {
        double beta = cs.beta[c];
        cs.den_norm[c] = beta * sqrt(pp);
        bool indef;
        double alpha = alpha_from_parts(vp, vv, pp, indef);
        if (indef) {
            cs.err[c] = 1;
            cs.active[c] = 0;
            return;
        }
        cs.alpha[c] = alpha;
        cs.inv_alpha[c] = alpha > 0.0 ? 1.0 / alpha : 0.0;
        cs.beta_prev_inv[c] = 1.0;
        if (alpha == 0.0) {
            cs.stop[c] = WX_STOP_INITIAL_NORMAL_EQ_ZERO;
            cs.active[c] = 0;
            cs.phi_bar[c] = beta;
            cs.zeta_bar[c] = 0.0;
            cs.zeta0[c] = 1.0;
            return;
        }
        cs.alpha_bar[c] = alpha;
        cs.phi_bar[c] = beta;
        cs.c_bar[c] = 1.0;
        cs.s_bar[c] = 0.0;
        double zb = alpha * beta;
        cs.zeta_bar[c] = zb;
        cs.zeta0[c] = fmax(fabs(zb), 2.2250738585072014e-308);
        cs.a_norm_sq[c] = alpha * alpha;
        cs.prev_rho[c] = 1.0;
        cs.prev_rho_bar[c] = 1.0;
        cs.gscale[c] = -(alpha * cs.beta_inv[c]); // u_1 is kept as b = beta_1 * u_1-hat
    }
}

// v <- v / alpha_1, h = v, x = 0, h_bar = 0 (SolutionState::init, recurrence.rs:134-141).
__global__ __launch_bounds__(BLK) void k_init_solution(uint64_t N, double *__restrict__ v, double *__restrict__ h,
                                                       double *__restrict__ hb, double *__restrict__ x, ColState cs,
                                                       uint32_t k, uint32_t kshift, const double *__restrict__ pt,
                                                       double *__restrict__ Vpush, double *__restrict__ Ppush,
                                                       const uint32_t *__restrict__ blk_col) {
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint32_t c = (blk_col ? blk_col[blockIdx.x] : 0u) + j;
    const bool act = cs.active[c];
    const double ia = cs.inv_alpha[c];
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t dof = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; dof < N; dof += rstride) {
        const uint64_t q = dof * k + j;
        x[q] = 0.0;
        hb[q] = 0.0;
        if (act) {
            double z = v[q] * ia;
            v[q] = z;
            h[q] = z;
            if (Vpush) { // ring push of v and p~ / alpha (bidiag.rs:459-462)
                Vpush[q] = z;
                Ppush[q] = pt[q] * ia;
            }
        } else {
            h[q] = 0.0;
        }
    }
}

// One LSMR scalar step per column: alpha from <v,p~>, the two Givens chains, the solution
// coefficients and the dual stopping test.  Follows lsmr_from_bidiag (lsmr.rs:323-361),
// LsmrRecurrenceState::step (recurrence.rs:81-107), SolutionState::update's coefficients
// (recurrence.rs:143-161) and ConvergenceState::{observe, check} (recurrence.rs:239-255).
// The iteration count is kept per column on the device, so the launch is identical every
// iteration and a captured graph can replay it.
__device__ inline float alpha_step_column(const double *part, uint32_t nb, uint32_t kp, const ColState &cs, uint32_t k,
                                         Ranges rg, uint32_t c, double *sh) {
    const uint32_t j = c % k;
    double vp = reduce_partials(part, nb, kp, 3, 0, j, sh, false, rg, c / k);
    double vv = reduce_partials(part, nb, kp, 3, 1, j, sh, false, rg, c / k);
    double pp = reduce_partials(part, nb, kp, 3, 2, j, sh, false, rg, c / k);
    if (threadIdx.x != 0) return 0.0f;
    if (!cs.active[c]) {
        cs.upd[c] = 0;
        return 0.0f;
    }
    double alpha, beta = cs.beta[c];
    if (cs.brk[c]) {
        alpha = 0.0;
    } else {
        bool indef;
        alpha = alpha_from_parts(vp, vv, pp, indef);
        if (indef) {
            cs.err[c] = 1;
            cs.active[c] = 0;
            cs.upd[c] = 0;
            cs.iters[c] += 1; // steps taken (reset by set_beta1)
            return 0.0f;
        }
    }
    cs.inv_alpha[c] = alpha > 0.0 ? 1.0 / alpha : 0.0;
    // convergence.observe
    double a_norm_sq = cs.a_norm_sq[c] + alpha * alpha + beta * beta;
    cs.a_norm_sq[c] = a_norm_sq;
    // recurrence.step
    double ph_c, ph_s, ph_r;
    givens(cs.alpha_bar[c], beta, ph_c, ph_s, ph_r);
    const double theta_new = ph_s * alpha;
    const double alpha_bar_new = -ph_c * alpha;
    const double phi_bar_new = ph_s * cs.phi_bar[c];
    const double theta_bar = cs.s_bar[c] * ph_r;
    double pb_c, pb_s, pb_r;
    givens(cs.c_bar[c] * ph_r, theta_new, pb_c, pb_s, pb_r);
    const double zeta_bar = cs.zeta_bar[c];
    const double zeta = pb_c * zeta_bar;
    const double zeta_bar_new = -pb_s * zeta_bar;
    cs.alpha_bar[c] = alpha_bar_new;
    cs.phi_bar[c] = phi_bar_new;
    cs.c_bar[c] = pb_c;
    cs.s_bar[c] = pb_s;
    cs.zeta_bar[c] = zeta_bar_new;
    // solution.update coefficients
    const double eps = 2.220446049250313e-16;
    const double t_x_denom = ph_r * pb_r;
    cs.t_x[c] = fabs(t_x_denom) > eps ? zeta / t_x_denom : 0.0;
    const double t_hbar_denom = cs.prev_rho[c] * cs.prev_rho_bar[c];
    cs.t_hbar[c] = fabs(t_hbar_denom) > eps ? theta_bar * ph_r / t_hbar_denom : 0.0;
    cs.t_h[c] = fabs(ph_r) > eps ? theta_new / ph_r : 0.0;
    cs.prev_rho[c] = ph_r;
    cs.prev_rho_bar[c] = pb_r;
    // bidiagonal state for the next step
    cs.gscale[c] = -(alpha * cs.beta_inv[c]); // -(alpha_k / beta_k); 0 after a breakdown
    cs.alpha[c] = alpha;
    // convergence.check
    const double residual = fabs(phi_bar_new);
    int32_t stop = -1;
    double progress = 0.0; // min over the two tests of value / tolerance (scheduling only)
    if (residual <= cs.abs_tol[c]) stop = WX_STOP_RESIDUAL_TOLERANCE;
    else {
        const double a_norm = fmax(sqrt(a_norm_sq), 2.2250738585072014e-308);
        const double normar = fabs(zeta_bar_new);
        const double test2 = normar / (a_norm * fmax(residual, 2.2250738585072014e-308));
        if (test2 <= cs.rel_tol[c]) stop = WX_STOP_NORMAL_EQ_TOLERANCE;
        else progress = fmin(residual / cs.abs_tol[c], test2 / cs.rel_tol[c]);
    }
    cs.upd[c] = 1;
    cs.iters[c] += 1; // steps taken (reset by set_beta1)
    if (stop >= 0) {
        cs.stop[c] = stop;
        cs.active[c] = 0;
    }
    return (float)progress;
}

// One workgroup per column; the column's new "active" flag and its progress towards the stopping
// test are also stored to pinned host memory the host polls (mirror, progress), so no copy of
// the flags is enqueued after an iteration.
__global__ __launch_bounds__(BLK) void k_alpha_step(const double *part, uint32_t nb, uint32_t kp, ColState cs,
                                                    uint32_t k, Ranges rg, int32_t *mirror, float *progress) {
    __shared__ double sh[BLK];
    const uint32_t c = blockIdx.x;
    const float pg = alpha_step_column(part, nb, kp, cs, k, rg, c, sh);
    if (threadIdx.x == 0 && mirror) {
        mirror[c] = cs.active[c];
        progress[c] = pg;
    }
}

// x/h/h_bar update with the normalized v (recurrence.rs:163-190); v <- v / alpha first
// (bidiag.rs:505-507).
__global__ __launch_bounds__(BLK) void k_update_solution(uint64_t N, double *__restrict__ v, double *__restrict__ h,
                                                         double *__restrict__ hb, double *__restrict__ x, ColState cs,
                                                         uint32_t k, uint32_t kshift, const double *__restrict__ pt,
                                                         double *__restrict__ Vpush, double *__restrict__ Ppush,
                                                         const uint32_t *__restrict__ blk_col) {
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    const uint32_t c = (blk_col ? blk_col[blockIdx.x] : 0u) + j;
    if (j >= k || !cs.upd[c]) return;
    const double ia = cs.inv_alpha[c], t_x = cs.t_x[c], t_hbar = cs.t_hbar[c], t_h = cs.t_h[c];
    const bool scale = ia > 0.0;
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t dof = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; dof < N; dof += rstride) {
        const uint64_t q = dof * k + j;
        double vq = v[q];
        if (scale) {
            vq *= ia;
            v[q] = vq;
        }
        if (Vpush) { // ring push of the normalised v and p~ / alpha (bidiag.rs:501-510)
            Vpush[q] = vq;
            Ppush[q] = pt[q] * ia;
        }
        const double h_old = h[q];
        const double hbq = h_old - t_hbar * hb[q];
        hb[q] = hbq;
        x[q] += t_x * hbq;
        h[q] = vq - t_h * h_old;
    }
}

// dm_int[i, j] = y_int[i, j] - (D x)[i, j] in internal order (solver.rs:469-473; the gather
// is unweighted and folds each term from -0.0 like the operator's).
template <int TT, bool PLAIN>
__global__ __launch_bounds__(BLK) void k_demean(DesignArgs d, const double *__restrict__ x,
                                                const double *__restrict__ y_int, double *__restrict__ dm_int,
                                                uint32_t k, uint32_t kshift) {
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    constexpr int R = 4;
    for (uint64_t i0 = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; i0 < d.n; i0 += R * rstride) {
        uint64_t rows[R];
        bool ok[R];
        double yv[R], acc[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            ok[q] = i0 + q * rstride < d.n;
            rows[q] = ok[q] ? i0 + q * rstride : i0; // clamped: loads stay unconditional
            yv[q] = y_int[rows[q] * k + j];
        }
        gather_rows<R, TT, PLAIN>(d, x, rows, k, j, acc);
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (ok[q]) dm_int[rows[q] * k + j] = yv[q] - acc[q];
    }
}

// k_demean with two columns per lane (plain designs, even width); per column the arithmetic is
// that of k_demean.
template <int TT>
__global__ __launch_bounds__(BLK) void k_demean2(DesignArgs d, const uint4 *__restrict__ codes4, const double *__restrict__ x,
                                                 const double *__restrict__ y_int, double *__restrict__ dm_int,
                                                 uint32_t k, uint32_t kshift) {
    const uint32_t kv = (1u << kshift) >> 1, j0 = 2 * (threadIdx.x & (kv - 1));
    if (j0 >= k) return;
    const uint32_t rstride = (uint32_t)(((uint64_t)gridDim.x * BLK) / kv);
    constexpr int R = 4;
    for (uint32_t i0 = (blockIdx.x * BLK + threadIdx.x) / kv; i0 < d.n; i0 += R * rstride) {
        uint32_t rows[R];
        bool ok[R];
        dbl2 yv[R];
        uint32_t code[TT][R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            ok[q] = i0 + q * rstride < d.n;
            rows[q] = ok[q] ? i0 + q * rstride : i0; // clamped: loads stay unconditional
            yv[q] = *(const dbl2 *)&y_int[(uint64_t)rows[q] * k + j0];
            const uint4 c4 = codes4[rows[q]];
            const uint32_t cc[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
            for (int t = 0; t < TT; ++t) code[t][q] = cc[t];
        }
        dbl2 val[TT][R];
#pragma unroll
        for (int t = 0; t < TT; ++t)
#pragma unroll
            for (int q = 0; q < R; ++q) val[t][q] = *(const dbl2 *)&x[(d.terms[t].offset + code[t][q]) * k + j0];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if (!ok[q]) continue;
            double a0 = 0.0, a1 = 0.0;
#pragma unroll
            for (int t = 0; t < TT; ++t) {
                a0 += val[t][q].x;
                a1 += val[t][q].y;
            }
            dbl2 o;
            o.x = yv[q].x - a0;
            o.y = yv[q].y - a1;
            *(dbl2 *)&dm_int[(uint64_t)rows[q] * k + j0] = o;
        }
    }
}

// Internal-order interleaved dm_int -> caller-order column-major Dm (solver.rs:483), through
// 64-row LDS tiles (k+1 doubles per tile row) so the column-major writes are coalesced.
template <uint32_t RM>
__global__ __launch_bounds__(BLK) void k_store_caller(uint64_t n, const uint32_t *__restrict__ inv_perm,
                                                      const double *__restrict__ dm_int, double *__restrict__ Dm,
                                                      uint64_t ldd, uint32_t k, uint32_t kshift, uint32_t kv) {
    extern __shared__ double tile[]; // 64 x (k + 1)
    const uint32_t ld = k + 1;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1), grp = threadIdx.x >> kshift;
    const uint32_t ngrp = BLK >> kshift;
    const uint64_t n_tiles = (n + 63) / 64;
    // a tile's gathers into registers: this thread's rows grp + m * ngrp (m < RM, RM * ngrp >=
    // 64), column j; the next tile's are issued before the current tile's stores
    double v[RM];
    auto fetch = [&](uint64_t tb) {
        const uint64_t r0 = tb * 64;
#pragma unroll
        for (uint32_t m = 0; m < RM; ++m) {
            const uint32_t rr = grp + m * ngrp;
            const uint64_t r = r0 + rr;
            v[m] = j < k && rr < 64 && r < n ? dm_int[(inv_perm ? (uint64_t)inv_perm[r] : r) * k + j] : 0.0;
        }
    };
    if (blockIdx.x < n_tiles) fetch(blockIdx.x);
    for (uint64_t tb = blockIdx.x; tb < n_tiles; tb += gridDim.x) {
        const uint64_t r0 = tb * 64;
#pragma unroll
        for (uint32_t m = 0; m < RM; ++m) {
            const uint32_t rr = grp + m * ngrp;
            if (j < k && rr < 64 && r0 + rr < n) tile[rr * ld + j] = v[m];
        }
        __syncthreads();
        if (tb + gridDim.x < n_tiles) fetch(tb + gridDim.x);
        for (uint32_t q = threadIdx.x; q < 64 * k; q += BLK) {
            const uint32_t c = q >> 6, rr = q & 63;
            const uint64_t r = r0 + rr;
            if (r < n && c < kv) Dm[r + c * ldd] = tile[rr * ld + c];
        }
        __syncthreads();
    }
}

// The demeaned output y - D x straight into caller order, column-major (k_demean2 followed by
// k_store_caller, without the internal-order round trip), per DEMEAN_TILE-row tile of caller
// rows: the tile of the caller's response Y (the same values y_int holds in internal order) is
// read column by column into LDS, each row's code record in caller order gives its coefficient
// rows, the terms are added in order from 0.0 and subtracted from y, as k_demean2 does (so the
// values are the same bits), and the tile is written column by column.  Every stream but the
// coefficient rows is sequential.  For plain designs of up to four terms (code records present).
// (Tiles of 128 and 256 rows measured slower.)
constexpr uint32_t DEMEAN_TILE = 64;
template <int TT>
__global__ __launch_bounds__(BLK) void k_demean_store(DesignArgs d, const uint4 *__restrict__ codes4c,
                                                      const double *__restrict__ x, const double *__restrict__ Y,
                                                      uint64_t ldy, double *__restrict__ Dm, uint64_t ldd, uint32_t k,
                                                      uint32_t kshift) {
    extern __shared__ double tile[]; // DEMEAN_TILE x (k + 1)
    const uint32_t ld = k + 1;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1), grp = threadIdx.x >> kshift;
    const uint32_t ngrp = BLK >> kshift;
    const uint64_t n = d.n, n_tiles = (n + DEMEAN_TILE - 1) / DEMEAN_TILE;
    constexpr int R = DEMEAN_TILE / 16 < 4 ? 4 : DEMEAN_TILE / 16;
    for (uint64_t tb = blockIdx.x; tb < n_tiles; tb += gridDim.x) {
        const uint64_t r0 = tb * DEMEAN_TILE;
        // (D x)_r of rows rb + q * ngrp (q < R), the terms added in order; ok[q] = the row exists
        bool ok[R];
        double dx[R];
        auto rows_dx = [&](uint32_t rb) {
            uint32_t code[TT][R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const uint64_t r = r0 + rb + q * ngrp;
                ok[q] = rb + q * ngrp < DEMEAN_TILE && r < n;
                const uint4 c4 = codes4c[ok[q] ? r : 0];
                const uint32_t cc[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
                for (int t = 0; t < TT; ++t) code[t][q] = cc[t];
            }
            double val[TT][R];
#pragma unroll
            for (int t = 0; t < TT; ++t)
#pragma unroll
                for (int q = 0; q < R; ++q) val[t][q] = x[(d.terms[t].offset + code[t][q]) * k + j];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                double a = 0.0;
#pragma unroll
                for (int t = 0; t < TT; ++t) a += val[t][q];
                dx[q] = a;
            }
        };
        // the first rows' codes and coefficient gathers in flight with the tile's loads
        if (j < k) rows_dx(grp);
        for (uint32_t q = threadIdx.x; q < DEMEAN_TILE * k; q += BLK) {
            const uint32_t c = q / DEMEAN_TILE, rr = q % DEMEAN_TILE;
            const uint64_t r = r0 + rr;
            if (r < n) tile[rr * ld + c] = Y[r + c * ldy];
        }
        __syncthreads();
        if (j < k)
            for (uint32_t rb = grp; rb < DEMEAN_TILE; rb += R * ngrp) {
                if (rb != grp) rows_dx(rb);
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    if (!ok[q]) continue;
                    double &yv = tile[(rb + q * ngrp) * ld + j];
                    yv = yv - dx[q];
                }
            }
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < DEMEAN_TILE * k; q += BLK) {
            const uint32_t c = q / DEMEAN_TILE, rr = q % DEMEAN_TILE;
            const uint64_t r = r0 + rr;
            if (r < n) Dm[r + c * ldd] = tile[rr * ld + c];
        }
        __syncthreads();
    }
}

// X[dof + j*ldx] = x[dof*k + j]
__global__ void k_store_x(uint64_t N, const double *__restrict__ x, double *__restrict__ X, uint64_t ldx, uint32_t k) {
    const uint64_t idx = (uint64_t)blockIdx.x * BLK + threadIdx.x;
    if (idx >= N * k) return;
    const uint64_t dof = idx / k, j = idx % k;
    X[dof + j * ldx] = x[idx];
}

// Packed x -> each member's X (column-major, its own DOF order, leading dimension own_ld[m]);
// padding DOFs (own index -1) are skipped.
__global__ void k_store_x_packed(uint64_t N, const double *__restrict__ x, uint32_t k,
                                 const uint32_t *__restrict__ dof_member, const int32_t *__restrict__ dof_own,
                                 double *const *__restrict__ X, const uint64_t *__restrict__ own_ld,
                                 const uint32_t *__restrict__ own_kv) {
    const uint64_t idx = (uint64_t)blockIdx.x * BLK + threadIdx.x;
    if (idx >= N * k) return;
    const uint64_t dof = idx / k, j = idx % k;
    const int32_t own = dof_own[dof];
    if (own < 0) return;
    const uint32_t m = dof_member[dof];
    if (j >= own_kv[m]) return;
    double *Xm = X[m];
    if (Xm) Xm[(uint64_t)own + j * own_ld[m]] = x[idx];
}

// sum of squares per column of an N x k interleaved array (block partials)
__global__ __launch_bounds__(BLK) void k_sumsq(uint64_t N, const double *__restrict__ a, double *__restrict__ part,
                                               uint32_t k, uint32_t kshift) {
    __shared__ double sh[BLK];
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    double sq = 0.0;
    if (j < k)
        for (uint64_t dof = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; dof < N; dof += rstride) {
            double q = a[dof * k + j];
            sq += q * q;
        }
    double tot = block_reduce_cols(sq, kp, sh);
    if (threadIdx.x < kp) part[(uint64_t)blockIdx.x * kp + threadIdx.x] = tot;
}

__global__ __launch_bounds__(BLK) void k_reduce_to(const double *part, uint32_t nb, uint32_t kp, double *outv,
                                                   uint32_t k = 1, Ranges rg = {}) {
    __shared__ double sh[BLK];
    const uint32_t c = blockIdx.x, j = rg.ptr ? c % k : c;
    double s = reduce_partials(part, nb, kp, 1, 0, j, sh, false, rg, rg.ptr ? c / k : 0u);
    if (threadIdx.x == 0) outv[c] = s;
}

uint32_t grid_for(uint64_t rows, uint32_t kshift, uint32_t cap = MAX_PART_BLOCKS) {
    const uint64_t rows_per_block = BLK >> kshift;
    uint64_t g = (rows + rows_per_block - 1) / rows_per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

} // namespace

DesignArgs DevDesign::args() const {
    DesignArgs a{};
    a.n = n;
    a.T = (uint32_t)terms.size();
    a.codes = codes.as<uint32_t>();
    a.slopes = slopes.as<double>();
    for (size_t t = 0; t < terms.size(); ++t) a.terms[t] = terms[t].info;
    return a;
}

// ---------------------------------------------------------------------------
// Design upload: codes, slopes, weights and the per-term D' segment plans
// ---------------------------------------------------------------------------
namespace {
// D' plan of one term (or of the pieces of a fused pair) over positions in its level-sorted
// order, level l at positions [start[l], start[l + 1]): a work item per level of at most
// SEG_CHUNK positions, larger levels split into SEG_CHUNK chunks whose partial sums are
// combined in a fixed order (a one-stage fixup per level, or for single-column terms a
// sequence of segmented-sum stages of at most FIX_CHUNK partials each).  Returns the partial
// slots the plan needs; `items` receives the host copy of the work items.
uint64_t build_seg_plan(DevTerm &dt, const std::vector<uint64_t> &start, uint64_t n_levels,
                        std::vector<SegItem> &items, uint32_t chunk = SEG_CHUNK) {
    std::vector<SegFix> fixes;
    uint32_t partials = 0;
    items.clear();
    items.reserve(n_levels + start[n_levels] / chunk + 1);
    for (uint64_t l = 0; l < n_levels; ++l) {
        uint64_t b = start[l], e = start[l + 1];
        if (e - b <= chunk) {
            items.push_back({(uint32_t)l, (uint32_t)b, (uint32_t)e, (int32_t)l});
        } else {
            uint32_t first = partials;
            uint32_t count = 0;
            for (uint64_t p = b; p < e; p += chunk) {
                items.push_back({(uint32_t)l, (uint32_t)p, (uint32_t)std::min<uint64_t>(e, p + chunk),
                                 -(int32_t)(partials + 1)});
                ++partials;
                ++count;
            }
            fixes.push_back({(uint32_t)l, first, count, 0});
        }
    }
    dt.items.upload(items.data(), items.size());
    dt.n_items = (uint32_t)items.size();
    dt.fixes.upload(fixes.data(), fixes.size());
    dt.n_fixes = (uint32_t)fixes.size();
    dt.n_partials = partials;
    uint64_t slots = (uint64_t)partials * dt.info.ncols;
    dt.stages.clear();
    if (dt.info.ncols == 1) {
        // combine stages: a split level with <= FIX_CHUNK partials is summed into its output,
        // a larger one into ceil(count / FIX_CHUNK) partials of the next region
        std::vector<SegFix> cur = fixes;
        uint64_t in_base = 0, out_base = partials;
        while (!cur.empty()) {
            std::vector<SegItem> st_items;
            std::vector<SegFix> next;
            uint32_t np = 0;
            for (const SegFix &f : cur) {
                if (f.count <= FIX_CHUNK) {
                    st_items.push_back({f.level, f.first, f.first + f.count, (int32_t)f.level});
                    continue;
                }
                const uint32_t first = np;
                uint32_t cnt = 0;
                for (uint32_t q = 0; q < f.count; q += FIX_CHUNK) {
                    st_items.push_back({f.level, f.first + q, f.first + std::min(f.count, q + FIX_CHUNK),
                                        -(int32_t)(np + 1)});
                    ++np;
                    ++cnt;
                }
                next.push_back({f.level, first, cnt, 0});
            }
            DevTerm::FixStage stage;
            stage.items.upload(st_items.data(), st_items.size());
            stage.n_items = (uint32_t)st_items.size();
            stage.src_base = in_base;
            stage.dst_base = out_base;
            dt.stages.push_back(std::move(stage));
            in_base = out_base;
            out_base += np;
            cur.swap(next);
        }
        slots = out_base;
    }
    return slots;
}

// a row's level codes as one 16-byte record (terms beyond T read as 0)
__global__ __launch_bounds__(BLK) void k_codes4(uint64_t n, uint32_t T, const uint32_t *__restrict__ codes,
                                                uint4 *__restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLK) {
        uint4 r;
        r.x = codes[i];
        r.y = T > 1 ? codes[n + i] : 0u;
        r.z = T > 2 ? codes[2 * n + i] : 0u;
        r.w = T > 3 ? codes[3 * n + i] : 0u;
        out[i] = r;
    }
}

// the records in caller row order: out[r] = codes4[inv_perm[r]]
__global__ __launch_bounds__(BLK) void k_codes4_caller(uint64_t n, const uint32_t *__restrict__ inv_perm,
                                                       const uint4 *__restrict__ codes4, uint4 *__restrict__ out) {
    for (uint64_t r = (uint64_t)blockIdx.x * BLK + threadIdx.x; r < n; r += (uint64_t)gridDim.x * BLK)
        out[r] = codes4[inv_perm[r]];
}

// The fused D' of two unsorted single-intercept terms of a plain design (k_segreduce2_pair):
// the primary term P (the unsorted term with the most levels) is reduced over its rows in
// (P, Q) order, Q being the unsorted term with the fewest levels.  A run of equal Q codes
// inside one of P's work items is a piece; the pass that sums P's levels also writes every
// piece's sum, and Q's D' is then a segmented sum over the pieces in Q order.  So u is read
// once for both terms, and the extra traffic is one write and one read of the pieces.  Used
// only when the pieces number at most a quarter of the rows.
std::pair<int32_t, int32_t> pair_candidate(const Design &d, bool plain) {
    int32_t P = -1, Q = -1;
    if (!plain) return {-1, -1};
    for (size_t t = 0; t < d.terms.size(); ++t) {
        if (d.terms[t].sorted) continue;
        if (P < 0 || d.terms[t].n_levels > d.terms[P].n_levels) P = (int32_t)t;
        if (Q < 0 || d.terms[t].n_levels <= d.terms[Q].n_levels) Q = (int32_t)t;
    }
    if (P < 0 || Q < 0 || P == Q) return {-1, -1};
    return {P, Q};
}

// An unsorted term's code-sorted permutation (stable) and its position-ordered weights, on the
// device from the uploaded codes and weights.
void term_permutation(DevDesign &dv, DevTerm &dt, uint32_t t, uint64_t n_levels) {
    const uint64_t n = dv.n;
    dt.perm.alloc(n * 4);
    dev_stable_argsort(dv.codes.as<uint32_t>() + (uint64_t)t * n, n, n_levels, dt.perm.as<uint32_t>(), nullptr);
    if (dv.w.p) { // weights in this term's sorted order (sequential reads)
        dt.w_pos.alloc(n * 8);
        dt.sw_pos.alloc(n * 8);
        dev_gather(dv.w.as<double>(), dt.perm.as<uint32_t>(), n, dt.w_pos.as<double>(), nullptr);
        dev_gather(dv.sw.as<double>(), dt.perm.as<uint32_t>(), n, dt.sw_pos.as<double>(), nullptr);
        WX_HIP(hipStreamSynchronize(nullptr));
    }
}

// Returns false when the design has no qualifying pair (the terms keep their own plans).
bool build_pair_plan(const Design &d, const std::vector<std::vector<uint64_t>> &starts, DevDesign &dv) {
    dv.pair = PairPlan{};
    const auto [P, Q] = pair_candidate(d, dv.plain);
    if (P < 0) return false;
    const uint64_t n = d.n_obs, LQ = d.terms[Q].n_levels, LP = d.terms[P].n_levels;
    const uint32_t *cp = dv.codes.as<uint32_t>() + (uint64_t)P * n, *cq = dv.codes.as<uint32_t>() + (uint64_t)Q * n;
    hipStream_t st = nullptr;
    // rows in (P, Q) order (a stable sort of the composite key) and Q's code at each position
    DevBuf perm, qpos;
    perm.alloc(n * 4);
    qpos.alloc(n * 4);
    dev_stable_argsort2(cp, cq, LP, LQ, n, perm.as<uint32_t>(), st);
    dev_gather_u32(cq, perm.as<uint32_t>(), n, qpos.as<uint32_t>(), st);
    // P's work items of up to PAIR_CHUNK rows, and the pieces they hold (the kernel's rule)
    DevTerm pplan;
    pplan.info = dv.terms[P].info;
    std::vector<SegItem> items;
    const uint64_t pslots = build_seg_plan(pplan, starts[P], LP, items, PAIR_CHUNK);
    DevBuf piece0, piece_q;
    piece0.alloc((items.size() + 1) * 4);
    const uint64_t n_pieces =
        dev_pieces(qpos.as<uint32_t>(), n, pplan.items.as<SegItem>(), pplan.n_items, piece0.as<uint32_t>(), piece_q, st);
    if (n_pieces > n / 4 || n_pieces > 0xFFFFFFFFull) return false;
    PairPlan &pp = dv.pair;
    pp.p = P;
    pp.q = Q;
    pp.n_pieces = (uint32_t)n_pieces;
    DevTerm &pt = dv.terms[P];
    pt.items = std::move(pplan.items);
    pt.n_items = pplan.n_items;
    pt.fixes = std::move(pplan.fixes);
    pt.n_fixes = pplan.n_fixes;
    pt.n_partials = pplan.n_partials;
    pt.stages = std::move(pplan.stages);
    dv.max_partials = std::max<uint32_t>(dv.max_partials, (uint32_t)pslots);
    pt.perm = std::move(perm);
    if (dv.w.p) {
        pt.w_pos.alloc(n * 8);
        pt.sw_pos.alloc(n * 8);
        dev_gather(dv.w.as<double>(), pt.perm.as<uint32_t>(), n, pt.w_pos.as<double>(), st);
        dev_gather(dv.sw.as<double>(), pt.perm.as<uint32_t>(), n, pt.sw_pos.as<double>(), st);
    }
    pp.qcode_pos = std::move(qpos);
    pp.item_piece0 = std::move(piece0);
    // the pieces are stored in Q order (piece -> its slot there), so Q's reduction reads them
    // contiguously
    DevTerm &qt = pp.qt;
    qt.info = dv.terms[Q].info;
    DevBuf pperm;
    pperm.alloc(n_pieces * 4);
    dev_stable_argsort(piece_q.as<uint32_t>(), n_pieces, LQ, pperm.as<uint32_t>(), st);
    pp.piece_slot.alloc(n_pieces * 4);
    dev_invert_perm(pperm.as<uint32_t>(), n_pieces, pp.piece_slot.as<uint32_t>(), st);
    std::vector<uint32_t> hq(n_pieces);
    WX_HIP(hipMemcpyAsync(hq.data(), piece_q.p, n_pieces * 4, hipMemcpyDeviceToHost, st));
    WX_HIP(hipStreamSynchronize(st));
    std::vector<uint64_t> start(LQ + 1, 0);
    for (uint32_t c : hq) start[c + 1]++;
    for (uint64_t l = 0; l < LQ; ++l) start[l + 1] += start[l];
    std::vector<SegItem> qitems;
    const uint64_t slots = build_seg_plan(qt, start, LQ, qitems);
    dv.max_partials = std::max<uint32_t>(dv.max_partials, (uint32_t)slots);
    return true;
}

std::vector<SegItem> download_items(const DevBuf &b, uint32_t n) {
    std::vector<SegItem> h(n);
    if (n) WX_HIP(hipMemcpy(h.data(), b.p, (size_t)n * sizeof(SegItem), hipMemcpyDeviceToHost));
    return h;
}

// `in`'s items in absolute terms: positions in region `src`, partial slots in region `dst`,
// output rows at `offset` (complete items name their output row as their destination)
void append_absolute(std::vector<SegItem> &out, const std::vector<SegItem> &in, uint64_t src, uint64_t dst,
                     uint64_t offset) {
    for (SegItem it : in) {
        it.begin += (uint32_t)src;
        it.end += (uint32_t)src;
        it.level += (uint32_t)offset;
        if (it.dest >= 0) it.dest = (int32_t)it.level;
        else if (it.dest != SEG_SKIP) it.dest = -(int32_t)(dst + (uint64_t)(-(int64_t)it.dest - 1)) - 1;
        out.push_back(it);
    }
}

// The merged fix-up launches (MergedFix) of a plain design of exactly the fused sorted term S and
// the fused pair (P, Q); off when the slots or the output rows would not fit the items' 31 bits.
void build_merged_fix(const Design &d, DevDesign &dv) {
    dv.mfix = MergedFix{};
    const PairPlan &pp = dv.pair;
    if (!dv.plain || d.terms.size() != 3 || pp.p < 0 || dv.fuse_term < 0 || dv.fuse_term == pp.p ||
        dv.fuse_term == pp.q)
        return;
    const DevTerm &S = dv.terms[dv.fuse_term], &P = dv.terms[pp.p], &Q = pp.qt;
    auto slots_of = [](const DevTerm &t) {
        return t.stages.empty() ? (uint64_t)t.n_partials : t.stages.back().dst_base;
    };
    // every stage's output region follows its input region, so a term's slots end where its last
    // stage's output region starts (that stage writes complete levels only)
    const uint64_t s_slots = std::max<uint64_t>(slots_of(S), S.n_partials);
    const uint64_t p_slots = std::max<uint64_t>(slots_of(P), P.n_partials);
    const uint64_t q_slots = std::max<uint64_t>(slots_of(Q), Q.n_partials);
    MergedFix m;
    m.base_p = s_slots;
    m.base_pieces = m.base_p + p_slots;
    const uint64_t base_q = m.base_pieces + pp.n_pieces;
    m.slots = base_q + q_slots;
    if (m.slots >= (1ull << 31) || dv.N >= (1ull << 31)) return;
    const size_t L = std::max({S.stages.size(), P.stages.size(), Q.stages.size() + 1});
    for (size_t lv = 0; lv < L; ++lv) {
        std::vector<SegItem> items;
        if (lv < S.stages.size()) {
            const DevTerm::FixStage &g = S.stages[lv];
            append_absolute(items, download_items(g.items, g.n_items), g.src_base, g.dst_base, S.info.offset);
        }
        if (lv < P.stages.size()) {
            const DevTerm::FixStage &g = P.stages[lv];
            append_absolute(items, download_items(g.items, g.n_items), m.base_p + g.src_base, m.base_p + g.dst_base,
                            P.info.offset);
        }
        if (lv == 0) {
            append_absolute(items, download_items(Q.items, Q.n_items), m.base_pieces, base_q, Q.info.offset);
        } else if (lv - 1 < Q.stages.size()) {
            const DevTerm::FixStage &g = Q.stages[lv - 1];
            append_absolute(items, download_items(g.items, g.n_items), base_q + g.src_base, base_q + g.dst_base,
                            Q.info.offset);
        }
        m.levels.emplace_back();
        m.levels.back().upload(items.data(), items.size());
        m.n_items.push_back((uint32_t)items.size());
    }
    dv.max_partials = std::max<uint32_t>(dv.max_partials, (uint32_t)m.slots);
    dv.mfix = std::move(m);
}
} // namespace

void upload_design(Solver &s, const std::vector<double> &weights_internal, uint32_t tile) {
    const Design &d = s.design;
    DevDesign &dv = s.dev;
    // WITHIN_AMD_VERBOSE: per-phase wall times of the upload on stderr
    const bool verbose = std::getenv("WITHIN_AMD_VERBOSE") != nullptr;
    auto t_last = std::chrono::steady_clock::now();
    auto phase = [&](const char *name) {
        if (!verbose) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[within_amd] device design: %s %.1f ms\n", name,
                     std::chrono::duration<double, std::milli>(now - t_last).count());
        t_last = now;
    };
    if (d.terms.size() > MAX_TERMS)
        fail(WX_ERR_INVALID_INPUT, "the device path supports at most " + std::to_string(MAX_TERMS) + " terms");
    if (tile < 16 || tile > 4096) fail(WX_ERR_INVALID_INPUT, "work-item tile must be 16..4096 rows");
    dv.n = d.n_obs;
    dv.N = d.n_dofs;
    dv.tile = tile;
    const uint64_t n = d.n_obs;
    hipStream_t st0 = nullptr;
    // codes, weights and the row permutation go up once; everything derived from them (square
    // roots, inverse permutation, the terms' sort permutations, position-ordered weights, code
    // records) is formed on the device
    dv.codes.upload(d.codes.data(), d.codes.size());
    dv.slopes.upload(d.slopes.data(), d.slopes.size());
    if (!weights_internal.empty()) {
        dv.w.upload(weights_internal.data(), weights_internal.size());
        dv.sw.alloc(n * 8);
        dev_sqrt(dv.w.as<double>(), n, dv.sw.as<double>(), st0);
    }
    if (!d.obs_perm.empty()) {
        dv.obs_perm.upload(d.obs_perm.data(), d.obs_perm.size());
        dv.inv_perm.alloc(n * 4);
        dev_invert_perm(dv.obs_perm.as<uint32_t>(), n, dv.inv_perm.as<uint32_t>(), st0);
        if (!weights_internal.empty()) { // sqrt-weights in caller order for the right-hand-side load
            dv.sw_caller.alloc(n * 8);
            dev_gather(dv.sw.as<double>(), dv.inv_perm.as<uint32_t>(), n, dv.sw_caller.as<double>(), st0);
        }
    }
    WX_HIP(hipStreamSynchronize(st0));
    if (s.precond && s.precond->kind == Precond::DIAGONAL)
        dv.inv_diag.upload(s.precond->inv_diag.data(), s.precond->inv_diag.size());
    phase("codes, weights, permutation");
    dv.terms.clear();
    dv.terms.resize(d.terms.size());
    dv.max_partials = 0;
    std::vector<std::vector<SegItem>> host_items(d.terms.size()); // every term's D' work items
    dv.plain = true;
    for (const Term &tm : d.terms) dv.plain = dv.plain && tm.ncols() == 1 && tm.loads[0] < 0;
    // level starts of every term (counts in parallel over terms)
    std::vector<std::vector<uint64_t>> starts(d.terms.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t t = 0; t < (int64_t)d.terms.size(); ++t) {
        const uint32_t *col = d.level_column((uint64_t)t);
        std::vector<uint64_t> &start = starts[t];
        start.assign(d.terms[t].n_levels + 1, 0);
        for (uint64_t i = 0; i < n; ++i) start[col[i] + 1]++;
        for (uint64_t l = 0; l < d.terms[t].n_levels; ++l) start[l + 1] += start[l];
    }
    const int32_t pair_p = pair_candidate(d, dv.plain).first;
    for (size_t t = 0; t < d.terms.size(); ++t) {
        const Term &tm = d.terms[t];
        DevTerm &dt = dv.terms[t];
        if (tm.ncols() > MAX_COLS)
            fail(WX_ERR_INVALID_INPUT, "the device path supports at most " + std::to_string(MAX_COLS) +
                                           " coefficient columns per effect");
        dt.info.n_levels = tm.n_levels;
        dt.info.offset = tm.offset;
        dt.info.ncols = (uint32_t)tm.ncols();
        for (uint32_t c = 0; c < MAX_COLS; ++c) dt.info.load[c] = c < tm.ncols() ? (int32_t)tm.loads[c] : -1;
        // the fused pair's primary term gets its permutation from the pair plan
        if (!tm.sorted && (int32_t)t != pair_p) term_permutation(dv, dt, (uint32_t)t, tm.n_levels);
        const uint64_t slots = build_seg_plan(dt, starts[t], tm.n_levels, host_items[t], tile);
        dv.max_partials = std::max<uint32_t>(dv.max_partials, (uint32_t)slots);
    }
    phase("D' plans");
    if (!build_pair_plan(d, starts, dv) && pair_p >= 0) // no fused pair: P's own permutation
        term_permutation(dv, dv.terms[pair_p], (uint32_t)pair_p, d.terms[pair_p].n_levels);
    phase("fused pair plan");
    // a row's codes as one aligned 16-byte record: the fused gather and the demean read them with
    // one load instead of one per term
    dv.codes4.release();
    dv.codes4_caller.release();
    if (dv.plain && d.terms.size() <= 4 && !d.terms.empty()) {
        dv.codes4.alloc(n * 16);
        const uint32_t g = (uint32_t)std::min<uint64_t>((n + BLK - 1) / BLK, 1u << 20);
        hipLaunchKernelGGL(k_codes4, dim3(g), dim3(BLK), 0, st0, n, (uint32_t)d.terms.size(),
                           dv.codes.as<uint32_t>(), dv.codes4.as<uint4>());
        if (dv.inv_perm.p) { // the demeaned output's caller-order pass reads them row by row
            dv.codes4_caller.alloc(n * 16);
            hipLaunchKernelGGL(k_codes4_caller, dim3(g), dim3(BLK), 0, st0, n, dv.inv_perm.as<uint32_t>(),
                               dv.codes4.as<uint4>(), dv.codes4_caller.as<uint4>());
        }
        WX_HIP(hipGetLastError());
        WX_HIP(hipStreamSynchronize(st0));
    }
    // the sorted single-intercept term with the most work items gets its D' fused into the gather
    dv.fuse_term = -1;
    for (size_t t = 0; t < d.terms.size(); ++t) {
        const Term &tm = d.terms[t];
        if (!tm.sorted || tm.ncols() != 1 || tm.loads[0] >= 0) continue;
        if (dv.fuse_term < 0 || dv.terms[t].n_items > dv.terms[dv.fuse_term].n_items) dv.fuse_term = (int32_t)t;
    }
    if (dv.fuse_term >= 0) {
        // the fused gather's items: runs of consecutive complete levels (empty ones included)
        // merged up to SEG_CHUNK rows and GATHER_SPAN levels; chunks of split levels keep their
        // partial slots
        constexpr uint32_t GATHER_SPAN = 256;
        const std::vector<SegItem> &src = host_items[dv.fuse_term];
        std::vector<SegItem> g;
        g.reserve(src.size());
        for (const SegItem &it : src) {
            if (!g.empty() && it.dest >= 0) {
                SegItem &m = g.back();
                if (m.dest >= 0 && it.level == (uint32_t)m.dest + 1 && it.begin == m.end &&
                    (it.end - m.begin) <= tile && (it.level - m.level) < GATHER_SPAN) {
                    m.end = it.end;
                    m.dest = it.dest;
                    continue;
                }
            }
            g.push_back(it);
        }
        DevTerm &ft = dv.terms[dv.fuse_term];
        ft.gitems.upload(g.data(), g.size());
        ft.n_gitems = (uint32_t)g.size();
    }
    build_merged_fix(d, dv);
    phase("code records and gather items");
}

namespace {
void clone_buf(DevBuf &dst, const DevBuf &src) {
    if (!src.p) {
        dst.release();
        return;
    }
    dst.alloc(src.bytes);
    WX_HIP(hipMemcpy(dst.p, src.p, src.bytes, hipMemcpyDeviceToDevice));
}

// A term's D' plan (everything but its position-ordered weights) copied on the device.
void clone_term_plan(DevTerm &dt, const DevTerm &bt) {
    dt.info = bt.info;
    clone_buf(dt.perm, bt.perm);
    clone_buf(dt.items, bt.items);
    dt.n_items = bt.n_items;
    clone_buf(dt.fixes, bt.fixes);
    dt.n_fixes = bt.n_fixes;
    dt.n_partials = bt.n_partials;
    clone_buf(dt.gitems, bt.gitems);
    dt.n_gitems = bt.n_gitems;
    dt.stages.resize(bt.stages.size());
    for (size_t q = 0; q < bt.stages.size(); ++q) {
        clone_buf(dt.stages[q].items, bt.stages[q].items);
        dt.stages[q].n_items = bt.stages[q].n_items;
        dt.stages[q].src_base = bt.stages[q].src_base;
        dt.stages[q].dst_base = bt.stages[q].dst_base;
    }
}
} // namespace

// The device design of `base` with new observation weights (internal order; empty:
// unweighted): every index structure is copied on the device, only the weight arrays (and the
// diagonal preconditioner) are formed anew.
void reweight_design(Solver &s, const Solver &base, const std::vector<double> &w) {
    const DevDesign &b = base.dev;
    DevDesign &dv = s.dev;
    dv.n = b.n;
    dv.N = b.N;
    clone_buf(dv.codes, b.codes);
    clone_buf(dv.slopes, b.slopes);
    clone_buf(dv.obs_perm, b.obs_perm);
    clone_buf(dv.inv_perm, b.inv_perm);
    clone_buf(dv.codes4, b.codes4);
    clone_buf(dv.codes4_caller, b.codes4_caller);
    dv.plain = b.plain;
    dv.fuse_term = b.fuse_term;
    dv.max_partials = b.max_partials;
    dv.sw.release();
    dv.w.release();
    dv.sw_caller.release();
    std::vector<double> sw(w.size());
    for (size_t i = 0; i < w.size(); ++i) sw[i] = std::sqrt(w[i]);
    if (!w.empty()) {
        dv.sw.upload(sw.data(), sw.size());
        dv.w.upload(w.data(), w.size());
        if (!s.design.obs_perm.empty()) {
            std::vector<double> swc(w.size());
            for (uint64_t i = 0; i < w.size(); ++i) swc[s.design.obs_perm[i]] = sw[i];
            dv.sw_caller.upload(swc.data(), swc.size());
        }
    }
    dv.inv_diag.release();
    if (s.precond && s.precond->kind == Precond::DIAGONAL)
        dv.inv_diag.upload(s.precond->inv_diag.data(), s.precond->inv_diag.size());
    dv.terms.clear();
    dv.terms.resize(b.terms.size());
    std::vector<uint32_t> perm;
    // the fused pair's index structures (its primary term's weights follow from that term's
    // permutation below)
    dv.pair = PairPlan{};
    dv.pair.p = b.pair.p;
    dv.pair.q = b.pair.q;
    dv.pair.n_pieces = b.pair.n_pieces;
    clone_buf(dv.pair.qcode_pos, b.pair.qcode_pos);
    clone_buf(dv.pair.item_piece0, b.pair.item_piece0);
    clone_buf(dv.pair.piece_slot, b.pair.piece_slot);
    clone_term_plan(dv.pair.qt, b.pair.qt);
    dv.mfix = MergedFix{};
    dv.mfix.base_p = b.mfix.base_p;
    dv.mfix.base_pieces = b.mfix.base_pieces;
    dv.mfix.slots = b.mfix.slots;
    dv.mfix.n_items = b.mfix.n_items;
    dv.mfix.levels.resize(b.mfix.levels.size());
    for (size_t q = 0; q < b.mfix.levels.size(); ++q) clone_buf(dv.mfix.levels[q], b.mfix.levels[q]);
    for (size_t t = 0; t < b.terms.size(); ++t) {
        const DevTerm &bt = b.terms[t];
        DevTerm &dt = dv.terms[t];
        clone_term_plan(dt, bt);
        if (bt.perm.p && !w.empty()) { // weights in this term's sorted order
            perm.resize(b.n);
            WX_HIP(hipMemcpy(perm.data(), bt.perm.p, b.n * sizeof(uint32_t), hipMemcpyDeviceToHost));
            std::vector<double> wp(b.n), swp(b.n);
            for (uint64_t q = 0; q < b.n; ++q) {
                wp[q] = w[perm[q]];
                swp[q] = sw[perm[q]];
            }
            dt.w_pos.upload(wp.data(), wp.size());
            dt.sw_pos.upload(swp.data(), swp.size());
        }
    }
}

namespace {

// Combines the split levels of one term's D' reduction.
void launch_segfix(const DevTerm &dt, double *partials, double *out, const int32_t *active, uint32_t k,
                   uint32_t kshift, hipStream_t st) {
    if (!dt.n_fixes) return;
    if (!dt.stages.empty()) {
        const uint32_t groups_per_block = BLK >> kshift;
        TermInfo fi = dt.info; // partials are already weighted by the column's load: plain sums
        fi.load[0] = -1;
        for (const DevTerm::FixStage &sg : dt.stages) {
            const uint32_t grid = std::max<uint32_t>(1, (sg.n_items + groups_per_block - 1) / groups_per_block);
            if (k >= 2 && (k & 1) == 0) {
                const uint32_t grid2 = std::max<uint32_t>(1, (sg.n_items + 2 * groups_per_block - 1) /
                                                                 (2 * groups_per_block));
                hipLaunchKernelGGL((k_segreduce2<false>), dim3(grid2), dim3(BLK), 0, st, fi, nullptr,
                                   sg.items.as<SegItem>(), sg.n_items, nullptr, partials + sg.src_base * k, out,
                                   partials + sg.dst_base * k, active, k, kshift);
            } else {
                hipLaunchKernelGGL((k_segreduce<1, false>), dim3(grid), dim3(BLK), 0, st, DesignArgs{}, fi, nullptr,
                                   sg.items.as<SegItem>(), sg.n_items, nullptr, partials + sg.src_base * k, out,
                                   partials + sg.dst_base * k, active, k, kshift);
            }
        }
        return;
    }
    const uint32_t fgrid = dt.n_fixes;
    switch (dt.info.ncols) {
    case 1: hipLaunchKernelGGL((k_segfix<1>), dim3(fgrid), dim3(BLK), 0, st, dt.info, dt.fixes.as<SegFix>(), dt.n_fixes, partials, out, active, k, kshift); break;
    case 2: hipLaunchKernelGGL((k_segfix<2>), dim3(fgrid), dim3(BLK), 0, st, dt.info, dt.fixes.as<SegFix>(), dt.n_fixes, partials, out, active, k, kshift); break;
    case 3: hipLaunchKernelGGL((k_segfix<3>), dim3(fgrid), dim3(BLK), 0, st, dt.info, dt.fixes.as<SegFix>(), dt.n_fixes, partials, out, active, k, kshift); break;
    default: hipLaunchKernelGGL((k_segfix<4>), dim3(fgrid), dim3(BLK), 0, st, dt.info, dt.fixes.as<SegFix>(), dt.n_fixes, partials, out, active, k, kshift); break;
    }
}

// Launches the D' reduction of every term (but `skip`): out (N x k) = D' (f .* src) with
// f = 1 (which = 0), sqrt(w) (which = 1) or w (which = 2).
// `pieces` (n_pieces x k, may be NULL) holds the fused pair's piece sums; the pair runs fused
// when it is given and the width is even.
void launch_adjoint(const DevDesign &dv, const DesignArgs &da, int which, const double *src, double *out,
                    double *partials, const int32_t *active, uint32_t k, uint32_t kshift, hipStream_t st,
                    int32_t skip = -1, double *pieces = nullptr) {
    const PairPlan &pp = dv.pair;
    const bool pair = pp.p >= 0 && pieces && k >= 2 && (k & 1) == 0 && pp.p != skip && pp.q != skip;
    for (size_t t = 0; t < dv.terms.size(); ++t) {
        if ((int32_t)t == skip || (pair && (int32_t)t == pp.q)) continue;
        const DevTerm &dt = dv.terms[t];
        const uint32_t groups_per_block = BLK >> kshift;
        if (pair && (int32_t)t == pp.p) {
            // P's levels and the pieces in one pass over u, then Q's levels from the pieces
            const double *f = which == 1 ? dt.sw_pos.as<double>() : which == 2 ? dt.w_pos.as<double>() : nullptr;
            // one item per row group
            const uint32_t grid2 = std::max<uint32_t>(
                1, std::min<uint32_t>((dt.n_items + 2 * groups_per_block - 1) / (2 * groups_per_block) * 2,
                                      65535u * 8));
            if (f)
                hipLaunchKernelGGL((k_segreduce2_pair<true>), dim3(grid2), dim3(BLK), 0, st, dt.info,
                                   dt.perm.as<uint32_t>(), dt.items.as<SegItem>(), dt.n_items,
                                   pp.item_piece0.as<uint32_t>(), pp.piece_slot.as<uint32_t>(),
                                   pp.qcode_pos.as<uint32_t>(), f, src, out, partials, pieces, active, k, kshift,
                                   ColState{}, BetaSide{});
            else
                hipLaunchKernelGGL((k_segreduce2_pair<false>), dim3(grid2), dim3(BLK), 0, st, dt.info,
                                   dt.perm.as<uint32_t>(), dt.items.as<SegItem>(), dt.n_items,
                                   pp.item_piece0.as<uint32_t>(), pp.piece_slot.as<uint32_t>(),
                                   pp.qcode_pos.as<uint32_t>(), f, src, out, partials, pieces, active, k, kshift,
                                   ColState{}, BetaSide{});
            launch_segfix(dt, partials, out, active, k, kshift, st);
            const DevTerm &qt = pp.qt;
            const uint32_t gq = std::max<uint32_t>(
                1, std::min<uint32_t>((qt.n_items + 2 * groups_per_block - 1) / (2 * groups_per_block), 65535u * 8));
            hipLaunchKernelGGL((k_segreduce2<false>), dim3(gq), dim3(BLK), 0, st, qt.info, nullptr,
                               qt.items.as<SegItem>(), qt.n_items, nullptr, pieces, out, partials, active, k, kshift);
            launch_segfix(qt, partials, out, active, k, kshift, st);
            continue;
        }
        uint32_t grid = (dt.n_items + groups_per_block - 1) / groups_per_block;
        grid = std::max<uint32_t>(1, std::min<uint32_t>(grid, 65535u * 8));
        const uint32_t *perm = dt.perm.p ? dt.perm.as<uint32_t>() : nullptr;
        // the factor in the term's sorted order: the plain arrays for the sorted term
        const double *f = nullptr;
        if (which == 1) f = perm ? dt.sw_pos.as<double>() : dv.sw.as<double>();
        if (which == 2) f = perm ? dt.w_pos.as<double>() : dv.w.as<double>();
#define WX_SEG(NC, SC)                                                                                           \
    hipLaunchKernelGGL((k_segreduce<NC, SC>), dim3(grid), dim3(BLK), 0, st, da, dt.info, perm,                   \
                       dt.items.as<SegItem>(), dt.n_items, f, src, out, partials, active, k, kshift)
#define WX_SEG_NC(SC)                                                                                            \
    switch (dt.info.ncols) {                                                                                     \
    case 1: WX_SEG(1, SC); break;                                                                                \
    case 2: WX_SEG(2, SC); break;                                                                                \
    case 3: WX_SEG(3, SC); break;                                                                                \
    default: WX_SEG(4, SC); break;                                                                               \
    }
        if (dt.info.ncols == 1 && dt.info.load[0] < 0 && k >= 2 && (k & 1) == 0) {
            // single intercept column, even width: two columns per lane
            const uint32_t grid2 = std::max<uint32_t>(
                1, std::min<uint32_t>((dt.n_items + 2 * groups_per_block - 1) / (2 * groups_per_block), 65535u * 8));
            if (f)
                hipLaunchKernelGGL((k_segreduce2<true>), dim3(grid2), dim3(BLK), 0, st, dt.info, perm,
                                   dt.items.as<SegItem>(), dt.n_items, f, src, out, partials, active, k, kshift);
            else
                hipLaunchKernelGGL((k_segreduce2<false>), dim3(grid2), dim3(BLK), 0, st, dt.info, perm,
                                   dt.items.as<SegItem>(), dt.n_items, f, src, out, partials, active, k, kshift);
        } else if (f) { WX_SEG_NC(true) }
        else { WX_SEG_NC(false) }
#undef WX_SEG_NC
#undef WX_SEG
        launch_segfix(dt, partials, out, active, k, kshift, st);
    }
    WX_HIP(hipGetLastError());
}

// The fused pair's D' and the merged fix-ups (MergedFix) after the fused gather has left the
// sorted term's level sums in `out` and its partials at the start of `partials`; P's pass writes
// its partials and the pieces into their regions of the same buffer.
// With bs.ncol > 0 the pair launch also runs the beta reduction (BetaSide) on column state cs.
void launch_adjoint_merged(const DevDesign &dv, int which, const double *src, double *out, double *partials,
                           const int32_t *active, uint32_t k, uint32_t kshift, hipStream_t st,
                           const ColState &cs = ColState{}, const BetaSide &bs = BetaSide{}) {
    const PairPlan &pp = dv.pair;
    const MergedFix &m = dv.mfix;
    const DevTerm &dt = dv.terms[pp.p];
    const uint32_t groups_per_block = BLK >> kshift;
    const double *f = which == 1 ? dt.sw_pos.as<double>() : which == 2 ? dt.w_pos.as<double>() : nullptr;
    const uint32_t grid2 = std::max<uint32_t>(
        1, std::min<uint32_t>((dt.n_items + 2 * groups_per_block - 1) / (2 * groups_per_block) * 2, 65535u * 8));
    double *pparts = partials + m.base_p * k, *pieces = partials + m.base_pieces * k;
    const uint32_t gtot = grid2 + bs.ncol;
    if (f)
        hipLaunchKernelGGL((k_segreduce2_pair<true>), dim3(gtot), dim3(BLK), 0, st, dt.info, dt.perm.as<uint32_t>(),
                           dt.items.as<SegItem>(), dt.n_items, pp.item_piece0.as<uint32_t>(),
                           pp.piece_slot.as<uint32_t>(), pp.qcode_pos.as<uint32_t>(), f, src, out, pparts, pieces,
                           active, k, kshift, cs, bs);
    else
        hipLaunchKernelGGL((k_segreduce2_pair<false>), dim3(gtot), dim3(BLK), 0, st, dt.info,
                           dt.perm.as<uint32_t>(), dt.items.as<SegItem>(), dt.n_items,
                           pp.item_piece0.as<uint32_t>(), pp.piece_slot.as<uint32_t>(),
                           pp.qcode_pos.as<uint32_t>(), f, src, out, pparts, pieces, active, k, kshift, cs, bs);
    TermInfo z = dt.info; // absolute output rows, plain sums
    z.offset = 0;
    z.load[0] = -1;
    for (size_t lv = 0; lv < m.levels.size(); ++lv) {
        const uint32_t n = m.n_items[lv];
        if (!n) continue;
        const uint32_t g =
            std::max<uint32_t>(1, std::min<uint32_t>((n + 2 * groups_per_block - 1) / (2 * groups_per_block), 65535u * 8));
        hipLaunchKernelGGL((k_segreduce2<false>), dim3(g), dim3(BLK), 0, st, z, nullptr, m.levels[lv].as<SegItem>(), n,
                           nullptr, partials, out, partials, active, k, kshift);
    }
    WX_HIP(hipGetLastError());
}

// D v gather launch, specialised on the term count when every term is one intercept column.
template <bool W, bool F>
void launch_gather_t(const DevDesign &dv, const DesignArgs &da, uint32_t grid, const double *sw, const double *v,
                     const double *uin, double *u, const ColState &cs, double *part, uint32_t k, uint32_t kshift,
                     hipStream_t st) {
#define WX_G(TT, PL)                                                                                             \
    hipLaunchKernelGGL((k_gather_axpy<W, F, TT, PL>), dim3(grid), dim3(BLK), 0, st, da, sw, v, uin, u, cs, part, \
                       k, kshift)
    switch (dv.plain ? da.T : 0) {
    case 1: WX_G(1, true); break;
    case 2: WX_G(2, true); break;
    case 3: WX_G(3, true); break;
    case 4: WX_G(4, true); break;
    default: WX_G(0, false); break;
    }
#undef WX_G
}

void launch_gather(const DevDesign &dv, const DesignArgs &da, bool first, uint32_t grid, const double *v,
                   const double *uin, double *u, const ColState &cs, double *part, uint32_t k, uint32_t kshift,
                   hipStream_t st) {
    const double *sw = dv.sw.p ? dv.sw.as<double>() : nullptr;
    if (sw) {
        if (first) launch_gather_t<true, true>(dv, da, grid, sw, v, uin, u, cs, part, k, kshift, st);
        else launch_gather_t<true, false>(dv, da, grid, sw, v, uin, u, cs, part, k, kshift, st);
    } else {
        if (first) launch_gather_t<false, true>(dv, da, grid, sw, v, uin, u, cs, part, k, kshift, st);
        else launch_gather_t<false, false>(dv, da, grid, sw, v, uin, u, cs, part, k, kshift, st);
    }
}

// Gather fused with the D' of the sorted term dv.fuse_term (k_gather_seg); returns the grid,
// i.e. the number of norm partials written.
// A packed design (pk) launches exactly one item per row group, so that every workgroup's
// items belong to one member, and names the member's column-state base per workgroup.
template <bool W, bool F>
uint32_t launch_gather_seg_t(const DevDesign &dv, const DesignArgs &da, const double *sw, const double *v,
                             const double *uin, double *u, double *atu, double *partials, const ColState &cs,
                             double *part, uint32_t k, uint32_t kshift, hipStream_t st, const Pack *pk) {
    const DevTerm &dt = dv.terms[dv.fuse_term];
    const uint32_t groups_per_block = BLK >> kshift;
    // narrow rows (fewer than 8 lanes): one level per item, so there are enough row groups to
    // fill the device, and ordinary loads (non-temporal ones re-fetch lines for 8-byte rows)
    const bool wide = kshift >= 3;
    const bool use_g = pk || wide; // a pack's padded item list is always the gather's
    const uint32_t n_it = use_g ? dt.n_gitems : dt.n_items;
    const uint32_t grid = pk ? pk->n_item_blocks
                             : std::max<uint32_t>(1, std::min<uint32_t>((n_it + groups_per_block - 1) /
                                                                            groups_per_block,
                                                                        MAX_FUSED_BLOCKS));
    const uint32_t *blk_col = pk ? pk->item_col.as<uint32_t>() : nullptr;
    const SegItem *items = use_g ? dt.gitems.as<SegItem>() : dt.items.as<SegItem>();
    const uint32_t *fcodes = da.codes + (uint64_t)dv.fuse_term * da.n;
    if (!pk && wide && dv.codes4.p && (k & 1) == 0) {
        // one wavefront per item, two columns per lane
        const uint32_t wpb = BLK / 64;
        const uint32_t gridw = std::max<uint32_t>(1, std::min<uint32_t>((n_it + wpb - 1) / wpb, GATHER_WAVE_BLOCKS));
#define WX_GW(TT)                                                                                                \
    hipLaunchKernelGGL((k_gather_wave<W, F, TT, GATHER_WAVE_STEPS>), dim3(gridw), dim3(BLK), 0, st, da, dt.info,  \
                       dv.codes4.as<uint4>(), (uint32_t)dv.fuse_term, items, n_it, sw, v, uin, u, atu, partials,      \
                       cs, part, k, kshift)
        switch (da.T) {
        case 1: WX_GW(1); break;
        case 2: WX_GW(2); break;
        case 3: WX_GW(3); break;
        default: WX_GW(4); break;
        }
#undef WX_GW
        return gridw;
    }
#define WX_GS(TT, PL)                                                                                            \
    if (wide)                                                                                                    \
        hipLaunchKernelGGL((k_gather_seg<W, F, TT, PL, true>), dim3(grid), dim3(BLK), 0, st, da, dt.info, fcodes, \
                           items, n_it, sw, v, uin, u, atu, partials, cs, part, k, kshift, blk_col);              \
    else                                                                                                         \
        hipLaunchKernelGGL((k_gather_seg<W, F, TT, PL, false>), dim3(grid), dim3(BLK), 0, st, da, dt.info,        \
                           fcodes, items, n_it, sw, v, uin, u, atu, partials, cs, part, k, kshift, blk_col)
    switch (dv.plain ? da.T : 0) {
    case 1: WX_GS(1, true); break;
    case 2: WX_GS(2, true); break;
    case 3: WX_GS(3, true); break;
    case 4: WX_GS(4, true); break;
    default: WX_GS(0, false); break;
    }
#undef WX_GS
    return grid;
}

uint32_t launch_gather_seg(const DevDesign &dv, const DesignArgs &da, bool first, const double *v, const double *uin,
                           double *u, double *atu, double *partials, const ColState &cs, double *part, uint32_t k,
                           uint32_t kshift, hipStream_t st, const Pack *pk = nullptr) {
    const double *sw = dv.sw.p ? dv.sw.as<double>() : nullptr;
    if (sw)
        return first ? launch_gather_seg_t<true, true>(dv, da, sw, v, uin, u, atu, partials, cs, part, k, kshift, st, pk)
                     : launch_gather_seg_t<true, false>(dv, da, sw, v, uin, u, atu, partials, cs, part, k, kshift, st,
                                                        pk);
    return first ? launch_gather_seg_t<false, true>(dv, da, sw, v, uin, u, atu, partials, cs, part, k, kshift, st, pk)
                 : launch_gather_seg_t<false, false>(dv, da, sw, v, uin, u, atu, partials, cs, part, k, kshift, st, pk);
}

void launch_demean(const DevDesign &dv, const DesignArgs &da, uint32_t grid, const double *x, const double *y_int,
                   double *dm_int, uint32_t k, uint32_t kshift, hipStream_t st) {
    if (dv.codes4.p && k >= 2 && (k & 1) == 0 && da.n < (1ull << 31)) {
        const uint32_t g2 = std::max<uint32_t>(1, grid / 2);
#define WX_DM2(TT)                                                                                               \
    hipLaunchKernelGGL((k_demean2<TT>), dim3(g2), dim3(BLK), 0, st, da, dv.codes4.as<uint4>(), x, y_int, dm_int, k, \
                       kshift)
        switch (da.T) {
        case 1: WX_DM2(1); break;
        case 2: WX_DM2(2); break;
        case 3: WX_DM2(3); break;
        default: WX_DM2(4); break;
        }
#undef WX_DM2
        return;
    }
#define WX_DM(TT, PL)                                                                                            \
    hipLaunchKernelGGL((k_demean<TT, PL>), dim3(grid), dim3(BLK), 0, st, da, x, y_int, dm_int, k, kshift)
    switch (dv.plain ? da.T : 0) {
    case 1: WX_DM(1, true); break;
    case 2: WX_DM(2, true); break;
    case 3: WX_DM(3, true); break;
    case 4: WX_DM(4, true); break;
    default: WX_DM(0, false); break;
    }
#undef WX_DM
}

// HIP-event timer of the solve phases (enabled by wx_lsmr_options::timing).
struct PhaseTimer {
    static constexpr int NP = 4;
    bool on = false;
    hipStream_t st = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[NP];
    hipEvent_t open[NP] = {};
    void begin(int p) {
        if (!on) return;
        hipEvent_t e;
        WX_HIP(hipEventCreate(&e));
        WX_HIP(hipEventRecord(e, st));
        open[p] = e;
    }
    void end(int p) {
        if (!on) return;
        hipEvent_t e;
        WX_HIP(hipEventCreate(&e));
        WX_HIP(hipEventRecord(e, st));
        ev[p].push_back({open[p], e});
    }
    void collect(std::vector<double> &out) {
        out.assign(2 * NP, 0.0);
        for (int p = 0; p < NP; ++p) {
            for (auto &pr : ev[p]) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) out[2 * p] += ms;
                out[2 * p + 1] += 1.0;
                (void)hipEventDestroy(pr.first);
                (void)hipEventDestroy(pr.second);
            }
            ev[p].clear();
        }
    }
};

// Device workspace of one batch width, reused across solves of a Solver (a pool keyed by
// width; a solve takes one out and returns it, so concurrent solves never share one).
struct Workspace {
    uint32_t k = 0;
    uint32_t cap = MAX_BATCH; // column-state entries (packed members x k)
    DevBuf u, v, pt, atu, h, hb, x, yint, part, segpart, cols, ints, bad, red;
    DevBuf pieces; // the fused pair's piece sums (n_pieces x k)
    double *pieces_or_null() { return pieces.p ? pieces.as<double>() : nullptr; }
    ColState cs{};
    std::unique_ptr<SchwarzRun> runs[2];
    // hbad (pinned): the right-hand-side load's (non-finite count, needs-scaled-norm) flags;
    // the events that mark the last two iteration groups complete (the host polls one group
    // behind the device)
    int32_t *hbad = nullptr, *dbad = nullptr; // host and device addresses
    // pinned copies of the column state and the true-residual numerators read at the end
    double *hcols = nullptr, *hnum = nullptr;
    int32_t *hints = nullptr;
    hipEvent_t flag_ev[2] = {nullptr, nullptr};
    // host-coherent pinned copy of the "active" flags that k_alpha_step stores every
    // iteration (host and device addresses); a flag, once 0, stays 0 for the rest of a solve.
    // progress: each column's stopping-test value over its tolerance (0 once stopped)
    int32_t *hmirror = nullptr, *dmirror = nullptr;
    float *hprogress = nullptr, *dprogress = nullptr;
    // captured iteration group (replayed by later solves on this workspace), its key
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_key = 0;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }
    ~Workspace() {
        drop_graph();
        if (hbad) (void)hipHostFree(hbad);
        if (hcols) (void)hipHostFree(hcols);
        if (hnum) (void)hipHostFree(hnum);
        if (hints) (void)hipHostFree(hints);
        if (hmirror) (void)hipHostFree(hmirror);
        for (auto &e : flag_ev)
            if (e) (void)hipEventDestroy(e);
    }
};

void setup_colstate(Workspace &ws) {
    const uint32_t cap = ws.cap;
    ws.cols.alloc(sizeof(double) * cap * N_DOUBLE_FIELDS);
    ws.ints.alloc(sizeof(int32_t) * cap * N_INT_FIELDS);
    double *dp = ws.cols.as<double>();
    double **fields[N_DOUBLE_FIELDS] = {&ws.cs.alpha, &ws.cs.beta, &ws.cs.beta_prev_inv, &ws.cs.beta_inv,
                                        &ws.cs.p_coeff, &ws.cs.gscale, &ws.cs.alpha_bar, &ws.cs.phi_bar,
                                        &ws.cs.c_bar, &ws.cs.s_bar, &ws.cs.zeta_bar, &ws.cs.zeta0,
                                        &ws.cs.a_norm_sq, &ws.cs.abs_tol, &ws.cs.prev_rho, &ws.cs.prev_rho_bar,
                                        &ws.cs.t_x, &ws.cs.t_hbar, &ws.cs.t_h, &ws.cs.inv_alpha,
                                        &ws.cs.bnorm_scale, &ws.cs.den_norm, &ws.cs.rel_tol};
    for (int f = 0; f < N_DOUBLE_FIELDS; ++f) *fields[f] = dp + (uint64_t)f * cap;
    int32_t *ip = ws.ints.as<int32_t>();
    int32_t **ifields[N_INT_FIELDS] = {&ws.cs.active, &ws.cs.upd, &ws.cs.brk, &ws.cs.stop, &ws.cs.iters, &ws.cs.err};
    for (int f = 0; f < N_INT_FIELDS; ++f) *ifields[f] = ip + (uint64_t)f * cap;
}

// The column state of packed member m (entries m * k ...) seen as a state of k columns.
ColState colstate_view(const ColState &cs, uint32_t off) {
    ColState v = cs;
    double **d[N_DOUBLE_FIELDS] = {&v.alpha, &v.beta, &v.beta_prev_inv, &v.beta_inv, &v.p_coeff, &v.gscale,
                                   &v.alpha_bar, &v.phi_bar, &v.c_bar, &v.s_bar, &v.zeta_bar, &v.zeta0,
                                   &v.a_norm_sq, &v.abs_tol, &v.prev_rho, &v.prev_rho_bar, &v.t_x, &v.t_hbar,
                                   &v.t_h, &v.inv_alpha, &v.bnorm_scale, &v.den_norm, &v.rel_tol};
    for (auto *f : d) *f += off;
    int32_t **i[N_INT_FIELDS] = {&v.active, &v.upd, &v.brk, &v.stop, &v.iters, &v.err};
    for (auto *f : i) *f += off;
    return v;
}

// cap: column-state entries (k, or members x k of a pack); part_blocks: partial blocks of the
// largest exact-grid launch of a pack (0: the grid caps of the unpacked kernels)
std::shared_ptr<Workspace> new_workspace(const DevDesign &dv, uint32_t k, uint32_t cap = MAX_BATCH,
                                         uint64_t part_blocks = 0) {
    auto w = std::make_shared<Workspace>();
    const uint64_t n = dv.n, N = dv.N, kp = 1u << pow2_shift(k);
    w->k = k;
    w->cap = std::max(cap, MAX_BATCH);
    setup_colstate(*w);
    w->u.alloc(n * k * 8);
    w->yint.alloc(n * k * 8);
    for (DevBuf *b : {&w->v, &w->pt, &w->atu, &w->h, &w->hb, &w->x}) b->alloc(N * k * 8);
    w->part.alloc((size_t)std::max<uint64_t>(std::max(MAX_PART_BLOCKS, MAX_FUSED_BLOCKS), part_blocks) * kp * 3 * 8);
    w->segpart.alloc((size_t)std::max<uint32_t>(dv.max_partials, 1) * k * 8);
    if (dv.pair.p >= 0) w->pieces.alloc((size_t)std::max<uint32_t>(dv.pair.n_pieces, 1) * k * 8);
    w->bad.alloc(8);
    w->red.alloc(2 * (size_t)w->cap * 8);
    WX_HIP(hipHostMalloc((void **)&w->hbad, 2 * sizeof(int32_t), hipHostMallocCoherent));
    WX_HIP(hipHostMalloc((void **)&w->hcols, (size_t)N_DOUBLE_FIELDS * w->cap * 8, hipHostMallocDefault));
    WX_HIP(hipHostMalloc((void **)&w->hnum, (size_t)w->cap * 8, hipHostMallocDefault));
    WX_HIP(hipHostMalloc((void **)&w->hints, (size_t)N_INT_FIELDS * w->cap * 4, hipHostMallocDefault));
    WX_HIP(hipHostGetDevicePointer((void **)&w->dbad, w->hbad, 0));
    WX_HIP(hipHostMalloc((void **)&w->hmirror, (size_t)w->cap * 8, hipHostMallocCoherent));
    std::fill(w->hmirror, w->hmirror + w->cap, 0);
    w->hprogress = reinterpret_cast<float *>(w->hmirror + w->cap);
    std::fill(w->hprogress, w->hprogress + w->cap, 0.0f);
    WX_HIP(hipHostGetDevicePointer((void **)&w->dmirror, w->hmirror, 0));
    w->dprogress = reinterpret_cast<float *>(w->dmirror + w->cap);
    for (auto &e : w->flag_ev) WX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return w;
}

// Takes a workspace of width k out of the solver's pool (or makes one) and returns it on scope exit.
struct PooledWorkspace {
    const Solver &s;
    std::shared_ptr<Workspace> ws;
    PooledWorkspace(const Solver &solver, uint32_t k, uint32_t cap = MAX_BATCH, uint64_t part_blocks = 0)
        : s(solver) {
        {
            std::lock_guard<std::mutex> lk(s.pool_mu);
            for (size_t i = 0; i < s.pool.size(); ++i)
                if (s.pool_k[i] == k) {
                    ws = std::static_pointer_cast<Workspace>(s.pool[i]);
                    s.pool.erase(s.pool.begin() + i);
                    s.pool_k.erase(s.pool_k.begin() + i);
                    break;
                }
        }
        if (!ws) ws = new_workspace(s.dev, k, cap, part_blocks);
    }
    ~PooledWorkspace() {
        std::lock_guard<std::mutex> lk(s.pool_mu);
        if (s.pool.size() < 4) {
            s.pool.push_back(ws);
            s.pool_k.push_back(ws->k);
        }
    }
};

} // namespace

// ---------------------------------------------------------------------------
// Batched solve (Solver::solve_batch, crates/within/src/solver.rs:423-577)
// ---------------------------------------------------------------------------
namespace {

// One chunk (<= MAX_BATCH right-hand sides) of a batched solve as a resumable run, so a
// driver can interleave many independent runs on several streams (solve_many_device):
//   launch_load()  right-hand side load and beta_1 partials (asynchronous)
//   finish_init()  waits for the load's validation flags, then the bidiagonalisation init
//   launch_group() one group of `group` iterations: the first iteration directly, later
//                  groups as one replay of a captured hipGraph; k_alpha_step stores the
//                  per-column active flags to pinned memory, and an event marks the group
//   poll(wait)     once the group before the last one (lag one) is complete, the flags: 1
//                  when no column is active any more, 0 when one is, -1 when not yet known
//   finish()       demeaned output, x, FP64 true residual, statistics (synchronous)
class ChunkRun {
  public:
    // one chunk (columns c0 .. c0 + k) of one problem
    ChunkRun(const Solver &s, const double *Y, uint64_t ldy, uint64_t c0, uint32_t k, const wx_lsmr_options &opts,
             double *X, double *Dm, wx_rhs_stats *stats, hipStream_t st, bool x_solver_basis, PhaseTimer *timer,
             int group, bool use_graph)
        : ChunkRun(s, nullptr, k, opts, st, timer, group, use_graph) {
        Y_ = Y + c0 * ldy;
        ldy_ = ldy;
        c0_ = c0;
        X_ = X;
        Dm_ = Dm;
        stats_ = stats;
        x_solver_basis_ = x_solver_basis;
    }
    // a pack: member m solves mY[m] (n_m x k column-major, leading dimension mld[m]) and
    // writes mX[m], mDm[m] (either may be null) and mStats[m][0 .. k)
    ChunkRun(const Pack &pk, std::vector<const double *> mY, std::vector<uint64_t> mld, const wx_lsmr_options &opts,
             std::vector<double *> mX, std::vector<double *> mDm, std::vector<wx_rhs_stats *> mStats, hipStream_t st,
             int group, bool use_graph)
        : ChunkRun(*pk.merged, &pk, pk.k, opts, st, nullptr, group, use_graph) {
        mY_ = std::move(mY);
        mld_ = std::move(mld);
        mX_ = std::move(mX);
        mDm_ = std::move(mDm);
        mStats_ = std::move(mStats);
    }

    void launch_load() {
        WX_HIP(hipMemsetAsync(ws_->bad.p, 0, 8, st_));
        int32_t *badp = ws_->bad.as<int32_t>();
        if (pk_) {
            // each member's rows from its own caller order into its slice of the merged rows;
            // its load partials land in its own block range
            for (size_t m = 0; m < pk_->members.size(); ++m) {
                const PackMember &mb = pk_->members[m];
                const DevDesign &md = mb.s->dev;
                const uint32_t *ip = md.inv_perm.p ? md.inv_perm.as<uint32_t>() : nullptr;
                double *yint = ws_->yint.as<double>() + mb.row_off * k_;
                double *pp = part_ + (uint64_t)mb.tile_off * kp_ * 2;
                launch_load_rhs(mb.n_tiles, mb.n, ip,
                                weighted_ ? (ip ? md.sw_caller.as<double>() : md.sw.as<double>()) : nullptr, mY_[m],
                                mld_[m], yint, pp, badp, mb.kv);
            }
            hipLaunchKernelGGL(k_beta1, dim3(ncol_), dim3(BLK), 0, st_, part_, pk_->n_tiles, kp_, cs_, opts_.tol,
                               badp + 1, k_, rg_tile_, badp, nullptr);
            WX_HIP(hipMemcpyAsync(ws_->hbad, ws_->bad.p, 8, hipMemcpyDeviceToHost, st_));
        } else {
            launch_load_rhs(gt_, n_, inv_perm_,
                            weighted_ ? (inv_perm_ ? dv_.sw_caller.as<double>() : dv_.sw.as<double>()) : nullptr, Y_,
                            ldy_, ws_->yint.as<double>(), part_, badp, k_);
            // the non-finite count goes straight to pinned memory; finish() reads it
            hipLaunchKernelGGL(k_beta1, dim3(k_), dim3(BLK), 0, st_, part_, gt_, kp_, cs_, opts_.tol, badp + 1, k_,
                               Ranges{}, badp, ws_->dbad);
        }
        WX_HIP(hipGetLastError());
    }

    void finish_init() {
        if (!pk_) {
            // no wait for the load: a non-finite response has stopped every column (finish()
            // reports it), and the scaled norm pass returns at once unless a column needs it
            const uint32_t g = grid_for(n_, kshift_);
            int32_t *badp = ws_->bad.as<int32_t>();
            if (weighted_)
                hipLaunchKernelGGL((k_scaled_sumsq<true>), dim3(g), dim3(BLK), 0, st_, n_, dv_.sw.as<double>(),
                                   ws_->yint.as<double>(), cs_, part_, k_, kshift_, badp + 1);
            else
                hipLaunchKernelGGL((k_scaled_sumsq<false>), dim3(g), dim3(BLK), 0, st_, n_, nullptr,
                                   ws_->yint.as<double>(), cs_, part_, k_, kshift_, badp + 1);
            hipLaunchKernelGGL(k_beta1_scaled, dim3(k_), dim3(BLK), 0, st_, part_, g, kp_, cs_, opts_.tol, badp + 1,
                               badp);
        } else {
            WX_HIP(hipStreamSynchronize(st_));
            check_load();
        }
        if (pk_ && ws_->hbad[1]) { // a column outside the safe range of the unscaled sum of squares
            for (size_t m = 0; m < pk_->members.size(); ++m) {
                const uint64_t r0 = pk_->members[m].row_off, nr = pk_->members[m].n;
                const ColState v = colstate_view(cs_, (uint32_t)(m * k_));
                const uint32_t g = grid_for(nr, kshift_);
                const double *y = ws_->yint.as<double>() + r0 * k_;
                if (weighted_)
                    hipLaunchKernelGGL((k_scaled_sumsq<true>), dim3(g), dim3(BLK), 0, st_, nr,
                                       dv_.sw.as<double>() + r0, y, v, part_, k_, kshift_, nullptr);
                else
                    hipLaunchKernelGGL((k_scaled_sumsq<false>), dim3(g), dim3(BLK), 0, st_, nr, nullptr, y, v, part_,
                                       k_, kshift_, nullptr);
                hipLaunchKernelGGL(k_beta1_scaled, dim3(k_), dim3(BLK), 0, st_, part_, g, kp_, v, opts_.tol, nullptr,
                                   nullptr);
            }
        }
        // init: p~ = A'b / beta_1, v = M^{-1} p~, alpha_1
        launch_adjoint(dv_, da_, weighted_ ? 2 : 0, ws_->yint.as<double>(), ws_->pt.as<double>(),
                       ws_->segpart.as<double>(), gate(), k_, kshift_, st_, -1, ws_->pieces_or_null());
        const uint64_t N = N_;
        if (fused_) {
            // p~ = A'b / beta_1 inside the restriction, the alpha_1 dot products inside the
            // prolongation
            ApplyFusion fz;
            fz.atu = ws_->pt.as<double>();
            fz.init = true;
            fz.beta_inv = cs_.beta_inv;
            fz.p_coeff = cs_.p_coeff;
            fz.brk = cs_.brk;
            fz.active = cs_.active;
            fz.dots = part_;
            fz.dots_blocks = gN_;
            fz.blk_col = dof_col_;
            schwarz_->apply_fused(ws_->pt.as<double>(), ws_->v.as<double>(), k_, st_, gate(), fz);
        } else if (prec_kind_ == 2) {
            hipLaunchKernelGGL((k_ptilde<2, true>), dim3(gN_), dim3(BLK), 0, st_, N, nullptr, ws_->pt.as<double>(),
                               ws_->v.as<double>(), nullptr, cs_, part_, k_, kshift_, dof_col_);
            schwarz_->apply(ws_->pt.as<double>(), ws_->v.as<double>(), k_, st_, gate());
            hipLaunchKernelGGL(k_dots, dim3(gN_), dim3(BLK), 0, st_, N, ws_->pt.as<double>(), ws_->v.as<double>(),
                               cs_, part_, k_, kshift_, dof_col_);
        } else if (prec_kind_ == 1) {
            hipLaunchKernelGGL((k_ptilde<1, true>), dim3(gN_), dim3(BLK), 0, st_, N, nullptr, ws_->pt.as<double>(),
                               ws_->v.as<double>(), inv_diag_, cs_, part_, k_, kshift_, dof_col_);
        } else {
            hipLaunchKernelGGL((k_ptilde<0, true>), dim3(gN_), dim3(BLK), 0, st_, N, nullptr, ws_->pt.as<double>(),
                               ws_->v.as<double>(), nullptr, cs_, part_, k_, kshift_, dof_col_);
        }
        hipLaunchKernelGGL(k_alpha1, dim3(ncol_), dim3(BLK), 0, st_, part_, gN_, kp_, cs_, k_, rg_dof_);
        // windowed reorthogonalisation ring: capacity min(local_size, n_obs, n_dofs)
        // (bidiag.rs:121-133); the write index is the same for every column
        if (cap_) {
            ringV_.alloc(cap_ * N_ * k_ * 8);
            ringP_.alloc(cap_ * N_ * k_ * 8);
            ringc_.alloc(MAX_BATCH * 8);
        }
        const uint64_t push = cap_ ? ring_advance() : 0;
        hipLaunchKernelGGL(k_init_solution, dim3(gN_), dim3(BLK), 0, st_, N, ws_->v.as<double>(), ws_->h.as<double>(),
                           ws_->hb.as<double>(), ws_->x.as<double>(), cs_, k_, kshift_, ws_->pt.as<double>(),
                           cap_ ? ring_v(push) : nullptr, cap_ ? ring_p(push) : nullptr, dof_col_);
        WX_HIP(hipGetLastError());
    }

    // the load's flags in pinned memory: a non-finite response is an error
    void check_load() const {
        if (ws_->hbad[0]) fail(WX_ERR_INVALID_INPUT, "Solver::solve: response contains non-finite values");
    }

    bool exhausted() const { return itn_ >= maxiter_; }
    int launches() const { return launches_; }

    void launch_group() {
        if (itn_ == 0) {
            iteration(true);
            ++itn_;
        } else if (use_graph_ && itn_ + group_ <= maxiter_) {
            replay_graph();
            itn_ += group_;
        } else {
            const int g = std::min(group_, maxiter_ - itn_);
            for (int q = 0; q < g; ++q) iteration(false);
            itn_ += g;
        }
        // the group's k_alpha_step launches store the flags to the pinned mirror; the event
        // marks the group complete
        const int slot = launches_ & 1;
        WX_HIP(hipEventRecord(ws_->flag_ev[slot], st_));
        ++launches_;
        WX_HIP(hipGetLastError());
    }

    // lag 2: the group before the last one; lag 1: the last group itself
    int poll(bool wait, int lag = 2) {
        if (launches_ < lag) return 0;
        const int slot = (launches_ - lag) & 1;
        if (wait) WX_HIP(hipEventSynchronize(ws_->flag_ev[slot]));
        else {
            const hipError_t q = hipEventQuery(ws_->flag_ev[slot]);
            if (q == hipErrorNotReady) return -1;
            WX_HIP(q);
        }
        load_done_ = true;
        // the mirror holds the flags of that group or of a later one: a column that reads 0
        // has stopped, so "all 0" is final either way
        const volatile int32_t *f = ws_->hmirror;
        for (uint32_t c = 0; c < ncol_; ++c)
            if (f[c]) return 0;
        return 1;
    }

    // True when the slowest active column of a completed group is expected to meet its test
    // within the next group, from its progress (k_alpha_step's test value over the tolerance,
    // <= 1 once met) and the rate at which that fell since the previous reading.  A driver
    // then waits for the last group (lag 1) rather than launch one more that would find every
    // column stopped.  Only the launch schedule depends on it, never a result.
    bool near_done() {
        const volatile int32_t *f = ws_->hmirror;
        const volatile float *pg = ws_->hprogress;
        float m = 0.0f;
        for (uint32_t c = 0; c < ncol_; ++c)
            if (f[c]) m = std::max(m, (float)pg[c]);
        const float prev = prev_progress_;
        prev_progress_ = m;
        if (!(m > 0.0f)) return false;
        const float rate = prev > m ? prev / m : 1.0f;
        return m / rate <= 2.0f;
    }

    // caller-order rows (kv of the k columns) -> internal order, sqrt(w) y partials; sw null:
    // unweighted
    void launch_load_rhs(uint64_t grid, uint64_t n, const uint32_t *inv_perm, const double *sw, const double *Y,
                         uint64_t ldy, double *y_int, double *part, int32_t *bad, uint32_t kv) {
        const dim3 g((uint32_t)grid), b(BLK);
        if (sw && k_ <= 16)
            hipLaunchKernelGGL((k_load_rhs<true, 4>), g, b, tile_bytes_, st_, n, inv_perm, sw, Y, ldy, y_int, part,
                               bad, k_, kshift_, kv);
        else if (sw)
            hipLaunchKernelGGL((k_load_rhs<true, 16>), g, b, tile_bytes_, st_, n, inv_perm, sw, Y, ldy, y_int, part,
                               bad, k_, kshift_, kv);
        else if (k_ <= 16)
            hipLaunchKernelGGL((k_load_rhs<false, 4>), g, b, tile_bytes_, st_, n, inv_perm, nullptr, Y, ldy, y_int,
                               part, bad, k_, kshift_, kv);
        else
            hipLaunchKernelGGL((k_load_rhs<false, 16>), g, b, tile_bytes_, st_, n, inv_perm, nullptr, Y, ldy, y_int,
                               part, bad, k_, kshift_, kv);
    }

    // internal-order rows -> the caller's column-major output (kv of the k columns)
    void launch_store_caller(uint64_t grid, uint64_t n, const uint32_t *inv_perm, const double *dm_int, double *Dm,
                             uint64_t ldd, uint32_t kv) {
        if (k_ <= 16)
            hipLaunchKernelGGL((k_store_caller<4>), dim3((uint32_t)grid), dim3(BLK), tile_bytes_, st_, n, inv_perm,
                               dm_int, Dm, ldd, k_, kshift_, kv);
        else
            hipLaunchKernelGGL((k_store_caller<16>), dim3((uint32_t)grid), dim3(BLK), tile_bytes_, st_, n, inv_perm,
                               dm_int, Dm, ldd, k_, kshift_, kv);
    }

    void finish() {
        if (!pk_) { // the load's validation (finish_init did not wait for it)
            if (!load_done_) WX_HIP(hipStreamSynchronize(st_));
            check_load();
        }
        const DesignArgs &da = da_;
        const uint64_t n = n_, N = N_;
        const uint32_t k = k_, kshift = kshift_, kp = kp_;
        hipStream_t st = st_;
        Workspace &ws = *ws_;
        double *dm_int = ws.u.as<double>();
        const bool true_res = opts_.true_residual != 0;
        // without the true residual nothing needs y - D x in internal order: a plain design's
        // demeaned output goes straight to caller order
        const bool direct = !pk_ && Dm_ && !true_res && dv_.codes4.p && da.T >= 1 && da.T <= 4;
        if (direct) {
            // a tile per workgroup where the grid allows (4,096 workgroups looping over tiles
            // measured 455 against 415 us at n = 5M, k = 16)
            const uint32_t gd = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>((n + DEMEAN_TILE - 1) / DEMEAN_TILE, 65535));
#define WX_DS(TT)                                                                                                \
    hipLaunchKernelGGL((k_demean_store<TT>), dim3(gd), dim3(BLK), (size_t)DEMEAN_TILE * (k + 1) * sizeof(double), \
                       st, da,                                                                                     \
                       (dv_.codes4_caller.p ? dv_.codes4_caller : dv_.codes4).as<uint4>(), ws.x.as<double>(), Y_,  \
                       ldy_, Dm_ + c0_ * n, n, k, kshift)
            switch (da.T) {
            case 1: WX_DS(1); break;
            case 2: WX_DS(2); break;
            case 3: WX_DS(3); break;
            default: WX_DS(4); break;
            }
#undef WX_DS
        } else {
            const uint32_t gd = grid_for(n, kshift, 65535);
            launch_demean(dv_, da, gd, ws.x.as<double>(), ws.yint.as<double>(), dm_int, k, kshift, st);
        }
        DevBuf xtab;
        if (pk_) {
            bool any_x = false;
            for (size_t m = 0; m < pk_->members.size(); ++m) {
                const PackMember &mb = pk_->members[m];
                const DevDesign &md = mb.s->dev;
                if (mDm_[m])
                    launch_store_caller(mb.n_tiles, mb.n, md.inv_perm.p ? md.inv_perm.as<uint32_t>() : nullptr,
                                        dm_int + mb.row_off * k, mDm_[m], mb.n, mb.kv);
                any_x = any_x || mX_[m];
            }
            if (any_x) {
                xtab.upload(mX_.data(), mX_.size(), st);
                const uint64_t tot = N * k;
                hipLaunchKernelGGL(k_store_x_packed, dim3((tot + BLK - 1) / BLK), dim3(BLK), 0, st, N,
                                   ws.x.as<double>(), k, pk_->dof_member.as<uint32_t>(), pk_->dof_own.as<int32_t>(),
                                   xtab.as<double *>(), pk_->own_ld.as<uint64_t>(), pk_->own_kv.as<uint32_t>());
            }
        } else {
            if (Dm_ && !direct) launch_store_caller(gt_, n, inv_perm_, dm_int, Dm_ + c0_ * n, n, k);
            if (X_ && (!s_.reparam || x_solver_basis_)) {
                const uint64_t tot = N * k;
                hipLaunchKernelGGL(k_store_x, dim3((tot + BLK - 1) / BLK), dim3(BLK), 0, st, N, ws.x.as<double>(),
                                   X_ + c0_ * N, N, k);
            }
        }
        // residual numerator ||D'W(y - Dx)||; the denominator ||D'W y|| = beta_1 ||p~_1|| was
        // kept at initialisation
        if (true_res) {
            launch_adjoint(dv_, da, weighted_ ? 2 : 0, dm_int, ws.h.as<double>(), ws.segpart.as<double>(), nullptr,
                           k, kshift, st, -1, ws.pieces_or_null());
            hipLaunchKernelGGL(k_sumsq, dim3(gN_), dim3(BLK), 0, st, N, ws.h.as<double>(), part_, k, kshift);
            hipLaunchKernelGGL(k_reduce_to, dim3(ncol_), dim3(BLK), 0, st, part_, gN_, kp, ws.red.as<double>(), k,
                               rg_dof_);
        }
        WX_HIP(hipGetLastError());
        const size_t cap = ws.cap;
        // the column state through the workspace's pinned buffers (asynchronous copies, one wait)
        if (true_res)
            WX_HIP(hipMemcpyAsync(ws.hnum, ws.red.as<double>(), ncol_ * 8, hipMemcpyDeviceToHost, st));
        WX_HIP(hipMemcpyAsync(ws.hcols, ws.cols.p, N_DOUBLE_FIELDS * cap * 8, hipMemcpyDeviceToHost, st));
        WX_HIP(hipMemcpyAsync(ws.hints, ws.ints.p, N_INT_FIELDS * cap * 4, hipMemcpyDeviceToHost, st));
        WX_HIP(hipStreamSynchronize(st));
        std::vector<double> num(ncol_, std::nan("")), host_d(ws.hcols, ws.hcols + N_DOUBLE_FIELDS * cap);
        std::vector<int32_t> host_i(ws.hints, ws.hints + N_INT_FIELDS * cap);
        if (true_res) std::copy(ws.hnum, ws.hnum + ncol_, num.begin());
        if (!pk_ && X_ && s_.reparam && !x_solver_basis_) {
            std::vector<double> xi(N * k), xo(N);
            WX_HIP(hipMemcpy(xi.data(), ws.x.p, N * k * 8, hipMemcpyDeviceToHost));
            for (uint32_t j = 0; j < k; ++j) {
                for (uint64_t q = 0; q < N; ++q) xo[q] = xi[q * k + j];
                s_.reparam->back_transform(s_.design, xo.data());
                WX_HIP(hipMemcpy(X_ + (c0_ + j) * N, xo.data(), N * 8, hipMemcpyHostToDevice));
            }
        }
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
        const double *phi_bar = host_d.data() + 7 * cap, *zeta_bar = host_d.data() + 10 * cap,
                     *zeta0 = host_d.data() + 11 * cap, *den_norm = host_d.data() + 21 * cap;
        const int32_t *active = host_i.data(), *stop = host_i.data() + 3 * cap, *iters = host_i.data() + 4 * cap,
                      *err = host_i.data() + 5 * cap;
        for (uint32_t c = 0; c < ncol_; ++c) {
            if (err[c]) fail(WX_ERR_NOT_POSITIVE_DEF, "mlsmr: preconditioner not positive definite (<v, Mv> < 0)");
            const uint32_t j = c % k;
            if (pk_ && j >= pk_->members[c / k].kv) continue; // a narrower member's padding column
            wx_rhs_stats &o = pk_ ? mStats_[c / k][j] : stats_[c0_ + j];
            int32_t sr = stop[c];
            if (active[c] || sr < 0) {
                sr = WX_STOP_MAX_ITERATIONS;
                o.converged = 0;
                o.iterations = maxiter_;
            } else {
                o.converged = 1;
                o.iterations = iters[c];
            }
            o.stop_reason = sr;
            if (sr == WX_STOP_ZERO_RHS) {
                o.residual = 0.0;
                o.residual_norm = 0.0;
            } else if (sr == WX_STOP_INITIAL_NORMAL_EQ_ZERO) {
                o.residual = 0.0;
                o.residual_norm = phi_bar[c];
            } else {
                o.residual = std::fabs(zeta_bar[c]) / zeta0[c];
                o.residual_norm = std::fabs(phi_bar[c]);
            }
            o.true_residual = !true_res ? std::nan("") : den_norm[c] > 0.0 ? std::sqrt(num[c]) / den_norm[c] : 0.0;
            o.time_solve = secs / k;
        }
    }

  private:
    ChunkRun(const Solver &s, const Pack *pk, uint32_t k, const wx_lsmr_options &opts, hipStream_t st,
             PhaseTimer *timer, int group, bool use_graph)
        : s_(s), pk_(pk), dv_(s.dev), da_(s.dev.args()), k_(k), opts_(opts), st_(st), timer_(timer),
          pw_(s, k, pk ? (uint32_t)(pk->members.size() * k) : MAX_BATCH, pk ? pk->part_blocks : 0) {
        n_ = dv_.n;
        N_ = dv_.N;
        weighted_ = dv_.sw.p != nullptr;
        inv_perm_ = dv_.inv_perm.p ? dv_.inv_perm.as<uint32_t>() : nullptr;
        maxiter_ = (int)std::max<int64_t>(0, opts.maxiter);
        const Precond *pc = s.precond.get();
        prec_kind_ = pc == nullptr ? 0 : (pc->kind == Precond::DIAGONAL ? 1 : 2);
        inv_diag_ = prec_kind_ == 1 ? dv_.inv_diag.as<double>() : nullptr;
        fused_ = prec_kind_ == 2 && g_fused_steps.load();
        const int prec_slot = opts.precision == WX_PRECISION_MIXED ? 1 : 0;
        kshift_ = pow2_shift(k);
        kp_ = 1u << kshift_;
        t0_ = std::chrono::steady_clock::now();
        ws_ = pw_.ws.get();
        if (prec_kind_ == 2 && !ws_->runs[prec_slot]) ws_->runs[prec_slot] = make_schwarz_run(*pc, opts.precision, st);
        schwarz_ = prec_kind_ == 2 ? ws_->runs[prec_slot].get() : nullptr;
        cs_ = ws_->cs;
        gn_ = grid_for(n_, kshift_);
        part_ = ws_->part.as<double>();
        gt_ = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n_ + 63) / 64, MAX_PART_BLOCKS));
        tile_bytes_ = (size_t)64 * (k + 1) * sizeof(double);
        group_ = std::max(1, group);
        if (pk) {
            // exact grids over the padded DOF blocks; reductions over each member's ranges
            ncol_ = (uint32_t)(pk->members.size() * k);
            gN_ = pk->n_dof_blocks;
            dof_col_ = pk->dof_col.as<uint32_t>();
            rg_dof_ = Ranges{pk->dof_rng_ptr.as<uint32_t>(), pk->dof_rng.as<uint32_t>()};
            rg_item_ = Ranges{pk->item_rng_ptr.as<uint32_t>(), pk->item_rng.as<uint32_t>()};
            rg_tile_ = Ranges{pk->tile_rng_ptr.as<uint32_t>(), pk->tile_rng.as<uint32_t>()};
            cap_ = 0;
        } else {
            ncol_ = k;
            gN_ = grid_for(N_, kshift_);
            cap_ = opts.local_size > 0 ? std::min<uint64_t>((uint64_t)opts.local_size, std::min(n_, N_)) : 0;
        }
        // a graph replays fixed launch arguments: not with the reorthogonalisation ring (its
        // slot changes every iteration) nor with the event timer
        use_graph_ = use_graph && cap_ == 0 && !(timer_ && timer_->on);
        graph_key_ = ((uint64_t)group_ * 4 + (uint64_t)prec_slot) * 2 + (fused_ ? 1 : 0);
    }

    // convergence gate of the kernels: the run's active flags (a pack's members converge
    // separately, so its kernels run ungated and the per-column state decides)
    const int32_t *gate() const { return pk_ ? nullptr : cs_.active; }

    // one LSMR iteration (u update, beta, D', p~, M^{-1}, reorthogonalisation, alpha, x update)
    void iteration(bool first) {
        Workspace &ws = *ws_;
        const uint64_t N = N_;
        const uint32_t k = k_, kshift = kshift_;
        hipStream_t st = st_;
        const bool timed = timer_ && timer_->on;
        if (timed) timer_->begin(3), timer_->begin(0);
        const double *uin = first ? ws.yint.as<double>() : ws.u.as<double>();
        uint32_t gpart = gn_;
        if (dv_.fuse_term >= 0)
            gpart = launch_gather_seg(dv_, da_, first, ws.v.as<double>(), uin, ws.u.as<double>(), ws.atu.as<double>(),
                                      ws.segpart.as<double>(), cs_, part_, k, kshift, st, pk_);
        else
            launch_gather(dv_, da_, first, gn_, ws.v.as<double>(), uin, ws.u.as<double>(), cs_, part_, k, kshift, st);
        if (timed) timer_->end(0);
        const bool merged = dv_.fuse_term >= 0 && dv_.mfix.on() && k >= 2 && (k & 1) == 0;
        if (!merged) hipLaunchKernelGGL(k_beta, dim3(ncol_), dim3(BLK), 0, st, part_, gpart, kp_, cs_, k, rg_item_);
        if (timed) timer_->begin(1);
        if (merged) {
            // the fused pair (its launch also reduces beta), then every term's split levels in
            // merged launches
            launch_adjoint_merged(dv_, weighted_ ? 1 : 0, ws.u.as<double>(), ws.atu.as<double>(),
                                  ws.segpart.as<double>(), gate(), k, kshift, st, cs_,
                                  BetaSide{part_, gpart, kp_, ncol_, k, rg_item_});
        } else {
            // split levels of the fused term first: the other terms reuse the partial buffer
            if (dv_.fuse_term >= 0)
                launch_segfix(dv_.terms[dv_.fuse_term], ws.segpart.as<double>(), ws.atu.as<double>(), gate(), k,
                              kshift, st);
            launch_adjoint(dv_, da_, weighted_ ? 1 : 0, ws.u.as<double>(), ws.atu.as<double>(),
                           ws.segpart.as<double>(), gate(), k, kshift, st, dv_.fuse_term, ws.pieces_or_null());
        }
        if (timed) timer_->end(1);
        if (fused_ && cap_ == 0) {
            // p~ update folded into the restriction, the alpha dot products into the prolongation
            ApplyFusion fz;
            fz.blk_col = dof_col_;
            fz.atu = ws.atu.as<double>();
            fz.beta_inv = cs_.beta_inv;
            fz.p_coeff = cs_.p_coeff;
            fz.brk = cs_.brk;
            fz.active = cs_.active;
            fz.dots = part_;
            fz.dots_blocks = gN_;
            if (timed) timer_->begin(2);
            schwarz_->apply_fused(ws.pt.as<double>(), ws.v.as<double>(), k, st, gate(), fz);
            if (timed) timer_->end(2);
        } else if (prec_kind_ == 2) {
            hipLaunchKernelGGL((k_ptilde<2, false>), dim3(gN_), dim3(BLK), 0, st, N, ws.atu.as<double>(),
                               ws.pt.as<double>(), ws.v.as<double>(), nullptr, cs_, part_, k, kshift, dof_col_);
            if (timed) timer_->begin(2);
            schwarz_->apply(ws.pt.as<double>(), ws.v.as<double>(), k, st, gate());
            if (timed) timer_->end(2);
            if (cap_) ring_mgs();
            hipLaunchKernelGGL(k_dots, dim3(gN_), dim3(BLK), 0, st, N, ws.pt.as<double>(), ws.v.as<double>(), cs_,
                               part_, k, kshift, dof_col_);
        } else {
            if (prec_kind_ == 1)
                hipLaunchKernelGGL((k_ptilde<1, false>), dim3(gN_), dim3(BLK), 0, st, N, ws.atu.as<double>(),
                                   ws.pt.as<double>(), ws.v.as<double>(), inv_diag_, cs_, part_, k, kshift, dof_col_);
            else
                hipLaunchKernelGGL((k_ptilde<0, false>), dim3(gN_), dim3(BLK), 0, st, N, ws.atu.as<double>(),
                                   ws.pt.as<double>(), ws.v.as<double>(), nullptr, cs_, part_, k, kshift, dof_col_);
            if (cap_) { // reorthogonalise, then the alpha dot products of the updated pair
                ring_mgs();
                hipLaunchKernelGGL(k_dots, dim3(gN_), dim3(BLK), 0, st, N, ws.pt.as<double>(), ws.v.as<double>(), cs_,
                                   part_, k, kshift, dof_col_);
            }
        }
        hipLaunchKernelGGL(k_alpha_step, dim3(ncol_), dim3(BLK), 0, st, part_, gN_, kp_, cs_, k, rg_dof_,
                           ws.dmirror, ws.dprogress);
        const uint64_t push = cap_ ? ring_advance() : 0;
        hipLaunchKernelGGL(k_update_solution, dim3(gN_), dim3(BLK), 0, st, N, ws.v.as<double>(), ws.h.as<double>(),
                           ws.hb.as<double>(), ws.x.as<double>(), cs_, k, kshift, ws.pt.as<double>(),
                           cap_ ? ring_v(push) : nullptr, cap_ ? ring_p(push) : nullptr, dof_col_);
        if (timed) timer_->end(3);
    }

    // `group_` iterations captured once per workspace into a graph and replayed
    void replay_graph() {
        Workspace &ws = *ws_;
        if (!ws.graph_exec || ws.graph_key != graph_key_) {
            ws.drop_graph();
            hipGraph_t g = nullptr;
            WX_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeRelaxed));
            for (int q = 0; q < group_; ++q) iteration(false);
            WX_HIP(hipStreamEndCapture(st_, &g));
            ws.graph = g;
            WX_HIP(hipGraphInstantiate(&ws.graph_exec, g, nullptr, nullptr, 0));
            ws.graph_key = graph_key_;
        }
        WX_HIP(hipGraphLaunch(ws.graph_exec, st_));
    }

    double *ring_v(uint64_t slot) { return ringV_.as<double>() + slot * N_ * k_; }
    double *ring_p(uint64_t slot) { return ringP_.as<double>() + slot * N_ * k_; }
    uint64_t ring_advance() {
        const uint64_t slot = ring_next_;
        ring_next_ = (ring_next_ + 1) % cap_;
        if (ring_count_ < cap_) ++ring_count_;
        return slot;
    }
    // one MGS sweep of v (and p~ in lockstep) over the stored slots, oldest first
    void ring_mgs() {
        const uint64_t start = ring_count_ < cap_ ? 0 : ring_next_;
        for (uint64_t q = 0; q < ring_count_; ++q) {
            const uint64_t slot = (start + q) % cap_;
            hipLaunchKernelGGL(k_ring_dot, dim3(gN_), dim3(BLK), 0, st_, N_, ws_->v.as<double>(), ring_p(slot), cs_,
                               part_, k_, kshift_);
            hipLaunchKernelGGL(k_reduce_to, dim3(k_), dim3(BLK), 0, st_, part_, gN_, kp_, ringc_.as<double>(), 1u,
                               Ranges{});
            hipLaunchKernelGGL(k_ring_axpy, dim3(gN_), dim3(BLK), 0, st_, N_, ws_->v.as<double>(), ws_->pt.as<double>(),
                               ring_v(slot), ring_p(slot), ringc_.as<double>(), cs_, k_, kshift_);
        }
    }

    const Solver &s_;
    const Pack *pk_ = nullptr;
    const DevDesign &dv_;
    const DesignArgs da_;
    const double *Y_ = nullptr;
    uint64_t ldy_ = 0, c0_ = 0;
    uint32_t k_, kshift_ = 0, kp_ = 1;
    wx_lsmr_options opts_;
    double *X_ = nullptr, *Dm_ = nullptr;
    wx_rhs_stats *stats_ = nullptr;
    std::vector<const double *> mY_;
    std::vector<uint64_t> mld_;
    std::vector<double *> mX_, mDm_;
    std::vector<wx_rhs_stats *> mStats_;
    hipStream_t st_;
    bool x_solver_basis_ = false;
    PhaseTimer *timer_;
    PooledWorkspace pw_;
    Workspace *ws_ = nullptr;
    SchwarzRun *schwarz_ = nullptr;
    ColState cs_{};
    uint64_t n_ = 0, N_ = 0;
    bool weighted_ = false;
    const uint32_t *inv_perm_ = nullptr;
    int maxiter_ = 0, prec_kind_ = 0;
    bool fused_ = false; // the LSMR steps around the Schwarz apply folded into it
    const double *inv_diag_ = nullptr;
    uint32_t gn_ = 1, gN_ = 1, gt_ = 1, ncol_ = 1;
    const uint32_t *dof_col_ = nullptr;
    Ranges rg_dof_{}, rg_item_{}, rg_tile_{};
    size_t tile_bytes_ = 0;
    double *part_ = nullptr;
    std::chrono::steady_clock::time_point t0_;
    uint64_t cap_ = 0, ring_next_ = 0, ring_count_ = 0;
    DevBuf ringV_, ringP_, ringc_;
    int group_ = 1, itn_ = 0, launches_ = 0;
    bool load_done_ = false; // an event recorded after the load has completed
    float prev_progress_ = -1.0f;
    bool use_graph_ = false;
    uint64_t graph_key_ = 0;
};

void check_solve_args(const Solver &s, uint64_t ldy, const wx_lsmr_options &opts) {
    if (!std::isfinite(opts.tol) || opts.tol < 0.0)
        fail(WX_ERR_INVALID_INPUT, "lsmr: tolerance must be finite and nonnegative, got " + std::to_string(opts.tol));
    if (ldy < s.dev.n) fail(WX_ERR_INVALID_INPUT, "leading dimension of Y is smaller than n_obs");
}

} // namespace

void solve_batch_device(const Solver &s, const double *Y, uint64_t ldy, uint64_t k_total, const wx_lsmr_options &opts,
                        double *X, double *Dm, wx_rhs_stats *stats, hipStream_t st, bool x_solver_basis) {
    check_solve_args(s, ldy, opts);
    if (k_total == 0) return;
    PhaseTimer timer;
    timer.on = opts.timing != 0;
    timer.st = st;
    // check_every > 1 launches that many iterations back to back between polls
    const int group = opts.check_every > 0 ? opts.check_every : 1;
    for (uint64_t c0 = 0; c0 < k_total; c0 += MAX_BATCH) {
        const uint32_t k = (uint32_t)std::min<uint64_t>(MAX_BATCH, k_total - c0);
        ChunkRun run(s, Y, ldy, c0, k, opts, X, Dm, stats, st, x_solver_basis, &timer, group, false);
        run.launch_load();
        run.finish_init();
        while (!run.exhausted()) {
            run.launch_group();
            if (run.poll(true) == 1) break;
            // near convergence, wait for the last group too rather than launch one more
            if (run.near_done() && run.poll(true, 1) == 1) break;
        }
        run.finish();
    }
    if (timer.on) {
        WX_HIP(hipStreamSynchronize(st));
        timer.collect(last_timing());
    }
}

// ---------------------------------------------------------------------------
// Many independent problems at once.  Every (problem, chunk) run is resumable; the driver
// spreads the runs over `streams` default-priority streams, replays each run's iterations
// as captured graphs of `group` iterations, and polls every run's convergence flags without
// blocking, so converged problems stop launching while the others continue.  A run whose
// columns have all converged produces its outputs at once.  The (streams, group) shape is
// chosen by a small autotuner on the first call for a problem-count class and then cached.
// ---------------------------------------------------------------------------
namespace {
struct ManyShape {
    int streams = 1, group = 1;
    bool graph = false;
    bool pack = false; // problems of one shape packed into one block-diagonal run
    uint32_t tile = DEFAULT_TILE; // rows per D' / gather work item of the packs
};
std::mutex g_tune_mu;
std::vector<std::pair<uint64_t, ManyShape>> g_tuned; // problem-count class -> shape
ManyShape g_last_shape;
ManyShape g_forced{0, 1, false, false}; // streams == 0: the autotuner decides
uint32_t g_forced_tile = 0;               // 0: the autotuner decides

// Streams of the many-problem runs.  A call leases its streams for its whole duration, so
// concurrent calls never share one: a run that captures its iteration graph on a stream would
// otherwise also capture another call's launches on it.  Returned streams are reused by later
// calls on the same device and live until the process ends.
std::mutex g_stream_mu;
std::vector<std::pair<int, hipStream_t>> g_free_streams; // (device, stream)

struct StreamLease {
    std::vector<hipStream_t> s;
    int dev = 0;
    explicit StreamLease(int n) {
        WX_HIP(hipGetDevice(&dev));
        {
            std::lock_guard<std::mutex> lk(g_stream_mu);
            for (size_t i = g_free_streams.size(); i-- > 0 && (int)s.size() < n;)
                if (g_free_streams[i].first == dev) {
                    s.push_back(g_free_streams[i].second);
                    g_free_streams.erase(g_free_streams.begin() + (std::ptrdiff_t)i);
                }
        }
        while ((int)s.size() < n) {
            hipStream_t st;
            if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
                give_back();
                fail(WX_ERR_DEVICE, "hipStreamCreateWithFlags failed");
            }
            s.push_back(st);
        }
    }
    void give_back() {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        for (hipStream_t st : s) g_free_streams.emplace_back(dev, st);
        s.clear();
    }
    ~StreamLease() { give_back(); }
};

void run_many(const std::vector<const Solver *> &solvers, const std::vector<const double *> &Y,
              const std::vector<uint64_t> &ldy, const std::vector<uint64_t> &k, const wx_lsmr_options &opts,
              const std::vector<double *> &X, const std::vector<double *> &Dm, const std::vector<wx_rhs_stats *> &stats,
              const ManyShape &shape, hipStream_t caller) {
    // the caller's stream work comes first
    StreamLease lease(std::max(1, shape.streams));
    hipEvent_t ready;
    WX_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    WX_HIP(hipEventRecord(ready, caller));
    std::vector<std::unique_ptr<ChunkRun>> runs;
    std::vector<std::shared_ptr<Pack>> packs; // kept alive for the runs
    std::vector<char> packed(solvers.size(), 0);
    int next_stream = 0;
    if (shape.pack && opts.local_size <= 0) {
        // problems of one signature and one power-of-two width class share a pack run at the
        // class's widest member; narrower members carry zero padding columns
        std::map<std::string, std::vector<size_t>> groups;
        for (size_t p = 0; p < solvers.size(); ++p) {
            std::string sig;
            if (k[p] < 1 || k[p] > MAX_BATCH || !pack_signature(*solvers[p], sig)) continue;
            groups[sig + "/" + std::to_string(pow2_shift((uint32_t)k[p]))].push_back(p);
        }
        for (auto &g : groups) {
            const std::vector<size_t> &idx = g.second;
            if (idx.size() < 2) continue;
            std::vector<const Solver *> mem;
            std::vector<uint32_t> kv;
            std::vector<const double *> mY;
            std::vector<uint64_t> mld;
            std::vector<double *> mX, mDm;
            std::vector<wx_rhs_stats *> mSt;
            for (size_t p : idx) {
                mem.push_back(solvers[p]);
                kv.push_back((uint32_t)k[p]);
                mY.push_back(Y[p]);
                mld.push_back(ldy[p]);
                mX.push_back(X[p]);
                mDm.push_back(Dm[p]);
                mSt.push_back(stats[p]);
                packed[p] = 1;
            }
            packs.push_back(get_pack(mem, kv, shape.tile));
            hipStream_t st = lease.s[next_stream];
            next_stream = (next_stream + 1) % shape.streams;
            WX_HIP(hipStreamWaitEvent(st, ready, 0));
            runs.push_back(std::make_unique<ChunkRun>(*packs.back(), mY, mld, opts, mX, mDm, mSt, st, shape.group,
                                                      shape.graph));
        }
    }
    for (size_t p = 0; p < solvers.size(); ++p)
        for (uint64_t c0 = 0; c0 < k[p] && !packed[p]; c0 += MAX_BATCH) {
            const uint32_t kk = (uint32_t)std::min<uint64_t>(MAX_BATCH, k[p] - c0);
            hipStream_t st = lease.s[next_stream];
            next_stream = (next_stream + 1) % shape.streams;
            WX_HIP(hipStreamWaitEvent(st, ready, 0));
            runs.push_back(std::make_unique<ChunkRun>(*solvers[p], Y[p], ldy[p], c0, kk, opts, X[p], Dm[p], stats[p], st,
                                                      false, nullptr, shape.group, shape.graph));
        }
    for (auto &r : runs) r->launch_load();
    for (auto &r : runs) r->finish_init();
    std::vector<ChunkRun *> active;
    for (auto &r : runs) active.push_back(r.get());
    while (!active.empty()) {
        bool progressed = false;
        for (size_t i = 0; i < active.size();) {
            ChunkRun &r = *active[i];
            const int state = r.poll(false);
            if (state == 1 || (r.exhausted() && state != -1)) {
                r.finish();
                active.erase(active.begin() + i);
                progressed = true;
                continue;
            }
            if (state == 0 && !r.exhausted()) {
                r.launch_group();
                progressed = true;
            }
            ++i;
        }
        if (!progressed) std::this_thread::yield();
    }
    WX_HIP(hipEventDestroy(ready));
}
} // namespace

void solve_many_device(const std::vector<const Solver *> &solvers, const std::vector<const double *> &Y,
                       const std::vector<uint64_t> &ldy, const std::vector<uint64_t> &k, const wx_lsmr_options &opts,
                       const std::vector<double *> &X, const std::vector<double *> &Dm,
                       const std::vector<wx_rhs_stats *> &stats, hipStream_t caller) {
    for (size_t p = 0; p < solvers.size(); ++p) check_solve_args(*solvers[p], ldy[p], opts);
    uint64_t n_runs = 0;
    for (uint64_t kp : k) n_runs += (kp + MAX_BATCH - 1) / MAX_BATCH;
    if (n_runs == 0) return;
    uint64_t cls = 1;
    while (cls < n_runs) cls <<= 1;
    ManyShape shape;
    bool tuned = false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (g_forced.streams > 0) {
            shape = g_forced;
            shape.tile = g_forced_tile ? g_forced_tile : DEFAULT_TILE;
            tuned = true;
        } else {
            for (auto &e : g_tuned)
                if (e.first == cls) {
                    shape = e.second;
                    tuned = true;
                    break;
                }
        }
    }
    if (!tuned) {
        // autotuner: time each candidate shape on the real problems with the iteration count
        // capped (the final run below overwrites every output)
        static const ManyShape cand[] = {{1, 1, false, false}, {4, 1, true, false}, {8, 2, true, false},
                                         {16, 4, true, false}, {1, 1, false, true},  {4, 1, false, true},
                                         {4, 1, true, true},   {4, 2, true, true},   {8, 2, true, true}};
        wx_lsmr_options probe = opts;
        probe.maxiter = std::min<int64_t>(opts.maxiter, 24);
        probe.timing = 0;
        double best = 1e300;
        auto trial = [&](const ManyShape &c) {
            // the first run of a shape builds what later calls reuse (packs, captured graphs,
            // workspaces); the second is timed
            run_many(solvers, Y, ldy, k, probe, X, Dm, stats, c, caller);
            WX_HIP(hipStreamSynchronize(caller));
            const auto t0 = std::chrono::steady_clock::now();
            run_many(solvers, Y, ldy, k, probe, X, Dm, stats, c, caller);
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (std::getenv("WITHIN_AMD_VERBOSE"))
                fprintf(stderr, "[within_amd] autotune streams %d group %d graph %d pack %d tile %u: %.2f ms\n",
                        c.streams, c.group, (int)c.graph, (int)c.pack, c.tile, dt * 1e3);
            if (dt < best) {
                best = dt;
                shape = c;
            }
        };
        for (const ManyShape &c : cand) trial(c);
        // then the packs' work-item tile: rows per D' and gather item (more items and more
        // split-level partials when small, fewer and longer items when large)
        if (shape.pack) {
            const ManyShape base = shape;
            for (uint32_t t : {32u, 128u, 256u}) {
                ManyShape c = base;
                c.tile = t;
                trial(c);
            }
        }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned.push_back({cls, shape});
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_last_shape = shape;
    }
    run_many(solvers, Y, ldy, k, opts, X, Dm, stats, shape, caller);
}

void force_many_tile(int32_t rows) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_forced_tile = rows <= 0 ? 0u : (uint32_t)std::max(16, std::min(rows, 4096));
}

int32_t last_many_tile() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return (int32_t)g_last_shape.tile;
}

void force_many_shape(int32_t streams, int32_t group, int32_t graph, int32_t pack) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_forced.streams = std::max(0, std::min(streams, 64));
    g_forced.group = std::max(1, std::min(group, 64));
    g_forced.graph = graph != 0;
    g_forced.pack = pack != 0;
}

void last_many_shape(int32_t *streams, int32_t *group, int32_t *graph, int32_t *pack) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (streams) *streams = g_last_shape.streams;
    if (group) *group = g_last_shape.group;
    if (graph) *graph = g_last_shape.graph ? 1 : 0;
    if (pack) *pack = g_last_shape.pack ? 1 : 0;
}

std::vector<double> &last_timing() {
    static thread_local std::vector<double> t;
    return t;
}

// ---------------------------------------------------------------------------
// Adjoint of the projection (reverse-mode differentiation of R = Z - D x(Z, w))
// ---------------------------------------------------------------------------
namespace {
// Y[r + j*n] = G[r + j*ldg] / w_r   (caller order; w is stored in internal order)
__global__ void k_div_weights(uint64_t n, const uint32_t *__restrict__ inv_perm, const double *__restrict__ w,
                              const double *__restrict__ G, uint64_t ldg, double *__restrict__ Y, uint32_t k) {
    const uint64_t idx = (uint64_t)blockIdx.x * BLK + threadIdx.x;
    if (idx >= n * k) return;
    const uint64_t j = idx / n, r = idx - j * n;
    const double wr = w ? w[inv_perm ? inv_perm[r] : r] : 1.0;
    Y[r + j * n] = G[r + j * ldg] / wr;
}

// One fused pass per caller row r: (D lambda)_r for every column (the gather of the
// operator, unweighted, column-major lambda), then grad_z = G - w .* (D lambda) and
// grad_w = -sum_j R_j .* (D lambda)_j.
__global__ __launch_bounds__(BLK) void k_projection_grad(DesignArgs d, const uint32_t *__restrict__ inv_perm,
                                                          const double *__restrict__ w,
                                                          const double *__restrict__ lam, uint64_t N,
                                                          const double *__restrict__ G, uint64_t ldg,
                                                          const double *__restrict__ R, uint64_t ldr,
                                                          double *__restrict__ gz, double *__restrict__ gw,
                                                          uint32_t k) {
    const uint64_t stride = (uint64_t)gridDim.x * BLK;
    for (uint64_t r = (uint64_t)blockIdx.x * BLK + threadIdx.x; r < d.n; r += stride) {
        const uint64_t i = inv_perm ? inv_perm[r] : r;
        const double wi = w ? w[i] : 1.0;
        double gacc = 0.0;
        for (uint32_t j = 0; j < k; ++j) {
            double dl = 0.0;
            for (uint32_t t = 0; t < d.T; ++t) {
                const TermInfo &ti = d.terms[t];
                const uint64_t code = d.codes[(uint64_t)t * d.n + i];
                double tv = -0.0;
                for (uint32_t c = 0; c < ti.ncols; ++c)
                    tv = tv + lam[ti.offset + c * ti.n_levels + code + (uint64_t)j * N] * term_load(d, ti.load[c], i);
                dl += tv;
            }
            gz[r + (uint64_t)j * d.n] = G[r + (uint64_t)j * ldg] - wi * dl;
            gacc -= R[r + (uint64_t)j * ldr] * dl;
        }
        if (gw) gw[r] = gacc;
    }
}
} // namespace

void projection_grad_device(const Solver &s, const double *G, uint64_t ldg, const double *R, uint64_t ldr, uint64_t k,
                            const wx_lsmr_options &opts, double *grad_z, double *grad_w, wx_rhs_stats *stats,
                            hipStream_t st) {
    const DevDesign &dv = s.dev;
    const uint64_t n = dv.n, N = dv.N;
    if (ldg < n || ldr < n) fail(WX_ERR_INVALID_INPUT, "leading dimension smaller than n_obs");
    if (!(s.min_weight > 0.0))
        fail(WX_ERR_INVALID_INPUT, "the projection's adjoint needs strictly positive weights");
    if (k == 0) return;
    const uint32_t *inv_perm = dv.inv_perm.p ? dv.inv_perm.as<uint32_t>() : nullptr;
    const double *w = dv.w.p ? dv.w.as<double>() : nullptr;
    // adjoint solve: D'WD lambda = D'G is the normal equation of the projection of G ./ w
    DevBuf Y, lam;
    Y.alloc(n * k * 8);
    lam.alloc(N * k * 8);
    hipLaunchKernelGGL(k_div_weights, dim3((uint32_t)((n * k + BLK - 1) / BLK)), dim3(BLK), 0, st, n, inv_perm, w, G,
                       ldg, Y.as<double>(), (uint32_t)k);
    WX_HIP(hipGetLastError());
    solve_batch_device(s, Y.as<double>(), n, k, opts, lam.as<double>(), nullptr, stats, st, true);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + BLK - 1) / BLK, 65535);
    hipLaunchKernelGGL(k_projection_grad, dim3(std::max<uint32_t>(grid, 1)), dim3(BLK), 0, st, dv.args(), inv_perm, w,
                       lam.as<double>(), N, G, ldg, R, ldr, grad_z, grad_w, (uint32_t)k);
    WX_HIP(hipGetLastError());
    WX_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// Standalone operator applies (the Operator seam, crates/schwarz-precond/src/lib.rs:48-57,
// implemented by DesignOperator, crates/within/src/operator/design.rs:92-117), in caller row order.
// ---------------------------------------------------------------------------
namespace {
// xi[d*k + j] = X[d + j*ld]
__global__ void k_interleave(uint64_t rows, const double *__restrict__ X, uint64_t ld,
                             const uint32_t *__restrict__ perm, double *__restrict__ xi, uint32_t k) {
    const uint64_t idx = (uint64_t)blockIdx.x * BLK + threadIdx.x;
    if (idx >= rows * k) return;
    const uint64_t r = idx / k, j = idx % k;
    xi[idx] = X[(perm ? perm[r] : r) + j * ld];
}
// Y[perm(r) + j*ld] = yi[r*k + j]
__global__ void k_deinterleave(uint64_t rows, const double *__restrict__ yi, double *__restrict__ Y, uint64_t ld,
                               const uint32_t *__restrict__ perm, uint32_t k) {
    const uint64_t idx = (uint64_t)blockIdx.x * BLK + threadIdx.x;
    if (idx >= rows * k) return;
    const uint64_t r = idx / k, j = idx % k;
    Y[(perm ? perm[r] : r) + j * ld] = yi[idx];
}
// y = sqrt(w) .* (D x), no update term
template <bool WEIGHTED>
__global__ __launch_bounds__(BLK) void k_gather_plain(DesignArgs d, const double *__restrict__ sw,
                                                      const double *__restrict__ x, double *__restrict__ y, uint32_t k,
                                                      uint32_t kshift) {
    const uint32_t kp = 1u << kshift;
    const uint32_t j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint64_t rstride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t i = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; i < d.n; i += rstride) {
        double acc = 0.0;
        for (uint32_t t = 0; t < d.T; ++t) {
            const TermInfo &ti = d.terms[t];
            const uint64_t code = d.codes[(uint64_t)t * d.n + i];
            double tv = -0.0;
            for (uint32_t c = 0; c < ti.ncols; ++c)
                tv = tv + x[(ti.offset + c * ti.n_levels + code) * k + j] * term_load(d, ti.load[c], i);
            acc += tv;
        }
        if (WEIGHTED) acc *= sw[i];
        y[i * k + j] = acc;
    }
}
} // namespace

void design_apply_device(const Solver &s, bool adjoint, const double *X, uint64_t k_total, double *Y,
                         hipStream_t st) {
    const DevDesign &dv = s.dev;
    const uint64_t n = dv.n, N = dv.N;
    const DesignArgs da = dv.args();
    const bool weighted = dv.sw.p != nullptr;
    const uint32_t *obs_perm = dv.obs_perm.p ? dv.obs_perm.as<uint32_t>() : nullptr;
    for (uint64_t c0 = 0; c0 < k_total; c0 += MAX_BATCH) {
        const uint32_t k = (uint32_t)std::min<uint64_t>(MAX_BATCH, k_total - c0);
        const uint32_t kshift = pow2_shift(k);
        DevBuf in, out, part, pieces;
        in.alloc((adjoint ? n : N) * k * 8);
        out.alloc((adjoint ? N : n) * k * 8);
        part.alloc((uint64_t)std::max<uint32_t>(dv.max_partials, 1) * k * 8);
        if (adjoint && dv.pair.p >= 0) pieces.alloc((uint64_t)std::max<uint32_t>(dv.pair.n_pieces, 1) * k * 8);
        if (!adjoint) {
            hipLaunchKernelGGL(k_interleave, dim3((N * k + BLK - 1) / BLK), dim3(BLK), 0, st, N, X + c0 * N, N,
                               nullptr, in.as<double>(), k);
            const uint32_t g = grid_for(n, kshift, 65535);
            if (weighted)
                hipLaunchKernelGGL((k_gather_plain<true>), dim3(g), dim3(BLK), 0, st, da, dv.sw.as<double>(),
                                   in.as<double>(), out.as<double>(), k, kshift);
            else
                hipLaunchKernelGGL((k_gather_plain<false>), dim3(g), dim3(BLK), 0, st, da, nullptr, in.as<double>(),
                                   out.as<double>(), k, kshift);
            hipLaunchKernelGGL(k_deinterleave, dim3((n * k + BLK - 1) / BLK), dim3(BLK), 0, st, n, out.as<double>(),
                               Y + c0 * n, n, obs_perm, k);
        } else {
            hipLaunchKernelGGL(k_interleave, dim3((n * k + BLK - 1) / BLK), dim3(BLK), 0, st, n, X + c0 * n, n,
                               obs_perm, in.as<double>(), k);
            launch_adjoint(dv, da, weighted ? 1 : 0, in.as<double>(), out.as<double>(), part.as<double>(), nullptr, k,
                           kshift, st, -1, pieces.p ? pieces.as<double>() : nullptr);
            hipLaunchKernelGGL(k_deinterleave, dim3((N * k + BLK - 1) / BLK), dim3(BLK), 0, st, N, out.as<double>(),
                               Y + c0 * N, N, nullptr, k);
        }
        WX_HIP(hipGetLastError());
        WX_HIP(hipStreamSynchronize(st));
    }
}

} // namespace wx
