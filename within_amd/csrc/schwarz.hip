// schwarz.hip -- device apply of the additive Schwarz preconditioner (gfx950).
//
// z = sum_c R_c' W_c A_c^+ W_c R_c r over factor-pair components, for k interleaved columns.
// The local solve of a component follows crates/within/src/block_elim/solver.rs:350-466 with
// the subdomain wrapper crates/schwarz-precond/src/local_solve.rs:106-147:
//   restrict + fold (negate the kept block) -> subtract the component mean -> eliminate the
//   larger block (1/D_e scaling, C' SpMV into the kept block) -> kept-block solve ->
//   back-substitution x_e = D_e^{-1}(b_e + C x_k) -> subtract the mean -> unfold -> prolong.
// The combine is a deterministic gather over a DOF -> slot map (no atomics), replacing the
// reference's atomic / per-worker-buffer reductions (schwarz/executor.rs:25-165).
// Coefficients are stored in the precision of the run (double, or float in mixed mode);
// every vector and every accumulation is FP64, so the operator stays exactly linear and
// symmetric up to FP64 rounding of fixed coefficients.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "schwarz.hpp"

namespace wx {

namespace {

template <typename T> inline T *dp(const DevBuf &b) { return static_cast<T *>(b.p); }

constexpr int BLK = 256;
constexpr int SP_BLK = 1024; // threads of the per-component sparse triangular solve (one workgroup per component)
constexpr uint32_t ROW_CHUNK = 128;   // CSR entries per work item before a row is split
                                      // (256: too few items for the C' SpMV's long rows)
constexpr uint32_t MEAN_CHUNK = 64;   // slots per partial-sum chunk
constexpr uint32_t MEAN_ITEM = 16;    // chunks per workgroup of the component means
constexpr uint32_t MEAN_SMALL = 16;   // one-chunk components per workgroup of the means
constexpr uint32_t DENSE_TILED_MIN = 256; // kept blocks at least this large use the tiled product
constexpr uint64_t DENSE_TILE_TARGET = 8192; // (row tile, column segment) pairs wanted in flight
constexpr uint64_t LONG_ROW = 64;         // factor rows with more entries are summed by a whole workgroup
constexpr uint64_t BIG_LEVEL = 256;       // factor levels with at least this many rows use grid-wide launches
constexpr uint32_t TAIL_LEVEL_ROWS = 64;  // factor levels this small at the end of the schedule ...
constexpr uint32_t TAIL_MAX_ROWS = 1024;  // ... up to this many rows ...
constexpr uint32_t TAIL_MIN_LEVELS = 8;   // ... and at least this many levels form the dense tail,
constexpr uint64_t TAIL_FILL = 8;         // ... whose nt^2 entries stay within this many times the
                                          // component's factor entries and rows
constexpr uint32_t TAIL_CHUNK = 32;       // tail-row entries per partial sum

struct RowItem {
    uint32_t row;   // slot id of the output row
    uint32_t pad;
    uint64_t begin, end; // CSR entry range
    int64_t dest;   // >= 0 final row, < 0 partial slot -dest-1
};
struct RowFix {
    uint32_t row;
    uint32_t count;
    uint64_t first;
};
struct Chunk {
    uint32_t comp;
    uint32_t begin, end; // slot range
    uint32_t pad;
};
// A work item of the component means.  kind 0: chunks [chunk0, chunk0 + n) (n <= MEAN_ITEM) of
// component comp, whose items are first .. first + nitems - 1 of the list; kind 1: the n chunks
// from chunk0 (n <= MEAN_SMALL), each the only chunk of its component, a row group per chunk.
struct MeanItem {
    uint32_t kind, comp, chunk0, n, first, nitems;
};

template <typename CT>
struct SchwarzDev {
    uint64_t ns = 0, nc = 0, N = 0;
    DevBuf glob, sgnpw, inv_de, slot_comp;
    // C' over kept rows (entry index = eliminated slot) and C over eliminated rows (entry = kept slot)
    DevBuf ct_idx, ct_val, ct_items, ct_fix;
    uint32_t n_ct_items = 0, n_ct_fix = 0, n_ct_part = 0, n_ct_fix_small = 0;
    DevBuf c_idx, c_val, c_items, c_fix;
    uint32_t n_c_items = 0, n_c_fix = 0, n_c_part = 0, n_c_fix_small = 0;
    // means: chunks over all slots (per component) and over kept slots of sparse components
    DevBuf all_chunks, kept_chunks;
    uint32_t n_all_chunks = 0, n_kept_chunks = 0;
    // the component means (sk_mean): the work items over each chunk list
    DevBuf all_items, kept_items;
    uint32_t n_all_items = 0, n_kept_items = 0;
    // the components of several items in each list (sk_mean_combine): comp, first, nitems
    DevBuf all_comb, kept_comb;
    uint32_t n_all_comb = 0, n_kept_comb = 0;
    DevBuf comp_count, comp_kcount;
    // dense kept blocks: one entry per kept row: (matrix row offset, first kept slot, nk)
    DevBuf drow_off, drow_base, drow_nk, drow_slot, dmat;
    uint32_t n_drows = 0;
    // large dense kept blocks: split-K tiles, per-component segment counts and row starts
    DevBuf dtiles, dcomps, drow_start;
    uint32_t n_dtiles = 0, n_dcomps = 0;
    // FP32 runs: the large blocks as their upper triangle of 64 x 64 blocks (sk_dense_upper),
    // one work item per super-tile of 4 x 4 blocks
    DevBuf dsym, stiles;
    uint32_t n_stiles = 0;
    uint64_t dense_part_rows = 0;
    uint64_t dense_tiled_rows = 0; // kept rows of the tiled dense blocks (sk_dense_reduce's outputs)
    // sparse kept blocks (one workgroup per component): levels and factor rows
    DevBuf sp_comp, sp_lvl_f_ptr, sp_lvl_b_ptr, sp_lvl_f_rows, sp_lvl_b_rows, sp_meta;
    DevBuf sp_lvl_f_long, sp_lvl_b_long; // per level: first long row
    DevBuf grid_f_rows, grid_b_rows;     // GridRow lists of the grid-wide steps
    std::vector<uint64_t> grid_f_ptr, grid_b_ptr; // step s: rows [ptr[s], ptr[s+1])
    DevBuf cover_rows;                   // auxiliary slots of the Gremban covers
    uint64_t n_cover_rows = 0;
    DevBuf diag_rows;                    // every factor row (GridRow), for the pivots
    uint64_t n_diag_rows = 0;
    DevBuf l_ptr, l_idx, l_val, lt_ptr, lt_idx, lt_val, piv;
    DevBuf tail_rows, tail_cptr, tail_chunks, tail_idx, tail_val, tail_mat; // dense tails
    DevBuf tail_prod; // TailProdRow of each tail row
    uint64_t n_tail_rows = 0, n_tail_chunks = 0;
    uint32_t n_sparse = 0;
    // auxiliary slots [ns, ns + n_aux) of the covers: real kept slot | sheet bit (second sheet)
    DevBuf aux_src;
    uint64_t n_aux = 0;
    // prolongation: DOF -> slots
    DevBuf d_ptr, d_slot;
    DevBuf d_w, d_comp; // per d_slot entry: its slot's sgnpw and component (no dependent lookup)
    bool has_sparse = false;
};

// per-sparse-component metadata
struct SpMeta {
    uint32_t comp;
    uint32_t kept_base; // first factor row's slot: a kept slot, or an auxiliary slot of a cover
    uint32_t nk;        // factor rows (twice the kept block for a cover)
    uint32_t nlev_f;
    uint32_t nlev_b;
    uint32_t cover;     // 1: the factor is a Gremban cover's, on auxiliary slots
    uint64_t lvl_f_off; // into lvl_f_ptr
    uint64_t lvl_b_off;
    uint64_t row_off;   // first row pointer of the component in l_ptr / lt_ptr
    uint64_t piv_off;   // first pivot
    uint32_t nbig_f;    // leading forward levels solved by grid-wide launches (sk_sparse_grid)
    uint32_t nbig_b;    // trailing backward levels solved by grid-wide launches
    // dense tail: the last-eliminated nt factor rows (the last levels, a chain of hubs) are
    // solved as one dense product x_T = M_T (rhs_T - L_TF y_F); nt = 0: none
    uint32_t nt;
    uint32_t n_tchunks; // chunks of the tail rows' entries on earlier rows
    uint64_t tail_off;  // into tail_rows / tail_cptr (rows) and the tail scratch
    uint64_t tchunk_off; // into tail_chunks and the chunk scratch
    uint64_t tmat_off;  // M_T, nt x nt row-major
};

// A run of at most TAIL_CHUNK entries of one tail row (its entries on rows before the tail)
struct TailChunk {
    uint32_t row;     // tail-local row
    uint32_t pad;
    uint64_t p0, p1;  // into tail_idx / tail_val
};

// A tail row's place in its component's dense tail: its row of M_T (offset into tail_mat), the
// offset of the component's tail rows (into tail_rows and tsc), the tail size and the
// component's SpMeta
struct TailProdRow {
    uint64_t mrow;
    uint64_t toff;
    uint32_t nt, meta;
};

// Dense tails of the sparse factors (device pointers) and their scratch: per tail row the
// reduced right-hand side (tsc), per chunk a partial sum (csc)
template <typename CT> struct TailArgs {
    const uint32_t *rows;  // slot of each tail row
    const uint64_t *cptr;  // first chunk of each tail row (one more entry at the end)
    const TailChunk *chunks;
    const uint32_t *idx;
    const CT *val;
    const CT *mat;         // M_T of each component, row-major
    CT *tsc, *csc;         // reduced tail right-hand sides, chunk partials (slot-space precision)
};

// A row of a large level, solved by a grid-wide launch: the component's SpMeta and the row slot.
struct GridRow {
    uint32_t meta;
    uint32_t row;
};

// A solve passes its per-column "active" flags: when no column is active (the iteration
// launched after the last column converged) every kernel returns at once.
__device__ inline bool gate_closed(const int32_t *gate, uint32_t k) {
    if (!gate) return false;
    int32_t a = 0;
    for (uint32_t j = 0; j < k; ++j) a |= gate[j];
    return a == 0;
}

// Sum over the threads of a workgroup that share (tid & (kp-1)), by a fixed-order LDS tree;
// every thread of the workgroup must call it.
__device__ inline double cols_block_sum(double v, uint32_t kp, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t s = BLK / 2; s >= kp && s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[threadIdx.x];
}

// ---------------------------------------------------------------------------
// kernels (kp lanes per row; lane j = RHS column)
// ---------------------------------------------------------------------------
template <typename CT>
__global__ __launch_bounds__(BLK) void sk_restrict(uint64_t ns, const uint32_t *__restrict__ glob,
                                                   const CT *__restrict__ sgnpw, const double *__restrict__ r,
                                                   CT *__restrict__ loc, uint32_t k, uint32_t kshift, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint64_t stride = ((uint64_t)gridDim.x * BLK) >> kshift;
    uint64_t s = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift;
    for (; s + stride < ns; s += 2 * stride) { // two slots' loads in flight
        const uint64_t s1 = s + stride;
        const uint64_t g0 = glob[s], g1 = glob[s1];
        const double w0 = (double)sgnpw[s], w1 = (double)sgnpw[s1];
        const double r0 = r[g0 * k + j], r1 = r[g1 * k + j];
        loc[s * k + j] = (CT)(w0 * r0);
        loc[s1 * k + j] = (CT)(w1 * r1);
    }
    if (s < ns) loc[s * k + j] = (CT)((double)sgnpw[s] * r[(uint64_t)glob[s] * k + j]);
}

// The restriction by DOF with the p~ update of the LSMR step folded in (ApplyFusion): a row
// group per DOF forms p~ for its active columns (an inactive column keeps its p~), stores it,
// and writes each of the DOF's slots, loc[s] = sgnpw[s] * p~.  Every slot belongs to one DOF,
// so every slot is written once, with the value sk_restrict gives from the stored p~.
template <typename CT>
__global__ __launch_bounds__(BLK) void sk_restrict_dof(uint64_t N, const uint64_t *__restrict__ d_ptr,
                                                       const uint32_t *__restrict__ d_slot,
                                                       const CT *__restrict__ d_w, const double *__restrict__ atu,
                                                       double *__restrict__ pt, ApplyFusion fz, CT *__restrict__ loc,
                                                       uint32_t k, uint32_t kshift, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint32_t c = (fz.blk_col ? fz.blk_col[blockIdx.x] : 0u) + j;
    const bool act = fz.active[c] != 0, brk = fz.brk[c] != 0;
    const double bi = fz.beta_inv[c], pc = fz.p_coeff[c];
    const uint64_t stride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t d = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; d < N; d += stride) {
        const uint64_t q = d * k + j;
        double p;
        if (!act) p = pt[q];
        else {
            if (fz.init) p = bi * pt[q];
            else p = brk ? 0.0 : bi * atu[q] + (-pc) * pt[q];
            pt[q] = p;
        }
        const uint64_t e = d_ptr[d + 1];
        uint64_t i = d_ptr[d];
        for (; i + 2 <= e; i += 2) { // two slots' stores in flight
            const uint64_t s0 = d_slot[i], s1 = d_slot[i + 1];
            const double w0 = (double)d_w[i], w1 = (double)d_w[i + 1];
            loc[s0 * k + j] = (CT)(w0 * p);
            loc[s1 * k + j] = (CT)(w1 * p);
        }
        if (i < e) {
            const uint64_t s = d_slot[i];
            loc[s * k + j] = (CT)((double)d_w[i] * p);
        }
    }
}

// Mean of component c over the item sums part[q0 .. q1), by the whole workgroup: row groups
// stride over the items with eight accumulators (item q goes to (q - q0) / ngrp mod 8, eight
// loads in flight), combined in a fixed order, then a fixed-order LDS tree across the groups.
__device__ inline void comp_mean_one(uint32_t c, uint32_t q0, uint32_t q1, const double *__restrict__ part,
                                     const uint32_t *__restrict__ count, double *__restrict__ mu, uint32_t k,
                                     uint32_t kshift, double *sh) {
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1), grp = threadIdx.x >> kshift;
    const uint32_t ngrp = BLK >> kshift;
    double a8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (j < k) {
        uint32_t q = q0 + grp;
        for (; q + 7 * ngrp < q1; q += 8 * ngrp) {
#pragma unroll
            for (int m = 0; m < 8; ++m) a8[m] += part[(uint64_t)(q + m * ngrp) * k + j];
        }
#pragma unroll
        for (int m = 0; m < 7; ++m)
            if (q + m * ngrp < q1) a8[m] += part[(uint64_t)(q + m * ngrp) * k + j];
    }
    const double acc = ((a8[0] + a8[1]) + (a8[2] + a8[3])) + ((a8[4] + a8[5]) + (a8[6] + a8[7]));
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t s = BLK / 2; s >= kp && s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < k) {
        const uint32_t n = count[c];
        mu[(uint64_t)c * k + threadIdx.x] = n ? sh[threadIdx.x] / (double)n : 0.0;
    }
    __syncthreads();
}

// acc plus column j of the chunk's slots, added in slot order with four loads in flight
template <typename VT>
__device__ inline double chunk_add(double acc, const Chunk &C, const VT *__restrict__ src, uint32_t k, uint32_t j) {
    uint32_t q = C.begin;
    for (; q + 4 <= C.end; q += 4) {
        double t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) t[u] = (double)src[(uint64_t)(q + u) * k + j];
        acc += t[0];
        acc += t[1];
        acc += t[2];
        acc += t[3];
    }
    for (; q < C.end; ++q) acc += (double)src[(uint64_t)q * k + j];
    return acc;
}

// Component means, workgroup b taking work item b (MeanItem).  A kind-0 item's row groups add
// their chunks' slots in order and a fixed LDS tree adds the groups.  When the item is its
// component's only one, that is the component's sum and the item writes the mean; otherwise the
// item's sums go to part[b] and sk_mean_combine adds them.  A kind-1 item's row groups each form
// the mean of a one-chunk component.  Every path divides the same sum, (total + 0.0), by the
// count.
template <typename VT>
__global__ __launch_bounds__(BLK) void sk_mean(const MeanItem *__restrict__ items, const Chunk *__restrict__ chunks,
                                               const VT *__restrict__ src, const uint32_t *__restrict__ count,
                                               double *__restrict__ part, double *__restrict__ mu, uint32_t k,
                                               uint32_t kshift, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    __shared__ double sh[BLK];
    const MeanItem it = items[blockIdx.x];
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1), grp = threadIdx.x >> kshift;
    const uint32_t ngrp = BLK >> kshift;
    if (it.kind == 1) {
        if (j < k)
            for (uint32_t q = grp; q < it.n; q += ngrp) {
                const Chunk C = chunks[it.chunk0 + q];
                const double tot = chunk_add(0.0, C, src, k, j);
                const uint32_t n = count[C.comp];
                mu[(uint64_t)C.comp * k + j] = n ? (tot + 0.0) / (double)n : 0.0;
            }
        return;
    }
    double acc = 0.0;
    if (j < k)
        for (uint32_t ch = it.chunk0 + grp; ch < it.chunk0 + it.n; ch += ngrp) acc = chunk_add(acc, chunks[ch], src, k, j);
    const double tot = cols_block_sum(acc, kp, sh);
    if (threadIdx.x < kp && j < k) {
        if (it.nitems == 1) {
            const uint32_t n = count[it.comp];
            mu[(uint64_t)it.comp * k + j] = n ? (tot + 0.0) / (double)n : 0.0;
        } else {
            part[(uint64_t)blockIdx.x * k + j] = tot;
        }
    }
}

// The means of the components with several work items: workgroup b adds component
// comb[b].comp's item sums in item order (comp_mean_one).
__global__ __launch_bounds__(BLK) void sk_mean_combine(const MeanItem *__restrict__ comb, const double *__restrict__ part,
                                                       const uint32_t *__restrict__ count, double *__restrict__ mu,
                                                       uint32_t k, uint32_t kshift, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    __shared__ double sh[BLK];
    const MeanItem it = comb[blockIdx.x];
    comp_mean_one(it.comp, it.first, it.first + it.nitems, part, count, mu, k, kshift, sh);
}

// MODE 0: reduced rhs of kept rows:  red[row] = (loc[row]-mu) + sum_p ct[p]*(loc[e]-mu), the C'
//         values ct already scaled by their eliminated row's 1 / D_e
// MODE 1: back-substitution of eliminated rows: xs[row] = inv_de[row]*((loc[row]-mu) + sum_p c[p]*xs[kept])
template <typename CT, int MODE>
__device__ inline double row_term(const uint32_t *idx, const CT *val, uint64_t p, const CT *loc,
                                  const CT *xs, const CT *inv_de, const double *mu_c, uint32_t k, uint32_t j) {
    const uint64_t e = idx[p];
    if (MODE == 0) return (double)val[p] * ((double)loc[e * k + j] - mu_c[j]);
    return (double)val[p] * (double)xs[e * k + j];
}

// the output of an elimination SpMV (the reduced right-hand side, MODE 0, or the back-substituted
// slots, MODE 1): a slot-space vector in the coefficient precision
template <typename CT, int MODE> using OutT = CT;

template <typename CT, int MODE>
__device__ inline void row_finish(uint32_t row, double sum, const CT *loc, OutT<CT, MODE> *out, const CT *inv_de,
                                  const double *mu_c, uint32_t k, uint32_t j) {
    const double b = (double)loc[(uint64_t)row * k + j] - mu_c[j];
    if (MODE == 0) out[(uint64_t)row * k + j] = (CT)(b + sum);
    else {
        const double id = (double)inv_de[row];
        out[(uint64_t)row * k + j] = (CT)(id * b + id * sum);
    }
}

template <typename CT, int MODE>
__global__ __launch_bounds__(BLK) void sk_rows(const RowItem *__restrict__ items, uint32_t n_items,
                                               const uint32_t *__restrict__ idx, const CT *__restrict__ val,
                                               const uint32_t *__restrict__ slot_comp, const CT *__restrict__ loc,
                                               const CT *__restrict__ xs, const CT *__restrict__ inv_de,
                                               const double *__restrict__ mu, OutT<CT, MODE> *__restrict__ out,
                                               CT *__restrict__ part, uint32_t k, uint32_t kshift, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint32_t groups = (gridDim.x * BLK) >> kshift;
    for (uint32_t it = (blockIdx.x * BLK + threadIdx.x) >> kshift; it < n_items; it += groups) {
        const RowItem item = items[it];
        const double *mu_c = mu + (uint64_t)slot_comp[item.row] * k;
        double acc = 0.0;
        uint64_t p = item.begin;
        for (; p + 4 <= item.end; p += 4) { // four independent gathers in flight, summed in order
            double t[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) t[q] = row_term<CT, MODE>(idx, val, p + q, loc, xs, inv_de, mu_c, k, j);
            acc += t[0];
            acc += t[1];
            acc += t[2];
            acc += t[3];
        }
        for (; p < item.end; ++p) acc += row_term<CT, MODE>(idx, val, p, loc, xs, inv_de, mu_c, k, j);
        if (item.dest >= 0) row_finish<CT, MODE>(item.row, acc, loc, out, inv_de, mu_c, k, j);
        else part[(uint64_t)(-item.dest - 1) * k + j] = (CT)acc;
    }
}

// sk_rows with two columns per lane (one load of the two columns of a gathered vector row,
// half the lanes per row) for an even number of columns; per column the arithmetic and the
// summation order are those of sk_rows.
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt2 __attribute__((ext_vector_type(2)));
template <typename CT> struct Vec2;
template <> struct Vec2<double> { typedef dbl2 type; };
template <> struct Vec2<float> { typedef flt2 type; };

// two consecutive columns of a slot-space vector, widened to FP64 / narrowed for a store
template <typename CT> __device__ inline dbl2 load2(const CT *p) {
    const typename Vec2<CT>::type v = *(const typename Vec2<CT>::type *)p;
    dbl2 r;
    r.x = (double)v.x;
    r.y = (double)v.y;
    return r;
}
template <typename CT> __device__ inline void store2(CT *p, double x, double y) {
    typename Vec2<CT>::type v;
    v.x = (CT)x;
    v.y = (CT)y;
    *(typename Vec2<CT>::type *)p = v;
}

template <typename CT, int MODE>
__device__ inline dbl2 row_term2(const uint32_t *idx, const CT *val, uint64_t p, const CT *loc, const CT *xs,
                                 const CT *inv_de, dbl2 mu, uint32_t k, uint32_t j0) {
    const uint64_t e = idx[p];
    const double v = (double)val[p];
    dbl2 r;
    if (MODE == 0) {
        const dbl2 l = load2<CT>(&loc[e * k + j0]);
        r.x = v * (l.x - mu.x);
        r.y = v * (l.y - mu.y);
    } else {
        const dbl2 x = load2<CT>(&xs[e * k + j0]);
        r.x = v * x.x;
        r.y = v * x.y;
    }
    return r;
}

template <typename CT, int MODE>
__global__ __launch_bounds__(BLK) void sk_rows2(const RowItem *__restrict__ items, uint32_t n_items,
                                                const uint32_t *__restrict__ idx, const CT *__restrict__ val,
                                                const uint32_t *__restrict__ slot_comp, const CT *__restrict__ loc,
                                                const CT *__restrict__ xs, const CT *__restrict__ inv_de,
                                                const double *__restrict__ mu, OutT<CT, MODE> *__restrict__ out,
                                                CT *__restrict__ part, uint32_t k, uint32_t kshift,
                                                const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kv = (1u << kshift) >> 1, j0 = 2 * (threadIdx.x & (kv - 1));
    if (j0 >= k) return;
    const uint32_t groups = (gridDim.x * BLK) / kv;
    for (uint32_t it = (blockIdx.x * BLK + threadIdx.x) / kv; it < n_items; it += groups) {
        const RowItem item = items[it];
        const double *mu_c = mu + (uint64_t)slot_comp[item.row] * k;
        const dbl2 m = *(const dbl2 *)&mu_c[j0];
        double a0 = 0.0, a1 = 0.0;
        uint64_t p = item.begin;
        for (; p + 8 <= item.end; p += 8) { // eight independent gathers in flight, summed in order
            dbl2 t[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) t[q] = row_term2<CT, MODE>(idx, val, p + q, loc, xs, inv_de, m, k, j0);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                a0 += t[q].x;
                a1 += t[q].y;
            }
        }
        for (; p < item.end; ++p) {
            const dbl2 t = row_term2<CT, MODE>(idx, val, p, loc, xs, inv_de, m, k, j0);
            a0 += t.x;
            a1 += t.y;
        }
        if (item.dest >= 0) {
            const dbl2 b = load2<CT>(&loc[(uint64_t)item.row * k + j0]);
            const double b0 = b.x - m.x, b1 = b.y - m.y;
            if (MODE == 0) {
                store2<OutT<CT, MODE>>(&out[(uint64_t)item.row * k + j0], b0 + a0, b1 + a1);
            } else {
                const double id = (double)inv_de[item.row];
                store2<OutT<CT, MODE>>(&out[(uint64_t)item.row * k + j0], id * b0 + id * a0, id * b1 + id * a1);
            }
        } else {
            store2<CT>(&part[(uint64_t)(-item.dest - 1) * k + j0], a0, a1);
        }
    }
}

// sk_rows with four columns per lane for FP32 slot-space vectors (one 16-byte load of four
// columns of a gathered row, a quarter of the lanes per row, so a wave has twice the rows of
// sk_rows2 in flight) for a width divisible by four; per column the arithmetic and the
// summation order are those of sk_rows.
typedef float flt4 __attribute__((ext_vector_type(4)));

template <int MODE>
__global__ __launch_bounds__(BLK) void sk_rows4(const RowItem *__restrict__ items, uint32_t n_items,
                                                const uint32_t *__restrict__ idx, const float *__restrict__ val,
                                                const uint32_t *__restrict__ slot_comp, const float *__restrict__ loc,
                                                const float *__restrict__ xs, const float *__restrict__ inv_de,
                                                const double *__restrict__ mu, OutT<float, MODE> *__restrict__ out,
                                                float *__restrict__ part, uint32_t k, uint32_t kshift,
                                                const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kv = (1u << kshift) >> 2, j0 = 4 * (threadIdx.x & (kv - 1));
    if (j0 >= k) return;
    const uint32_t groups = (gridDim.x * BLK) / kv;
    for (uint32_t it = (blockIdx.x * BLK + threadIdx.x) / kv; it < n_items; it += groups) {
        const RowItem item = items[it];
        const double *mu_c = mu + (uint64_t)slot_comp[item.row] * k + j0;
        const double m0 = mu_c[0], m1 = mu_c[1], m2 = mu_c[2], m3 = mu_c[3];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        uint64_t p = item.begin;
        for (; p + 8 <= item.end; p += 8) { // eight independent gathers in flight, summed in order
            double v[8];
            flt4 g[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint64_t e = idx[p + q];
                v[q] = (double)val[p + q];
                g[q] = *(const flt4 *)&(MODE == 0 ? loc : xs)[e * k + j0];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (MODE == 0) {
                    a0 += v[q] * ((double)g[q].x - m0);
                    a1 += v[q] * ((double)g[q].y - m1);
                    a2 += v[q] * ((double)g[q].z - m2);
                    a3 += v[q] * ((double)g[q].w - m3);
                } else {
                    a0 += v[q] * (double)g[q].x;
                    a1 += v[q] * (double)g[q].y;
                    a2 += v[q] * (double)g[q].z;
                    a3 += v[q] * (double)g[q].w;
                }
            }
        }
        if (p < item.end) { // the last entries: their gathers in flight together, summed in order
            const uint32_t rem = (uint32_t)(item.end - p);
            double v[7];
            flt4 g[7];
#pragma unroll
            for (int q = 0; q < 7; ++q)
                if ((uint32_t)q < rem) {
                    const uint64_t e = idx[p + q];
                    v[q] = (double)val[p + q];
                    g[q] = *(const flt4 *)&(MODE == 0 ? loc : xs)[e * k + j0];
                }
#pragma unroll
            for (int q = 0; q < 7; ++q)
                if ((uint32_t)q < rem) {
                    if (MODE == 0) {
                        a0 += v[q] * ((double)g[q].x - m0);
                        a1 += v[q] * ((double)g[q].y - m1);
                        a2 += v[q] * ((double)g[q].z - m2);
                        a3 += v[q] * ((double)g[q].w - m3);
                    } else {
                        a0 += v[q] * (double)g[q].x;
                        a1 += v[q] * (double)g[q].y;
                        a2 += v[q] * (double)g[q].z;
                        a3 += v[q] * (double)g[q].w;
                    }
                }
        }
        if (item.dest >= 0) {
            const flt4 b = *(const flt4 *)&loc[(uint64_t)item.row * k + j0];
            const double b0 = (double)b.x - m0, b1 = (double)b.y - m1, b2 = (double)b.z - m2, b3 = (double)b.w - m3;
            if (MODE == 0) {
                flt4 o;
                o.x = (float)(b0 + a0);
                o.y = (float)(b1 + a1);
                o.z = (float)(b2 + a2);
                o.w = (float)(b3 + a3);
                *(flt4 *)&out[(uint64_t)item.row * k + j0] = o;
            } else {
                const double id = (double)inv_de[item.row];
                flt4 o;
                o.x = (float)(id * b0 + id * a0);
                o.y = (float)(id * b1 + id * a1);
                o.z = (float)(id * b2 + id * a2);
                o.w = (float)(id * b3 + id * a3);
                *(flt4 *)&out[(uint64_t)item.row * k + j0] = o;
            }
        } else {
            flt4 o;
            o.x = (float)a0;
            o.y = (float)a1;
            o.z = (float)a2;
            o.w = (float)a3;
            *(flt4 *)&part[(uint64_t)(-item.dest - 1) * k + j0] = o;
        }
    }
}

// A split row's partials added by one workgroup: its row groups stride over the partials (a
// hub row can have thousands), then a fixed LDS tree adds the groups' sums.  Block bid of nblk.
template <typename CT, int MODE>
__device__ inline void rows_fix_wide(const RowFix *__restrict__ fixes, uint32_t n_fix,
                                     const uint32_t *__restrict__ slot_comp, const CT *__restrict__ loc,
                                     const CT *__restrict__ inv_de, const double *__restrict__ mu,
                                     const CT *__restrict__ part, OutT<CT, MODE> *__restrict__ out, uint32_t k,
                                     uint32_t kshift, uint32_t bid, uint32_t nblk) {
    __shared__ double sh[BLK];
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    const uint32_t grp = threadIdx.x >> kshift, ngrp = BLK >> kshift;
    for (uint32_t fi = bid; fi < n_fix; fi += nblk) {
        const RowFix f = fixes[fi];
        double acc = 0.0;
        if (j < k)
            for (uint32_t q = grp; q < f.count; q += ngrp) acc += (double)part[(uint64_t)(f.first + q) * k + j];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (uint32_t s = BLK / 2; s >= kp && s > 0; s >>= 1) {
            if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x < kp && j < k)
            row_finish<CT, MODE>(f.row, sh[threadIdx.x], loc, out, inv_de, mu + (uint64_t)slot_comp[f.row] * k, k, j);
        __syncthreads();
    }
}

// The same for split rows of at most FIX_SMALL partials, one row group per row instead of one
// workgroup.  The workgroup version gives row group g the value 0.0 + partial g and then adds
// group g + s into group g for s = ngrp/2, ..., 1; for at most FIX_SMALL <= ngrp partials the
// levels above the partial count only add zeros, so this applies the remaining levels to the
// same values and returns the same bits.
constexpr uint32_t FIX_SMALL = 8;
template <typename CT, int MODE>
__device__ inline void rows_fix_narrow(const RowFix *__restrict__ fixes, uint32_t n_fix,
                                       const uint32_t *__restrict__ slot_comp, const CT *__restrict__ loc,
                                       const CT *__restrict__ inv_de, const double *__restrict__ mu,
                                       const CT *__restrict__ part, OutT<CT, MODE> *__restrict__ out, uint32_t k,
                                       uint32_t kshift, uint32_t bid, uint32_t nblk) {
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint32_t stride = (nblk * BLK) >> kshift;
    for (uint32_t fi = (bid * BLK + threadIdx.x) >> kshift; fi < n_fix; fi += stride) {
        const RowFix f = fixes[fi];
        double v[FIX_SMALL];
#pragma unroll
        for (uint32_t g = 0; g < FIX_SMALL; ++g)
            v[g] = g < f.count ? 0.0 + (double)part[(f.first + g) * k + j] : 0.0;
        uint32_t w = 1;
        while (w < f.count) w <<= 1;
#pragma unroll
        for (uint32_t h = FIX_SMALL / 2; h >= 1; h >>= 1)
            if (h < w)
#pragma unroll
                for (uint32_t g = 0; g < h; ++g) v[g] += v[g + h];
        row_finish<CT, MODE>(f.row, v[0], loc, out, inv_de, mu + (uint64_t)slot_comp[f.row] * k, k, j);
    }
}

// Every split row in one launch: blocks [0, g_small) take the first n_small rows (at most
// FIX_SMALL partials each) a row group per row, the other blocks the rest a workgroup per row.
template <typename CT, int MODE>
__global__ __launch_bounds__(BLK) void sk_rows_fix(const RowFix *__restrict__ fixes, uint32_t n_small,
                                                   uint32_t n_fix, uint32_t g_small,
                                                   const uint32_t *__restrict__ slot_comp,
                                                   const CT *__restrict__ loc, const CT *__restrict__ inv_de,
                                                   const double *__restrict__ mu, const CT *__restrict__ part,
                                                   OutT<CT, MODE> *__restrict__ out, uint32_t k, uint32_t kshift,
                                                   const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    if (blockIdx.x < g_small)
        rows_fix_narrow<CT, MODE>(fixes, n_small, slot_comp, loc, inv_de, mu, part, out, k, kshift, blockIdx.x,
                                  g_small);
    else
        rows_fix_wide<CT, MODE>(fixes + n_small, n_fix - n_small, slot_comp, loc, inv_de, mu, part, out, k, kshift,
                                blockIdx.x - g_small, gridDim.x - g_small);
}

// Large dense kept blocks: xs = S^+ red as a split-K product that streams S^+ once.  S^+ is
// symmetric, so column i of S^+ is row i: one wave owns 64 consecutive output rows (one per
// lane) and walks a segment [l0, l1) of l, reading S^+[l][row0 + lane] -- 64 consecutive
// coefficients, one coalesced load per l -- while red[l][c0 .. c0+KC) is the same for every
// lane (a wave-uniform load).  Each lane keeps KC FP64 accumulators.  Segment partials are
// summed in a fixed order by sk_dense_reduce, so the result is deterministic.
struct DenseTile {
    uint64_t mat_off;   // S^+ of the component (row-major nk x nk, symmetric, rows ld apart)
    uint32_t nk, ld, kept_base;
    uint32_t row0, l0, l1;
    uint32_t seg;       // segment index of [l0, l1)
    uint64_t part_off;  // partial rows of this component: part[(part_off + seg*nk + i)*k + c]
};
struct DenseComp {
    uint32_t nk, kept_base, n_seg, pad;
    uint64_t part_off;
};

template <typename CT, int KC>
__global__ __launch_bounds__(64) void sk_dense_sym(const DenseTile *__restrict__ tiles, const CT *__restrict__ mat,
                                                   const CT *__restrict__ red, CT *__restrict__ part,
                                                   uint32_t kstride, uint32_t c0, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, kstride)) return;
    const DenseTile T = tiles[blockIdx.x];
    const uint32_t row = T.row0 + threadIdx.x;
    const bool on = row < T.nk;
    const CT *M = mat + T.mat_off + (on ? row : 0);
    const CT *R = red + (uint64_t)T.kept_base * kstride + c0;
    // the segment's products accumulate in the coefficient precision: FP64 runs in FP64; FP32
    // runs in FP32 (full-rate FMAs on the wave-uniform right-hand-side operands, a segment being
    // a few hundred terms), their segment partials summed in FP64 by sk_dense_reduce
    CT acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] = (CT)0;
    uint32_t l = T.l0;
    for (; l + 8 <= T.l1; l += 8) { // eight coefficient loads in flight per lane
        CT m[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) m[q] = M[(uint64_t)(l + q) * T.ld];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const CT *rl = R + (uint64_t)(l + q) * kstride;
#pragma unroll
            for (int c = 0; c < KC; ++c) acc[c] += m[q] * rl[c];
        }
    }
    for (; l < T.l1; ++l) {
        const CT m = M[(uint64_t)l * T.ld];
        const CT *rl = R + (uint64_t)l * kstride;
#pragma unroll
        for (int c = 0; c < KC; ++c) acc[c] += m * rl[c];
    }
    if (on) {
        CT *o = part + (T.part_off + (uint64_t)T.seg * T.nk + row) * kstride + c0;
#pragma unroll
        for (int c = 0; c < KC; ++c) o[c] = acc[c];
    }
}

// Large dense blocks of an FP32 run by their upper triangle.  S^+ is symmetric, so a component
// stores only its 64 x 64 blocks (I, J) with I <= J (row-major inside a block, zero past nk),
// block row by block row; the product reads each stored block once and applies it twice:
// y_I += B x_J and, off the diagonal, y_J += B' x_I.  A workgroup takes a super-tile of 4 x 4
// blocks (super-row P, super-column Q, P <= Q), a wavefront per block row.  Each wavefront
// stages its block in LDS (coalesced loads, XOR-swizzled so both the row and the column reads
// are free of bank conflicts), accumulates y_I over the super-tile's block columns, and the
// four wavefronts' y_J are added in wavefront order through LDS.  Every super-tile writes one
// partial per block row (source Q + 1) and one per block column (source P) of its component,
// so output row i of block K has nt + 1 partials (nt super-rows), which sk_dense_reduce adds in
// source order: the result is deterministic, and exactly symmetric in the stored coefficients.
struct SymTile {
    uint64_t blk_off;  // the component's first stored block, in floats
    uint32_t nk, nb;   // kept levels, 64-row blocks
    uint32_t kept_base;
    uint32_t P, Q;     // super-row, super-column
    uint32_t pad;
    uint64_t part_off; // partial rows of the component: part[(part_off + src * nk + i) * k + c]
};

__device__ inline uint64_t upper_block(uint32_t I, uint32_t J, uint32_t nb) {
    return (uint64_t)I * nb - (uint64_t)I * (I - 1) / 2 + (J - I);
}
// float index of element (r, c) of a block staged in LDS: rows of 64 floats, the 16-byte chunks
// of row r permuted by r mod 8, so the 16-byte row reads and writes of eight lanes, and the
// column reads, fall in distinct banks
__device__ inline uint32_t sym_lds(uint32_t r, uint32_t c) { return r * 64 + ((((c >> 2) ^ (r & 7)) << 2) | (c & 3)); }

typedef float sym_f4 __attribute__((ext_vector_type(4)));

// One super-tile per workgroup, a wavefront per block row I; the products on the matrix cores
// (v_mfma_f32_16x16x4f32: D[i][j] += sum_k A[i][k] B[k][j], lane l holding A[l % 16][l / 16],
// B[l / 16][l % 16] and D[4 (l / 16) + v][l % 16] in register v).  The 64-term inner dimension
// is taken as 16 steps s, lane group g = l / 16 supplying term 16 g + s, so a lane's A values
// for y_I = B x_J are 16 consecutive coefficients of one row of B, and for y_J = B' x_I 16
// consecutive coefficients of one column.  kc <= 16 right-hand-side columns from c0.
__global__ __launch_bounds__(256) void sk_dense_upper(const SymTile *__restrict__ tiles,
                                                      const float *__restrict__ blocks,
                                                      const float *__restrict__ red, float *__restrict__ part,
                                                      uint32_t kstride, uint32_t c0, uint32_t kc,
                                                      const int32_t *__restrict__ gate) {
    if (gate_closed(gate, kstride)) return;
    __shared__ float lds[4 * 64 * 64];
    const SymTile T = tiles[blockIdx.x];
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint32_t li = lane & 15, lg = lane >> 4; // MFMA row / column index, lane group
    float *B = lds + w * 4096;
    const uint32_t I = 4 * T.P + w;
    const bool row_ok = I < T.nb;
    const uint32_t rows_I = row_ok ? min(64u, T.nk - 64 * I) : 0u;
    const float *X = red + (uint64_t)T.kept_base * kstride + c0;
    // x_I as the B operand: lane holds x[64 I + 16 lg + s][li], s = 0 .. 15 (0 past nk or kc)
    float xi[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint32_t r = 16 * lg + q;
        xi[q] = (r < rows_I && li < kc) ? X[(uint64_t)(64 * I + r) * kstride + li] : 0.0f;
    }
    sym_f4 acc_row[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) acc_row[m] = sym_f4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t jj = 0; jj < 4; ++jj) {
        const uint32_t J = 4 * T.Q + jj;
        const bool blk = row_ok && J < T.nb && I <= J;
        const uint32_t cols_J = J < T.nb ? min(64u, T.nk - 64 * J) : 0u;
        float xj[16];
        if (blk) {
            // stage block (I, J): 16 coalesced loads of 64 x 16 bytes, chunk m = rows 4m .. 4m + 3
            const float4 *src = (const float4 *)(blocks + T.blk_off + upper_block(I, J, T.nb) * 4096);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float4 v[8];
#pragma unroll
                for (int m = 0; m < 8; ++m) v[m] = src[(8 * h + m) * 64 + lane];
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const uint32_t r = 4 * (8 * h + m) + (lane >> 4), c = 4 * (lane & 15);
                    *(float4 *)&B[sym_lds(r, c)] = v[m];
                }
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const uint32_t r = 16 * lg + q;
                xj[q] = (r < cols_J && li < kc) ? X[(uint64_t)(64 * J + r) * kstride + li] : 0.0f;
            }
        }
        __syncthreads();
        sym_f4 acc_col[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) acc_col[m] = sym_f4{0.0f, 0.0f, 0.0f, 0.0f};
        if (blk) {
#pragma unroll
            for (int m = 0; m < 4; ++m) { // y_I rows 16 m .. 16 m + 15: A = B[16 m + li][16 lg + s]
                float a[16];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 t = *(const float4 *)&B[sym_lds(16 * m + li, 16 * lg + 4 * q)];
                    a[4 * q] = t.x;
                    a[4 * q + 1] = t.y;
                    a[4 * q + 2] = t.z;
                    a[4 * q + 3] = t.w;
                }
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    acc_row[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q], xj[q], acc_row[m], 0, 0, 0);
            }
            if (J != I) {
#pragma unroll
                for (int m = 0; m < 4; ++m) { // y_J rows 16 m .. 16 m + 15: A = B[16 lg + s][16 m + li]
#pragma unroll
                    for (int q = 0; q < 16; ++q)
                        acc_col[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(B[sym_lds(16 * lg + q, 16 * m + li)], xi[q],
                                                                          acc_col[m], 0, 0, 0);
                }
            }
        }
        __syncthreads(); // every wavefront is done with its block
        // this wavefront's y_J to LDS as [column][row], columns 65 floats apart
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) B[li * 65 + 16 * m + 4 * lg + v] = acc_col[m][v];
        __syncthreads();
        if (J < T.nb) { // block column J's partial from this super-row: the four rows' sums in order
            for (uint32_t q = threadIdx.x; q < 64 * kc; q += 256) {
                const uint32_t c = q >> 6, l = q & 63;
                const float sum = ((lds[c * 65 + l] + lds[4096 + c * 65 + l]) + lds[8192 + c * 65 + l]) +
                                  lds[12288 + c * 65 + l];
                if (l < cols_J) part[(T.part_off + (uint64_t)T.P * T.nk + 64 * J + l) * kstride + c0 + c] = sum;
            }
        }
        __syncthreads();
    }
    if (li < kc) { // block row I's partial from this super-column
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint32_t r = 16 * m + 4 * lg + v;
                if (r < rows_I)
                    part[(T.part_off + (uint64_t)(T.Q + 1) * T.nk + 64 * I + r) * kstride + c0 + li] = acc_row[m][v];
            }
    }
}

// Packs component blocks: dst block b (I <= J, block-row order) from the row-major nk x nk src.
template <typename CT>
__global__ __launch_bounds__(256) void sk_pack_upper(const CT *__restrict__ src, uint32_t nk, uint32_t nb,
                                                     float *__restrict__ dst) {
    const uint64_t b = blockIdx.x;
    uint32_t I = 0;
    while (upper_block(I + 1, I + 1, nb) <= b && I + 1 < nb) ++I;
    const uint32_t J = (uint32_t)(b - upper_block(I, I, nb)) + I;
    for (uint32_t e = threadIdx.x; e < 4096; e += 256) {
        const uint32_t r = 64 * I + (e >> 6), c = 64 * J + (e & 63);
        dst[b * 4096 + e] = (r < nk && c < nk) ? (float)src[(uint64_t)r * nk + c] : 0.0f;
    }
}

// xs[kept_base + i] = sum over segments (in order) of the partials
template <typename CT>
__global__ __launch_bounds__(BLK) void sk_dense_reduce(const DenseComp *__restrict__ comps, uint32_t n_comps,
                                                       const uint64_t *__restrict__ row_start,
                                                       const CT *__restrict__ part, CT *__restrict__ xs,
                                                       uint32_t k, uint32_t kshift, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint64_t total = row_start[n_comps];
    const uint64_t stride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t q = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; q < total; q += stride) {
        uint32_t lo = 0, hi = n_comps - 1;
        while (lo < hi) { // component owning global dense row q
            const uint32_t mid = (lo + hi + 1) / 2;
            if (row_start[mid] <= q) lo = mid;
            else hi = mid - 1;
        }
        const DenseComp C = comps[lo];
        const uint64_t i = q - row_start[lo];
        // the segments' partials added in segment order, eight loads in flight
        double acc = 0.0;
        uint32_t s = 0;
        for (; s + 8 <= C.n_seg; s += 8) {
            double t[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) t[q] = (double)part[(C.part_off + (uint64_t)(s + q) * C.nk + i) * k + j];
#pragma unroll
            for (int q = 0; q < 8; ++q) acc += t[q];
        }
        for (; s < C.n_seg; ++s) acc += (double)part[(C.part_off + (uint64_t)s * C.nk + i) * k + j];
        xs[(uint64_t)(C.kept_base + i) * k + j] = (CT)acc;
    }
}

// Dense kept-block solve for small blocks: xs[row] = sum_l S^+[i, l] * red[l]   (S^+ = P S_g^{-1} P)
template <typename CT>
__global__ __launch_bounds__(BLK) void sk_dense(uint32_t n_rows, const uint64_t *__restrict__ row_off,
                                                const uint32_t *__restrict__ row_base,
                                                const uint32_t *__restrict__ row_nk,
                                                const uint32_t *__restrict__ row_slot, const CT *__restrict__ mat,
                                                const CT *__restrict__ red, CT *__restrict__ xs, uint32_t k,
                                                uint32_t kshift, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint32_t groups = (gridDim.x * BLK) >> kshift;
    for (uint32_t r = (blockIdx.x * BLK + threadIdx.x) >> kshift; r < n_rows; r += groups) {
        const CT *m = mat + row_off[r];
        const CT *rd = red + row_base[r] * k + j;
        const uint32_t nk = row_nk[r];
        double acc = 0.0;
        uint32_t l = 0;
        for (; l + 8 <= nk; l += 8) { // eight products' loads in flight, summed in order
            double mv[8], rv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                mv[q] = (double)m[l + q];
                rv[q] = (double)rd[(uint64_t)(l + q) * k];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) acc += mv[q] * rv[q];
        }
        for (; l < nk; ++l) acc += (double)m[l] * (double)rd[(uint64_t)l * k];
        xs[(uint64_t)row_slot[r] * k + j] = (CT)acc;
    }
}

// Right-hand side of a forward-substitution row: the reduced rhs minus the component mean, or
// for a Gremban cover's auxiliary slot the antisymmetric embed of its kept slot's.
template <typename CT>
__device__ inline double sparse_rhs(const SpMeta &m, uint32_t row, const CT *red, const uint32_t *aux_src,
                                    uint64_t ns, double mu, uint32_t k, uint32_t j) {
    if (m.cover) { // antisymmetric embed [b, -b] of the kept block's reduced rhs
        const uint32_t a = aux_src[row - ns];
        const double b = (double)red[(uint64_t)(a & 0x7FFFFFFFu) * k + j];
        return (a >> 31) ? -b : b;
    }
    return (double)red[(uint64_t)row * k + j] - mu;
}

// sum(val[p] * xs[idx[p]]) over [p0, p1) for one column, accumulated into acc in entry order
// (acc -= each term when SUB, else acc += each), with eight entries' loads in flight at a time:
// the chain index -> gathered value is a memory round trip, so one entry at a time would
// leave a hub row of thousands of entries latency-bound.
template <typename CT, bool SUB>
__device__ inline double dot_entries(double acc, uint64_t p0, uint64_t p1, uint64_t step, const uint32_t *idx,
                                     const CT *val, const CT *xs, uint32_t k, uint32_t j) {
    constexpr int U = 8;
    uint64_t p = p0;
    for (; p + (U - 1) * step < p1; p += U * step) {
        uint32_t e[U];
        double v[U], x[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            e[q] = idx[p + q * step];
            v[q] = (double)val[p + q * step];
        }
#pragma unroll
        for (int q = 0; q < U; ++q) x[q] = (double)xs[(uint64_t)e[q] * k + j];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            if (SUB) acc -= v[q] * x[q];
            else acc += v[q] * x[q];
        }
    }
    for (; p < p1; p += step) {
        if (SUB) acc -= (double)val[p] * (double)xs[(uint64_t)idx[p] * k + j];
        else acc += (double)val[p] * (double)xs[(uint64_t)idx[p] * k + j];
    }
    return acc;
}

// One level of a triangular solve.  Rows [b, el) are short: one row group each, entries
// summed in order.  Rows [el, e) are long (more than LONG_ROW entries, typically the hubs
// eliminated last): the whole workgroup splits each one's entries over its row groups and
// adds the partial sums in a fixed LDS tree.  FORWARD: x = rhs - L x; else x = x - L' x.
template <typename CT, bool FORWARD>
__device__ inline void sparse_level(const SpMeta &m, uint64_t b, uint64_t el, uint64_t e, const uint32_t *rows,
                                    const uint64_t *ptr, const uint32_t *idx, const CT *val, const CT *red,
                                    const uint32_t *aux_src, uint64_t ns, double mu, CT *xs, double *sh,
                                    uint32_t k, uint32_t kshift) {
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    const uint32_t grp = threadIdx.x >> kshift, ngrp = SP_BLK >> kshift;
    const bool on = j < k;
    if (on)
        for (uint64_t q = b + grp; q < el; q += ngrp) {
            const uint32_t row = rows[q];
            const uint64_t lr = (uint64_t)(row - m.kept_base) + m.row_off;
            double acc = FORWARD ? sparse_rhs<CT>(m, row, red, aux_src, ns, mu, k, j)
                                 : (double)xs[(uint64_t)row * k + j];
            acc = dot_entries<CT, true>(acc, ptr[lr], ptr[lr + 1], 1, idx, val, xs, k, j);
            xs[(uint64_t)row * k + j] = (CT)acc;
        }
    for (uint64_t q = el; q < e; ++q) {
        const uint32_t row = rows[q];
        const uint64_t lr = (uint64_t)(row - m.kept_base) + m.row_off;
        double part = 0.0;
        if (on) part = dot_entries<CT, false>(0.0, ptr[lr] + grp, ptr[lr + 1], ngrp, idx, val, xs, k, j);
        sh[threadIdx.x] = part;
        __syncthreads();
        for (uint32_t s = SP_BLK / 2; s >= kp && s > 0; s >>= 1) {
            if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x < kp && on) {
            const double base = FORWARD ? sparse_rhs<CT>(m, row, red, aux_src, ns, mu, k, j)
                                        : (double)xs[(uint64_t)row * k + j];
            xs[(uint64_t)row * k + j] = (CT)(base - sh[threadIdx.x]);
        }
        __syncthreads();
    }
}

// Sparse kept-block solve, one workgroup per component: forward L y = red - mu2 level by
// level, y /= d (a zero pivot -- the floating gauge -- gives 0), backward L' x = y level by
// level.  Rows within a level are independent; __syncthreads() orders the levels.  Row and
// column ids are global slot ids; y and x live in xs (in place).  A Gremban cover's factor
// runs on auxiliary slots: its right-hand side is the embed [b, -b] and the kept slots
// receive (y+ - y-) / 2 at the end.
template <typename CT>
__global__ __launch_bounds__(SP_BLK) void sk_sparse(const SpMeta *__restrict__ meta, const uint64_t *__restrict__ lvl_f_ptr,
                                                 const uint64_t *__restrict__ lvl_f_long,
                                                 const uint32_t *__restrict__ lvl_f_rows,
                                                 const uint64_t *__restrict__ lvl_b_ptr,
                                                 const uint64_t *__restrict__ lvl_b_long,
                                                 const uint32_t *__restrict__ lvl_b_rows,
                                                 const uint64_t *__restrict__ l_ptr, const uint32_t *__restrict__ l_idx,
                                                 const CT *__restrict__ l_val, const uint64_t *__restrict__ lt_ptr,
                                                 const uint32_t *__restrict__ lt_idx, const CT *__restrict__ lt_val,
                                                 const CT *__restrict__ piv, const CT *__restrict__ red,
                                                 const double *__restrict__ mu2, const uint32_t *__restrict__ aux_src,
                                                 uint64_t ns, CT *xs, uint32_t k,
                                                 uint32_t kshift, const int32_t *__restrict__ gate, uint32_t phase) {
    if (gate_closed(gate, k)) return;
    __shared__ double sh[SP_BLK];
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    const SpMeta m = meta[blockIdx.x];
    const bool on = j < k;
    const double mu = on ? mu2[(uint64_t)m.comp * k + j] : 0.0;
    if (phase == 0) {
        // forward substitution; l_ptr / piv are indexed by (row slot - kept_base) + row offset;
        // the first nbig_f levels were solved by grid-wide launches before this kernel
        for (uint32_t lv = m.nbig_f; lv < m.nlev_f; ++lv) {
            const uint64_t o = m.lvl_f_off + lv;
            sparse_level<CT, true>(m, lvl_f_ptr[o], lvl_f_long[o], lvl_f_ptr[o + 1], lvl_f_rows, l_ptr, l_idx, l_val,
                                   red, aux_src, ns, mu, xs, sh, k, kshift);
            __syncthreads();
        }
        return;
    }
    // (pivots applied by sk_sparse_diag in between) backward substitution with L' (rows of L' =
    // columns of L); the last nbig_b levels are solved by grid-wide launches after this kernel
    for (uint32_t lv = 0; lv + m.nbig_b < m.nlev_b; ++lv) {
        const uint64_t o = m.lvl_b_off + lv;
        sparse_level<CT, false>(m, lvl_b_ptr[o], lvl_b_long[o], lvl_b_ptr[o + 1], lvl_b_rows, lt_ptr, lt_idx, lt_val,
                                red, aux_src, ns, mu, xs, sh, k, kshift);
        __syncthreads();
    }
}

// The dense tails' chunk partials, all components across the grid (one row group per chunk,
// entries summed in order): csc[chunk] = sum of L entries x xs over the chunk's entries on rows
// before the tail, which the forward phase has solved.
template <typename CT>
__global__ __launch_bounds__(BLK) void sk_tail_chunks(TailArgs<CT> ta, uint64_t n_chunks, const CT *xs, uint32_t k,
                                                      uint32_t kshift, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint64_t stride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t c = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; c < n_chunks; c += stride) {
        const TailChunk ch = ta.chunks[c];
        ta.csc[c * k + j] = (CT)dot_entries<CT, false>(0.0, ch.p0, ch.p1, 1, ta.idx, ta.val, xs, k, j);
    }
}

// The dense tails' reduced right-hand sides tsc = rhs_T - L_TF y_F of every component across the
// grid: one wave per (tail row, column); lane t adds the row's chunk partials t, t + 64, ...
// (sk_tail_chunks), and a fixed shuffle tree adds the lanes' sums.
template <typename CT>
__global__ __launch_bounds__(BLK) void sk_tail_rhs(TailArgs<CT> ta, const TailProdRow *__restrict__ prod,
                                                   uint64_t n_rows, const SpMeta *__restrict__ meta,
                                                   const CT *__restrict__ red, const double *__restrict__ mu2,
                                                   const uint32_t *__restrict__ aux_src, uint64_t ns, uint32_t k,
                                                   const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t task = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> 6; // wave-uniform
    if (task >= n_rows * k) return;
    const uint64_t ra = task / k;
    const uint32_t j = (uint32_t)(task % k);
    const uint64_t c0 = ta.cptr[ra], c1 = ta.cptr[ra + 1];
    double part = 0.0;
    uint64_t c = c0 + lane;
    for (; c + 7 * 64 < c1; c += 8 * 64) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = (double)ta.csc[(c + q * 64) * k + j];
#pragma unroll
        for (int q = 0; q < 8; ++q) part += v[q];
    }
    for (; c < c1; c += 64) part += (double)ta.csc[c * k + j];
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    if (lane == 0) {
        const SpMeta m = meta[prod[ra].meta];
        const double mu = mu2[(uint64_t)m.comp * k + j];
        ta.tsc[ra * k + j] = (CT)(sparse_rhs<CT>(m, ta.rows[ra], red, aux_src, ns, mu, k, j) - part);
    }
}

// The dense tails' products x_T = M_T tsc of every component across the grid (final: no pivot
// step follows): one wave per (tail row, column); lane t takes the products b = t, t + 64, ...
// (eight loads in flight), and a fixed shuffle tree adds the lanes' sums.
template <typename CT>
__global__ __launch_bounds__(BLK) void sk_tail_product(TailArgs<CT> ta, const TailProdRow *__restrict__ prod,
                                                       uint64_t n_rows, CT *__restrict__ xs, uint32_t k,
                                                       const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t task = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> 6; // wave-uniform
    if (task >= n_rows * k) return;
    const uint64_t ra = task / k;
    const uint32_t j = (uint32_t)(task % k);
    const TailProdRow p = prod[ra];
    const CT *Mr = ta.mat + p.mrow;
    const CT *r = ta.tsc + p.toff * k + j;
    double acc = 0.0;
    uint32_t b = lane;
    for (; b + 7 * 64 < p.nt; b += 8 * 64) {
        double mv[8], rv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            mv[q] = (double)Mr[b + q * 64];
            rv[q] = (double)r[(uint64_t)(b + q * 64) * k];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) acc += mv[q] * rv[q];
    }
    for (; b < p.nt; b += 64) acc += (double)Mr[b] * (double)r[(uint64_t)b * k];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) xs[(uint64_t)ta.rows[ra] * k + j] = (CT)acc;
}

// One large level of every component that has it, all rows in parallel across the grid (one
// row group per row, entries summed in order).  FORWARD: x = rhs - L x; else x = x - L' x.
template <typename CT, bool FORWARD>
__global__ __launch_bounds__(BLK) void sk_sparse_grid(const GridRow *__restrict__ rows, uint64_t n_rows,
                                                      const SpMeta *__restrict__ meta, const uint64_t *__restrict__ ptr,
                                                      const uint32_t *__restrict__ idx, const CT *__restrict__ val,
                                                      const CT *__restrict__ red, const double *__restrict__ mu2,
                                                      const uint32_t *__restrict__ aux_src, uint64_t ns, CT *xs,
                                                      uint32_t k, uint32_t kshift, const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint64_t stride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t q = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; q < n_rows; q += stride) {
        const GridRow gr = rows[q];
        const SpMeta &m = meta[gr.meta];
        const uint64_t lr = (uint64_t)(gr.row - m.kept_base) + m.row_off;
        double acc = FORWARD ? sparse_rhs<CT>(m, gr.row, red, aux_src, ns, mu2[(uint64_t)m.comp * k + j], k, j)
                             : (double)xs[(uint64_t)gr.row * k + j];
        acc = dot_entries<CT, true>(acc, ptr[lr], ptr[lr + 1], 1, idx, val, xs, k, j);
        xs[(uint64_t)gr.row * k + j] = (CT)acc;
    }
}

// y /= d on every factor row (a zero pivot -- the floating gauge -- gives 0), all components
// across the grid, between the forward and the backward substitutions.
template <typename CT>
__global__ __launch_bounds__(BLK) void sk_sparse_diag(const GridRow *__restrict__ rows, uint64_t n_rows,
                                                      const SpMeta *__restrict__ meta, const CT *__restrict__ piv,
                                                      CT *xs, uint32_t k, uint32_t kshift,
                                                      const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint64_t stride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t q = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; q < n_rows; q += stride) {
        const GridRow gr = rows[q];
        const SpMeta &m = meta[gr.meta];
        const double dv = (double)piv[m.piv_off + (gr.row - m.kept_base)];
        const uint64_t x = (uint64_t)gr.row * k + j;
        xs[x] = dv != 0.0 ? (CT)((double)xs[x] / dv) : (CT)0;
    }
}

// Gremban covers: read back x = (y+ - y-) / 2 into the real kept slots, the first sheet's
// rows assigning and then the second sheet's subtracting (two launches, sheet 0 then 1).
template <typename CT>
__global__ __launch_bounds__(BLK) void sk_cover_readback(const uint32_t *__restrict__ cover_rows, uint64_t n,
                                                         const uint32_t *__restrict__ aux_src, uint64_t ns, CT *xs,
                                                         uint32_t sheet, uint32_t k, uint32_t kshift,
                                                         const int32_t *__restrict__ gate) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    if (j >= k) return;
    const uint64_t stride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t q = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; q < n; q += stride) {
        const uint32_t r = cover_rows[q];
        const uint32_t a = aux_src[r - ns];
        if ((a >> 31) != sheet) continue;
        const double y = 0.5 * (double)xs[(uint64_t)r * k + j];
        CT &dst = xs[(uint64_t)(a & 0x7FFFFFFFu) * k + j];
        dst = sheet ? (CT)((double)dst - y) : (CT)y;
    }
}

// z[dof] = sum over the DOF's slots of sgnpw[s] * (xs[s] - mu3[comp(s)]); uncovered DOFs -> 0.
// sgnpw[s] and comp(s) come per d_slot entry (d_w, d_comp), so only xs and mu3 are gathered.
// DOTS: also the block partials of <z,r>, <z,z>, <r,r> (r = the apply's input p~), each lane
// adding its DOFs in loop order, then a fixed-order tree over the lanes of a column.
template <typename CT, bool DOTS>
__global__ __launch_bounds__(BLK) void sk_prolong(uint64_t N, const uint64_t *__restrict__ d_ptr,
                                                  const uint32_t *__restrict__ d_slot, const CT *__restrict__ d_w,
                                                  const uint32_t *__restrict__ d_comp,
                                                  const CT *__restrict__ xs, const double *__restrict__ mu3,
                                                  double *__restrict__ z, uint32_t k, uint32_t kshift, const int32_t *__restrict__ gate,
                                                  const double *__restrict__ r, double *__restrict__ dots) {
    if (gate_closed(gate, k)) return;
    const uint32_t kp = 1u << kshift, j = threadIdx.x & (kp - 1);
    double vp = 0.0, vv = 0.0, pp = 0.0;
    const uint64_t stride = ((uint64_t)gridDim.x * BLK) >> kshift;
    for (uint64_t d = ((uint64_t)blockIdx.x * BLK + threadIdx.x) >> kshift; j < k && d < N; d += stride) {
        double acc = 0.0;
        const uint64_t q1 = d_ptr[d + 1];
        uint64_t q = d_ptr[d];
        for (; q + 2 <= q1; q += 2) { // two slots' loads in flight, added in slot order
            const uint64_t s0 = d_slot[q], s1 = d_slot[q + 1];
            const double w0 = (double)d_w[q], w1 = (double)d_w[q + 1];
            const double x0 = (double)xs[s0 * k + j], x1 = (double)xs[s1 * k + j];
            const double m0 = mu3[(uint64_t)d_comp[q] * k + j], m1 = mu3[(uint64_t)d_comp[q + 1] * k + j];
            acc += w0 * (x0 - m0);
            acc += w1 * (x1 - m1);
        }
        for (; q < q1; ++q) {
            const uint64_t s = d_slot[q];
            acc += (double)d_w[q] * ((double)xs[s * k + j] - mu3[(uint64_t)d_comp[q] * k + j]);
        }
        z[d * k + j] = acc;
        if (DOTS) {
            const double p = r[d * k + j];
            vp += acc * p;
            vv += acc * acc;
            pp += p * p;
        }
    }
    if (DOTS) {
        __shared__ double sh[BLK];
        const double a = cols_block_sum(vp, kp, sh);
        __syncthreads();
        const double b = cols_block_sum(vv, kp, sh);
        __syncthreads();
        const double c = cols_block_sum(pp, kp, sh);
        if (threadIdx.x < kp) {
            double *o = dots + ((uint64_t)blockIdx.x * kp + threadIdx.x) * 3;
            o[0] = a;
            o[1] = b;
            o[2] = c;
        }
    }
}

uint32_t grid_rows(uint64_t rows, uint32_t kshift, uint32_t cap = 65535) {
    const uint64_t per = BLK >> kshift;
    uint64_t g = (rows + per - 1) / per;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

__global__ __launch_bounds__(256) void sk_zero_words(uint32_t *__restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = 0u;
}

__global__ __launch_bounds__(256) void sk_cast_f32(uint64_t n, const double *__restrict__ src, float *__restrict__ dst) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        dst[i] = (float)src[i];
}

template <typename T, typename V> std::vector<T> cast_vec(const V &v) {
    std::vector<T> o(v.size());
#pragma omp parallel for schedule(static) if (v.size() > (1u << 20))
    for (int64_t i = 0; i < (int64_t)v.size(); ++i) o[i] = (T)v[i];
    return o;
}

// Splits CSR rows into work items of at most ROW_CHUNK entries.
void make_items(const std::vector<uint32_t> &rows, const std::vector<uint64_t> &begin,
                const std::vector<uint64_t> &end, std::vector<RowItem> &items, std::vector<RowFix> &fixes,
                uint32_t &n_part) {
    n_part = 0;
    for (size_t r = 0; r < rows.size(); ++r) {
        const uint64_t b = begin[r], e = end[r];
        if (e - b <= ROW_CHUNK) {
            items.push_back({rows[r], 0, b, e, (int64_t)rows[r]});
        } else {
            RowFix f{rows[r], 0, n_part};
            for (uint64_t p = b; p < e; p += ROW_CHUNK) {
                items.push_back({rows[r], 0, p, std::min<uint64_t>(e, p + ROW_CHUNK), -(int64_t)(n_part + 1)});
                ++n_part;
                ++f.count;
            }
            fixes.push_back(f);
        }
    }
}

template <typename CT>
std::shared_ptr<SchwarzDev<CT>> build_dev(const Precond::Schwarz &S) {
    auto D = std::make_shared<SchwarzDev<CT>>();
    const uint64_t ns = S.n_slots(), nc = S.n_comp();
    D->ns = ns;
    D->nc = nc;
    D->N = S.n_dofs;
    if (ns >= 0xFFFFFFFFull) fail(WX_ERR_INVALID_INPUT, "too many Schwarz slots for the device path");
    std::vector<uint32_t> slot_comp(ns);
    std::vector<CT> sgnpw(ns);
    for (uint64_t c = 0; c < nc; ++c)
        for (uint64_t s = S.slot_start[c]; s < S.slot_start[c + 1]; ++s) {
            slot_comp[s] = (uint32_t)c;
            sgnpw[s] = (CT)(S.fold[s] * S.pw[s]);
        }
    D->glob.upload(S.glob.data(), ns);
    D->sgnpw.upload(sgnpw.data(), ns);
    {
        auto v = cast_vec<CT>(S.inv_de);
        D->inv_de.upload(v.data(), v.size());
    }
    D->slot_comp.upload(slot_comp.data(), ns);
    // C over eliminated rows: entries -> kept slot ids; C' over kept rows: entries -> elim slot ids
    std::vector<uint32_t> c_idx(S.c_idx.size());
    std::vector<uint32_t> c_rows;
    std::vector<uint64_t> c_b, c_e;
    uint64_t ct_nnz = S.c_idx.size();
    std::vector<uint64_t> ct_ptr(ns + 1, 0);
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t base = S.slot_start[c], kb = base + S.ne[c];
        for (uint64_t e = base; e < kb; ++e) {
            c_rows.push_back((uint32_t)e);
            c_b.push_back(S.c_ptr[e]);
            c_e.push_back(S.c_ptr[e + 1]);
            for (uint64_t p = S.c_ptr[e]; p < S.c_ptr[e + 1]; ++p) {
                c_idx[p] = (uint32_t)(kb + S.c_idx[p]);
                ct_ptr[kb + S.c_idx[p] + 1]++;
            }
        }
    }
    for (uint64_t s = 0; s < ns; ++s) ct_ptr[s + 1] += ct_ptr[s];
    std::vector<uint32_t> ct_idx(ct_nnz);
    std::vector<double> ct_val(ct_nnz);
    {
        std::vector<uint64_t> cur(ct_ptr.begin(), ct_ptr.end() - 1);
        for (uint64_t c = 0; c < nc; ++c) {
            const uint64_t base = S.slot_start[c], kb = base + S.ne[c];
            for (uint64_t e = base; e < kb; ++e)
                for (uint64_t p = S.c_ptr[e]; p < S.c_ptr[e + 1]; ++p) {
                    const uint64_t row = kb + S.c_idx[p];
                    // C' entries carry the eliminated row's 1 / D_e: the reduced right-hand side's
                    // terms are then ct_val * (loc - mu) with no second gather per entry
                    ct_idx[cur[row]] = (uint32_t)e;
                    ct_val[cur[row]] = S.c_val[p] * S.inv_de[e];
                    cur[row]++;
                }
        }
    }
    std::vector<uint32_t> ct_rows;
    std::vector<uint64_t> ct_b, ct_e;
    for (uint64_t c = 0; c < nc; ++c)
        for (uint64_t s = S.slot_start[c] + S.ne[c]; s < S.slot_start[c + 1]; ++s) {
            ct_rows.push_back((uint32_t)s);
            ct_b.push_back(ct_ptr[s]);
            ct_e.push_back(ct_ptr[s + 1]);
        }
    {
        std::vector<RowItem> items;
        std::vector<RowFix> fixes;
        // split rows with few partials first (a row group each), then the rest (a workgroup each)
        auto small_first = [](std::vector<RowFix> &fx) {
            std::stable_partition(fx.begin(), fx.end(), [](const RowFix &f) { return f.count <= FIX_SMALL; });
            return (uint32_t)std::count_if(fx.begin(), fx.end(), [](const RowFix &f) { return f.count <= FIX_SMALL; });
        };
        make_items(c_rows, c_b, c_e, items, fixes, D->n_c_part);
        D->c_items.upload(items.data(), items.size());
        D->n_c_items = (uint32_t)items.size();
        D->n_c_fix_small = small_first(fixes);
        D->c_fix.upload(fixes.data(), fixes.size());
        D->n_c_fix = (uint32_t)fixes.size();
        items.clear();
        fixes.clear();
        make_items(ct_rows, ct_b, ct_e, items, fixes, D->n_ct_part);
        D->ct_items.upload(items.data(), items.size());
        D->n_ct_items = (uint32_t)items.size();
        D->n_ct_fix_small = small_first(fixes);
        D->ct_fix.upload(fixes.data(), fixes.size());
        D->n_ct_fix = (uint32_t)fixes.size();
    }
    D->c_idx.upload(c_idx.data(), c_idx.size());
    {
        auto v = cast_vec<CT>(S.c_val);
        D->c_val.upload(v.data(), v.size());
    }
    D->ct_idx.upload(ct_idx.data(), ct_idx.size());
    {
        auto v = cast_vec<CT>(ct_val);
        D->ct_val.upload(v.data(), v.size());
    }
    // mean chunks
    std::vector<Chunk> all, kept;
    std::vector<uint32_t> all_cptr(nc + 1, 0), kept_cptr(nc + 1, 0), cnt(nc), kcnt(nc);
    for (uint64_t c = 0; c < nc; ++c) {
        const uint32_t b = (uint32_t)S.slot_start[c], e = (uint32_t)S.slot_start[c + 1];
        for (uint32_t s = b; s < e; s += MEAN_CHUNK) all.push_back({(uint32_t)c, s, std::min(e, s + MEAN_CHUNK), 0});
        all_cptr[c + 1] = (uint32_t)all.size();
        // a count of 0 makes the mean 0: only floating components are gauged by their means
        const bool floating = S.gauge[c] == 0;
        cnt[c] = floating ? e - b : 0;
        const uint32_t kb = b + S.ne[c];
        if (S.fac_kind[c] == 1 && floating)
            for (uint32_t s = kb; s < e; s += MEAN_CHUNK) kept.push_back({(uint32_t)c, s, std::min(e, s + MEAN_CHUNK), 0});
        kept_cptr[c + 1] = (uint32_t)kept.size();
        kcnt[c] = floating ? e - kb : 0;
    }
    D->all_chunks.upload(all.data(), all.size());
    D->n_all_chunks = (uint32_t)all.size();
    D->kept_chunks.upload(kept.data(), kept.size());
    D->n_kept_chunks = (uint32_t)kept.size();
    D->comp_count.upload(cnt.data(), cnt.size());
    D->comp_kcount.upload(kcnt.data(), kcnt.size());
    {
        // work items of the means: components of one chunk MEAN_SMALL to an item, larger ones
        // MEAN_ITEM chunks to an item
        auto mean_items = [nc](const std::vector<uint32_t> &cp) {
            std::vector<MeanItem> it;
            for (uint64_t c = 0; c < nc; ++c) {
                const uint32_t c0 = cp[c], nch = cp[c + 1] - c0;
                if (nch == 1) {
                    MeanItem *o = it.empty() ? nullptr : &it.back();
                    if (o && o->kind == 1 && o->chunk0 + o->n == c0 && o->n < MEAN_SMALL) ++o->n;
                    else it.push_back({1u, (uint32_t)c, c0, 1u, 0u, 0u});
                    continue;
                }
                const uint32_t ni = (nch + MEAN_ITEM - 1) / MEAN_ITEM, first = (uint32_t)it.size();
                for (uint32_t i = 0; i < ni; ++i)
                    it.push_back({0u, (uint32_t)c, c0 + i * MEAN_ITEM, std::min(MEAN_ITEM, nch - i * MEAN_ITEM), first, ni});
            }
            return it;
        };
        auto combined = [](const std::vector<MeanItem> &it) {
            std::vector<MeanItem> cb;
            for (size_t i = 0; i < it.size(); ++i)
                if (it[i].kind == 0 && it[i].nitems > 1 && it[i].first == (uint32_t)i) cb.push_back(it[i]);
            return cb;
        };
        const std::vector<MeanItem> ai = mean_items(all_cptr), ki = mean_items(kept_cptr);
        const std::vector<MeanItem> ac = combined(ai), kc = combined(ki);
        D->all_items.upload(ai.data(), ai.size());
        D->kept_items.upload(ki.data(), ki.size());
        D->n_all_items = (uint32_t)ai.size();
        D->n_kept_items = (uint32_t)ki.size();
        D->all_comb.upload(ac.data(), ac.size());
        D->kept_comb.upload(kc.data(), kc.size());
        D->n_all_comb = (uint32_t)ac.size();
        D->n_kept_comb = (uint32_t)kc.size();
    }
    // dense rows: small blocks row by row, blocks with nk >= DENSE_TILED_MIN as split-K tiles
    std::vector<uint64_t> drow_off;
    std::vector<uint32_t> drow_base, drow_nk, drow_slot;
    std::vector<DenseComp> dcomps;
    std::vector<uint64_t> drow_start{0};
    uint64_t row_tiles = 0;
    for (uint64_t c = 0; c < nc; ++c)
        if (S.fac_kind[c] == 0 && S.nk[c] >= DENSE_TILED_MIN) row_tiles += (S.nk[c] + 63) / 64;
    // device layout of the dense operators: in an FP64 run a tiled block's rows get a leading
    // dimension padded to a multiple of 64 coefficients and the block a 64-coefficient-aligned
    // start, so each wave's load of 64 consecutive coefficients covers whole 128-byte lines (an
    // FP32 run keeps its tiled blocks in dsym instead); the other blocks stay packed in the host
    // order
    std::vector<uint32_t> dorder; // dense components by host offset
    for (uint64_t c = 0; c < nc; ++c)
        if (S.fac_kind[c] == 0) dorder.push_back((uint32_t)c);
    std::sort(dorder.begin(), dorder.end(), [&](uint32_t a, uint32_t b) { return S.dense_off[a] < S.dense_off[b]; });
    // FP32 runs keep the tiled blocks only as their upper block triangle (dsym), outside dmat
    constexpr bool sym = std::is_same<CT, float>::value;
    std::vector<uint64_t> doff(nc, 0);
    std::vector<uint32_t> dld(nc, 0);
    uint64_t dtotal = 0;
    bool relayout = false;
    for (uint32_t c : dorder) {
        const uint32_t nk = S.nk[c];
        const bool tiled = nk >= DENSE_TILED_MIN;
        if (tiled && sym) { // dld 0: not in dmat
            relayout = true;
            continue;
        }
        dld[c] = tiled ? (nk + 63) / 64 * 64 : nk;
        if (tiled) dtotal = (dtotal + 63) / 64 * 64;
        doff[c] = dtotal;
        dtotal += (uint64_t)nk * dld[c];
        relayout = relayout || doff[c] != S.dense_off[c] || dld[c] != nk;
    }
    std::vector<DenseTile> tiles;
    std::vector<SymTile> stiles;
    std::vector<uint64_t> sym_off(nc, 0);
    uint64_t part_rows = 0, sym_total = 0;
    for (uint64_t c = 0; c < nc; ++c) {
        if (S.fac_kind[c] != 0) continue;
        const uint32_t nk = S.nk[c];
        const uint32_t kb = (uint32_t)(S.slot_start[c] + S.ne[c]);
        if (nk >= DENSE_TILED_MIN && sym) {
            // super-tiles (P, Q), P <= Q, of 4 x 4 blocks; nt + 1 partial sources per row
            const uint32_t nb = (nk + 63) / 64, nt = (nb + 3) / 4;
            sym_off[c] = sym_total;
            for (uint32_t P = 0; P < nt; ++P)
                for (uint32_t Q = P; Q < nt; ++Q) stiles.push_back({sym_total, nk, nb, kb, P, Q, 0, part_rows});
            sym_total += (uint64_t)nb * (nb + 1) / 2 * 4096;
            dcomps.push_back(DenseComp{nk, kb, nt + 1, 0, part_rows});
            part_rows += (uint64_t)(nt + 1) * nk;
            drow_start.push_back(drow_start.back() + nk);
            continue;
        }
        if (nk >= DENSE_TILED_MIN) {
            // enough (row tile, column segment) pairs to fill the device, segments >= 256 columns
            const uint32_t max_seg = (nk + 255) / 256;
            const uint32_t want = (uint32_t)std::max<uint64_t>(1, (DENSE_TILE_TARGET + row_tiles - 1) / row_tiles);
            const uint32_t n_seg = std::min(max_seg, want);
            const uint32_t seg_len = ((nk + n_seg - 1) / n_seg + 63) / 64 * 64;
            DenseComp dc{nk, kb, 0, 0, part_rows};
            for (uint32_t s = 0; s * seg_len < nk; ++s) {
                for (uint32_t r0 = 0; r0 < nk; r0 += 64)
                    tiles.push_back({doff[c], nk, dld[c], kb, r0, s * seg_len, std::min(nk, (s + 1) * seg_len), s,
                                     part_rows});
                dc.n_seg++;
            }
            part_rows += (uint64_t)dc.n_seg * nk;
            dcomps.push_back(dc);
            drow_start.push_back(drow_start.back() + nk);
            continue;
        }
        for (uint32_t i = 0; i < nk; ++i) {
            drow_off.push_back(doff[c] + (uint64_t)i * nk);
            drow_base.push_back(kb);
            drow_nk.push_back(nk);
            drow_slot.push_back(kb + i);
        }
    }
    D->n_dtiles = (uint32_t)tiles.size();
    D->dtiles.upload(tiles.data(), tiles.size());
    D->n_stiles = (uint32_t)stiles.size();
    D->stiles.upload(stiles.data(), stiles.size());
    D->n_dcomps = (uint32_t)dcomps.size();
    D->dcomps.upload(dcomps.data(), dcomps.size());
    D->drow_start.upload(drow_start.data(), drow_start.size());
    D->dense_part_rows = part_rows;
    D->dense_tiled_rows = drow_start.back();
    D->n_drows = (uint32_t)drow_off.size();
    D->drow_off.upload(drow_off.data(), drow_off.size());
    D->drow_base.upload(drow_base.data(), drow_base.size());
    D->drow_nk.upload(drow_nk.data(), drow_nk.size());
    D->drow_slot.upload(drow_slot.data(), drow_slot.size());
    // the dense operators: host-built blocks uploaded, GPU-built ones copied (and for FP32
    // converted) on the device
    {
        const uint64_t total = S.dense.size();
        DevBuf raw; // the host layout
        raw.alloc(total * sizeof(CT));
        std::vector<std::pair<uint64_t, uint64_t>> dev_ranges;
        for (const auto &b : S.dense_dev) dev_ranges.push_back({b.off, b.count});
        std::sort(dev_ranges.begin(), dev_ranges.end());
        uint64_t at = 0;
        auto upload_host = [&](uint64_t from, uint64_t to) {
            if (to <= from) return;
            if constexpr (std::is_same<CT, double>::value) {
                WX_HIP(hipMemcpy(raw.template as<CT>() + from, S.dense.data() + from, (to - from) * 8,
                                 hipMemcpyHostToDevice));
            } else {
                std::vector<CT> v(to - from);
#pragma omp parallel for schedule(static) if (to - from > (1u << 20))
                for (int64_t i = 0; i < (int64_t)(to - from); ++i) v[i] = (CT)S.dense[from + i];
                WX_HIP(hipMemcpy(raw.template as<CT>() + from, v.data(), v.size() * sizeof(CT),
                                 hipMemcpyHostToDevice));
            }
        };
        for (const auto &r : dev_ranges) {
            upload_host(at, r.first);
            at = r.first + r.second;
        }
        upload_host(at, total);
        for (const auto &b : S.dense_dev) {
            if constexpr (std::is_same<CT, double>::value) {
                WX_HIP(hipMemcpy(raw.template as<CT>() + b.off, b.dev.get(), b.count * 8, hipMemcpyDeviceToDevice));
            } else {
                const uint32_t g = (uint32_t)std::min<uint64_t>((b.count + 255) / 256, 1u << 20);
                hipLaunchKernelGGL(sk_cast_f32, dim3(g), dim3(256), 0, 0, b.count, b.dev.get(),
                                   raw.template as<CT>() + b.off);
                WX_HIP(hipGetLastError());
            }
        }
        WX_HIP(hipDeviceSynchronize());
        if (sym_total) { // the tiled blocks' upper block triangles
            D->dsym.alloc(sym_total * sizeof(float));
            for (uint64_t c = 0; c < nc; ++c) {
                if (S.fac_kind[c] != 0 || S.nk[c] < DENSE_TILED_MIN) continue;
                const uint32_t nk = S.nk[c], nb = (nk + 63) / 64;
                hipLaunchKernelGGL((sk_pack_upper<CT>), dim3((uint32_t)((uint64_t)nb * (nb + 1) / 2)), dim3(256), 0, 0,
                                   raw.template as<CT>() + S.dense_off[c], nk, nb,
                                   D->dsym.template as<float>() + sym_off[c]);
                WX_HIP(hipGetLastError());
            }
            WX_HIP(hipDeviceSynchronize());
        }
        if (!relayout) {
            D->dmat = std::move(raw);
        } else {
            // packed runs copied whole, tiled blocks row by row into their padded rows
            D->dmat.alloc(std::max<uint64_t>(dtotal, 1) * sizeof(CT));
            CT *dst = D->dmat.template as<CT>();
            const CT *src = raw.template as<CT>();
            for (size_t q = 0; q < dorder.size();) {
                const uint32_t c = dorder[q];
                const uint64_t nk = S.nk[c];
                if (dld[c] == 0) { // kept in dsym only
                    ++q;
                    continue;
                }
                if (dld[c] != nk) {
                    WX_HIP(hipMemcpy2D(dst + doff[c], dld[c] * sizeof(CT), src + S.dense_off[c], nk * sizeof(CT),
                                       nk * sizeof(CT), nk, hipMemcpyDeviceToDevice));
                    ++q;
                    continue;
                }
                uint64_t len = nk * nk;
                size_t r = q + 1;
                for (; r < dorder.size(); ++r) {
                    const uint32_t e = dorder[r];
                    if (dld[e] != S.nk[e] || doff[e] != doff[c] + len || S.dense_off[e] != S.dense_off[c] + len) break;
                    len += (uint64_t)S.nk[e] * S.nk[e];
                }
                if (len)
                    WX_HIP(hipMemcpy(dst + doff[c], src + S.dense_off[c], len * sizeof(CT), hipMemcpyDeviceToDevice));
                q = r;
            }
            WX_HIP(hipDeviceSynchronize());
        }
    }
    // prolongation map
    std::vector<uint64_t> d_ptr(S.n_dofs + 1, 0);
    for (uint64_t s = 0; s < ns; ++s) d_ptr[S.glob[s] + 1]++;
    for (uint64_t d = 0; d < S.n_dofs; ++d) d_ptr[d + 1] += d_ptr[d];
    std::vector<uint32_t> d_slot(ns);
    {
        std::vector<uint64_t> cur(d_ptr.begin(), d_ptr.end() - 1);
        for (uint64_t s = 0; s < ns; ++s) d_slot[cur[S.glob[s]]++] = (uint32_t)s;
    }
    D->d_ptr.upload(d_ptr.data(), d_ptr.size());
    D->d_slot.upload(d_slot.data(), d_slot.size());
    {
        std::vector<CT> dw(ns);
        std::vector<uint32_t> dc(ns);
        for (uint64_t i = 0; i < ns; ++i) {
            dw[i] = sgnpw[d_slot[i]];
            dc[i] = slot_comp[d_slot[i]];
        }
        D->d_w.upload(dw.data(), ns);
        D->d_comp.upload(dc.data(), ns);
    }
    for (uint64_t c = 0; c < nc; ++c) D->has_sparse |= S.fac_kind[c] == 1 || S.fac_kind[c] == 2;
    if (D->has_sparse) {
        // level schedules of L (forward) and L' (backward) per sparse component
        std::vector<SpMeta> metas;
        std::vector<uint64_t> lptr, ltptr, lvf_ptr, lvb_ptr, lvf_long, lvb_long;
        std::vector<uint32_t> lidx, ltidx, lvf_rows, lvb_rows;
        // dense tails: rows (slots), per-row first chunk, chunks, entries on earlier rows, M_T
        std::vector<uint32_t> t_rows, t_idx;
        std::vector<TailProdRow> t_prod;
        std::vector<uint64_t> t_cptr;
        std::vector<TailChunk> t_chunks;
        std::vector<CT> t_val, t_mat;
        std::vector<std::vector<uint32_t>> tail_local; // per sparse component: its tail rows (local)
        std::vector<CT> lval, ltval, piv;
        uint64_t max_depth = 0, total_rows = 0;
        std::vector<uint32_t> aux_src;
        for (uint64_t c = 0; c < nc; ++c) {
            if (S.fac_kind[c] != 1 && S.fac_kind[c] != 2) continue;
            const bool cover = S.fac_kind[c] == 2;
            // factor rows: the kept slots (in elimination order), or auxiliary slots holding
            // the cover's vertices in its elimination order
            const uint32_t nk = cover ? 2 * S.nk[c] : S.nk[c];
            const uint32_t real_kb = (uint32_t)(S.slot_start[c] + S.ne[c]);
            const uint32_t kb = cover ? (uint32_t)(ns + aux_src.size()) : real_kb;
            if (cover)
                for (uint32_t t = 0; t < nk; ++t) {
                    const uint32_t v = S.cover_perm[S.cover_off[c] + t];
                    aux_src.push_back(v < S.nk[c] ? real_kb + v : ((real_kb + v - S.nk[c]) | 0x80000000u));
                }
            const uint64_t ro = S.l_row_off[c];
            SpMeta m{};
            m.comp = (uint32_t)c;
            m.kept_base = kb;
            m.nk = nk;
            m.cover = cover ? 1u : 0u;
            m.row_off = lptr.size();
            m.piv_off = piv.size();
            // L rows
            const uint64_t lbase = lidx.size();
            for (uint32_t i = 0; i <= nk; ++i) lptr.push_back(lbase + (S.l_ptr[ro + i] - S.l_ptr[ro]));
            for (uint64_t p = S.l_ptr[ro]; p < S.l_ptr[ro + nk]; ++p) {
                lidx.push_back(kb + S.l_idx[p]);
                lval.push_back((CT)S.l_val[p]);
            }
            for (uint32_t i = 0; i < nk; ++i) piv.push_back((CT)S.piv[S.piv_off[c] + i]);
            // L' rows (column structure of L)
            std::vector<uint64_t> tp(nk + 1, 0);
            for (uint64_t p = S.l_ptr[ro]; p < S.l_ptr[ro + nk]; ++p) tp[S.l_idx[p] + 1]++;
            for (uint32_t i = 0; i < nk; ++i) tp[i + 1] += tp[i];
            std::vector<uint32_t> ti(tp[nk]);
            std::vector<CT> tv(tp[nk]);
            {
                std::vector<uint64_t> cur(tp.begin(), tp.end() - 1);
                for (uint32_t i = 0; i < nk; ++i)
                    for (uint64_t p = S.l_ptr[ro + i]; p < S.l_ptr[ro + i + 1]; ++p) {
                        const uint32_t col = S.l_idx[p];
                        ti[cur[col]] = kb + i;
                        tv[cur[col]] = (CT)S.l_val[p];
                        cur[col]++;
                    }
            }
            const uint64_t tbase = ltidx.size();
            for (uint32_t i = 0; i <= nk; ++i) ltptr.push_back(tbase + tp[i]);
            ltidx.insert(ltidx.end(), ti.begin(), ti.end());
            ltval.insert(ltval.end(), tv.begin(), tv.end());
            // levels: forward lvl(i) = 1 + max lvl(j), j in row i of L; the backward schedule is
            // its mirror, depth - 1 - lvl(i) (a row's dependents in L' sit on higher forward
            // levels), so the large leading forward levels become the trailing backward ones
            std::vector<uint32_t> lf(nk, 0), lb(nk, 0);
            uint32_t df = 0;
            for (uint32_t i = 0; i < nk; ++i) {
                uint32_t v = 0;
                for (uint64_t p = S.l_ptr[ro + i]; p < S.l_ptr[ro + i + 1]; ++p) v = std::max(v, lf[S.l_idx[p]] + 1);
                lf[i] = v;
                df = std::max(df, v + 1);
            }
            // dense tail: the trailing levels with at most TAIL_LEVEL_ROWS rows each (a chain of
            // hubs eliminated last), at most TAIL_MAX_ROWS rows, when they span TAIL_MIN_LEVELS
            // levels or more and the tail's nt^2 dense entries stay within TAIL_FILL times the
            // component's factor entries and rows (so no tail outweighs the factor it replaces);
            // the schedules then cover the rows before the cut only
            uint32_t cut = df;
            {
                std::vector<uint32_t> per(df, 0);
                for (uint32_t i = 0; i < nk; ++i) per[lf[i]]++;
                const uint64_t fill_cap = TAIL_FILL * ((S.l_ptr[ro + nk] - S.l_ptr[ro]) + nk);
                uint64_t rows_t = 0;
                while (cut > 1 && per[cut - 1] <= TAIL_LEVEL_ROWS && rows_t + per[cut - 1] <= TAIL_MAX_ROWS &&
                       (rows_t + per[cut - 1]) * (rows_t + per[cut - 1]) <= fill_cap)
                    rows_t += per[--cut];
                if (df - cut < TAIL_MIN_LEVELS) cut = df;
            }
            constexpr uint32_t SKIP = UINT32_MAX;
            const uint32_t df_all = df;
            df = cut;
            const uint32_t db = df;
            for (uint32_t i = 0; i < nk; ++i) {
                if (lf[i] >= cut) lf[i] = lb[i] = SKIP;
                else lb[i] = df - 1 - lf[i];
            }
            m.nt = 0;
            m.n_tchunks = 0;
            m.tail_off = t_rows.size();
            m.tchunk_off = t_chunks.size();
            m.tmat_off = t_mat.size();
            tail_local.emplace_back();
            if (cut < df_all) {
                std::vector<uint32_t> &trow = tail_local.back();
                std::vector<int32_t> tix(nk, -1);
                for (uint32_t i = 0; i < nk; ++i)
                    if (lf[i] == SKIP) {
                        tix[i] = (int32_t)trow.size();
                        trow.push_back(i);
                    }
                const uint32_t nt = (uint32_t)trow.size();
                m.nt = nt;
                // entries of the tail rows on rows before the tail, in entry order, chunked
                for (uint32_t a = 0; a < nt; ++a) {
                    const uint32_t i = trow[a];
                    t_prod.push_back({m.tmat_off + (uint64_t)a * nt, m.tail_off, nt, (uint32_t)metas.size()});
                    t_rows.push_back(kb + i);
                    t_cptr.push_back(t_chunks.size());
                    uint64_t p0 = t_idx.size();
                    for (uint64_t p = S.l_ptr[ro + i]; p < S.l_ptr[ro + i + 1]; ++p)
                        if (tix[S.l_idx[p]] < 0) {
                            t_idx.push_back(kb + S.l_idx[p]);
                            t_val.push_back((CT)S.l_val[p]);
                        }
                    for (; p0 < t_idx.size(); p0 += TAIL_CHUNK)
                        t_chunks.push_back({a, 0u, p0, std::min<uint64_t>(t_idx.size(), p0 + TAIL_CHUNK)});
                }
                m.n_tchunks = (uint32_t)(t_chunks.size() - m.tchunk_off);
                // M_T = (I + L_TT')^{-1} D_T^+ (I + L_TT)^{-1}, one column per unit vector: the
                // tail's forward substitution, the pivots (zero: the gauge, giving 0) and the
                // backward substitution, exactly the operator the sparse solve applies
                // L_TT: each tail row's entries on tail rows (tail-local column, value), in entry order
                std::vector<uint32_t> tt_ptr(nt + 1, 0), tt_col;
                std::vector<double> tt_val;
                for (uint32_t a = 0; a < nt; ++a) {
                    const uint32_t i = trow[a];
                    for (uint64_t p = S.l_ptr[ro + i]; p < S.l_ptr[ro + i + 1]; ++p)
                        if (tix[S.l_idx[p]] >= 0) {
                            tt_col.push_back((uint32_t)tix[S.l_idx[p]]);
                            tt_val.push_back(S.l_val[p]);
                        }
                    tt_ptr[a + 1] = (uint32_t)tt_col.size();
                }
                std::vector<double> M((size_t)nt * nt), y(nt);
                for (uint32_t sidx = 0; sidx < nt; ++sidx) {
                    std::fill(y.begin(), y.end(), 0.0);
                    y[sidx] = 1.0;
                    for (uint32_t a = 0; a < nt; ++a) {
                        double v = y[a];
                        for (uint32_t q = tt_ptr[a]; q < tt_ptr[a + 1]; ++q) v -= tt_val[q] * y[tt_col[q]];
                        y[a] = v;
                    }
                    for (uint32_t a = 0; a < nt; ++a) {
                        const double dv = S.piv[S.piv_off[c] + trow[a]];
                        y[a] = dv != 0.0 ? y[a] / dv : 0.0;
                    }
                    for (uint32_t a = nt; a-- > 0;) // column-oriented: row a is final here
                        for (uint32_t q = tt_ptr[a]; q < tt_ptr[a + 1]; ++q) y[tt_col[q]] -= tt_val[q] * y[a];
                    for (uint32_t a = 0; a < nt; ++a) M[(size_t)a * nt + sidx] = y[a];
                }
                for (double v : M) t_mat.push_back((CT)v);
                if (std::getenv("WITHIN_AMD_VERBOSE"))
                    std::fprintf(stderr, "[within_amd] sparse factor of component %u: dense tail of %u rows (%u of %u levels)\n",
                                 (unsigned)c, nt, df_all - cut, df_all);
            }
            // rows bucketed by level, short rows first and rows with more than LONG_ROW entries
            // last; `lng` (aligned with `ptr`) marks where a level's long rows start
            auto bucket = [&](const std::vector<uint32_t> &lvl, uint32_t depth, const std::vector<uint64_t> &rnnz,
                              std::vector<uint64_t> &ptr, std::vector<uint64_t> &lng, std::vector<uint32_t> &rows) {
                const uint64_t off = ptr.size();
                std::vector<uint64_t> cnt(depth + 1, 0), nlong(depth + 1, 0);
                for (uint32_t i = 0; i < nk; ++i) {
                    if (lvl[i] == SKIP) continue; // a dense-tail row
                    cnt[lvl[i] + 1]++;
                    if (rnnz[i] > LONG_ROW) nlong[lvl[i]]++;
                }
                for (uint32_t l = 0; l < depth; ++l) cnt[l + 1] += cnt[l];
                const uint64_t rbase = rows.size();
                rows.resize(rbase + cnt[depth]);
                std::vector<uint64_t> cs(depth), cl(depth);
                for (uint32_t l = 0; l < depth; ++l) {
                    cs[l] = cnt[l];
                    cl[l] = cnt[l + 1] - nlong[l];
                }
                for (uint32_t i = 0; i < nk; ++i) {
                    const uint32_t l = lvl[i];
                    if (l == SKIP) continue;
                    rows[rbase + (rnnz[i] > LONG_ROW ? cl[l]++ : cs[l]++)] = kb + i;
                }
                for (uint32_t l = 0; l <= depth; ++l) {
                    ptr.push_back(rbase + cnt[l]);
                    lng.push_back(rbase + (l < depth ? cnt[l + 1] - nlong[l] : cnt[depth]));
                }
                return off;
            };
            std::vector<uint64_t> nnz_f(nk), nnz_b(nk);
            for (uint32_t i = 0; i < nk; ++i) {
                nnz_f[i] = S.l_ptr[ro + i + 1] - S.l_ptr[ro + i];
                nnz_b[i] = tp[i + 1] - tp[i];
            }
            m.nlev_f = df;
            m.nlev_b = db;
            m.lvl_f_off = bucket(lf, df, nnz_f, lvf_ptr, lvf_long, lvf_rows);
            m.lvl_b_off = bucket(lb, db, nnz_b, lvb_ptr, lvb_long, lvb_rows);
            // leading forward and trailing backward levels with at least BIG_LEVEL rows are solved
            // by grid-wide launches; the workgroup of the component takes the rest
            auto level_rows = [&](const std::vector<uint64_t> &ptr, uint64_t off, uint32_t l) {
                return ptr[off + l + 1] - ptr[off + l];
            };
            m.nbig_f = 0;
            while (m.nbig_f < df && level_rows(lvf_ptr, m.lvl_f_off, m.nbig_f) >= BIG_LEVEL) ++m.nbig_f;
            m.nbig_b = 0;
            while (m.nbig_b < db && level_rows(lvb_ptr, m.lvl_b_off, db - 1 - m.nbig_b) >= BIG_LEVEL) ++m.nbig_b;
            max_depth = std::max<uint64_t>(max_depth, std::max(df, db));
            total_rows += nk;
            metas.push_back(m);
        }
        D->n_sparse = (uint32_t)metas.size();
        D->sp_meta.upload(metas.data(), metas.size());
        // grid-wide steps: forward step g solves level g of every component with nbig_f > g;
        // backward steps are aligned on the last level, step h solving level nlev_b - maxB + h
        uint32_t max_f = 0, max_b = 0;
        for (const SpMeta &m : metas) {
            max_f = std::max(max_f, m.nbig_f);
            max_b = std::max(max_b, m.nbig_b);
        }
        std::vector<GridRow> gf, gb;
        D->grid_f_ptr.assign(1, 0);
        D->grid_b_ptr.assign(1, 0);
        for (uint32_t g = 0; g < max_f; ++g) {
            for (uint32_t mi = 0; mi < metas.size(); ++mi) {
                const SpMeta &m = metas[mi];
                if (g >= m.nbig_f) continue;
                const uint64_t o = m.lvl_f_off + g;
                for (uint64_t q = lvf_ptr[o]; q < lvf_ptr[o + 1]; ++q) gf.push_back({mi, lvf_rows[q]});
            }
            D->grid_f_ptr.push_back(gf.size());
        }
        for (uint32_t h = 0; h < max_b; ++h) {
            for (uint32_t mi = 0; mi < metas.size(); ++mi) {
                const SpMeta &m = metas[mi];
                if (h + m.nbig_b < max_b) continue;
                const uint64_t o = m.lvl_b_off + (m.nlev_b - max_b + h);
                for (uint64_t q = lvb_ptr[o]; q < lvb_ptr[o + 1]; ++q) gb.push_back({mi, lvb_rows[q]});
            }
            D->grid_b_ptr.push_back(gb.size());
        }
        D->grid_f_rows.upload(gf.data(), gf.size());
        D->grid_b_rows.upload(gb.data(), gb.size());
        std::vector<uint32_t> crow;
        std::vector<GridRow> drow;
        for (uint32_t mi = 0; mi < metas.size(); ++mi) {
            const SpMeta &m = metas[mi];
            std::vector<uint8_t> in_tail(m.nk, 0); // tail rows are final after the dense product
            for (uint32_t i : tail_local[mi]) in_tail[i] = 1;
            for (uint32_t r = 0; r < m.nk; ++r)
                if (!in_tail[r]) drow.push_back({mi, m.kept_base + r});
            if (m.cover)
                for (uint32_t r = 0; r < m.nk; ++r) crow.push_back(m.kept_base + r);
        }
        D->n_diag_rows = drow.size();
        D->diag_rows.upload(drow.data(), drow.size());
        D->n_cover_rows = crow.size();
        D->cover_rows.upload(crow.data(), crow.size());
        D->n_aux = aux_src.size();
        if (ns + D->n_aux >= 0x7FFFFFFFull) fail(WX_ERR_INVALID_INPUT, "too many Schwarz slots for the device path");
        D->aux_src.upload(aux_src.data(), aux_src.size());
        D->l_ptr.upload(lptr.data(), lptr.size());
        D->l_idx.upload(lidx.data(), lidx.size());
        D->l_val.upload(lval.data(), lval.size());
        D->lt_ptr.upload(ltptr.data(), ltptr.size());
        D->lt_idx.upload(ltidx.data(), ltidx.size());
        D->lt_val.upload(ltval.data(), ltval.size());
        D->piv.upload(piv.data(), piv.size());
        t_cptr.push_back(t_chunks.size());
        D->n_tail_rows = t_rows.size();
        D->n_tail_chunks = t_chunks.size();
        D->tail_rows.upload(t_rows.data(), t_rows.size());
        D->tail_cptr.upload(t_cptr.data(), t_cptr.size());
        D->tail_chunks.upload(t_chunks.data(), t_chunks.size());
        D->tail_idx.upload(t_idx.data(), t_idx.size());
        D->tail_val.upload(t_val.data(), t_val.size());
        D->tail_mat.upload(t_mat.data(), t_mat.size());
        D->tail_prod.upload(t_prod.data(), t_prod.size());
        if (std::getenv("WITHIN_AMD_VERBOSE"))
            std::fprintf(stderr, "[within_amd] sparse factors: %llu rows, %llu factor entries, dense tails of %llu rows "
                                 "(%llu entries, %.1f MB)\n",
                         (unsigned long long)piv.size(), (unsigned long long)lidx.size(),
                         (unsigned long long)t_rows.size(),
                         (unsigned long long)t_mat.size(), (double)(t_mat.size() * sizeof(CT)) / 1e6);
        D->sp_lvl_f_ptr.upload(lvf_ptr.data(), lvf_ptr.size());
        D->sp_lvl_f_long.upload(lvf_long.data(), lvf_long.size());
        D->sp_lvl_b_long.upload(lvb_long.data(), lvb_long.size());
        D->sp_lvl_f_rows.upload(lvf_rows.data(), lvf_rows.size());
        D->sp_lvl_b_ptr.upload(lvb_ptr.data(), lvb_ptr.size());
        D->sp_lvl_b_rows.upload(lvb_rows.data(), lvb_rows.size());
        if (std::getenv("WITHIN_AMD_VERBOSE"))
            std::fprintf(stderr, "[within_amd] sparse kept blocks: %u components, %lu rows, nnz(L)=%zu, max depth %lu\n",
                         D->n_sparse, (unsigned long)total_rows, lidx.size(), (unsigned long)max_depth);
    }
    return D;
}

template <typename CT>
struct SchwarzRunImpl : SchwarzRun {
    std::shared_ptr<SchwarzDev<CT>> D;
    uint32_t kcap = 0;
    DevBuf loc, red, xs, part, rowpart, mu1, mu2, mu3, dpart, tsc, csc;

    void ensure(uint32_t k, hipStream_t st) {
        if (k <= kcap) return;
        kcap = k;
        const uint64_t ns = D->ns;
        // slot-space vectors in the coefficient precision (FP32 in mixed runs); the component
        // means and the chunk partials behind them stay FP64
        constexpr uint64_t V = sizeof(CT);
        loc.alloc(ns * k * V);
        red.alloc(ns * k * V);
        xs.alloc((ns + D->n_aux) * k * V);
        part.alloc((uint64_t)std::max(D->n_all_chunks, D->n_kept_chunks) * k * 8 + 8);
        rowpart.alloc((uint64_t)std::max(D->n_c_part, D->n_ct_part) * k * V + 8);
        mu1.alloc(D->nc * k * 8);
        mu2.alloc(D->nc * k * 8);
        mu3.alloc(D->nc * k * 8);
        // a component without chunks in a list (the kept list of a dense or grounded one) has
        // no sk_mean item and keeps the mean 0 set here.  Cleared by a kernel on the run's
        // stream: a memset on the default stream while another thread captures a graph can end
        // up in that graph.
        const uint64_t nmu = (uint64_t)D->nc * k * 2; // 32-bit words
        for (DevBuf *m : {&mu1, &mu2, &mu3})
            hipLaunchKernelGGL(sk_zero_words, dim3((uint32_t)std::min<uint64_t>((nmu + 255) / 256 + 1, 65535)),
                               dim3(256), 0, st, m->as<uint32_t>(), nmu);
        WX_HIP(hipGetLastError());
        dpart.alloc(D->dense_part_rows * k * V + 8);
        tsc.alloc(D->n_tail_rows * k * V + 8);
        csc.alloc(D->n_tail_chunks * k * V + 8);
    }

    // the elimination SpMVs (C' rows into the reduced right-hand side, C rows in the
    // back-substitution); two columns per lane for an even width
    template <int MODE>
    void rows_spmv(const DevBuf &items, uint32_t n_items, const DevBuf &idx, const DevBuf &val, CT *locp,
                   CT *xsp, OutT<CT, MODE> *out, uint32_t k, uint32_t kshift, hipStream_t st, const int32_t *gate) {
        const SchwarzDev<CT> &d = *D;
        if constexpr (std::is_same<CT, float>::value) {
            // FP32 slot vectors: four columns per lane.  The C' rows (MODE 0) are the long ones,
            // split into items of up to ROW_CHUNK entries that a lane walks serially: below 16
            // columns two columns per lane keep twice the lanes (and waves) on them (worker-firm
            // panel at k = 4: 103 -> 52 us; 64 uneven problems: +3 %)
            if ((k & 3) == 0 && (MODE == 1 ? k >= 4 : k >= 16)) {
                const uint64_t per = 4 * (uint64_t)(BLK >> kshift);
                const uint32_t g =
                    (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n_items + per - 1) / per, 65535));
                hipLaunchKernelGGL((sk_rows4<MODE>), dim3(g), dim3(BLK), 0, st, dp<RowItem>(items), n_items,
                                   dp<uint32_t>(idx), dp<float>(val), dp<uint32_t>(d.slot_comp), locp, xsp,
                                   dp<float>(d.inv_de), dp<double>(mu1), out, dp<float>(rowpart), k, kshift, gate);
                return;
            }
        }
        if (k >= 2 && (k & 1) == 0) {
            const uint64_t per = 2 * (uint64_t)(BLK >> kshift);
            const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n_items + per - 1) / per, 65535));
            hipLaunchKernelGGL((sk_rows2<CT, MODE>), dim3(g), dim3(BLK), 0, st, dp<RowItem>(items), n_items,
                               dp<uint32_t>(idx), dp<CT>(val), dp<uint32_t>(d.slot_comp), locp, xsp,
                               dp<CT>(d.inv_de), dp<double>(mu1), out, dp<CT>(rowpart), k, kshift, gate);
        } else {
            hipLaunchKernelGGL((sk_rows<CT, MODE>), dim3(grid_rows(n_items, kshift)), dim3(BLK), 0, st,
                               dp<RowItem>(items), n_items, dp<uint32_t>(idx), dp<CT>(val), dp<uint32_t>(d.slot_comp),
                               locp, xsp, dp<CT>(d.inv_de), dp<double>(mu1), out, dp<CT>(rowpart), k, kshift,
                               gate);
        }
    }

    // split rows: the ones with at most FIX_SMALL partials (stored first) a row group each when a
    // workgroup holds at least FIX_SMALL row groups, the rest a workgroup each; one launch
    template <int MODE>
    void fix_rows(const DevBuf &fixes, uint32_t n_fix, uint32_t n_small, CT *locp, const double *mu,
                  OutT<CT, MODE> *out, uint32_t k, uint32_t kshift, hipStream_t st, const int32_t *gate) {
        const SchwarzDev<CT> &d = *D;
        if ((BLK >> kshift) < FIX_SMALL) n_small = 0;
        if (!n_fix) return;
        // both parts loop over their rows, so either grid may be capped
        const uint32_t g_small = n_small ? grid_rows(n_small, kshift, 32768) : 0u;
        const uint32_t g_wide = std::min<uint32_t>(n_fix - n_small, 32767u);
        hipLaunchKernelGGL((sk_rows_fix<CT, MODE>), dim3(g_small + g_wide), dim3(BLK), 0, st, dp<RowFix>(fixes),
                           n_small, n_fix, g_small, dp<uint32_t>(d.slot_comp), locp, dp<CT>(d.inv_de), mu,
                           dp<CT>(rowpart), out, k, kshift, gate);
    }

    template <typename VT>
    void mean(const DevBuf &chunks, const DevBuf &items, uint32_t n_items, const DevBuf &comb, uint32_t n_comb,
              const DevBuf &count, const VT *src, double *mu, uint32_t k, uint32_t kshift, hipStream_t st,
              const int32_t *gate) {
        if (!n_items) return;
        hipLaunchKernelGGL((sk_mean<VT>), dim3(n_items), dim3(BLK), 0, st, dp<MeanItem>(items), dp<Chunk>(chunks),
                           src, dp<uint32_t>(count), dp<double>(part), mu, k, kshift, gate);
        if (n_comb)
            hipLaunchKernelGGL(sk_mean_combine, dim3(n_comb), dim3(BLK), 0, st, dp<MeanItem>(comb), dp<double>(part),
                               dp<uint32_t>(count), mu, k, kshift, gate);
    }

    void apply_fused(double *r, double *z, uint32_t k, hipStream_t st, const int32_t *gate,
                     const ApplyFusion &fz) override {
        ensure(k, st);
        const uint32_t kshift = pow2_shift(k);
        const SchwarzDev<CT> &d = *D;
        CT *locp = dp<CT>(loc), *redp = dp<CT>(red), *xsp = dp<CT>(xs);
        if (fz.atu)
            hipLaunchKernelGGL((sk_restrict_dof<CT>), dim3(fz.blk_col ? fz.dots_blocks : grid_rows(d.N, kshift)),
                               dim3(BLK), 0, st,
                               d.N, dp<uint64_t>(d.d_ptr), dp<uint32_t>(d.d_slot), dp<CT>(d.d_w), fz.atu, r, fz,
                               locp, k, kshift, gate);
        else
            hipLaunchKernelGGL((sk_restrict<CT>), dim3(grid_rows((d.ns + 1) / 2, kshift)), dim3(BLK), 0, st, d.ns,
                               dp<uint32_t>(d.glob), dp<CT>(d.sgnpw), r, locp, k, kshift, gate);
        mean(d.all_chunks, d.all_items, d.n_all_items, d.all_comb, d.n_all_comb, d.comp_count, locp, dp<double>(mu1),
             k, kshift, st, gate);
        // reduced right-hand side on kept rows
        rows_spmv<0>(d.ct_items, d.n_ct_items, d.ct_idx, d.ct_val, locp, xsp, redp, k, kshift, st, gate);
        fix_rows<0>(d.ct_fix, d.n_ct_fix, d.n_ct_fix_small, locp, dp<double>(mu1), redp, k, kshift, st, gate);
        // kept-block solves
        if (d.n_sparse) {
            mean(d.kept_chunks, d.kept_items, d.n_kept_items, d.kept_comb, d.n_kept_comb, d.comp_kcount, redp,
                 dp<double>(mu2), k, kshift, st, gate);
            // large leading forward levels across the grid, then each component's workgroup
            // (remaining forward levels, pivots, leading backward levels), then the large
            // trailing backward levels across the grid, then the covers' read-back
            for (size_t g = 0; g + 1 < d.grid_f_ptr.size(); ++g) {
                const uint64_t b = d.grid_f_ptr[g], n = d.grid_f_ptr[g + 1] - b;
                hipLaunchKernelGGL((sk_sparse_grid<CT, true>), dim3(grid_rows(n, kshift)), dim3(BLK), 0, st,
                                   dp<GridRow>(d.grid_f_rows) + b, n, dp<SpMeta>(d.sp_meta), dp<uint64_t>(d.l_ptr),
                                   dp<uint32_t>(d.l_idx), dp<CT>(d.l_val), redp, dp<double>(mu2),
                                   dp<uint32_t>(d.aux_src), d.ns, xsp, k, kshift, gate);
            }
            const TailArgs<CT> ta{dp<uint32_t>(d.tail_rows), dp<uint64_t>(d.tail_cptr), dp<TailChunk>(d.tail_chunks),
                                  dp<uint32_t>(d.tail_idx),  dp<CT>(d.tail_val),      dp<CT>(d.tail_mat),
                                  dp<CT>(tsc),               dp<CT>(csc)};
            auto sparse = [&](uint32_t phase) {
                hipLaunchKernelGGL((sk_sparse<CT>), dim3(d.n_sparse), dim3(SP_BLK), 0, st, dp<SpMeta>(d.sp_meta),
                                   dp<uint64_t>(d.sp_lvl_f_ptr), dp<uint64_t>(d.sp_lvl_f_long),
                                   dp<uint32_t>(d.sp_lvl_f_rows), dp<uint64_t>(d.sp_lvl_b_ptr),
                                   dp<uint64_t>(d.sp_lvl_b_long), dp<uint32_t>(d.sp_lvl_b_rows),
                                   dp<uint64_t>(d.l_ptr), dp<uint32_t>(d.l_idx), dp<CT>(d.l_val),
                                   dp<uint64_t>(d.lt_ptr), dp<uint32_t>(d.lt_idx), dp<CT>(d.lt_val), dp<CT>(d.piv),
                                   redp, dp<double>(mu2), dp<uint32_t>(d.aux_src), d.ns, xsp, k, kshift, gate, phase);
            };
            // forward levels (phase 0), the dense tails across the grid (chunk partials, reduced
            // right-hand sides, M_T products), the pivots, the backward levels (phase 1)
            sparse(0);
            if (d.n_tail_rows) {
                if (d.n_tail_chunks)
                    hipLaunchKernelGGL((sk_tail_chunks<CT>), dim3(grid_rows(d.n_tail_chunks, kshift)), dim3(BLK), 0,
                                       st, ta, d.n_tail_chunks, xsp, k, kshift, gate);
                const uint32_t tgrid = (uint32_t)((d.n_tail_rows * k + BLK / 64 - 1) / (BLK / 64));
                hipLaunchKernelGGL((sk_tail_rhs<CT>), dim3(tgrid), dim3(BLK), 0, st, ta, dp<TailProdRow>(d.tail_prod),
                                   d.n_tail_rows, dp<SpMeta>(d.sp_meta), redp, dp<double>(mu2),
                                   dp<uint32_t>(d.aux_src), d.ns, k, gate);
                hipLaunchKernelGGL((sk_tail_product<CT>), dim3(tgrid), dim3(BLK), 0, st, ta,
                                   dp<TailProdRow>(d.tail_prod), d.n_tail_rows, xsp, k, gate);
            }
            if (d.n_diag_rows)
                hipLaunchKernelGGL((sk_sparse_diag<CT>), dim3(grid_rows(d.n_diag_rows, kshift)), dim3(BLK), 0, st,
                                   dp<GridRow>(d.diag_rows), d.n_diag_rows, dp<SpMeta>(d.sp_meta), dp<CT>(d.piv), xsp,
                                   k, kshift, gate);
            sparse(1);
            for (size_t h = 0; h + 1 < d.grid_b_ptr.size(); ++h) {
                const uint64_t b = d.grid_b_ptr[h], n = d.grid_b_ptr[h + 1] - b;
                hipLaunchKernelGGL((sk_sparse_grid<CT, false>), dim3(grid_rows(n, kshift)), dim3(BLK), 0, st,
                                   dp<GridRow>(d.grid_b_rows) + b, n, dp<SpMeta>(d.sp_meta), dp<uint64_t>(d.lt_ptr),
                                   dp<uint32_t>(d.lt_idx), dp<CT>(d.lt_val), redp, dp<double>(mu2),
                                   dp<uint32_t>(d.aux_src), d.ns, xsp, k, kshift, gate);
            }
            if (d.n_cover_rows)
                for (uint32_t sheet = 0; sheet < 2; ++sheet)
                    hipLaunchKernelGGL((sk_cover_readback<CT>), dim3(grid_rows(d.n_cover_rows, kshift)), dim3(BLK), 0, st,
                                       dp<uint32_t>(d.cover_rows), d.n_cover_rows, dp<uint32_t>(d.aux_src), d.ns,
                                       xsp, sheet, k, kshift, gate);
        }
        if constexpr (std::is_same<CT, float>::value)
            for (uint32_t c0 = 0; d.n_stiles && c0 < k; c0 += 16)
                hipLaunchKernelGGL(sk_dense_upper, dim3(d.n_stiles), dim3(256), 0, st, dp<SymTile>(d.stiles),
                                   dp<float>(d.dsym), redp, dp<float>(dpart), k, c0, std::min<uint32_t>(16, k - c0),
                                   gate);
        if (d.n_dtiles) {
            // column windows of exact power-of-two widths (<= 16): the wave-uniform red loads
            // never read past column k
            for (uint32_t c0 = 0; c0 < k;) {
                uint32_t kc = 16;
                while (kc > k - c0) kc >>= 1;
#define WX_DS(N)                                                                                                   \
    hipLaunchKernelGGL((sk_dense_sym<CT, N>), dim3(d.n_dtiles), dim3(64), 0, st, dp<DenseTile>(d.dtiles),         \
                       dp<CT>(d.dmat), redp, dp<CT>(dpart), k, c0, gate)
                if (kc == 16) WX_DS(16);
                else if (kc == 8) WX_DS(8);
                else if (kc == 4) WX_DS(4);
                else if (kc == 2) WX_DS(2);
                else WX_DS(1);
#undef WX_DS
                c0 += kc;
            }
        }
        if (d.n_dtiles || d.n_stiles) {
            hipLaunchKernelGGL((sk_dense_reduce<CT>), dim3(grid_rows(d.dense_tiled_rows, kshift)), dim3(BLK), 0, st,
                               dp<DenseComp>(d.dcomps), d.n_dcomps, dp<uint64_t>(d.drow_start), dp<CT>(dpart), xsp,
                               k, kshift, gate);
        }
        if (d.n_drows)
            hipLaunchKernelGGL((sk_dense<CT>), dim3(grid_rows(d.n_drows, kshift)), dim3(BLK), 0, st, d.n_drows,
                               dp<uint64_t>(d.drow_off), dp<uint32_t>(d.drow_base), dp<uint32_t>(d.drow_nk),
                               dp<uint32_t>(d.drow_slot), dp<CT>(d.dmat), redp, xsp, k, kshift, gate);
        // back-substitution on eliminated rows
        rows_spmv<1>(d.c_items, d.n_c_items, d.c_idx, d.c_val, locp, xsp, xsp, k, kshift, st, gate);
        fix_rows<1>(d.c_fix, d.n_c_fix, d.n_c_fix_small, locp, dp<double>(mu1), xsp, k, kshift, st, gate);
        mean(d.all_chunks, d.all_items, d.n_all_items, d.all_comb, d.n_all_comb, d.comp_count, xsp, dp<double>(mu3),
             k, kshift, st, gate);
        if (fz.dots)
            hipLaunchKernelGGL((sk_prolong<CT, true>), dim3(fz.dots_blocks), dim3(BLK), 0, st, d.N,
                               dp<uint64_t>(d.d_ptr), dp<uint32_t>(d.d_slot), dp<CT>(d.d_w),
                               dp<uint32_t>(d.d_comp), xsp, dp<double>(mu3), z, k, kshift, gate, r, fz.dots);
        else
            hipLaunchKernelGGL((sk_prolong<CT, false>), dim3(grid_rows(d.N, kshift)), dim3(BLK), 0, st, d.N,
                               dp<uint64_t>(d.d_ptr), dp<uint32_t>(d.d_slot), dp<CT>(d.d_w),
                               dp<uint32_t>(d.d_comp), xsp, dp<double>(mu3), z, k, kshift, gate, nullptr,
                               nullptr);
        WX_HIP(hipGetLastError());
    }
};

template <typename CT> std::shared_ptr<SchwarzDev<CT>> get_dev(const Precond &p, int slot) {
    std::lock_guard<std::mutex> lk(p.dev_mu);
    if (!p.dev_cache[slot]) p.dev_cache[slot] = build_dev<CT>(*p.schwarz);
    return std::static_pointer_cast<SchwarzDev<CT>>(p.dev_cache[slot]);
}

// y[d + j*n] <-> interleaved conversions for the standalone apply
__global__ void sk_to_interleaved(uint64_t n, const double *__restrict__ x, double *__restrict__ xi, uint32_t k) {
    const uint64_t idx = (uint64_t)blockIdx.x * BLK + threadIdx.x;
    if (idx >= n * k) return;
    xi[idx] = x[(idx % k) * n + idx / k];
}
__global__ void sk_from_interleaved(uint64_t n, const double *__restrict__ yi, double *__restrict__ y, uint32_t k) {
    const uint64_t idx = (uint64_t)blockIdx.x * BLK + threadIdx.x;
    if (idx >= n * k) return;
    y[(idx % k) * n + idx / k] = yi[idx];
}
__global__ void sk_diag(uint64_t n, const double *__restrict__ inv, const double *__restrict__ x,
                        double *__restrict__ y, uint32_t k) {
    const uint64_t idx = (uint64_t)blockIdx.x * BLK + threadIdx.x;
    if (idx >= n * k) return;
    y[idx] = inv[idx % n] * x[idx];
}

} // namespace

std::unique_ptr<SchwarzRun> make_schwarz_run(const Precond &p, int32_t precision, hipStream_t st) {
    (void)st;
    if (precision == WX_PRECISION_MIXED) {
        auto r = std::make_unique<SchwarzRunImpl<float>>();
        r->D = get_dev<float>(p, 1);
        return r;
    }
    auto r = std::make_unique<SchwarzRunImpl<double>>();
    r->D = get_dev<double>(p, 0);
    return r;
}

void precond_apply_device(const Precond &p, const double *x, uint64_t k, double *y, hipStream_t st,
                          int32_t precision) {
    const uint64_t n = p.n;
    if (p.kind == Precond::DIAGONAL) {
        DevBuf inv;
        inv.upload(p.inv_diag.data(), n, st);
        const uint64_t tot = n * k;
        hipLaunchKernelGGL(sk_diag, dim3((tot + BLK - 1) / BLK), dim3(BLK), 0, st, n, dp<double>(inv), x, y,
                           (uint32_t)k);
        WX_HIP(hipStreamSynchronize(st));
        return;
    }
    auto run = make_schwarz_run(p, precision, st);
    for (uint64_t c0 = 0; c0 < k; c0 += MAX_BATCH) {
        const uint32_t kk = (uint32_t)std::min<uint64_t>(MAX_BATCH, k - c0);
        DevBuf xi, yi;
        xi.alloc(n * kk * 8);
        yi.alloc(n * kk * 8);
        const uint64_t tot = n * kk;
        hipLaunchKernelGGL(sk_to_interleaved, dim3((tot + BLK - 1) / BLK), dim3(BLK), 0, st, n, x + c0 * n,
                           dp<double>(xi), kk);
        run->apply(dp<double>(xi), dp<double>(yi), kk, st, nullptr);
        hipLaunchKernelGGL(sk_from_interleaved, dim3((tot + BLK - 1) / BLK), dim3(BLK), 0, st, n, dp<double>(yi),
                           y + c0 * n, kk);
        WX_HIP(hipStreamSynchronize(st));
    }
}

} // namespace wx
