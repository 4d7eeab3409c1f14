// solver.hpp -- the Solver / Preconditioner objects behind the C ABI.
#pragma once

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "device.hpp"

namespace wx {

constexpr uint32_t MAX_TERMS = 8;
constexpr uint32_t DEFAULT_TILE = 64; // rows per D' / fused-gather work item
constexpr uint32_t MAX_COLS = 4;   // coefficient columns per term on the device path
constexpr uint32_t MAX_BATCH = 64; // RHS columns interleaved per device pass
constexpr uint32_t KERNEL_BLOCK = 256;        // threads per workgroup of the solver kernels
constexpr uint32_t MAX_PART_BLOCKS = 4096;    // grid cap of the grid-stride kernels
constexpr uint32_t PACK_ALIGN = 256;          // DOF / work-item alignment of a packed problem
constexpr int32_t SEG_SKIP = INT32_MIN;       // SegItem::dest of a padding item (no rows, no output)

// Per-term description passed by value to kernels.
struct TermInfo {
    uint64_t n_levels;
    uint64_t offset;
    uint32_t ncols;
    int32_t load[MAX_COLS]; // -1 intercept, else slope column index
};

struct DesignArgs {
    uint64_t n;
    uint32_t T;
    const uint32_t *codes;  // T x n
    const double *slopes;   // S x n
    TermInfo terms[MAX_TERMS];
};

// Work item of the deterministic D' segmented reduction: rows at positions [begin, end)
// of the term's code-sorted order sum into `dest` (>= 0: the levels level .. dest, complete,
// usually one; < 0: partial slot -dest-1 of `level`, combined afterwards in a fixed order;
// SEG_SKIP: a padding item with no rows).
struct SegItem {
    uint32_t level;
    uint32_t begin;
    uint32_t end;
    int32_t dest;
};
struct SegFix {
    uint32_t level;
    uint32_t first;
    uint32_t count;
    uint32_t pad;
};

struct DevTerm {
    TermInfo info;
    DevBuf perm;   // position -> internal row, empty when the term is sorted
    DevBuf sw_pos, w_pos; // sqrt-weights and weights by position (unsorted weighted terms)
    DevBuf items;  // SegItem[n_items], one level (or one chunk of a level) each
    uint32_t n_items = 0;
    // fused gather's items: consecutive complete levels merged up to SEG_CHUNK rows, so every
    // row group gets a similar amount of work (the fused term only)
    DevBuf gitems;
    uint32_t n_gitems = 0;
    DevBuf fixes;  // SegFix[n_fixes] (one-stage combine, terms with several columns)
    uint32_t n_fixes = 0;
    uint32_t n_partials = 0;
    // single-column terms: the split levels' partials are combined in stages, each stage a
    // segmented sum over the previous stage's partial slots (<= 64 per item, in slot order)
    struct FixStage {
        DevBuf items;    // SegItem over partial slots of region src_base
        uint32_t n_items = 0;
        uint64_t src_base = 0, dst_base = 0; // partial-slot offsets of the input / output regions
    };
    std::vector<FixStage> stages;
};

// The fused D' of two unsorted terms P and Q of a plain design: P's rows in (P, Q) order, a
// piece per run of equal Q codes inside one of P's work items, and Q's D' as a segmented sum over
// the pieces (lsmr.hip, build_pair_plan).
struct PairPlan {
    int32_t p = -1, q = -1;
    DevBuf qcode_pos;   // u32[n]: Q's code at each of P's positions
    DevBuf item_piece0; // u32[P's n_items + 1]: first piece of each of P's work items
    DevBuf piece_slot;  // u32[n_pieces]: where each piece is stored (the pieces in Q order)
    uint32_t n_pieces = 0;
    DevTerm qt;         // Q's reduction over the stored pieces: items over Q's levels, fixups
};

// The fix-ups of a three-term plain design's D' after the fused gather and the fused pair,
// merged across the terms: launch d runs the sorted term's and P's combine stage d together
// with Q's reduction over the pieces (d = 0) or Q's stage d - 1.  Every item is addressed in one
// partial-slot buffer (the sorted term's slots first, then P's, the pieces and Q's) and names its
// absolute output row (lsmr.hip, build_merged_fix).
struct MergedFix {
    std::vector<DevBuf> levels; // SegItem[] per launch
    std::vector<uint32_t> n_items;
    uint64_t base_p = 0, base_pieces = 0, slots = 0;
    bool on() const { return !levels.empty(); }
};

struct DevDesign {
    uint64_t n = 0, N = 0;
    DevBuf codes, slopes, sw, w, obs_perm, inv_perm, inv_diag;
    DevBuf sw_caller; // sqrt-weights in caller row order (when the rows are permuted)
    DevBuf codes4;    // plain designs with at most 4 terms: the row's codes as one 16-byte record
    DevBuf codes4_caller; // the same records in caller row order (when the rows are permuted)
    std::vector<DevTerm> terms;
    uint32_t max_partials = 0;
    uint32_t tile = DEFAULT_TILE; // rows per D' work item (upload_design)
    bool plain = false; // every term is a single intercept column
    int32_t fuse_term = -1; // sorted single-intercept term whose D' is fused into the gather
    PairPlan pair;          // p < 0: no fused pair
    MergedFix mfix;         // off unless the design is the sorted term plus a fused pair
    DesignArgs args() const;
};

// Preconditioner (host data + lazily built device form).
struct Precond {
    enum Kind { DIAGONAL = 0, ADDITIVE = 1 } kind = DIAGONAL;
    uint64_t n = 0;
    std::vector<double> inv_diag;
    struct Schwarz;                 // defined in schwarz.hpp
    std::shared_ptr<Schwarz> schwarz;
    // device forms of the Schwarz data, built on first use per precision (0 FP64, 1 mixed)
    mutable std::mutex dev_mu;
    mutable std::shared_ptr<void> dev_cache[2];
    virtual ~Precond() = default;
};

struct Solver {
    uint64_t uid = 0; // process-unique id (cache key of packed groups)
    Design design;
    std::unique_ptr<SlopeReparam> reparam;
    bool weighted = false;
    double min_weight = 1.0; // smallest observation weight (1 when unweighted)
    std::vector<double> w_internal; // observation weights in internal order (empty: unweighted)
    std::shared_ptr<Precond> precond;
    std::vector<std::string> warnings;
    double time_setup = 0.0;
    DevDesign dev;
    // pool of reusable device workspaces (see lsmr.hip), keyed by batch width
    mutable std::mutex pool_mu;
    mutable std::vector<std::shared_ptr<void>> pool;
    mutable std::vector<uint32_t> pool_k;
};

// Device-side index structures (setup_gpu.hip): stable argsorts by one key or by a pair of keys
// (radix sorts of (key, row)), gathers, square roots, inverse permutations, and the pieces of a
// fused pair (runs of equal Q code inside P's work items; piece0 receives each item's first
// piece, piece_q each piece's Q code; returns the piece count).  Enqueued on `st`.
// Slope reparametrisation with the per-level moments and whitening on the device; the same
// contract as build_reparam.
std::unique_ptr<SlopeReparam> build_reparam_device(Design &d, const double *weights);
void dev_sqrt(const double *src, uint64_t n, double *out, hipStream_t st);
void dev_gather(const double *src, const uint32_t *perm, uint64_t n, double *out, hipStream_t st);
void dev_gather_u32(const uint32_t *src, const uint32_t *perm, uint64_t n, uint32_t *out, hipStream_t st);
void dev_invert_perm(const uint32_t *perm, uint64_t n, uint32_t *inv, hipStream_t st);
void dev_stable_argsort(const uint32_t *key, uint64_t n, uint64_t n_levels, uint32_t *perm, hipStream_t st);
void dev_stable_argsort2(const uint32_t *kp, const uint32_t *kq, uint64_t lp, uint64_t lq, uint64_t n, uint32_t *perm,
                         hipStream_t st);
uint64_t dev_pieces(const uint32_t *qpos, uint64_t n, const SegItem *items, uint32_t n_items, uint32_t *piece0,
                    DevBuf &piece_q, hipStream_t st);

// Uploads the design (codes, slopes, weights, D' segment plans) to the current device.
// `tile`: the most rows of one D' work item (a larger level is split into pieces of that many
// rows, whose partial sums are combined in order), and of one fused-gather item.
void upload_design(Solver &s, const std::vector<double> &weights_internal, uint32_t tile = DEFAULT_TILE);

// s.dev = base's device design with the observation weights w (internal order; empty:
// unweighted), the index structures copied on the device (s.design and s.precond set).
void reweight_design(Solver &s, const Solver &base, const std::vector<double> &w);

// Gram matrix of a column-major device matrix (fwl.hip)
uint32_t gram_max_cols();
uint64_t gram_scratch_doubles(uint64_t n, uint32_t m);
void gram_device(const double *A, uint64_t n, uint32_t m, uint64_t lda, double *G, double *part, hipStream_t st);
uint32_t gram_noalloc_max_cols();
double *gram_scratch_ptr(std::string &err);

uint64_t next_solver_uid();

// Independent problems of one shape packed into one block-diagonal design, so that every
// iteration kernel runs once for all of them.  Member m's rows are the merged rows
// [row_off, row_off + n) in the member's own internal order; its term-t levels are the
// merged levels [lvl_off[t], lvl_off[t] + L[t]) of term t, padded to PACK_ALIGN so that no
// workgroup of a DOF-space kernel straddles two members; the fused term's work items are
// padded the same way.  LSMR state of member m, column j lives at index m * k + j.
struct PackMember {
    const Solver *s = nullptr;
    uint64_t row_off = 0, n = 0, n_dofs = 0;
    uint32_t tile_off = 0, n_tiles = 0; // right-hand-side load partial blocks
    uint32_t kv = 1;                    // the member's own columns (<= the pack's k)
    uint64_t lvl_off[MAX_TERMS] = {}, L[MAX_TERMS] = {}, own_off[MAX_TERMS] = {};
};
struct Pack {
    std::vector<PackMember> members;
    uint32_t k = 1, kshift = 0;
    std::unique_ptr<Solver> merged;
    uint32_t n_dof_blocks = 0, n_item_blocks = 0, n_tiles = 0;
    DevBuf dof_col, item_col;         // u32 per DOF block / fused-item block: member * k
    DevBuf dof_rng_ptr, dof_rng;      // per member: CSR of [lo, hi) DOF-block ranges
    DevBuf item_rng_ptr, item_rng;    // per member: its fused-item blocks
    DevBuf tile_rng_ptr, tile_rng;    // per member: its load partial blocks
    DevBuf dof_member, dof_own;       // per merged DOF: member, own DOF index (-1: padding)
    DevBuf own_ld;                    // per member: own n_dofs (u64)
    DevBuf own_kv;                    // per member: own columns (u32)
    uint64_t part_blocks = 0;         // partial blocks the largest packed launch writes
};

// Whether the solvers can share one pack (same term count and sorted term, plain design,
// same weighting and preconditioner kind) and their key.
bool pack_signature(const Solver &s, std::string &sig);

// Builds (or returns from a small cache) the pack of `members` with kv[m] columns each, run at
// batch width k = max kv.
std::shared_ptr<Pack> get_pack(const std::vector<const Solver *> &members, const std::vector<uint32_t> &kv,
                               uint32_t tile = DEFAULT_TILE);

// Runs the batched solve with device-resident inputs/outputs on `stream`.  x_solver_basis:
// X is left in the solver's (reparametrised) coefficient basis instead of the caller's.
void solve_batch_device(const Solver &s, const double *Y, uint64_t ldy, uint64_t k, const wx_lsmr_options &opts,
                        double *X, double *Dm, wx_rhs_stats *stats, hipStream_t stream, bool x_solver_basis = false);

// Many independent batched solves at once (solvers[p] on Y[p], n_p x k[p] column-major,
// device arrays); outputs and statistics as for solve_batch_device, per problem.
void solve_many_device(const std::vector<const Solver *> &solvers, const std::vector<const double *> &Y,
                       const std::vector<uint64_t> &ldy, const std::vector<uint64_t> &k, const wx_lsmr_options &opts,
                       const std::vector<double *> &X, const std::vector<double *> &Dm,
                       const std::vector<wx_rhs_stats *> &stats, hipStream_t caller);

// Fixes the shape of later solve_many_device calls (streams 0: back to the autotuner).
void force_many_shape(int32_t streams, int32_t group, int32_t graph, int32_t pack);

// The (streams, iterations per graph, graph on, packed) shape the last solve_many_device used.
void last_many_shape(int32_t *streams, int32_t *group, int32_t *graph, int32_t *pack);
void force_many_tile(int32_t rows);
int32_t last_many_tile();

// Whether later solves fold the LSMR steps on either side of an additive Schwarz apply into its
// first and last launch (default) or run them as launches of their own; returns the previous
// setting.  The two forms compute the same values in the same order.
bool set_fused_steps(bool on);

// Adjoint of the weighted projection R = Z - D x(Z, w) for reverse-mode differentiation:
// grad_z = G - w .* (D lambda), grad_w = -sum_j R_j .* (D lambda)_j with D'WD lambda = D'G.
// Column-major device arrays in caller row order; grad_w may be null.
void projection_grad_device(const Solver &s, const double *G, uint64_t ldg, const double *R, uint64_t ldr, uint64_t k,
                            const wx_lsmr_options &opts, double *grad_z, double *grad_w, wx_rhs_stats *stats,
                            hipStream_t stream);

// Per-phase (ms, launches) pairs of this thread's last timed solve (see wx_last_timing).
std::vector<double> &last_timing();

// Y = sqrt(W) D X (adjoint: Y = D' sqrt(W) X) for column-major device arrays in caller
// row order (n x k and N x k).
void design_apply_device(const Solver &s, bool adjoint, const double *X, uint64_t k, double *Y, hipStream_t stream);

// y = M^{-1} x on the device for n x k column-major arrays (FP64 coefficients).
// y = M^{-1} x (n x k column-major device arrays); `precision` selects the apply of the FP64 or
// the mixed solve (FP32 coefficients and slot-space vectors).
void precond_apply_device(const Precond &p, const double *x, uint64_t k, double *y, hipStream_t stream,
                          int32_t precision = WX_PRECISION_FP64);

} // namespace wx
