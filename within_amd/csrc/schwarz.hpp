// schwarz.hpp -- additive Schwarz preconditioner: host representation and device apply.
#pragma once

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "solver.hpp"

namespace wx {

// Local solver settings (crates/within/src/config.rs:20-131).
struct LocalSolverSettings {
    uint64_t approx_chol_seed = 0;
    int32_t split_merge = 2;     // <= 0: None
    bool schur_exact = false;
    uint64_t schur_seed = 0;
    int32_t schur_split = 1;
    int64_t dense_threshold = 24; // dense exact factor iff n_kept <= dense_threshold (< 0: never)
    bool dense_device = true;     // additionally dense where the device cost model prefers it
    double scaling_tolerance = 1e-9;
    int64_t scaling_max_sweeps = 2048;
    bool scaling_error = false;
};

// Allocator whose resize leaves doubles uninitialised: the dense kept blocks (hundreds of MB) are
// written in full by their builders, so zero-filling them first would only cost time.
template <typename T> struct NoInitAlloc : std::allocator<T> {
    template <typename U> struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <typename U> NoInitAlloc(const NoInitAlloc<U> &) {}
    template <typename U, typename... A> void construct(U *p, A &&...a) {
        if constexpr (sizeof...(A) == 0) ::new ((void *)p) U;
        else ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
using DenseVec = std::vector<double, NoInitAlloc<double>>;

// Host form of a one-level additive Schwarz preconditioner over factor-pair components.
// Every component is stored eliminated-block first: slots [slot_start[c], +ne[c]) are the
// eliminated levels, the next nk[c] slots the kept levels.  A component's matrix is stored
// in its SDDM coordinates (crates/within/src/domain/factor_pairs/sddm.rs): the restriction
// multiplies by fold[s] (the coordinate map: +1 / -1 on the two blocks of a plain pair, sign
// times dominance scale of a general one), the prolongation by fold[s] again.
struct Precond::Schwarz {
    uint64_t n_dofs = 0;
    // components
    std::vector<uint64_t> slot_start; // n_comp + 1
    std::vector<uint32_t> ne, nk;
    // 0: dense kept-block operator (S^+ of a floating Laplacian, S^-1 otherwise);
    // 1: sparse LDL' over the kept slots (in elimination order);
    // 2: sparse LDL' of the Gremban cover of a signed kept block (2 nk vertices, cover_perm)
    std::vector<uint8_t> fac_kind;
    std::vector<uint8_t> gauge;       // 0 floating (component means removed), 1 none
    // slots
    std::vector<uint32_t> glob;       // global DOF
    std::vector<double> pw;           // partition weight
    std::vector<double> fold;         // coordinate-map factor
    std::vector<double> inv_de;       // 1/D_elim (0 on kept slots)
    // C (eliminated x kept, local kept index) as CSR over slots; kept slots have empty rows
    std::vector<uint64_t> c_ptr;      // n_slots + 1
    std::vector<uint32_t> c_idx;
    std::vector<double> c_val;
    // dense kept-block operators: S^+ (nk x nk row-major) at dense_off[c]
    std::vector<uint64_t> dense_off;  // n_comp
    DenseVec dense;
    // Blocks formed on the GPU stay resident there (FP64, `count` entries that belong at `off` in
    // `dense`); the device apply copies them from there, and their host copy in `dense` is
    // written only when the host needs it (serialisation, merging): host_dense().
    struct DeviceBlock {
        uint64_t off = 0, count = 0;
        std::shared_ptr<double> dev;
    };
    std::vector<DeviceBlock> dense_dev;
    std::shared_ptr<std::mutex> dense_mu = std::make_shared<std::mutex>();
    bool dense_on_host = true; // guarded by *dense_mu
    // Completes `dense` on the host (downloads device blocks once); safe from several threads.
    void host_dense() const;
    // sparse LDL' factors: L strictly lower, CSR by rows (local indices), pivots d
    std::vector<uint64_t> l_row_off;  // n_comp: offset of the component's first row pointer
    std::vector<uint64_t> l_ptr;
    std::vector<uint32_t> l_idx;
    std::vector<double> l_val;
    std::vector<uint64_t> piv_off;    // n_comp
    std::vector<double> piv;
    // Gremban covers (fac_kind 2): cover vertex eliminated t-th is cover_perm[cover_off[c] + t];
    // vertex v < nk is kept level v on the first sheet, v >= nk kept level v - nk on the second
    std::vector<uint64_t> cover_off;  // n_comp (unused unless fac_kind 2)
    std::vector<uint32_t> cover_perm;

    uint64_t n_comp() const { return ne.size(); }
    uint64_t n_slots() const { return glob.size(); }
};

// The LSMR steps on either side of a preconditioner apply that a solve can fold into it
// (column j's state at index j, or at blk_col[b] + j in a packed run):
//  - before: p~ <- beta_inv A'u - p_coeff p~ for every active column (0 after a breakdown),
//    formed per DOF inside the restriction, which then reads the new p~ (bidiag.rs:475-485);
//  - after:  per-block partial sums of <v,p~>, <v,v>, <p~,p~> (bidiag.rs:99-111), formed inside
//    the prolongation, laid out part[(b * kp + j) * 3 + f] over `dots_blocks` blocks.
struct ApplyFusion {
    const double *atu = nullptr; // A'u (null: p~ is the input as it stands)
    bool init = false;           // the first step instead: p~ <- beta_inv p~ (p~ holds A'b)
    const double *beta_inv = nullptr, *p_coeff = nullptr;
    const int32_t *brk = nullptr, *active = nullptr;
    double *dots = nullptr; // null: no partial sums
    uint32_t dots_blocks = 0;
    // a packed run: both launches take `dots_blocks` blocks, block b covering whole DOF blocks
    // of one member, whose column-state base is blk_col[b] (null: one problem, base 0)
    const uint32_t *blk_col = nullptr;
};

// Device apply of an additive Schwarz preconditioner for one batch width.
struct SchwarzRun {
    virtual ~SchwarzRun() = default;
    // v = M^{-1} pt for N x k interleaved FP64 arrays, enqueued on `st`.  `gate` (k flags on
    // the device, or null) lets every kernel return at once when no flag is set.
    void apply(const double *pt, double *v, uint32_t k, hipStream_t st, const int32_t *gate) {
        apply_fused(const_cast<double *>(pt), v, k, st, gate, ApplyFusion{});
    }
    // The same with the steps of `fz` folded in; pt is updated in place when fz.atu is set.
    virtual void apply_fused(double *pt, double *v, uint32_t k, hipStream_t st, const int32_t *gate,
                             const ApplyFusion &fz) = 0;
};

std::unique_ptr<SchwarzRun> make_schwarz_run(const Precond &p, int32_t precision, hipStream_t st);

// Builds the additive Schwarz preconditioner of a design (operator/schwarz.rs:232-271).
// Returns nullptr when the design has no factor pair (plain LSMR, schwarz.rs:249-252).
std::shared_ptr<Precond> build_additive(const Design &d, const std::vector<double> &weights,
                                        const LocalSolverSettings &cfg, std::vector<std::string> &warnings);

// A component's kept-block reduction on the GPU (set-up time): S = D_k - C' D_e^{-1} C with its
// diagonal formed as the host assembly forms it (Laplacian row sums, or the signed form's
// magnitudes, plus the grounded surplus), then S^+ (floating Laplacian: P G P over the grounded
// minor's inverse G) or S^{-1}, written row-major n_kept x n_kept to `out`.  False when the
// Cholesky factorisation meets a non-positive pivot.
struct DenseSchurArgs {
    uint32_t ne, nk;
    const uint64_t *cptr; // C: ne x nk CSR, columns ascending per row
    const uint32_t *cidx;
    const double *cval;
    uint64_t nnz;
    const double *diag;    // ne + nk
    const double *surplus; // ne + nk, or null (not grounded)
    const double *rowabs;  // ne row sums of |C| (signed form), or null
    const uint64_t *tp;    // C' CSR: kept row k's eliminated neighbours te/tv[tp[k] .. tp[k+1])
    const uint32_t *te;
    const double *tv;
    bool signed_form;
};
bool gpu_available();

// The cross-tab cells of a plain factor pair (r, c) in row-major order: key = level_r * L_c +
// level_c, val = the weight sum of the cell's rows added in observation order (cells with no rows
// are absent).  device_cross_tabs computes every listed pair's cells on the GPU, bitwise equal to
// the host accumulation; it returns false (and computes nothing) when no GPU is visible.
struct CellList {
    std::vector<uint64_t> key;
    std::vector<double> val;
};
bool device_cross_tabs(const Design &d, const std::vector<double> &w,
                       const std::vector<std::pair<uint32_t, uint32_t>> &pairs, std::vector<CellList> &out);
// The operator (n_kept x n_kept, row-major, FP64) stays on the device in `out`.
bool dense_schur_pinv_gpu(const DenseSchurArgs &a, bool floating, std::shared_ptr<double> &out);

// Diagonal preconditioner (operator/schwarz.rs:185-229).
std::shared_ptr<Precond> build_diagonal(const Design &d, const std::vector<double> &weights);

// Test hooks (include/within_amd_testing.h): the builder's component pipeline on one general
// bipartite block [[diag_r, B], [B', diag_c]] given directly.
std::shared_ptr<Precond> build_block_precond(uint32_t n_rows, uint32_t n_cols, const double *B, const double *diag,
                                             const LocalSolverSettings &cfg);
std::vector<double> block_reduced_laplacian(uint32_t n_rows, uint32_t n_cols, const double *B, const double *diag,
                                            const LocalSolverSettings &cfg, uint32_t &n_out);

// Versioned binary format of a preconditioner.
std::vector<uint8_t> serialize_precond(const Precond &p);
std::shared_ptr<Precond> deserialize_precond(const uint8_t *buf, uint64_t len);

} // namespace wx
