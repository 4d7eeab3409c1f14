// core.hpp -- host-side types shared by the design, preconditioner builders and solver.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/within_amd.h"

namespace wx {

// Error carrying a wx_status and, where one applies, the name of the reference error variant
// it stands for (e.g. "BuildError::InvalidWeight"); converted at the C boundary.
struct Error : std::runtime_error {
    wx_status code;
    const char *variant;
    Error(wx_status c, const std::string &msg, const char *v) : std::runtime_error(msg), code(c), variant(v) {}
};

// The reference variant a status stands for when the throw site names none.
inline const char *default_variant(wx_status code) {
    switch (code) {
    case WX_ERR_INVALID_INPUT: return "SolveError::InvalidInput";
    case WX_ERR_NOT_POSITIVE_DEF: return "SolveError::InvalidInput";
    case WX_ERR_BUILD: return "BuildError::LocalSolverBuild";
    default: return "";
    }
}

[[noreturn]] inline void fail(wx_status code, const std::string &msg, const char *variant = nullptr) {
    throw Error(code, msg, variant ? variant : default_variant(code));
}

// One term (effect) of the design.  Coefficient column c of level l lives at
// offset + c * n_levels + l (crates/within/src/domain.rs:80-103).
struct Term {
    uint64_t n_levels = 0;
    uint64_t offset = 0;
    bool sorted = true;
    // loads[c] = -1 for the intercept column, else the index of a slope column.
    std::vector<int64_t> loads;
    uint64_t ncols() const { return loads.size(); }
    uint64_t n_dofs() const { return ncols() * n_levels; }
};

// Fixed-effects design in internal row order (crates/within/src/domain.rs:143-153).
struct Design {
    uint64_t n_obs = 0;
    uint64_t n_dofs = 0;
    std::vector<Term> terms;
    std::vector<uint32_t> codes;    // term t at codes[t * n_obs + i]
    std::vector<double> slopes;     // slope column s at slopes[s * n_obs + i]
    uint64_t n_slopes = 0;
    std::vector<uint32_t> obs_perm; // empty when the internal order is the caller's
    bool has_slopes() const { return n_slopes > 0; }
    const uint32_t *level_column(uint64_t t) const { return codes.data() + t * n_obs; }
    const double *slope_column(uint64_t s) const { return slopes.data() + s * n_obs; }
};

// Builds a design from per-term code columns (caller order) and term structure.
// Restates Design::build (crates/within/src/domain.rs:187-247).
Design build_design(uint64_t n_obs, const std::vector<std::vector<uint32_t>> &columns,
                    const std::vector<std::vector<int64_t>> &structure,
                    const std::vector<std::vector<double>> &slopes, bool locality_sort = true);

// Stable ascending argsort by level code (crates/within/src/domain.rs:110-140).
std::vector<uint32_t> stable_argsort(const uint32_t *key, uint64_t n, uint64_t n_levels);

// Caller order <-> internal order (crates/within/src/domain.rs:282-303).
void permute_in(const Design &d, const double *src, double *dst);

// Validates weights: length and finite non-negative entries (domain.rs:261-279).
void validate_weights(const Design &d, const double *w, uint64_t len);

// Active levels per term (crates/within/src/domain/cross_tab.rs:25-39).
std::vector<std::vector<uint8_t>> find_active_levels(const Design &d);

// Per-level whitening of slope columns (crates/within/src/solver/reparam.rs:18-316).
struct SlopeReparam {
    struct TermBlock {
        uint64_t term = 0;
        uint64_t first_col = 0; // first slope column index inside the term's loads
        uint64_t n_slopes = 0;
        bool intercept = false;
        // per level: V x V upper-triangular back-transform T (x_orig = T x_whitened) and means mu
        std::vector<double> transform; // n_levels * V * V, row-major per level
        std::vector<double> mean;      // n_levels * V
    };
    std::vector<TermBlock> blocks;
    std::vector<uint64_t> unidentified; // triples (term, level, column)
    void back_transform(const Design &d, double *x) const;
};

// Builds the reparametrisation and rewrites the design's slope columns in place.
std::unique_ptr<SlopeReparam> build_reparam(Design &d, const double *weights);

} // namespace wx
