// capi.cpp -- the extern "C" boundary declared in include/within_amd.h.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/within_amd_testing.h"
#include "schwarz.hpp"
#include "solver.hpp"

using namespace wx;

struct wx_design {
    std::atomic<int> refs{1};
    uint64_t n_obs = 0;
    std::vector<std::vector<uint32_t>> columns;
    std::vector<std::vector<int64_t>> structure;
    std::vector<std::vector<double>> slopes;
};

struct wx_solver {
    std::atomic<int> refs{1};
    Solver s;
};

struct wx_precond {
    std::atomic<int> refs{1};
    std::shared_ptr<Precond> p;
};

namespace {

thread_local std::string g_last_error;
thread_local const char *g_last_variant = "";

template <typename F> wx_status guard(F &&f) {
    try {
        g_last_error.clear();
        g_last_variant = "";
        f();
        return WX_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        g_last_variant = e.variant;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "out of host memory";
        return WX_ERR_RUNTIME;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return WX_ERR_RUNTIME;
    }
}

void require_device() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        fail(WX_ERR_DEVICE, "no AMD GPU visible to the HIP runtime: within_amd runs its solves on the device only");
}

// The stream of a call: the caller's, or (NULL) a library-owned non-blocking stream leased
// from a per-device pool for the call's duration.  Pooled streams live until the process ends:
// creating and destroying one per call cost about 0.25 ms of host time, during which the
// device idled between back-to-back solves.
std::mutex g_own_mu;
std::vector<std::pair<int, hipStream_t>> g_own_free; // (device, idle stream)

struct Stream {
    hipStream_t s = nullptr;
    bool own = false;
    int dev = 0;
    explicit Stream(void *ext) {
        if (ext) {
            s = (hipStream_t)ext;
            return;
        }
        WX_HIP(hipGetDevice(&dev));
        {
            std::lock_guard<std::mutex> lk(g_own_mu);
            for (size_t i = 0; i < g_own_free.size(); ++i)
                if (g_own_free[i].first == dev) {
                    s = g_own_free[i].second;
                    g_own_free.erase(g_own_free.begin() + i);
                    break;
                }
        }
        if (!s) WX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        own = true;
    }
    ~Stream() {
        if (!own || !s) return;
        std::lock_guard<std::mutex> lk(g_own_mu);
        g_own_free.push_back({dev, s});
    }
    Stream(const Stream &) = delete;
    Stream &operator=(const Stream &) = delete;
};

LocalSolverSettings settings_from(const wx_precond_config &c) {
    LocalSolverSettings st;
    st.approx_chol_seed = c.approx_chol_seed;
    st.split_merge = c.split_merge;
    st.schur_exact = c.schur == WX_SCHUR_EXACT;
    st.schur_seed = c.schur_seed;
    st.schur_split = c.schur_split;
    st.dense_threshold = c.dense_threshold;
    st.dense_device = c.dense_policy != WX_DENSE_REFERENCE;
    st.scaling_tolerance = c.scaling_tolerance;
    st.scaling_max_sweeps = c.scaling_max_sweeps;
    st.scaling_error = c.scaling_error_on_failure != 0;
    return st;
}

} // namespace

extern "C" {

const char *wx_last_error(void) { return g_last_error.c_str(); }

const char *wx_last_error_variant(void) { return g_last_variant; }

int32_t wx_device_info(char *name, uint64_t name_len) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
    if (name && name_len) {
        name[0] = 0;
        if (count > 0) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, 0) == hipSuccess) {
                std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
            }
        }
    }
    return count;
}

void wx_precond_config_default(wx_precond_config *cfg) {
    // crates/within/src/config.rs:74-131 (LocalSolverConfig::default)
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->kind = WX_PRECOND_DEFAULT;
    cfg->schur = WX_SCHUR_APPROXIMATE;
    cfg->approx_chol_seed = 0;
    cfg->split_merge = 2;
    cfg->schur_seed = 0;
    cfg->schur_split = 1;
    cfg->dense_threshold = 24;
    cfg->dense_policy = WX_DENSE_DEVICE;
    cfg->scaling_tolerance = 1e-9;
    cfg->scaling_max_sweeps = 2048;
    cfg->scaling_error_on_failure = 0;
    cfg->reduction = 0;
}

void wx_lsmr_options_default(wx_lsmr_options *o) {
    // crates/within/src/config.rs:170-177
    o->tol = 1e-8;
    o->maxiter = 1000;
    o->local_size = 0;
    o->precision = WX_PRECISION_FP64;
    o->check_every = 0;
    o->timing = 0;
    o->true_residual = 1;
}

uint64_t wx_last_timing(double *out, uint64_t n) {
    const auto &t = last_timing();
    const uint64_t m = std::min<uint64_t>(n / 2, t.size() / 2);
    for (uint64_t i = 0; i < 2 * m; ++i) out[i] = t[i];
    return m;
}

wx_status wx_design_from_codes(const uint32_t *codes, uint64_t n_obs, uint64_t n_factors, int64_t rs, int64_t cs,
                               wx_design **out) {
    return guard([&] {
        if (!out) fail(WX_ERR_INVALID_INPUT, "null output pointer");
        auto *d = new wx_design();
        d->n_obs = n_obs;
        d->columns.assign(n_factors, std::vector<uint32_t>(n_obs));
        for (uint64_t f = 0; f < n_factors; ++f) {
            uint32_t *dst = d->columns[f].data();
            if (rs == 1) std::memcpy(dst, codes + (int64_t)f * cs, n_obs * 4);
            else
                for (uint64_t i = 0; i < n_obs; ++i) dst[i] = codes[(int64_t)i * rs + (int64_t)f * cs];
        }
        d->structure.assign(n_factors, std::vector<int64_t>{-1});
        *out = d;
    });
}

wx_status wx_design_from_effects(uint64_t n_obs, uint64_t n_effects, const uint32_t *const *levels,
                                 const uint8_t *intercept, const uint64_t *n_slopes, const double *const *slopes,
                                 wx_design **out) {
    return guard([&] {
        if (!out) fail(WX_ERR_INVALID_INPUT, "null output pointer");
        auto d = std::make_unique<wx_design>();
        d->n_obs = n_obs;
        uint64_t sp = 0;
        for (uint64_t e = 0; e < n_effects; ++e) {
            // Effect::new (crates/within/src/domain/effect.rs:13-42)
            std::vector<int64_t> cols;
            if (intercept[e]) cols.push_back(-1);
            for (uint64_t s = 0; s < n_slopes[e]; ++s) {
                const double *z = slopes[sp + s];
                for (uint64_t i = 0; i < n_obs; ++i)
                    if (!std::isfinite(z[i]))
                        fail(WX_ERR_BUILD, "slope " + std::to_string(s) + " loading at index " + std::to_string(i) +
                                               " must be finite, got " + std::to_string(z[i]),
                             "BuildError::InvalidLoading");
                cols.push_back((int64_t)d->slopes.size());
                d->slopes.emplace_back(z, z + n_obs);
            }
            sp += n_slopes[e];
            if (cols.empty())
                fail(WX_ERR_BUILD, "an effect must have an intercept or at least one slope", "BuildError::EmptyEffect");
            d->columns.emplace_back(levels[e], levels[e] + n_obs);
            d->structure.push_back(std::move(cols));
        }
        *out = d.release();
    });
}

void wx_design_release(wx_design *d) {
    if (d && --d->refs == 0) delete d;
}

namespace {
// WITHIN_AMD_VERBOSE: cumulative wall time of a Solver's set-up phases on stderr
void setup_phase(const char *name, std::chrono::steady_clock::time_point t0) {
    static const bool verbose = std::getenv("WITHIN_AMD_VERBOSE") != nullptr;
    if (!verbose) return;
    std::fprintf(stderr, "[within_amd] solver set-up: %s done at %.1f ms\n", name,
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
}
} // namespace

wx_status wx_solver_new(const wx_design *design, const double *weights, const wx_precond_config *cfg,
                        const wx_precond *prebuilt, wx_solver **out) {
    return guard([&] {
        if (!design || !out) fail(WX_ERR_INVALID_INPUT, "null design or output pointer");
        if (cfg && prebuilt) fail(WX_ERR_INVALID_INPUT, "pass either a preconditioner config or a prebuilt one");
        require_device();
        auto t0 = std::chrono::steady_clock::now();
        auto h = std::make_unique<wx_solver>();
        Solver &s = h->s;
        // crates/within/src/solver.rs:361-411
        s.design = build_design(design->n_obs, design->columns, design->structure, design->slopes, true);
        setup_phase("design", t0);
        std::vector<double> w;
        if (weights) {
            validate_weights(s.design, weights, s.design.n_obs);
            w.resize(s.design.n_obs);
            permute_in(s.design, weights, w.data());
        }
        s.weighted = weights != nullptr;
        if (weights) s.min_weight = *std::min_element(w.begin(), w.end());
        if (s.design.has_slopes()) {
            // the per-level moments and whitening run on the device (WITHIN_AMD_HOST_SETUP=1: host)
            static const bool host_setup = std::getenv("WITHIN_AMD_HOST_SETUP") != nullptr;
            s.reparam = host_setup ? build_reparam(s.design, weights ? w.data() : nullptr)
                                   : build_reparam_device(s.design, weights ? w.data() : nullptr);
            setup_phase("slope reparametrisation", t0);
        }
        if (prebuilt) {
            if (prebuilt->p->n != s.design.n_dofs)
                fail(WX_ERR_BUILD, "preconditioner dimension mismatch: expected " + std::to_string(s.design.n_dofs) +
                                       ", got " + std::to_string(prebuilt->p->n) + "x" +
                                       std::to_string(prebuilt->p->n),
                     "BuildError::PreconditionerDimensionMismatch");
            s.precond = prebuilt->p;
        } else {
            wx_precond_config c;
            if (cfg) c = *cfg;
            else wx_precond_config_default(&c);
            if (c.kind == WX_PRECOND_OFF) s.precond = nullptr;
            else if (c.kind == WX_PRECOND_DIAGONAL) s.precond = build_diagonal(s.design, w);
            else s.precond = build_additive(s.design, w, settings_from(c), s.warnings);
        }
        setup_phase("preconditioner", t0);
        upload_design(s, w);
        setup_phase("device design", t0);
        s.uid = next_solver_uid();
        s.w_internal = std::move(w);
        s.time_setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = h.release();
    });
}

wx_status wx_solver_reweight(const wx_solver *base, const double *weights, const wx_precond_config *cfg,
                             const wx_precond *prebuilt, wx_solver **out) {
    return guard([&] {
        if (!base || !out) fail(WX_ERR_INVALID_INPUT, "null solver or output pointer");
        if (cfg && prebuilt) fail(WX_ERR_INVALID_INPUT, "pass either a preconditioner config or a prebuilt one");
        const Solver &b = base->s;
        if (b.design.has_slopes())
            fail(WX_ERR_INVALID_INPUT, "reweighting keeps the design's whitened slope columns, which depend on the "
                                       "weights: build a new Solver for a design with slopes");
        require_device();
        auto t0 = std::chrono::steady_clock::now();
        auto h = std::make_unique<wx_solver>();
        Solver &s = h->s;
        s.design = b.design;
        std::vector<double> w;
        if (weights) {
            validate_weights(s.design, weights, s.design.n_obs);
            w.resize(s.design.n_obs);
            permute_in(s.design, weights, w.data());
        }
        s.weighted = weights != nullptr;
        if (weights) s.min_weight = *std::min_element(w.begin(), w.end());
        if (prebuilt) {
            if (prebuilt->p->n != s.design.n_dofs)
                fail(WX_ERR_BUILD, "preconditioner dimension mismatch: expected " + std::to_string(s.design.n_dofs) +
                                       ", got " + std::to_string(prebuilt->p->n),
                     "BuildError::PreconditionerDimensionMismatch");
            s.precond = prebuilt->p;
        } else {
            wx_precond_config c;
            if (cfg) c = *cfg;
            else wx_precond_config_default(&c);
            if (c.kind == WX_PRECOND_OFF) s.precond = nullptr;
            else if (c.kind == WX_PRECOND_DIAGONAL) s.precond = build_diagonal(s.design, w);
            else s.precond = build_additive(s.design, w, settings_from(c), s.warnings);
        }
        reweight_design(s, b, w);
        s.uid = next_solver_uid();
        s.w_internal = std::move(w);
        s.time_setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = h.release();
    });
}

wx_status wx_gram_device(const double *A, uint64_t n, uint64_t m, uint64_t lda, double *G, void *stream) {
    return guard([&] {
        if (!A || !G) fail(WX_ERR_INVALID_INPUT, "null matrix pointer");
        if (m == 0 || m > gram_max_cols())
            fail(WX_ERR_INVALID_INPUT, "the Gram matrix takes 1 to " + std::to_string(gram_max_cols()) + " columns");
        if (lda < n) fail(WX_ERR_INVALID_INPUT, "leading dimension smaller than the number of rows");
        require_device();
        hipStream_t st = (hipStream_t)stream;
        // per-thread partial-sum scratch, grown on demand and kept until the process ends (the
        // call synchronises, so a thread's calls never overlap on it)
        static thread_local DevBuf *part = new DevBuf();
        static thread_local int part_dev = -1;
        const uint64_t bytes = gram_scratch_doubles(n, (uint32_t)m) * 8;
        int dev = 0;
        WX_HIP(hipGetDevice(&dev));
        if (part->bytes < bytes || part_dev != dev) {
            if (part_dev != dev) part->p = nullptr; // another device's buffer: left to the process end
            part->alloc(bytes);
            part_dev = dev;
        }
        gram_device(A, n, (uint32_t)m, lda, G, part->as<double>(), st);
        WX_HIP(hipStreamSynchronize(st));
    });
}

wx_status wx_gram_device_noalloc(const double *A, uint64_t n, uint64_t m, uint64_t lda, double *G, void *stream) {
    return guard([&] {
        if (!A || !G) fail(WX_ERR_INVALID_INPUT, "null matrix pointer");
        if (m == 0 || m > gram_noalloc_max_cols())
            fail(WX_ERR_INVALID_INPUT, "the allocation-free Gram matrix takes 1 to " +
                                           std::to_string(gram_noalloc_max_cols()) + " columns");
        if (lda < n) fail(WX_ERR_INVALID_INPUT, "leading dimension smaller than the number of rows");
        require_device();
        std::string err;
        double *scratch = gram_scratch_ptr(err);
        if (!scratch) fail(WX_ERR_DEVICE, err);
        hipStream_t st = (hipStream_t)stream;
        // the same kernels and grid as wx_gram_device, with the partial triangles in this
        // device's static scratch; every partial the reduction reads is written by this call
        gram_device(A, n, (uint32_t)m, lda, G, scratch, st);
        WX_HIP(hipStreamSynchronize(st));
    });
}

wx_status wx_gram(const double *A, uint64_t n, uint64_t m, uint64_t lda, double *G) {
    return guard([&] {
        if (!A || !G) fail(WX_ERR_INVALID_INPUT, "null matrix pointer");
        if (m == 0 || m > gram_max_cols())
            fail(WX_ERR_INVALID_INPUT, "the Gram matrix takes 1 to " + std::to_string(gram_max_cols()) + " columns");
        if (lda < n) fail(WX_ERR_INVALID_INPUT, "leading dimension smaller than the number of rows");
        require_device();
        DevBuf a, g, part;
        a.alloc(std::max<uint64_t>(n, 1) * m * 8);
        if (n) WX_HIP(hipMemcpy2D(a.p, n * 8, A, lda * 8, n * 8, m, hipMemcpyHostToDevice));
        g.alloc(m * m * 8);
        part.alloc(gram_scratch_doubles(n, (uint32_t)m) * 8);
        gram_device(a.as<double>(), n, (uint32_t)m, n, g.as<double>(), part.as<double>(), nullptr);
        WX_HIP(hipMemcpy(G, g.p, m * m * 8, hipMemcpyDeviceToHost));
    });
}

wx_status wx_testing_design_index(const wx_design *design, uint64_t n_terms, uint64_t *n_levels, uint64_t *offset,
                                  uint8_t *sorted, int64_t *dominant, int32_t *permuted, uint32_t *obs_perm) {
    return guard([&] {
        if (!design) fail(WX_ERR_INVALID_INPUT, "null design");
        if (n_terms != design->columns.size())
            fail(WX_ERR_INVALID_INPUT, "n_terms disagrees with the design's term count");
        const Design d = build_design(design->n_obs, design->columns, design->structure, design->slopes, true);
        int64_t dom = -1;
        uint64_t best = 0;
        for (uint64_t t = 0; t < d.terms.size(); ++t) {
            if (n_levels) n_levels[t] = d.terms[t].n_levels;
            if (offset) offset[t] = d.terms[t].offset;
            if (sorted) sorted[t] = d.terms[t].sorted ? 1 : 0;
            if (dom < 0 || d.terms[t].n_dofs() >= best) {
                dom = (int64_t)t;
                best = d.terms[t].n_dofs();
            }
        }
        if (dominant) *dominant = dom;
        if (permuted) *permuted = d.obs_perm.empty() ? 0 : 1;
        if (obs_perm && !d.obs_perm.empty()) std::copy(d.obs_perm.begin(), d.obs_perm.end(), obs_perm);
    });
}

wx_status wx_testing_reparam(const wx_design *design, const double *weights, int32_t on_device, double *slopes_out,
                             double *transform_out, uint64_t transform_cap, double *mean_out, uint64_t mean_cap,
                             uint64_t *unidentified_out, uint64_t unidentified_cap, uint64_t *n_unidentified) {
    return guard([&] {
        if (!design) fail(WX_ERR_INVALID_INPUT, "null design");
        Design d = build_design(design->n_obs, design->columns, design->structure, design->slopes, true);
        std::vector<double> w;
        if (weights) {
            validate_weights(d, weights, d.n_obs);
            w.resize(d.n_obs);
            permute_in(d, weights, w.data());
        }
        if (on_device) require_device();
        std::unique_ptr<SlopeReparam> rp;
        if (d.has_slopes())
            rp = on_device ? build_reparam_device(d, weights ? w.data() : nullptr)
                           : build_reparam(d, weights ? w.data() : nullptr);
        const uint64_t n = d.n_obs;
        if (slopes_out) // caller row order
            for (uint64_t c = 0; c < d.slopes.size() / std::max<uint64_t>(n, 1); ++c)
                for (uint64_t i = 0; i < n; ++i)
                    slopes_out[c * n + (d.obs_perm.empty() ? i : d.obs_perm[i])] = d.slopes[c * n + i];
        uint64_t nt = 0, nm = 0;
        if (rp)
            for (const auto &blk : rp->blocks) {
                if (transform_out) {
                    if (nt + blk.transform.size() > transform_cap) fail(WX_ERR_INVALID_INPUT, "transform buffer too small");
                    std::copy(blk.transform.begin(), blk.transform.end(), transform_out + nt);
                }
                if (mean_out) {
                    if (nm + blk.mean.size() > mean_cap) fail(WX_ERR_INVALID_INPUT, "mean buffer too small");
                    std::copy(blk.mean.begin(), blk.mean.end(), mean_out + nm);
                }
                nt += blk.transform.size();
                nm += blk.mean.size();
            }
        const uint64_t nu = rp ? rp->unidentified.size() / 3 : 0;
        if (n_unidentified) *n_unidentified = nu;
        if (unidentified_out) {
            if (nu * 3 > unidentified_cap) fail(WX_ERR_INVALID_INPUT, "unidentified buffer too small");
            std::copy(rp->unidentified.begin(), rp->unidentified.end(), unidentified_out);
        }
    });
}

wx_status wx_testing_block_precond(uint32_t n_rows, uint32_t n_cols, const double *B, const double *diag,
                                   const wx_precond_config *cfg, wx_precond **out) {
    return guard([&] {
        if (!out || !diag || (!B && n_rows && n_cols)) fail(WX_ERR_INVALID_INPUT, "null pointer");
        if (n_rows + n_cols == 0) fail(WX_ERR_INVALID_INPUT, "empty block");
        wx_precond_config c;
        if (cfg) c = *cfg;
        else wx_precond_config_default(&c);
        auto h = std::make_unique<wx_precond>();
        h->p = build_block_precond(n_rows, n_cols, B, diag, settings_from(c));
        *out = h.release();
    });
}

wx_status wx_testing_reduced_laplacian(uint32_t n_rows, uint32_t n_cols, const double *B, const double *diag,
                                       const wx_precond_config *cfg, double *out, uint64_t cap, uint32_t *n_out) {
    return guard([&] {
        if (!out || !n_out || !diag || (!B && n_rows && n_cols)) fail(WX_ERR_INVALID_INPUT, "null pointer");
        wx_precond_config c;
        if (cfg) c = *cfg;
        else wx_precond_config_default(&c);
        uint32_t n = 0;
        const std::vector<double> L = block_reduced_laplacian(n_rows, n_cols, B, diag, settings_from(c), n);
        if ((uint64_t)n * n > cap) fail(WX_ERR_INVALID_INPUT, "output buffer too small");
        std::copy(L.begin(), L.end(), out);
        *n_out = n;
    });
}

wx_status wx_precond_build(const wx_design *design, const double *weights, const wx_precond_config *cfg,
                           wx_precond **out) {
    return guard([&] {
        if (!design || !out) fail(WX_ERR_INVALID_INPUT, "null design or output pointer");
        Design d = build_design(design->n_obs, design->columns, design->structure, design->slopes, true);
        std::vector<double> w;
        if (weights) {
            validate_weights(d, weights, d.n_obs);
            w.resize(d.n_obs);
            permute_in(d, weights, w.data());
        }
        std::unique_ptr<SlopeReparam> rp;
        if (d.has_slopes()) rp = build_reparam(d, weights ? w.data() : nullptr);
        wx_precond_config c;
        if (cfg) c = *cfg;
        else wx_precond_config_default(&c);
        std::shared_ptr<Precond> p;
        std::vector<std::string> warnings;
        if (c.kind == WX_PRECOND_DIAGONAL) p = build_diagonal(d, w);
        else if (c.kind != WX_PRECOND_OFF) p = build_additive(d, w, settings_from(c), warnings);
        if (!p) {
            *out = nullptr;
            return;
        }
        auto *h = new wx_precond();
        h->p = p;
        *out = h;
    });
}

void wx_solver_release(wx_solver *s) {
    if (s && --s->refs == 0) delete s;
}

uint64_t wx_solver_n_obs(const wx_solver *s) { return s ? s->s.design.n_obs : 0; }
uint64_t wx_solver_n_dofs(const wx_solver *s) { return s ? s->s.design.n_dofs : 0; }
double wx_solver_time_setup(const wx_solver *s) { return s ? s->s.time_setup : 0.0; }

wx_status wx_solver_solve_batch_device(const wx_solver *s, const double *Y, uint64_t ldy, uint64_t k,
                                       const wx_lsmr_options *opts, double *X, double *Dm, wx_rhs_stats *stats,
                                       void *stream) {
    return guard([&] {
        if (!s || !stats) fail(WX_ERR_INVALID_INPUT, "null solver or stats pointer");
        wx_lsmr_options o;
        if (opts) o = *opts;
        else wx_lsmr_options_default(&o);
        Stream st(stream);
        solve_batch_device(s->s, Y, ldy, k, o, X, Dm, stats, st.s);
        WX_HIP(hipStreamSynchronize(st.s));
    });
}

wx_status wx_solver_projection_grad_device(const wx_solver *s, const double *G, uint64_t ldg, const double *R,
                                           uint64_t ldr, uint64_t k, const wx_lsmr_options *opts, double *grad_z,
                                           double *grad_w, wx_rhs_stats *stats, void *stream) {
    return guard([&] {
        if (!s || !stats || (k && (!G || !R || !grad_z))) fail(WX_ERR_INVALID_INPUT, "null solver, array or stats pointer");
        wx_lsmr_options o;
        if (opts) o = *opts;
        else wx_lsmr_options_default(&o);
        Stream st(stream);
        projection_grad_device(s->s, G, ldg, R, ldr, k, o, grad_z, grad_w, stats, st.s);
    });
}

wx_status wx_solver_solve_batch(const wx_solver *s, const double *Y, uint64_t ldy, uint64_t k,
                                const wx_lsmr_options *opts, double *X, double *demeaned, wx_rhs_stats *stats) {
    return guard([&] {
        if (!s || !stats || (!Y && k)) fail(WX_ERR_INVALID_INPUT, "null solver, response or stats pointer");
        const uint64_t n = s->s.design.n_obs, N = s->s.design.n_dofs;
        if (ldy < n) fail(WX_ERR_INVALID_INPUT, "leading dimension of Y is smaller than n_obs");
        // validate the response on the host (solver.rs:424-442)
        for (uint64_t j = 0; j < k; ++j)
            for (uint64_t i = 0; i < n; ++i) {
                const double v = Y[j * ldy + i];
                if (!std::isfinite(v))
                    fail(WX_ERR_INVALID_INPUT, "Solver::solve: response at index " + std::to_string(i) +
                                                   " must be finite, got " + std::to_string(v));
            }
        if (k == 0) return;
        wx_lsmr_options o;
        if (opts) o = *opts;
        else wx_lsmr_options_default(&o);
        Stream st(nullptr);
        DevBuf dY, dX, dD;
        dY.alloc(n * k * 8);
        for (uint64_t j = 0; j < k; ++j)
            WX_HIP(hipMemcpyAsync(dY.as<double>() + j * n, Y + j * ldy, n * 8, hipMemcpyHostToDevice, st.s));
        if (X) dX.alloc(N * k * 8);
        if (demeaned) dD.alloc(n * k * 8);
        solve_batch_device(s->s, dY.as<double>(), n, k, o, X ? dX.as<double>() : nullptr,
                           demeaned ? dD.as<double>() : nullptr, stats, st.s);
        if (X) WX_HIP(hipMemcpyAsync(X, dX.p, N * k * 8, hipMemcpyDeviceToHost, st.s));
        if (demeaned) WX_HIP(hipMemcpyAsync(demeaned, dD.p, n * k * 8, hipMemcpyDeviceToHost, st.s));
        WX_HIP(hipStreamSynchronize(st.s));
    });
}

wx_status wx_solve_many_device(const wx_solver *const *solvers, uint64_t n_problems, const double *const *Y,
                               const uint64_t *ldy, const uint64_t *k, const wx_lsmr_options *opts, double *const *X,
                               double *const *demeaned, wx_rhs_stats *const *stats, void *stream) {
    return guard([&] {
        if (n_problems && (!solvers || !Y || !ldy || !k || !stats)) fail(WX_ERR_INVALID_INPUT, "null table pointer");
        std::vector<const Solver *> sv(n_problems);
        std::vector<const double *> yv(n_problems);
        std::vector<uint64_t> lv(n_problems), kv(n_problems);
        std::vector<double *> xv(n_problems), dv(n_problems);
        std::vector<wx_rhs_stats *> stv(n_problems);
        for (uint64_t p = 0; p < n_problems; ++p) {
            if (!solvers[p] || (!Y[p] && k[p]) || !stats[p])
                fail(WX_ERR_INVALID_INPUT, "null solver, response or stats pointer of problem " + std::to_string(p));
            sv[p] = &solvers[p]->s;
            yv[p] = Y[p];
            lv[p] = ldy[p];
            kv[p] = k[p];
            xv[p] = X ? X[p] : nullptr;
            dv[p] = demeaned ? demeaned[p] : nullptr;
            stv[p] = stats[p];
        }
        wx_lsmr_options o;
        if (opts) o = *opts;
        else wx_lsmr_options_default(&o);
        Stream st(stream);
        solve_many_device(sv, yv, lv, kv, o, xv, dv, stv, st.s);
    });
}

wx_status wx_solve_many(const wx_solver *const *solvers, uint64_t n_problems, const double *const *Y,
                        const uint64_t *ldy, const uint64_t *k, const wx_lsmr_options *opts, double *const *X,
                        double *const *demeaned, wx_rhs_stats *const *stats) {
    return guard([&] {
        if (n_problems && (!solvers || !Y || !ldy || !k || !stats)) fail(WX_ERR_INVALID_INPUT, "null table pointer");
        std::vector<DevBuf> dY(n_problems), dX(n_problems), dD(n_problems);
        std::vector<const wx_solver *> sv(n_problems);
        std::vector<const double *> yv(n_problems);
        std::vector<uint64_t> lv(n_problems);
        std::vector<double *> xv(n_problems, nullptr), dv(n_problems, nullptr);
        Stream st(nullptr);
        for (uint64_t p = 0; p < n_problems; ++p) {
            const wx_solver *s = solvers[p];
            if (!s || (!Y[p] && k[p]) || !stats[p])
                fail(WX_ERR_INVALID_INPUT, "null solver, response or stats pointer of problem " + std::to_string(p));
            const uint64_t n = s->s.design.n_obs, N = s->s.design.n_dofs;
            if (ldy[p] < n) fail(WX_ERR_INVALID_INPUT, "leading dimension of Y is smaller than n_obs");
            for (uint64_t j = 0; j < k[p]; ++j)
                for (uint64_t i = 0; i < n; ++i)
                    if (!std::isfinite(Y[p][j * ldy[p] + i]))
                        fail(WX_ERR_INVALID_INPUT, "Solver::solve: response at index " + std::to_string(i) +
                                                       " must be finite, got " + std::to_string(Y[p][j * ldy[p] + i]));
            dY[p].alloc(n * std::max<uint64_t>(k[p], 1) * 8);
            for (uint64_t j = 0; j < k[p]; ++j)
                WX_HIP(hipMemcpyAsync(dY[p].as<double>() + j * n, Y[p] + j * ldy[p], n * 8, hipMemcpyHostToDevice, st.s));
            if (X && X[p]) dX[p].alloc(N * std::max<uint64_t>(k[p], 1) * 8), xv[p] = dX[p].as<double>();
            if (demeaned && demeaned[p]) dD[p].alloc(n * std::max<uint64_t>(k[p], 1) * 8), dv[p] = dD[p].as<double>();
            sv[p] = s;
            yv[p] = dY[p].as<double>();
            lv[p] = n;
        }
        const wx_status rc = wx_solve_many_device(sv.data(), n_problems, yv.data(), lv.data(), k, opts, xv.data(),
                                                  dv.data(), stats, st.s);
        if (rc != WX_OK) fail(rc, std::string(wx_last_error()), g_last_variant);
        for (uint64_t p = 0; p < n_problems; ++p) {
            const uint64_t n = solvers[p]->s.design.n_obs, N = solvers[p]->s.design.n_dofs;
            if (xv[p]) WX_HIP(hipMemcpyAsync(X[p], xv[p], N * k[p] * 8, hipMemcpyDeviceToHost, st.s));
            if (dv[p]) WX_HIP(hipMemcpyAsync(demeaned[p], dv[p], n * k[p] * 8, hipMemcpyDeviceToHost, st.s));
        }
        WX_HIP(hipStreamSynchronize(st.s));
    });
}

void wx_set_many_shape(int32_t streams, int32_t iterations_per_graph, int32_t graph, int32_t pack) {
    force_many_shape(streams, iterations_per_graph, graph, pack);
}

void wx_set_many_tile(int32_t rows) { force_many_tile(rows); }

int32_t wx_last_many_tile(void) { return last_many_tile(); }

void wx_last_many_shape(int32_t *streams, int32_t *iterations_per_graph, int32_t *graph, int32_t *pack) {
    last_many_shape(streams, iterations_per_graph, graph, pack);
}

wx_status wx_solver_apply_operator(const wx_solver *s, int32_t adjoint, const double *X, uint64_t k, double *Y) {
    return guard([&] {
        if (!s || (k && (!X || !Y))) fail(WX_ERR_INVALID_INPUT, "null pointer");
        if (k == 0) return;
        const uint64_t n = s->s.design.n_obs, N = s->s.design.n_dofs;
        const uint64_t rin = adjoint ? n : N, rout = adjoint ? N : n;
        Stream st(nullptr);
        DevBuf dx, dy;
        dx.alloc(rin * k * 8);
        dy.alloc(rout * k * 8);
        WX_HIP(hipMemcpyAsync(dx.p, X, rin * k * 8, hipMemcpyHostToDevice, st.s));
        design_apply_device(s->s, adjoint != 0, dx.as<double>(), k, dy.as<double>(), st.s);
        WX_HIP(hipMemcpyAsync(Y, dy.p, rout * k * 8, hipMemcpyDeviceToHost, st.s));
        WX_HIP(hipStreamSynchronize(st.s));
    });
}

wx_status wx_solver_preconditioner(const wx_solver *s, wx_precond **out) {
    return guard([&] {
        if (!s || !out) fail(WX_ERR_INVALID_INPUT, "null pointer");
        if (!s->s.precond) {
            *out = nullptr;
            return;
        }
        auto *p = new wx_precond();
        p->p = s->s.precond;
        *out = p;
    });
}

uint64_t wx_solver_n_terms(const wx_solver *s) { return s ? s->s.design.terms.size() : 0; }

void wx_solver_layout(const wx_solver *s, uint64_t *offset, uint64_t *n_levels, uint64_t *n_columns) {
    if (!s) return;
    const auto &T = s->s.design.terms;
    for (size_t t = 0; t < T.size(); ++t) {
        if (offset) offset[t] = T[t].offset;
        if (n_levels) n_levels[t] = T[t].n_levels;
        if (n_columns) n_columns[t] = T[t].ncols();
    }
}

wx_status wx_solver_unidentified(const wx_solver *s, uint64_t *triples, uint64_t *count) {
    return guard([&] {
        if (!s || !count) fail(WX_ERR_INVALID_INPUT, "null pointer");
        const uint64_t have = s->s.reparam ? s->s.reparam->unidentified.size() / 3 : 0;
        if (triples) {
            const uint64_t m = std::min(have, *count);
            if (m) std::memcpy(triples, s->s.reparam->unidentified.data(), m * 3 * sizeof(uint64_t));
        }
        *count = have;
    });
}

uint64_t wx_solver_n_warnings(const wx_solver *s) { return s ? s->s.warnings.size() : 0; }
const char *wx_solver_warning(const wx_solver *s, uint64_t i) {
    return (s && i < s->s.warnings.size()) ? s->s.warnings[i].c_str() : "";
}

void wx_precond_release(wx_precond *p) {
    if (p && --p->refs == 0) delete p;
}

uint64_t wx_precond_n(const wx_precond *p) { return p ? p->p->n : 0; }

const char *wx_precond_variant(const wx_precond *p) {
    if (!p) return "";
    return p->p->kind == Precond::ADDITIVE ? "Additive" : "Diagonal";
}

namespace {
wx_status precond_apply_host(const wx_precond *p, const double *x, uint64_t k, double *y, int32_t precision) {
    return guard([&] {
        if (!p || (!x && k) || (!y && k)) fail(WX_ERR_INVALID_INPUT, "null pointer");
        if (precision != WX_PRECISION_FP64 && precision != WX_PRECISION_MIXED)
            fail(WX_ERR_INVALID_INPUT, "precision must be WX_PRECISION_FP64 or WX_PRECISION_MIXED");
        require_device();
        const uint64_t n = p->p->n;
        if (k == 0) return;
        Stream st(nullptr);
        DevBuf dx, dy;
        dx.alloc(n * k * 8);
        dy.alloc(n * k * 8);
        WX_HIP(hipMemcpyAsync(dx.p, x, n * k * 8, hipMemcpyHostToDevice, st.s));
        precond_apply_device(*p->p, dx.as<double>(), k, dy.as<double>(), st.s, precision);
        WX_HIP(hipMemcpyAsync(y, dy.p, n * k * 8, hipMemcpyDeviceToHost, st.s));
        WX_HIP(hipStreamSynchronize(st.s));
    });
}
} // namespace

wx_status wx_precond_apply(const wx_precond *p, const double *x, uint64_t k, double *y) {
    return precond_apply_host(p, x, k, y, WX_PRECISION_FP64);
}

wx_status wx_testing_precond_apply(const wx_precond *p, const double *x, uint64_t k, double *y, int32_t precision) {
    return precond_apply_host(p, x, k, y, precision);
}

int32_t wx_testing_set_fused_steps(int32_t on) { return set_fused_steps(on != 0) ? 1 : 0; }

wx_status wx_precond_serialize(const wx_precond *p, uint8_t **buf, uint64_t *len) {
    return guard([&] {
        if (!p || !buf || !len) fail(WX_ERR_INVALID_INPUT, "null pointer");
        std::vector<uint8_t> v = serialize_precond(*p->p);
        *buf = (uint8_t *)std::malloc(v.size() ? v.size() : 1);
        if (!*buf) throw std::bad_alloc();
        if (!v.empty()) std::memcpy(*buf, v.data(), v.size());
        *len = v.size();
    });
}

wx_status wx_precond_deserialize(const uint8_t *buf, uint64_t len, wx_precond **out) {
    return guard([&] {
        if (!buf || !out) fail(WX_ERR_INVALID_INPUT, "null pointer");
        auto *p = new wx_precond();
        try {
            p->p = deserialize_precond(buf, len);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

void wx_buffer_free(uint8_t *buf) { std::free(buf); }

} // extern "C"
