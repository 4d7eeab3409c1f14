// setup_gpu.hip -- device-side steps of the preconditioner set-up.
//
// Cross-tabs of plain factor pairs (crates/within/src/domain/cross_tab/accumulate.rs:86-273): for a
// pair of intercept terms (r, c) the cross-tab cell (j, l) is the weight sum of the rows with level j
// of r and level l of c.  On the device every row gets the key j * L_c + l; a stable radix sort of
// (key, weight) keeps the rows of each cell in observation order, and one thread per cell adds that
// cell's weights in that order -- the order of the host accumulation, so every cell equals the host's
// bitwise and the preconditioner built from it is the same whether or not a GPU built the cross-tab.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "device.hpp"
#include "schwarz.hpp"

namespace wx {

namespace {
constexpr int SB = 256;

__global__ __launch_bounds__(SB) void su_keys(uint64_t n, const uint32_t *__restrict__ cr, const uint32_t *__restrict__ cc,
                                              uint64_t lc, const double *__restrict__ w, uint64_t *__restrict__ key,
                                              double *__restrict__ val) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i >= n) return;
    key[i] = (uint64_t)cr[i] * lc + cc[i];
    val[i] = w ? w[i] : 1.0;
}

__global__ __launch_bounds__(SB) void su_heads(uint64_t n, const uint64_t *__restrict__ key, uint8_t *__restrict__ head) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// one thread per cell: its weights added in row (observation) order
__global__ __launch_bounds__(SB) void su_cell_sums(uint64_t n, uint64_t n_cells, const uint64_t *__restrict__ start,
                                                   const uint64_t *__restrict__ key, const double *__restrict__ val,
                                                   uint64_t *__restrict__ cell_key, double *__restrict__ cell_val) {
    const uint64_t q = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (q >= n_cells) return;
    const uint64_t b = start[q], e = q + 1 < n_cells ? start[q + 1] : n;
    double s = 0.0;
    for (uint64_t i = b; i < e; ++i) s += val[i];
    cell_key[q] = key[b];
    cell_val[q] = s;
}

uint32_t bits_for(uint64_t v) {
    uint32_t b = 1;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}
} // namespace

bool device_cross_tabs(const Design &d, const std::vector<double> &w, const std::vector<std::pair<uint32_t, uint32_t>> &pairs,
                       std::vector<CellList> &out) {
    out.assign(pairs.size(), CellList{});
    // WITHIN_AMD_HOST_SETUP=1 keeps the accumulation on the host (the same cells, bit for bit)
    const char *host_only = std::getenv("WITHIN_AMD_HOST_SETUP");
    if (host_only && host_only[0] == '1') return false;
    if (pairs.empty() || d.n_obs == 0 || !gpu_available()) return false;
    const uint64_t n = d.n_obs;
    hipStream_t st = nullptr;
    std::vector<DevBuf> col(d.terms.size());
    std::vector<uint8_t> need(d.terms.size(), 0);
    for (auto &pr : pairs) need[pr.first] = need[pr.second] = 1;
    for (size_t t = 0; t < d.terms.size(); ++t)
        if (need[t]) col[t].upload(d.level_column(t), n, st);
    DevBuf dw;
    if (!w.empty()) dw.upload(w.data(), n, st);
    DevBuf key, val, key2, val2, head, idx, start, nsel, ckey, cval, tmp;
    key.alloc(n * 8);
    val.alloc(n * 8);
    key2.alloc(n * 8);
    val2.alloc(n * 8);
    head.alloc(n);
    start.alloc(n * 8);
    nsel.alloc(8);
    const uint32_t grid = (uint32_t)((n + SB - 1) / SB);
    hipcub::CountingInputIterator<uint64_t> counting(0);
    for (size_t p = 0; p < pairs.size(); ++p) {
        const uint32_t r = pairs[p].first, c = pairs[p].second;
        const uint64_t lr = d.terms[r].n_levels, lc = d.terms[c].n_levels;
        hipLaunchKernelGGL(su_keys, dim3(grid), dim3(SB), 0, st, n, col[r].as<uint32_t>(), col[c].as<uint32_t>(), lc,
                           w.empty() ? nullptr : dw.as<double>(), key.as<uint64_t>(), val.as<double>());
        const uint32_t end_bit = bits_for(lr * lc - 1);
        size_t tb = 0;
        WX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.as<uint64_t>(), key2.as<uint64_t>(),
                                                  val.as<double>(), val2.as<double>(), (int)n, 0, (int)end_bit, st));
        size_t tb2 = 0;
        WX_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb2, counting, head.as<uint8_t>(), start.as<uint64_t>(),
                                             nsel.as<uint64_t>(), (int)n, st));
        tb = std::max(tb, tb2);
        if (tmp.bytes < tb) tmp.alloc(tb);
        WX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.as<uint64_t>(), key2.as<uint64_t>(), val.as<double>(),
                                                  val2.as<double>(), (int)n, 0, (int)end_bit, st));
        hipLaunchKernelGGL(su_heads, dim3(grid), dim3(SB), 0, st, n, key2.as<uint64_t>(), head.as<uint8_t>());
        WX_HIP(hipcub::DeviceSelect::Flagged(tmp.p, tb, counting, head.as<uint8_t>(), start.as<uint64_t>(),
                                             nsel.as<uint64_t>(), (int)n, st));
        uint64_t n_cells = 0;
        WX_HIP(hipMemcpyAsync(&n_cells, nsel.p, 8, hipMemcpyDeviceToHost, st));
        WX_HIP(hipStreamSynchronize(st));
        ckey.alloc(n_cells * 8);
        cval.alloc(n_cells * 8);
        hipLaunchKernelGGL(su_cell_sums, dim3((uint32_t)((n_cells + SB - 1) / SB)), dim3(SB), 0, st, n, n_cells,
                           start.as<uint64_t>(), key2.as<uint64_t>(), val2.as<double>(), ckey.as<uint64_t>(),
                           cval.as<double>());
        WX_HIP(hipGetLastError());
        CellList &cl = out[p];
        cl.key.resize(n_cells);
        cl.val.resize(n_cells);
        WX_HIP(hipMemcpyAsync(cl.key.data(), ckey.p, n_cells * 8, hipMemcpyDeviceToHost, st));
        WX_HIP(hipMemcpyAsync(cl.val.data(), cval.p, n_cells * 8, hipMemcpyDeviceToHost, st));
        WX_HIP(hipStreamSynchronize(st));
    }
    return true;
}

} // namespace wx

// ---------------------------------------------------------------------------
// Index structures of the device design, built where the design's codes already are
// (upload_design in lsmr.hip): stable argsorts by a term's codes (a radix sort of
// (code, row) pairs is stable), position-ordered weights, the inverse row permutation and the
// pieces of a fused pair.  The results equal the host counting sorts' exactly.
// ---------------------------------------------------------------------------
namespace wx {
namespace {
__global__ __launch_bounds__(SB) void su_iota(uint32_t *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}
__global__ __launch_bounds__(SB) void su_key2(uint64_t n, const uint32_t *__restrict__ kp, const uint32_t *__restrict__ kq,
                                              uint64_t lq, uint64_t *__restrict__ key) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i < n) key[i] = (uint64_t)kp[i] * lq + kq[i];
}
__global__ __launch_bounds__(SB) void su_gather_f64(uint64_t n, const double *__restrict__ src,
                                                    const uint32_t *__restrict__ perm, double *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i < n) out[i] = src[perm[i]];
}
__global__ __launch_bounds__(SB) void su_gather_u32(uint64_t n, const uint32_t *__restrict__ src,
                                                    const uint32_t *__restrict__ perm, uint32_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i < n) out[i] = src[perm[i]];
}
__global__ __launch_bounds__(SB) void su_sqrt(uint64_t n, const double *__restrict__ src, double *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i < n) out[i] = sqrt(src[i]);
}
__global__ __launch_bounds__(SB) void su_invert(uint64_t n, const uint32_t *__restrict__ perm, uint32_t *__restrict__ inv) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i < n) inv[perm[i]] = (uint32_t)i;
}
// pieces: a position starts a piece when it starts a non-empty work item or its Q code differs
// from the previous position's
__global__ __launch_bounds__(SB) void su_run_flags(uint64_t n, const uint32_t *__restrict__ q, uint32_t *__restrict__ f) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i < n) f[i] = (i > 0 && q[i] != q[i - 1]) ? 1u : 0u;
}
__global__ __launch_bounds__(SB) void su_item_flags(uint32_t n_items, const SegItem *__restrict__ items,
                                                    uint32_t *__restrict__ f) {
    const uint32_t it = blockIdx.x * SB + threadIdx.x;
    if (it < n_items && items[it].begin < items[it].end) f[items[it].begin] = 1u;
}
__global__ __launch_bounds__(SB) void su_item_piece0(uint32_t n_items, const SegItem *__restrict__ items,
                                                     const uint32_t *__restrict__ pid, uint32_t *__restrict__ piece0) {
    const uint32_t it = blockIdx.x * SB + threadIdx.x;
    if (it < n_items) piece0[it] = items[it].begin < items[it].end ? pid[items[it].begin] : 0u;
}
__global__ __launch_bounds__(SB) void su_piece_codes(uint64_t n, const uint32_t *__restrict__ f, const uint32_t *__restrict__ pid,
                                                     const uint32_t *__restrict__ q, uint32_t *__restrict__ piece_q) {
    const uint64_t i = (uint64_t)blockIdx.x * SB + threadIdx.x;
    if (i < n && f[i]) piece_q[pid[i]] = q[i];
}
inline uint32_t g1(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + SB - 1) / SB); }
} // namespace

void dev_sqrt(const double *src, uint64_t n, double *out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(su_sqrt, dim3(g1(n)), dim3(SB), 0, st, n, src, out);
}
void dev_gather(const double *src, const uint32_t *perm, uint64_t n, double *out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(su_gather_f64, dim3(g1(n)), dim3(SB), 0, st, n, src, perm, out);
}
void dev_gather_u32(const uint32_t *src, const uint32_t *perm, uint64_t n, uint32_t *out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(su_gather_u32, dim3(g1(n)), dim3(SB), 0, st, n, src, perm, out);
}
void dev_invert_perm(const uint32_t *perm, uint64_t n, uint32_t *inv, hipStream_t st) {
    if (n) hipLaunchKernelGGL(su_invert, dim3(g1(n)), dim3(SB), 0, st, n, perm, inv);
}

void dev_stable_argsort(const uint32_t *key, uint64_t n, uint64_t n_levels, uint32_t *perm, hipStream_t st) {
    if (!n) return;
    DevBuf iota, kout, tmp;
    iota.alloc(n * 4);
    kout.alloc(n * 4);
    hipLaunchKernelGGL(su_iota, dim3(g1(n)), dim3(SB), 0, st, iota.as<uint32_t>(), n);
    const int end_bit = (int)bits_for(n_levels > 0 ? n_levels - 1 : 0);
    size_t tb = 0;
    WX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, kout.as<uint32_t>(), iota.as<uint32_t>(), perm, (int)n,
                                              0, end_bit, st));
    tmp.alloc(tb);
    WX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key, kout.as<uint32_t>(), iota.as<uint32_t>(), perm, (int)n,
                                              0, end_bit, st));
    WX_HIP(hipStreamSynchronize(st));
}

void dev_stable_argsort2(const uint32_t *kp, const uint32_t *kq, uint64_t lp, uint64_t lq, uint64_t n, uint32_t *perm,
                         hipStream_t st) {
    if (!n) return;
    DevBuf key, kout, iota, tmp;
    key.alloc(n * 8);
    kout.alloc(n * 8);
    iota.alloc(n * 4);
    hipLaunchKernelGGL(su_key2, dim3(g1(n)), dim3(SB), 0, st, n, kp, kq, lq, key.as<uint64_t>());
    hipLaunchKernelGGL(su_iota, dim3(g1(n)), dim3(SB), 0, st, iota.as<uint32_t>(), n);
    const int end_bit = (int)bits_for(lp * lq > 0 ? lp * lq - 1 : 0);
    size_t tb = 0;
    WX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.as<uint64_t>(), kout.as<uint64_t>(), iota.as<uint32_t>(),
                                              perm, (int)n, 0, end_bit, st));
    tmp.alloc(tb);
    WX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.as<uint64_t>(), kout.as<uint64_t>(), iota.as<uint32_t>(),
                                              perm, (int)n, 0, end_bit, st));
    WX_HIP(hipStreamSynchronize(st));
}

uint64_t dev_pieces(const uint32_t *qpos, uint64_t n, const SegItem *items, uint32_t n_items, uint32_t *piece0,
                    DevBuf &piece_q, hipStream_t st) {
    if (!n) return 0;
    DevBuf flag, pid, tmp;
    flag.alloc(n * 4);
    pid.alloc(n * 4);
    hipLaunchKernelGGL(su_run_flags, dim3(g1(n)), dim3(SB), 0, st, n, qpos, flag.as<uint32_t>());
    hipLaunchKernelGGL(su_item_flags, dim3(g1(n_items)), dim3(SB), 0, st, n_items, items, flag.as<uint32_t>());
    size_t tb = 0;
    WX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag.as<uint32_t>(), pid.as<uint32_t>(), (int)n, st));
    tmp.alloc(tb);
    WX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flag.as<uint32_t>(), pid.as<uint32_t>(), (int)n, st));
    uint32_t last[2] = {0, 0};
    WX_HIP(hipMemcpyAsync(&last[0], pid.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, st));
    WX_HIP(hipMemcpyAsync(&last[1], flag.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, st));
    WX_HIP(hipStreamSynchronize(st));
    const uint64_t n_pieces = (uint64_t)last[0] + last[1];
    hipLaunchKernelGGL(su_item_piece0, dim3(g1(n_items)), dim3(SB), 0, st, n_items, items, pid.as<uint32_t>(), piece0);
    piece_q.alloc(std::max<uint64_t>(n_pieces, 1) * 4);
    hipLaunchKernelGGL(su_piece_codes, dim3(g1(n)), dim3(SB), 0, st, n, flag.as<uint32_t>(), pid.as<uint32_t>(), qpos,
                       piece_q.as<uint32_t>());
    WX_HIP(hipGetLastError());
    WX_HIP(hipStreamSynchronize(st));
    return n_pieces;
}

} // namespace wx
