// fwl.hip -- the Frisch-Waugh-Lovell step after demeaning: G = A'A for the demeaned block
// A = [y~, X~] (README.md:55-66 of the reference, where numpy's lstsq does it on the host).
// One pass over A in HBM: each workgroup stages chunks of up to 64 rows of A in LDS (the next
// chunk's loads in flight, in registers, while the current chunk's products run) and
// accumulates the upper triangle of A'A for them, each thread a 4 x 4 block of column pairs over
// one group of rows; the workgroups' partial triangles are then summed in workgroup order, so
// the result is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "device.hpp"

namespace wx {

namespace {
constexpr int GB = 256;               // threads per workgroup
constexpr uint32_t GRAM_LDS = 4160;   // doubles of the LDS chunk: rows x (m + 1)
constexpr uint32_t GRAM_MAX_COLS = 64;
constexpr uint32_t GRAM_MAX_BLOCKS = 2048;
constexpr uint32_t GRAM_NOALLOC_MAX_COLS = 32; // columns of the allocation-free entry point
constexpr uint32_t TB = 4;            // a thread's 4 x 4 block of column pairs

// rows per LDS chunk: the largest power of two up to 64 with R * (m + 1) doubles in GRAM_LDS
// (larger chunks measured slower: fewer workgroups fit a CU, and a chunk's load and its
// products do not overlap within a workgroup)
__host__ __device__ inline uint32_t gram_chunk_rows(uint32_t m) {
    uint32_t r = 64;
    while (r > 16 && r * (m + 1) > GRAM_LDS) r >>= 1;
    return r;
}

// pair p of the upper triangle (row-major over i <= j) -> (i, j)
__device__ inline void pair_of(uint32_t p, uint32_t m, uint32_t &i, uint32_t &j) {
    i = 0;
    uint32_t left = p;
    while (left >= m - i) {
        left -= m - i;
        ++i;
    }
    j = i + left;
}

// Per workgroup: the upper triangle of A'A over its chunks.  Columns are grouped by TB; a
// thread owns one block (bi <= bj) of TB x TB column pairs and one of `rg` row groups (rows
// g, g + rg, ... of each chunk), so a row's 2 TB values from LDS feed TB^2 products.  At the end
// the row groups' blocks are added in row-group order.
// PF: chunk elements per thread (R rows x 4 * ceil(m / 4) columns / GB, at most PF)
template <uint32_t PF>
__global__ __launch_bounds__(GB) void k_gram(const double *__restrict__ A, uint64_t n, uint32_t m, uint64_t lda,
                                             double *__restrict__ part) {
    __shared__ double sh[GRAM_LDS];
    const uint32_t ld = m + 1; // padded row stride
    const uint32_t R = gram_chunk_rows(m);
    const uint32_t mt = (m + TB - 1) / TB, nb = mt * (mt + 1) / 2;
    uint32_t rg = GB / nb; // row groups: a power of two dividing R
    rg = rg >= 64 ? 64 : rg >= 32 ? 32 : rg >= 16 ? 16 : rg >= 8 ? 8 : rg >= 4 ? 4 : rg >= 2 ? 2 : 1;
    if (rg > R) rg = R;
    const bool on = threadIdx.x < nb * rg;
    const uint32_t b = threadIdx.x % nb, g = threadIdx.x / nb;
    uint32_t bi = 0, bj = 0;
    if (on) pair_of(b, mt, bi, bj);
    double acc[TB][TB];
#pragma unroll
    for (uint32_t x = 0; x < TB; ++x)
#pragma unroll
        for (uint32_t y = 0; y < TB; ++y) acc[x][y] = 0.0;
    const uint64_t n_chunks = (n + R - 1) / R;
    // a thread's share of a chunk (element e = threadIdx.x + q * GB: column e / R, row e % R;
    // consecutive threads read consecutive rows of a column), held in registers so that the next
    // chunk's loads are in flight while this one's products run
    const uint32_t n_el = R * mt * TB;
    double pre[PF];
    auto fetch = [&](uint64_t ch) {
        const uint64_t r0 = ch * R;
        const uint32_t rows = n - r0 < R ? (uint32_t)(n - r0) : R;
#pragma unroll
        for (uint32_t q = 0; q < PF; ++q) {
            const uint32_t e = threadIdx.x + q * GB, c = e / R, r = e % R;
            pre[q] = (e < n_el && r < rows && c < m) ? A[c * lda + r0 + r] : 0.0;
        }
    };
    auto stage = [&]() { // columns past m (padding of the last column block) are zero
#pragma unroll
        for (uint32_t q = 0; q < PF; ++q) {
            const uint32_t e = threadIdx.x + q * GB, c = e / R, r = e % R;
            if (e < n_el && c < ld) sh[r * ld + c] = pre[q];
        }
    };
    uint64_t ch = blockIdx.x;
    if (ch < n_chunks) {
        fetch(ch);
        stage();
    }
    __syncthreads();
    for (; ch < n_chunks; ch += gridDim.x) {
        const bool more = ch + gridDim.x < n_chunks;
        if (more) fetch(ch + gridDim.x);
        if (on)
            for (uint32_t r = g; r < R; r += rg) {
                double vi[TB], vj[TB];
#pragma unroll
                for (uint32_t x = 0; x < TB; ++x) {
                    const uint32_t ci = bi * TB + x, cj = bj * TB + x;
                    vi[x] = ci < m ? sh[r * ld + ci] : 0.0;
                    vj[x] = cj < m ? sh[r * ld + cj] : 0.0;
                }
#pragma unroll
                for (uint32_t x = 0; x < TB; ++x)
#pragma unroll
                    for (uint32_t y = 0; y < TB; ++y) acc[x][y] += vi[x] * vj[y];
            }
        __syncthreads();
        if (more) stage();
        __syncthreads();
    }
    // row groups in order, through LDS (reused: the nb * TB^2 block entries fit)
    const uint32_t npairs = m * (m + 1) / 2;
    for (uint32_t step = 0; step < rg; ++step) {
        if (on && g == step)
#pragma unroll
            for (uint32_t x = 0; x < TB; ++x)
#pragma unroll
                for (uint32_t y = 0; y < TB; ++y)
                    sh[(b * TB + x) * TB + y] = (step == 0 ? 0.0 : sh[(b * TB + x) * TB + y]) + acc[x][y];
        __syncthreads();
    }
    for (uint32_t q = threadIdx.x; q < nb * TB * TB; q += GB) {
        const uint32_t bb = q / (TB * TB), x = (q / TB) % TB, y = q % TB;
        uint32_t ii, jj;
        pair_of(bb, mt, ii, jj);
        const uint32_t i = ii * TB + x, j = jj * TB + y;
        if (i < m && j < m && i <= j) {
            // index of pair (i, j) in the upper triangle
            const uint32_t p = i * m - i * (i - 1) / 2 + (j - i);
            part[(uint64_t)blockIdx.x * npairs + p] = sh[q];
        }
    }
}

// G[i, j] = G[j, i] = the workgroups' partials of pair (i, j) summed in a fixed order: one
// workgroup per pair, thread t adding workgroups t, t + 256, ... in turn, then an LDS tree
__global__ __launch_bounds__(GB) void k_gram_reduce(const double *__restrict__ part, uint32_t nblocks, uint32_t m,
                                                    double *__restrict__ G) {
    __shared__ double sh[GB];
    const uint32_t npairs = m * (m + 1) / 2;
    const uint32_t p = blockIdx.x;
    double a = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += GB) a += part[(uint64_t)b * npairs + p];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = GB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        uint32_t i, j;
        pair_of(p, m, i, j);
        G[(uint64_t)j * m + i] = sh[0];
        G[(uint64_t)i * m + j] = sh[0];
    }
}
} // namespace

uint32_t gram_max_cols() { return GRAM_MAX_COLS; }

// G (m x m, column-major, device) = A'A for the column-major n x m device matrix A; `part` is a
// device scratch of gram_scratch_doubles(n, m) doubles.  Enqueued on `st`.
uint64_t gram_scratch_doubles(uint64_t n, uint32_t m) {
    const uint32_t R = gram_chunk_rows(m);
    const uint64_t chunks = (n + R - 1) / R;
    return std::max<uint64_t>(1, std::min<uint64_t>(chunks, GRAM_MAX_BLOCKS)) * (uint64_t)m * (m + 1) / 2;
}

// Scratch of the allocation-free Gram entry point: the partial triangles of up to
// GRAM_MAX_BLOCKS workgroups for up to GRAM_NOALLOC_MAX_COLS columns (8.6 MB), one copy per
// GPU held until the process ends.  Host code never passes `gram_scratch` itself to a kernel:
// in host code the name is a stand-in, not a device address; gram_scratch_ptr() returns this
// device's address of its copy.
__device__ __align__(256) double gram_scratch[(uint64_t)GRAM_MAX_BLOCKS * GRAM_NOALLOC_MAX_COLS *
                                             (GRAM_NOALLOC_MAX_COLS + 1) / 2];

uint32_t gram_noalloc_max_cols() { return GRAM_NOALLOC_MAX_COLS; }

// This device's address of `gram_scratch`, looked up once per device (hipGetSymbolAddress with
// the device current) and kept; nullptr with `err` set when the lookup fails.
double *gram_scratch_ptr(std::string &err) {
    static std::mutex mu;
    static std::vector<double *> ptr; // per device
    int dev = 0, count = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceCount(&count) != hipSuccess || dev < 0 || dev >= count) {
        err = "no current HIP device";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    if ((int)ptr.size() < count) ptr.resize(count, nullptr);
    if (!ptr[dev]) {
        void *p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(gram_scratch));
        if (rc != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(gram_scratch) failed: ") + hipGetErrorString(rc);
            return nullptr;
        }
        ptr[dev] = (double *)p;
    }
    return ptr[dev];
}

void gram_device(const double *A, uint64_t n, uint32_t m, uint64_t lda, double *G, double *part, hipStream_t st) {
    const uint32_t R = gram_chunk_rows(m);
    const uint64_t chunks = (n + R - 1) / R;
    const uint32_t nb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(chunks, GRAM_MAX_BLOCKS));
    const uint32_t per_thread = (R * ((m + 3) / 4 * 4) + GB - 1) / GB;
    if (per_thread <= 4) hipLaunchKernelGGL(k_gram<4>, dim3(nb), dim3(GB), 0, st, A, n, m, lda, part);
    else if (per_thread <= 8) hipLaunchKernelGGL(k_gram<8>, dim3(nb), dim3(GB), 0, st, A, n, m, lda, part);
    else hipLaunchKernelGGL(k_gram<16>, dim3(nb), dim3(GB), 0, st, A, n, m, lda, part);
    const uint32_t npairs = m * (m + 1) / 2;
    hipLaunchKernelGGL(k_gram_reduce, dim3(npairs), dim3(GB), 0, st, part, nb, m, G);
    WX_HIP(hipGetLastError());
}

} // namespace wx
