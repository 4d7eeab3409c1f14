// dense_gpu.hip -- setup-time reduced blocks of large dense kept blocks on the GPU.
//
// A component's kept-block Schur complement S = D_k - C' D_e^{-1} C is assembled on the device
// (one workgroup per kept row, the row in LDS; the eliminated rows adjacent to it are added one
// after another, so every entry sums its terms in adjacency order), then its pseudo-inverse is
// formed there too.  For a connected Laplacian S (n x n) the operator is S^+ = P G P with G the
// inverse of the grounded minor S[0:n-1, 0:n-1] (padded with a zero row/column) and
// P = I - 11'/n; for a definite block (grounded or signed form) it is S^{-1}.  The inverse comes
// from rocSOLVER's Cholesky factorisation and inverse (dpotrf, dpotri).  This runs once per
// preconditioner build, off the solve path.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "schwarz.hpp"

namespace wx {

namespace {
#define WX_RB(call)                                                                                    \
    do {                                                                                               \
        rocblas_status s_ = (call);                                                                    \
        if (s_ != rocblas_status_success)                                                              \
            ::wx::fail(WX_ERR_DEVICE, std::string(#call " failed: ") + rocblas_status_to_string(s_));  \
    } while (0)

struct Handle {
    rocblas_handle h = nullptr;
    Handle() { WX_RB(rocblas_create_handle(&h)); }
    ~Handle() {
        if (h) rocblas_destroy_handle(h);
    }
};

constexpr int DB = 256;             // workgroup size of the set-up kernels
constexpr uint32_t ROW_CHUNK = 8192; // kept columns per assembly workgroup (64 KB of LDS)

// Off-diagonal part of kept row i, columns [c0, c0 + chunk): -sum over the eliminated rows e
// adjacent to i, in adjacency order, of (w_ie * C_ej) / D_e.  Entries of one eliminated row
// have distinct columns, so the threads of a workgroup update disjoint slots between barriers.
// Eliminated rows flagged in `heavy` are skipped here: their terms come from one dense product.
__global__ __launch_bounds__(DB) void k_schur_rows(uint32_t nk, const uint64_t *__restrict__ cptr,
                                                   const uint32_t *__restrict__ cidx, const double *__restrict__ cval,
                                                   const double *__restrict__ diag, const uint64_t *__restrict__ tp,
                                                   const uint32_t *__restrict__ te, const double *__restrict__ tv,
                                                   const uint8_t *__restrict__ heavy, double *__restrict__ S) {
    extern __shared__ double row[];
    const uint32_t i = blockIdx.x, c0 = blockIdx.y * ROW_CHUNK, c1 = min(nk, c0 + ROW_CHUNK), w = c1 - c0;
    for (uint32_t j = threadIdx.x; j < w; j += DB) row[j] = 0.0;
    __syncthreads();
    for (uint64_t q = tp[i]; q < tp[i + 1]; ++q) {
        const uint32_t e = te[q];
        if (heavy && heavy[e]) continue; // uniform over the workgroup
        const double wi = tv[q], inv = 1.0 / diag[e];
        for (uint64_t p = cptr[e] + threadIdx.x; p < cptr[e + 1]; p += DB) {
            const uint32_t j = cidx[p];
            if (j >= c0 && j < c1) row[j - c0] -= (wi * cval[p]) * inv;
        }
        __syncthreads();
    }
    for (uint32_t j = threadIdx.x; j < w; j += DB) S[(uint64_t)i * nk + c0 + j] = row[j];
}

// H (n_h x nk, row-major, zeroed): row h = C row heavy_rows[h] scaled by 1 / sqrt(D_e), so the
// heavy rows' part of the off-diagonal block is -H' H.
__global__ __launch_bounds__(DB) void k_heavy_rows(uint32_t nk, const uint32_t *__restrict__ rows,
                                                   const uint64_t *__restrict__ cptr,
                                                   const uint32_t *__restrict__ cidx, const double *__restrict__ cval,
                                                   const double *__restrict__ diag, double *__restrict__ H) {
    const uint32_t h = blockIdx.x, e = rows[h];
    const double s = 1.0 / sqrt(diag[e]);
    for (uint64_t p = cptr[e] + threadIdx.x; p < cptr[e + 1]; p += DB) H[(uint64_t)h * nk + cidx[p]] = cval[p] * s;
}

// Diagonal of kept row i: minus the off-diagonal row sum (a fixed-order tree; the magnitude
// of the signed form instead), plus the surplus carried to the kept block when grounded.
__global__ __launch_bounds__(DB) void k_schur_diag(uint32_t ne, uint32_t nk, const double *__restrict__ diag,
                                                   const double *__restrict__ surplus,
                                                   const double *__restrict__ rowabs, const uint64_t *__restrict__ tp,
                                                   const uint32_t *__restrict__ te, const double *__restrict__ tv,
                                                   int signed_form, double *__restrict__ S) {
    __shared__ double sh[DB];
    const uint32_t i = blockIdx.x;
    double a = 0.0;
    for (uint32_t j = threadIdx.x; j < nk; j += DB)
        if (j != i) a += S[(uint64_t)i * nk + j];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = DB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double sur = surplus ? surplus[ne + i] : 0.0, mag = 0.0;
        for (uint64_t q = tp[i]; q < tp[i + 1]; ++q) {
            const uint32_t e = te[q];
            const double wi = tv[q], inv = 1.0 / diag[e];
            if (surplus) sur += fabs(wi) * (inv * surplus[e]);
            if (signed_form) mag += fabs(wi) * (rowabs[e] - fabs(wi)) * inv;
        }
        S[(uint64_t)i * nk + i] = (signed_form ? mag : -sh[0]) + sur;
    }
}

// Full symmetric G from dpotri's lower triangle (column-major, leading dimension ld) on the
// leading m x m block; row and column m (the floating gauge's dropped vertex) set to zero when
// pad, i.e. n = m + 1.
__global__ __launch_bounds__(DB) void k_symmetrise(double *__restrict__ G, uint32_t m, uint32_t n, uint64_t ld) {
    const uint64_t idx = (uint64_t)blockIdx.x * DB + threadIdx.x;
    if (idx >= (uint64_t)n * n) return;
    const uint32_t r = (uint32_t)(idx % n), c = (uint32_t)(idx / n); // element (r, c) at r + c * ld
    if (r >= m || c >= m) G[r + c * ld] = 0.0;
    else if (r < c) G[r + c * ld] = G[c + r * ld];
}

// rmean[c] = (sum_r G[r, c]) / n, a fixed-order tree per column
__global__ __launch_bounds__(DB) void k_col_means(const double *__restrict__ G, uint32_t n, uint64_t ld,
                                                  double *__restrict__ rmean) {
    __shared__ double sh[DB];
    const uint32_t c = blockIdx.x;
    double a = 0.0;
    for (uint32_t r = threadIdx.x; r < n; r += DB) a += G[r + c * ld];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = DB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) rmean[c] = sh[0] / n;
}

// total = mean of the column means (one workgroup, fixed order); written to rmean[n]
__global__ __launch_bounds__(DB) void k_total_mean(double *__restrict__ rmean, uint32_t n) {
    __shared__ double sh[DB];
    double a = 0.0;
    for (uint32_t c = threadIdx.x; c < n; c += DB) a += rmean[c];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = DB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) rmean[n] = sh[0] / n;
}

// S^+ = P G P: G[r, c] - mean_r - mean_c + total
__global__ __launch_bounds__(DB) void k_project(double *__restrict__ G, uint32_t n, uint64_t ld,
                                                const double *__restrict__ rmean) {
    const uint64_t idx = (uint64_t)blockIdx.x * DB + threadIdx.x;
    if (idx >= (uint64_t)n * n) return;
    const uint32_t r = (uint32_t)(idx % n), c = (uint32_t)(idx / n);
    G[r + c * ld] = G[r + c * ld] - rmean[r] - rmean[c] + rmean[n];
}

template <typename T> void upload_vec(DevBuf &b, const T *p, size_t n) {
    b.alloc(std::max<size_t>(n, 1) * sizeof(T));
    if (n) WX_HIP(hipMemcpy(b.p, p, n * sizeof(T), hipMemcpyHostToDevice));
}
} // namespace

bool gpu_available() {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

namespace {
std::shared_ptr<double> adopt(DevBuf &b) {
    double *p = b.as<double>();
    b.p = nullptr;
    b.bytes = 0;
    return std::shared_ptr<double>(p, [](double *q) { (void)hipFree(q); });
}
} // namespace

void Precond::Schwarz::host_dense() const {
    std::lock_guard<std::mutex> lk(*dense_mu);
    if (dense_on_host) return;
    auto *self = const_cast<Precond::Schwarz *>(this);
    for (const DeviceBlock &b : dense_dev)
        WX_HIP(hipMemcpy(self->dense.data() + b.off, b.dev.get(), b.count * 8, hipMemcpyDeviceToHost));
    self->dense_on_host = true;
}

bool dense_schur_pinv_gpu(const DenseSchurArgs &a, bool floating, std::shared_ptr<double> &out) {
    const uint32_t n = a.nk;
    DevBuf cptr, cidx, cval, diag, sur, rabs, tp, te, tv, S, mean, info;
    if (floating && n <= 1) {
        S.alloc((size_t)n * n * 8);
        WX_HIP(hipMemset(S.p, 0, (size_t)n * n * 8));
        out = adopt(S);
        return true;
    }
    upload_vec(cptr, a.cptr, (size_t)a.ne + 1);
    upload_vec(cidx, a.cidx, a.nnz);
    upload_vec(cval, a.cval, a.nnz);
    upload_vec(diag, a.diag, (size_t)a.ne + a.nk);
    if (a.surplus) upload_vec(sur, a.surplus, (size_t)a.ne + a.nk);
    if (a.rowabs) upload_vec(rabs, a.rowabs, a.ne);
    upload_vec(tp, a.tp, (size_t)n + 1);
    upload_vec(te, a.te, a.nnz);
    upload_vec(tv, a.tv, a.nnz);
    S.alloc((size_t)n * n * 8);
    double *s = S.as<double>();
    // Eliminated rows touching many kept levels (deg >= n / 8) cost deg^2 scattered updates each in
    // the row kernel; they are gathered into a dense block whose product -H'H runs on the matrix
    // cores instead, and the row kernel adds only the others' terms.
    std::vector<uint32_t> heavy_rows;
    std::vector<uint8_t> heavy_flag(a.ne, 0);
    for (uint32_t e = 0; e < a.ne; ++e)
        if ((a.cptr[e + 1] - a.cptr[e]) * 8 >= n && n >= 1024) {
            heavy_rows.push_back(e);
            heavy_flag[e] = 1;
        }
    DevBuf hflag, hrows, hmat;
    if (!heavy_rows.empty()) {
        upload_vec(hflag, heavy_flag.data(), heavy_flag.size());
        upload_vec(hrows, heavy_rows.data(), heavy_rows.size());
    }
    const uint32_t chunks = (n + ROW_CHUNK - 1) / ROW_CHUNK;
    const size_t lds = (size_t)std::min(n, ROW_CHUNK) * 8;
    hipLaunchKernelGGL(k_schur_rows, dim3(n, chunks), dim3(DB), lds, 0, n, cptr.as<uint64_t>(), cidx.as<uint32_t>(),
                       cval.as<double>(), diag.as<double>(), tp.as<uint64_t>(), te.as<uint32_t>(), tv.as<double>(),
                       heavy_rows.empty() ? nullptr : hflag.as<uint8_t>(), s);
    WX_HIP(hipGetLastError());
    Handle H0;
    if (!heavy_rows.empty()) {
        const uint32_t nh = (uint32_t)heavy_rows.size();
        hmat.alloc((size_t)nh * n * 8);
        WX_HIP(hipMemsetAsync(hmat.p, 0, (size_t)nh * n * 8, 0));
        hipLaunchKernelGGL(k_heavy_rows, dim3(nh), dim3(DB), 0, 0, n, hrows.as<uint32_t>(), cptr.as<uint64_t>(),
                           cidx.as<uint32_t>(), cval.as<double>(), diag.as<double>(), hmat.as<double>());
        WX_HIP(hipGetLastError());
        // S (row-major n x n == column-major, symmetric) += -H' H; H row-major nh x n is the
        // column-major n x nh matrix A, so the product is A A'
        const double alpha = -1.0, beta = 1.0;
        WX_RB(rocblas_dgemm(H0.h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)n,
                            (rocblas_int)n, (rocblas_int)nh, &alpha, hmat.as<double>(), (rocblas_int)n, hmat.as<double>(),
                            (rocblas_int)n, &beta, s, (rocblas_int)n));
    }
    hipLaunchKernelGGL(k_schur_diag, dim3(n), dim3(DB), 0, 0, a.ne, n, diag.as<double>(),
                       a.surplus ? sur.as<double>() : nullptr, a.rowabs ? rabs.as<double>() : nullptr,
                       tp.as<uint64_t>(), te.as<uint32_t>(), tv.as<double>(), a.signed_form ? 1 : 0, s);
    WX_HIP(hipGetLastError());
    // Cholesky and inverse of the (grounded) minor in place: row-major == column-major
    const uint32_t m = floating ? n - 1 : n;
    Handle &H = H0;
    info.alloc(sizeof(rocblas_int));
    rocblas_int hinfo = 0;
    WX_RB(rocsolver_dpotrf(H.h, rocblas_fill_lower, (rocblas_int)m, s, (rocblas_int)n, info.as<rocblas_int>()));
    WX_HIP(hipMemcpy(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost));
    if (hinfo != 0) return false;
    WX_RB(rocsolver_dpotri(H.h, rocblas_fill_lower, (rocblas_int)m, s, (rocblas_int)n, info.as<rocblas_int>()));
    WX_HIP(hipMemcpy(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost));
    if (hinfo != 0) return false;
    const uint64_t nn = (uint64_t)n * n;
    const uint32_t eg = (uint32_t)((nn + DB - 1) / DB);
    hipLaunchKernelGGL(k_symmetrise, dim3(eg), dim3(DB), 0, 0, s, m, n, (uint64_t)n);
    if (floating) {
        mean.alloc(((size_t)n + 1) * 8);
        hipLaunchKernelGGL(k_col_means, dim3(n), dim3(DB), 0, 0, s, n, (uint64_t)n, mean.as<double>());
        hipLaunchKernelGGL(k_total_mean, dim3(1), dim3(DB), 0, 0, mean.as<double>(), n);
        hipLaunchKernelGGL(k_project, dim3(eg), dim3(DB), 0, 0, s, n, (uint64_t)n, mean.as<double>());
    }
    WX_HIP(hipGetLastError());
    WX_HIP(hipDeviceSynchronize());
    out = adopt(S);
    return true;
}

} // namespace wx
