// dense_gpu.cpp -- setup-time pseudo-inverse of large dense kept-block Laplacians on the GPU.
//
// For a connected Laplacian S (n x n) the kept-block operator is S^+ = P G P with G the
// inverse of the grounded minor S[0:n-1, 0:n-1] (padded with a zero row/column) and
// P = I - 11'/n.  The minor is factored by a blocked right-looking Cholesky: the small
// diagonal blocks on the host, the panel solves and trailing updates with rocBLAS level-3
// routines (dtrsm, dsyrk); then L^{-1} = dtrsm(L, I) and G = L^{-T} L^{-1} = dsyrk.
// This runs once per preconditioner build, off the solve path.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"

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

// host Cholesky of a small column-major block (lower), in place; false on a non-positive pivot
bool potrf_small(double *a, int nb) {
    for (int k = 0; k < nb; ++k) {
        double d = a[k + (size_t)k * nb];
        for (int q = 0; q < k; ++q) d -= a[k + (size_t)q * nb] * a[k + (size_t)q * nb];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        a[k + (size_t)k * nb] = d;
        for (int i = k + 1; i < nb; ++i) {
            double s = a[i + (size_t)k * nb];
            for (int q = 0; q < k; ++q) s -= a[i + (size_t)q * nb] * a[k + (size_t)q * nb];
            a[i + (size_t)k * nb] = s / d;
        }
        for (int i = 0; i < k; ++i) a[i + (size_t)k * nb] = 0.0; // clear the upper part
    }
    return true;
}

__global__ void k_set_identity(double *x, int64_t m) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= m * m) return;
    x[idx] = (idx % m == idx / m) ? 1.0 : 0.0;
}
} // namespace

bool gpu_available() {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

// S: n x n row-major symmetric Laplacian (connected); overwritten with S^+.
bool dense_laplacian_pinv_gpu(std::vector<double> &S, uint32_t n, bool floating) {
    if (floating && n <= 1) {
        std::fill(S.begin(), S.end(), 0.0);
        return true;
    }
    const int64_t m = floating ? (int64_t)n - 1 : (int64_t)n;
    Handle H;
    DevBuf A, X;
    A.alloc((size_t)m * m * 8);
    X.alloc((size_t)m * m * 8);
    // the minor is symmetric: row-major == column-major
    WX_HIP(hipMemcpy2D(A.p, m * 8, S.data(), (size_t)n * 8, m * 8, m, hipMemcpyHostToDevice));
    double *a = A.as<double>();
    const int64_t b = 256;
    std::vector<double> blk(b * b);
    const double one = 1.0, mone = -1.0, zero = 0.0;
    for (int64_t p = 0; p < m; p += b) {
        const int64_t nb = std::min<int64_t>(b, m - p);
        double *app = a + p + p * m;
        WX_HIP(hipMemcpy2D(blk.data(), nb * 8, app, m * 8, nb * 8, nb, hipMemcpyDeviceToHost));
        if (!potrf_small(blk.data(), (int)nb)) return false;
        WX_HIP(hipMemcpy2D(app, m * 8, blk.data(), nb * 8, nb * 8, nb, hipMemcpyHostToDevice));
        const int64_t rest = m - p - nb;
        if (rest > 0) {
            double *a21 = a + (p + nb) + p * m, *a22 = a + (p + nb) + (p + nb) * m;
            WX_RB(rocblas_dtrsm(H.h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                rocblas_diagonal_non_unit, (rocblas_int)rest, (rocblas_int)nb, &one, app, (rocblas_int)m,
                                a21, (rocblas_int)m));
            WX_RB(rocblas_dsyrk(H.h, rocblas_fill_lower, rocblas_operation_none, (rocblas_int)rest, (rocblas_int)nb,
                                &mone, a21, (rocblas_int)m, &one, a22, (rocblas_int)m));
        }
    }
    // X = L^{-1}
    hipLaunchKernelGGL(k_set_identity, dim3((unsigned)((m * m + 255) / 256)), dim3(256), 0, 0, X.as<double>(), m);
    WX_HIP(hipGetLastError());
    WX_RB(rocblas_dtrsm(H.h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                        (rocblas_int)m, (rocblas_int)m, &one, a, (rocblas_int)m, X.as<double>(), (rocblas_int)m));
    // G = X' X (lower triangle) into A
    WX_RB(rocblas_dsyrk(H.h, rocblas_fill_lower, rocblas_operation_transpose, (rocblas_int)m, (rocblas_int)m, &one,
                        X.as<double>(), (rocblas_int)m, &zero, a, (rocblas_int)m));
    std::vector<double> G((size_t)m * m);
    WX_HIP(hipMemcpy(G.data(), a, (size_t)m * m * 8, hipMemcpyDeviceToHost));
    if (!floating) { // positive definite block: the inverse itself, symmetrised
        for (int64_t i = 0; i < m; ++i)
            for (int64_t j = 0; j < m; ++j) S[(size_t)i * n + j] = i >= j ? G[i + j * m] : G[j + i * m];
        return true;
    }
    // symmetrise (column-major lower: G[i + j*m], i >= j), pad and project with P on both sides
    std::vector<double> rowmean(n, 0.0);
    double total = 0.0;
    for (int64_t j = 0; j < m; ++j)
        for (int64_t i = j; i < m; ++i) {
            const double g = G[i + j * m];
            rowmean[i] += g;
            if (i != j) rowmean[j] += g;
        }
    for (uint32_t i = 0; i < n; ++i) {
        total += rowmean[i];
        rowmean[i] /= n;
    }
    total /= (double)n * n;
    for (int64_t i = 0; i < (int64_t)n; ++i)
        for (int64_t j = 0; j < (int64_t)n; ++j) {
            double g = 0.0;
            if (i < m && j < m) g = i >= j ? G[i + j * m] : G[j + i * m];
            S[(size_t)i * n + j] = g - rowmean[i] - rowmean[j] + total;
        }
    return true;
}

} // namespace wx
