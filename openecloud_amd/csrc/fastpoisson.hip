// Direct FIT Poisson solve by the capacitance-matrix method.
//
// The operator A (poissonSolver.pyx:44-59) acts on the inside nodes, all of which lie in
// the box B of nodes 2 <= i <= nxExt-3, 2 <= j <= nyExt-3.  Let R be the box's uniform
// 5-point operator (x-edge coefficient cxu, y-edge cyu, zero Dirichlet data outside B) and M
// the box operator whose rows are A's rows on inside nodes and R's rows elsewhere.  M x = b
// with b = q on inside nodes reproduces A phi = q exactly on inside nodes (the inside rows
// never reference a non-inside node).  M and R differ only in the m rows S of inside nodes
// next to the wall: M = R + U W with U the unit columns of S.  By the Woodbury identity
//     x = R^-1 (b - U z),   z = C^-1 W R^-1 b,   C = I_m + W R^-1 U.
// R^-1 is applied exactly by the type-I discrete sine transform, written as four FP64
// GEMMs with the orthonormal DST matrices (MFMA through rocBLAS):
//     R^-1 b = Sx ((Sx b Sy) / L) Sy     (column-major box, x fastest),
// L[q, p] = cxu * lx_q + cyu * ly_p with the 1-D eigenvalues l_k = 2 - 2 cos(pi k / (n + 1)).
// C is built once on the device by m batched DST solves and inverted with rocSOLVER.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "common.h"

struct oecl_fastpoisson {
    uint32_t nx, ny, np;      // extended grid
    uint32_t bx, by, nb;      // box sizes, nb = bx * by
    uint32_t m;               // irregular rows
    rocblas_handle blas;
    double *Sx, *Sy, *Linv;   // bx*bx, by*by, nb
    int32_t *sIdx;            // [m] box index of row a
    int32_t *wPtr, *wCol;     // CSR of W over box columns, m rows
    double *wVal;
    double *Cinv;             // m*m, column-major
    double *b, *t1, *t2;      // nb each
    double *wy, *z;           // m each
    double *part;             // [nChunks][m] partial products of the capacitance GEMV
    uint32_t nChunks;
    const uint8_t *mask;      // inside flags of the extended grid (caller owned)
};

namespace {

__global__ void k_gather_box(uint32_t nx, uint32_t bx, uint32_t by, const uint8_t *__restrict__ mask,
                             const double *__restrict__ q, double *__restrict__ b)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= bx * by) return;
    uint32_t i = t % bx + 2, j = t / bx + 2;
    uint32_t p = i + j * nx;
    b[t] = mask[p] ? q[p] : 0.;
}

__global__ void k_scatter_phi(uint32_t nx, uint32_t ny, uint32_t bx, uint32_t by,
                              const uint8_t *__restrict__ mask, const double *__restrict__ x,
                              double *__restrict__ phi)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nx * ny) return;
    uint32_t i = p % nx, j = p / nx;
    double v = 0.;
    if (mask[p] && i >= 2 && j >= 2 && i - 2 < bx && j - 2 < by) v = x[(i - 2) + (j - 2) * bx];
    phi[p] = v;
}

__global__ void k_scale(uint32_t n, const double *__restrict__ a, double *__restrict__ x)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] *= a[t];
}

// wy[a] = sum_k W[a, k] y[k]
__global__ void k_wmul(uint32_t m, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                       const double *__restrict__ val, const double *__restrict__ y, double *__restrict__ wy)
{
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= m) return;
    double s = 0.;
    for (int32_t k = ptr[a]; k < ptr[a + 1]; ++k) s += val[k] * y[col[k]];
    wy[a] = s;
}

__global__ void k_sub_scatter(uint32_t m, const int32_t *__restrict__ sIdx, const double *__restrict__ z,
                              double *__restrict__ b)
{
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < m) b[sIdx[a]] -= z[a];
}

// Batch of unit right-hand sides after the first DST pass: column b of the batch holds
// (Sx e_s Sy) / L = outer(Sx[:, i_s], Sy[j_s, :]) / L.
__global__ void k_unit_spectra(uint32_t bx, uint32_t by, uint32_t a0, uint32_t cnt,
                               const int32_t *__restrict__ sIdx, const double *__restrict__ Sx,
                               const double *__restrict__ Sy, const double *__restrict__ Linv,
                               double *__restrict__ out)
{
    uint64_t nb = (uint64_t)bx * by;
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * cnt) return;
    uint32_t c = (uint32_t)(t / nb);
    uint32_t e = (uint32_t)(t % nb);
    uint32_t q = e % bx, p = e / bx;
    int32_t s = sIdx[a0 + c];
    uint32_t is = s % bx, js = s / bx;
    out[t] = Sx[q + (uint64_t)is * bx] * Sy[js + (uint64_t)p * by] * Linv[e];
}

// C[a, a0 + c] = delta + sum_k W[a, k] G_c[k]   (column-major m x m)
__global__ void k_cap_columns(uint32_t m, uint32_t a0, uint32_t cnt, uint64_t nb,
                              const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                              const double *__restrict__ val, const double *__restrict__ G,
                              double *__restrict__ C)
{
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)m * cnt) return;
    uint32_t a = (uint32_t)(t % m), c = (uint32_t)(t / m);
    const double *g = G + (uint64_t)c * nb;
    double s = 0.;
    for (int32_t k = ptr[a]; k < ptr[a + 1]; ++k) s += val[k] * g[col[k]];
    if (a == a0 + c) s += 1.;
    C[(uint64_t)(a0 + c) * m + a] = s;
}

// ------------------------------------------------------------------ per-solve kernels
// C = A * B (column-major, FP64) with 32 x 32 output tiles: 256 threads, 2 x 2 outputs per
// thread, K staged through LDS 32 at a time.  The box is only ~510 x 382, so small tiles
// are what fills the 256 CUs.  EPI: 0 store, 1 store times Linv (elementwise), 2 scatter
// into the extended-grid potential (masked; zero elsewhere is written by the caller).
constexpr int GT = 32;

template <int EPI>
__global__ void __launch_bounds__(256)
    k_gemm_nn(int M, int N, int K, const double *__restrict__ A, int lda, const double *__restrict__ B, int ldb,
              double *__restrict__ C, int ldc, const double *__restrict__ S, uint32_t nx,
              const uint8_t *__restrict__ mask)
{
    __shared__ double As[GT][GT + 1];
    __shared__ double Bs[GT][GT + 1];
    const int tid = threadIdx.x;
    const int tm = tid & 15, tn = tid >> 4;
    const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
    double c00 = 0., c01 = 0., c10 = 0., c11 = 0.;
    for (int k0 = 0; k0 < K; k0 += GT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int idx = tid + r * 256;
            int row = idx & (GT - 1), kk = idx >> 5;
            int gm = m0 + row, gk = k0 + kk;
            As[kk][row] = (gm < M && gk < K) ? A[gm + (size_t)gk * lda] : 0.;
            int kb = idx & (GT - 1), col = idx >> 5;
            int gkb = k0 + kb, gn = n0 + col;
            Bs[kb][col] = (gkb < K && gn < N) ? B[gkb + (size_t)gn * ldb] : 0.;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < GT; ++kk) {
            double a0 = As[kk][2 * tm], a1 = As[kk][2 * tm + 1];
            double b0 = Bs[kk][2 * tn], b1 = Bs[kk][2 * tn + 1];
            c00 = fma(a0, b0, c00);
            c01 = fma(a0, b1, c01);
            c10 = fma(a1, b0, c10);
            c11 = fma(a1, b1, c11);
        }
        __syncthreads();
    }
    double cv[2][2] = {{c00, c01}, {c10, c11}};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            int gm = m0 + 2 * tm + i, gn = n0 + 2 * tn + j;
            if (gm >= M || gn >= N) continue;
            size_t o = gm + (size_t)gn * ldc;
            if (EPI == 0) {
                C[o] = cv[i][j];
            } else if (EPI == 1) {
                C[o] = cv[i][j] * S[o];
            } else {
                uint32_t p = (uint32_t)(gm + 2) + (uint32_t)(gn + 2) * nx;
                C[p] = mask[p] ? cv[i][j] : 0.;
            }
        }
}

// Partial products of y = Cinv w over column chunks of 32: part[chunk * m + a].
__global__ void __launch_bounds__(256)
    k_gemv_part(uint32_t m, const double *__restrict__ Cm, const double *__restrict__ w,
                double *__restrict__ part)
{
    uint32_t a = blockIdx.x * 256 + threadIdx.x;
    uint32_t b0 = blockIdx.y * 32;
    if (a >= m) return;
    double s = 0.;
    uint32_t b1 = b0 + 32 < m ? b0 + 32 : m;
    for (uint32_t b = b0; b < b1; ++b) s = fma(Cm[a + (size_t)b * m], w[b], s);
    part[(size_t)blockIdx.y * m + a] = s;
}

// z[a] = sum of the partials; b[sIdx[a]] -= z[a].
__global__ void k_gemv_finish_sub(uint32_t m, uint32_t nChunks, const double *__restrict__ part,
                                  const int32_t *__restrict__ sIdx, double *__restrict__ b)
{
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= m) return;
    double s = 0.;
    for (uint32_t c = 0; c < nChunks; ++c) s += part[(size_t)c * m + a];
    b[sIdx[a]] -= s;
}

__global__ void k_zero_outside_box(uint32_t nx, uint32_t ny, double *__restrict__ phi)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nx * ny) return;
    uint32_t i = p % nx, j = p / nx;
    if (i < 2 || j < 2 || i >= nx - 2 || j >= ny - 2) phi[p] = 0.;
}

#define RB(call, where)                                         \
    do {                                                        \
        rocblas_status _s = (call);                             \
        if (_s != rocblas_status_success) {                     \
            char _m[160];                                       \
            snprintf(_m, sizeof(_m), "%s: rocblas status %d", where, (int)_s); \
            oecl_set_error_msg(_m);                             \
            return OECL_ERR_HIP;                                \
        }                                                       \
    } while (0)

// x (nb) <- R^-1 x via DST: t1 = Sx x ; x = t1 Sy ; x /= L ; t1 = Sx x ; x = t1 Sy
int dst_solve(oecl_fastpoisson *h, double *x, double *t1, hipStream_t s)
{
    const double one = 1., zero = 0.;
    int bx = h->bx, by = h->by;
    RB(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, bx, by, bx, &one, h->Sx,
                     bx, x, bx, &zero, t1, bx), "dgemm 1");
    RB(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, bx, by, by, &one, t1,
                     bx, h->Sy, by, &zero, x, bx), "dgemm 2");
    hipLaunchKernelGGL(k_scale, dim3(oecl_blocks(h->nb, 256)), dim3(256), 0, s, h->nb, h->Linv, x);
    RB(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, bx, by, bx, &one, h->Sx,
                     bx, x, bx, &zero, t1, bx), "dgemm 3");
    RB(rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, bx, by, by, &one, t1,
                     bx, h->Sy, by, &zero, x, bx), "dgemm 4");
    return OECL_OK;
}

}  // namespace

OECL_EXPORT int oecl_fastpoisson_create(uint32_t nx, uint32_t ny, const double *SxHost,
                                        const double *SyHost, const double *LinvHost, uint32_t m,
                                        const int32_t *sIdxHost, const int32_t *wPtrHost,
                                        const int32_t *wColHost, const double *wValHost,
                                        const uint8_t *maskDev, oecl_fastpoisson **out,
                                        double *capDiag)
{
    if (!out || !SxHost || !SyHost || !LinvHost || !maskDev || nx < 5 || ny < 5) return OECL_ERR_ARG;
    if (m > 0 && (!sIdxHost || !wPtrHost || !wColHost || !wValHost)) return OECL_ERR_ARG;
    oecl_fastpoisson *h = new oecl_fastpoisson();
    h->nx = nx;
    h->ny = ny;
    h->np = nx * ny;
    h->bx = nx - 4;
    h->by = ny - 4;
    h->nb = h->bx * h->by;
    h->m = m;
    h->mask = maskDev;
    uint32_t nnz = m ? (uint32_t)wPtrHost[m] : 0;
    if (rocblas_create_handle(&h->blas) != rocblas_status_success) {
        delete h;
        oecl_set_error_msg("rocblas_create_handle failed");
        return OECL_ERR_HIP;
    }
    uint64_t bytes = 8ull * ((uint64_t)h->bx * h->bx + (uint64_t)h->by * h->by + 4ull * h->nb +
                             2ull * m + (uint64_t)m * m + nnz) +
                     4ull * (m + (m + 1) + nnz) + 1024;
    char *mem = nullptr;
    OECL_HIP(hipMalloc(&mem, bytes), "fastpoisson alloc");
    char *cur = mem;
    h->Sx = (double *)cur; cur += 8ull * h->bx * h->bx;
    h->Sy = (double *)cur; cur += 8ull * h->by * h->by;
    h->Linv = (double *)cur; cur += 8ull * h->nb;
    h->b = (double *)cur; cur += 8ull * h->nb;
    h->t1 = (double *)cur; cur += 8ull * h->nb;
    h->t2 = (double *)cur; cur += 8ull * h->nb;
    h->wy = (double *)cur; cur += 8ull * m;
    h->z = (double *)cur; cur += 8ull * m;
    h->Cinv = (double *)cur; cur += 8ull * m * m;
    h->wVal = (double *)cur; cur += 8ull * nnz;
    h->sIdx = (int32_t *)cur; cur += 4ull * m;
    h->wPtr = (int32_t *)cur; cur += 4ull * (m + 1);
    h->wCol = (int32_t *)cur; cur += 4ull * nnz;
    h->nChunks = (m + 31) / 32;
    h->part = nullptr;
    if (m) OECL_HIP(hipMalloc(&h->part, 8ull * h->nChunks * m), "fastpoisson alloc");
    OECL_HIP(hipMemcpy(h->Sx, SxHost, 8ull * h->bx * h->bx, hipMemcpyHostToDevice), "copy Sx");
    OECL_HIP(hipMemcpy(h->Sy, SyHost, 8ull * h->by * h->by, hipMemcpyHostToDevice), "copy Sy");
    OECL_HIP(hipMemcpy(h->Linv, LinvHost, 8ull * h->nb, hipMemcpyHostToDevice), "copy L");
    if (m) {
        OECL_HIP(hipMemcpy(h->sIdx, sIdxHost, 4ull * m, hipMemcpyHostToDevice), "copy S");
        OECL_HIP(hipMemcpy(h->wPtr, wPtrHost, 4ull * (m + 1), hipMemcpyHostToDevice), "copy W");
        OECL_HIP(hipMemcpy(h->wCol, wColHost, 4ull * nnz, hipMemcpyHostToDevice), "copy W");
        OECL_HIP(hipMemcpy(h->wVal, wValHost, 8ull * nnz, hipMemcpyHostToDevice), "copy W");
        // Capacitance matrix, built in column chunks of `chunk` batched DST solves.
        uint32_t chunk = (uint32_t)((512ull << 20) / (16ull * h->nb));
        if (chunk < 1) chunk = 1;
        if (chunk > m) chunk = m;
        double *G = nullptr, *T = nullptr;
        OECL_HIP(hipMalloc(&G, 8ull * h->nb * chunk), "cap alloc");
        OECL_HIP(hipMalloc(&T, 8ull * h->nb * chunk), "cap alloc");
        const double one = 1., zero = 0.;
        int bx = h->bx, by = h->by;
        long long snb = h->nb;
        for (uint32_t a0 = 0; a0 < m; a0 += chunk) {
            uint32_t cnt = m - a0 < chunk ? m - a0 : chunk;
            uint64_t tot = (uint64_t)h->nb * cnt;
            hipLaunchKernelGGL(k_unit_spectra, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, h->bx,
                               h->by, a0, cnt, h->sIdx, h->Sx, h->Sy, h->Linv, G);
            RB(rocblas_dgemm_strided_batched(h->blas, rocblas_operation_none, rocblas_operation_none,
                                             bx, by, bx, &one, h->Sx, bx, 0, G, bx, snb, &zero, T, bx,
                                             snb, cnt), "cap dgemm 1");
            RB(rocblas_dgemm_strided_batched(h->blas, rocblas_operation_none, rocblas_operation_none,
                                             bx, by, by, &one, T, bx, snb, h->Sy, by, 0, &zero, G, bx,
                                             snb, cnt), "cap dgemm 2");
            uint64_t tc = (uint64_t)m * cnt;
            hipLaunchKernelGGL(k_cap_columns, dim3((unsigned)((tc + 255) / 256)), dim3(256), 0, 0, m, a0,
                               cnt, (uint64_t)h->nb, h->wPtr, h->wCol, h->wVal, G, h->Cinv);
        }
        OECL_CHECK_LAUNCH("capacitance build");
        OECL_HIP(hipDeviceSynchronize(), "capacitance build");
        if (capDiag) {
            // report max |C| for a scale check
            double *hc = (double *)malloc(8ull * m);
            OECL_HIP(hipMemcpy2D(hc, 8, h->Cinv, 8ull * (m + 1), 8, m, hipMemcpyDeviceToHost), "diag");
            double mx = 0.;
            for (uint32_t a = 0; a < m; ++a) mx = fabs(hc[a]) > mx ? fabs(hc[a]) : mx;
            *capDiag = mx;
            free(hc);
        }
        (void)hipFree(T);
        // In-place inverse: getrf then getri.
        rocblas_int *ipiv = nullptr, *info = nullptr;
        OECL_HIP(hipMalloc(&ipiv, 4ull * m), "ipiv");
        OECL_HIP(hipMalloc(&info, 4), "info");
        RB(rocsolver_dgetrf(h->blas, m, m, h->Cinv, m, ipiv, info), "dgetrf");
        RB(rocsolver_dgetri(h->blas, m, h->Cinv, m, ipiv, info), "dgetri");
        int hinfo = 0;
        OECL_HIP(hipMemcpy(&hinfo, info, 4, hipMemcpyDeviceToHost), "info");
        (void)hipFree(ipiv);
        (void)hipFree(info);
        (void)hipFree(G);
        if (hinfo != 0) {
            oecl_set_error_msg("capacitance matrix is singular");
            return OECL_ERR_NOCONV;
        }
    }
    OECL_HIP(hipDeviceSynchronize(), "fastpoisson create");
    *out = h;
    return OECL_OK;
}

OECL_EXPORT int oecl_fastpoisson_solve(oecl_fastpoisson *h, const double *q, double *phi, void *stream)
{
    if (!h || !q || !phi) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int bx = h->bx, by = h->by;
    dim3 gg((bx + GT - 1) / GT, (by + GT - 1) / GT), gb(256);
    hipLaunchKernelGGL(k_gather_box, dim3(oecl_blocks(h->nb, 256)), dim3(256), 0, s, h->nx, h->bx, h->by,
                       h->mask, q, h->b);
    hipLaunchKernelGGL(k_zero_outside_box, dim3(oecl_blocks(h->np, 256)), dim3(256), 0, s, h->nx, h->ny, phi);
    if (h->m) {
        // y = R^-1 b
        hipLaunchKernelGGL(k_gemm_nn<0>, gg, gb, 0, s, bx, by, bx, h->Sx, bx, h->b, bx, h->t1, bx, nullptr, 0u,
                           nullptr);
        hipLaunchKernelGGL(k_gemm_nn<1>, gg, gb, 0, s, bx, by, by, h->t1, bx, h->Sy, by, h->t2, bx, h->Linv, 0u,
                           nullptr);
        hipLaunchKernelGGL(k_gemm_nn<0>, gg, gb, 0, s, bx, by, bx, h->Sx, bx, h->t2, bx, h->t1, bx, nullptr, 0u,
                           nullptr);
        hipLaunchKernelGGL(k_gemm_nn<0>, gg, gb, 0, s, bx, by, by, h->t1, bx, h->Sy, by, h->t2, bx, nullptr, 0u,
                           nullptr);
        // b -= U C^-1 W y
        hipLaunchKernelGGL(k_wmul, dim3(oecl_blocks(h->m, 256)), dim3(256), 0, s, h->m, h->wPtr, h->wCol,
                           h->wVal, h->t2, h->wy);
        hipLaunchKernelGGL(k_gemv_part, dim3((h->m + 255) / 256, h->nChunks), dim3(256), 0, s, h->m, h->Cinv,
                           h->wy, h->part);
        hipLaunchKernelGGL(k_gemv_finish_sub, dim3(oecl_blocks(h->m, 256)), dim3(256), 0, s, h->m, h->nChunks,
                           h->part, h->sIdx, h->b);
    }
    // phi = R^-1 b, scattered into the extended grid
    hipLaunchKernelGGL(k_gemm_nn<0>, gg, gb, 0, s, bx, by, bx, h->Sx, bx, h->b, bx, h->t1, bx, nullptr, 0u,
                       nullptr);
    hipLaunchKernelGGL(k_gemm_nn<1>, gg, gb, 0, s, bx, by, by, h->t1, bx, h->Sy, by, h->b, bx, h->Linv, 0u,
                       nullptr);
    hipLaunchKernelGGL(k_gemm_nn<0>, gg, gb, 0, s, bx, by, bx, h->Sx, bx, h->b, bx, h->t1, bx, nullptr, 0u,
                       nullptr);
    hipLaunchKernelGGL(k_gemm_nn<2>, gg, gb, 0, s, bx, by, by, h->t1, bx, h->Sy, by, phi, 0, nullptr, h->nx,
                       h->mask);
    OECL_CHECK_LAUNCH("oecl_fastpoisson_solve");
    return OECL_OK;
}

OECL_EXPORT int oecl_fastpoisson_destroy(oecl_fastpoisson *h)
{
    if (!h) return OECL_OK;
    (void)hipFree(h->Sx);
    rocblas_destroy_handle(h->blas);
    delete h;
    return OECL_OK;
}
