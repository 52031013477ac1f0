// FIT Poisson solve on the device (poissonSolver.pyx:33-88).
//
// The reference assembles A = S~ M_eps S~^T restricted to inside nodes and back-substitutes
// a SuperLU factorisation.  Here A is applied matrix-free as a 5-point stencil on the full
// extended grid with edge coefficients cx, cy (non-inside nodes held at zero), and solved by
// Jacobi-preconditioned conjugate gradients in FP64.  Reductions are done block-wise in a
// fixed order, so a solve is bit-reproducible run to run.
#include "common.h"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    return v;
}

// Block sum of up to three values; results valid in thread 0.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double (*lds)[K])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.;
            for (int w = 0; w < kThreads / 64; ++w) s += lds[w][k];
            v[k] = s;
        }
    }
}

__device__ __forceinline__ double apply_a(const double *__restrict__ cx, const double *__restrict__ cy,
                                          const double *__restrict__ x, uint32_t p, uint32_t nx)
{
    double xp = x[p];
    return cx[p] * (xp - x[p + 1]) - cx[p - 1] * (x[p - 1] - xp) + cy[p] * (xp - x[p + nx]) -
           cy[p - nx] * (x[p - nx] - xp);
}

// r = q - A phi (inside), phi = 0 and r = 0 elsewhere; z = r / d; p = z.
// partials: [block][3] = (r.z, r.r, q.q)
__global__ void __launch_bounds__(kThreads)
    k_pcg_init(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
               const uint8_t *__restrict__ mask, const double *__restrict__ q, double *__restrict__ phi,
               double *__restrict__ r, double *__restrict__ z, double *__restrict__ pv,
               double *__restrict__ dinv, double *__restrict__ partials)
{
    __shared__ double lds[kThreads / 64][3];
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    double v[3] = {0., 0., 0.};
    if (i < np) {
        if (mask[i]) {
            double d = cx[i] + cx[i - 1] + cy[i] + cy[i - nx];
            double di = 1. / d;
            dinv[i] = di;
            double ri = q[i] - apply_a(cx, cy, phi, i, nx);
            double zi = ri * di;
            r[i] = ri;
            z[i] = zi;
            pv[i] = zi;
            v[0] = ri * zi;
            v[1] = ri * ri;
            v[2] = q[i] * q[i];
        } else {
            dinv[i] = 0.;
            r[i] = 0.;
            z[i] = 0.;
            pv[i] = 0.;
        }
    }
    block_sum<3>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) partials[3 * blockIdx.x + k] = v[k];
}

// y = A x on inside nodes with x taken as zero on the other nodes (the restriction of the
// operator to the inside unknowns, poissonSolver.pyx:55), y = 0 elsewhere.
__global__ void __launch_bounds__(kThreads)
    k_fit_apply(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
                const uint8_t *__restrict__ mask, const double *__restrict__ x, double *__restrict__ y)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= np) return;
    double a = 0.;
    if (mask[i]) {
        const double xp = x[i];
        const double xe = mask[i + 1] ? x[i + 1] : 0., xw = mask[i - 1] ? x[i - 1] : 0.;
        const double xn = mask[i + nx] ? x[i + nx] : 0., xs = mask[i - nx] ? x[i - nx] : 0.;
        a = cx[i] * (xp - xe) - cx[i - 1] * (xw - xp) + cy[i] * (xp - xn) - cy[i - nx] * (xs - xp);
    }
    y[i] = a;
}

// Zero phi on non-inside nodes before the first residual (runs before k_pcg_init).
__global__ void k_mask_phi(uint32_t np, const uint8_t *__restrict__ mask, double *__restrict__ phi)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np && !mask[i]) phi[i] = 0.;
}

// Scalars: s[0] = rz, s[1] = rr, s[2] = qq, s[3] = alpha, s[4] = beta, s[5] = pAp
__global__ void __launch_bounds__(kThreads)
    k_reduce_init(const double *__restrict__ partials, uint32_t nb, double *__restrict__ s)
{
    __shared__ double lds[kThreads / 64][3];
    double v[3] = {0., 0., 0.};
    for (uint32_t b = threadIdx.x; b < nb; b += kThreads)
        for (int k = 0; k < 3; ++k) v[k] += partials[3 * b + k];
    block_sum<3>(v, lds);
    if (threadIdx.x == 0) {
        s[0] = v[0];
        s[1] = v[1];
        s[2] = v[2];
    }
}

// ap = A p, partial p.ap
__global__ void __launch_bounds__(kThreads)
    k_pcg_ap(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
             const uint8_t *__restrict__ mask, const double *__restrict__ pv, double *__restrict__ ap,
             double *__restrict__ partials)
{
    __shared__ double lds[kThreads / 64][1];
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    double v[1] = {0.};
    if (i < np) {
        double a = 0.;
        if (mask[i]) {
            a = apply_a(cx, cy, pv, i, nx);
            v[0] = pv[i] * a;
        }
        ap[i] = a;
    }
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

__global__ void __launch_bounds__(kThreads)
    k_reduce_alpha(const double *__restrict__ partials, uint32_t nb, double *__restrict__ s)
{
    __shared__ double lds[kThreads / 64][1];
    double v[1] = {0.};
    for (uint32_t b = threadIdx.x; b < nb; b += kThreads) v[0] += partials[b];
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) {
        s[5] = v[0];
        s[3] = v[0] != 0. ? s[0] / v[0] : 0.;
    }
}

// phi += alpha p; r -= alpha ap; z = r * dinv; partial (r.z, r.r)
__global__ void __launch_bounds__(kThreads)
    k_pcg_update(uint32_t np, const double *__restrict__ s, double *__restrict__ phi,
                 double *__restrict__ r, double *__restrict__ z, const double *__restrict__ pv,
                 const double *__restrict__ ap, const double *__restrict__ dinv,
                 double *__restrict__ partials)
{
    __shared__ double lds[kThreads / 64][2];
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    double v[2] = {0., 0.};
    double alpha = s[3];
    if (i < np) {
        double ri = r[i] - alpha * ap[i];
        phi[i] += alpha * pv[i];
        double zi = ri * dinv[i];
        r[i] = ri;
        z[i] = zi;
        v[0] = ri * zi;
        v[1] = ri * ri;
    }
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = v[0];
        partials[2 * blockIdx.x + 1] = v[1];
    }
}

__global__ void __launch_bounds__(kThreads)
    k_reduce_beta(const double *__restrict__ partials, uint32_t nb, double *__restrict__ s)
{
    __shared__ double lds[kThreads / 64][2];
    double v[2] = {0., 0.};
    for (uint32_t b = threadIdx.x; b < nb; b += kThreads) {
        v[0] += partials[2 * b];
        v[1] += partials[2 * b + 1];
    }
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        s[4] = s[0] != 0. ? v[0] / s[0] : 0.;
        s[0] = v[0];
        s[1] = v[1];
    }
}

__global__ void k_pcg_dir(uint32_t np, const double *__restrict__ s, const double *__restrict__ z,
                          double *__restrict__ pv)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) pv[i] = z[i] + s[4] * pv[i];
}

__global__ void k_csr_spmv(uint32_t nRows, const int32_t *__restrict__ rowPtr,
                           const int32_t *__restrict__ col, const double *__restrict__ val,
                           const double *__restrict__ x, double *__restrict__ y)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nRows) return;
    double sum = 0.;
    for (int32_t k = rowPtr[r]; k < rowPtr[r + 1]; ++k) sum += val[k] * x[col[k]];
    y[r] = sum;
}

__global__ void k_axpy_rhs(uint32_t np, const double *rho, double scale, const double *prof,
                           double *out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) out[i] = prof ? rho[i] + scale * prof[i] : rho[i];
}

}  // namespace

OECL_EXPORT uint64_t oecl_pcg_work_bytes(uint32_t np)
{
    uint64_t nb = (np + kThreads - 1) / kThreads;
    return 8ull * (5ull * np + 3 * nb + 16) + 256;
}

OECL_EXPORT int oecl_pcg_solve(uint32_t nx, uint32_t np, const double *cx, const double *cy,
                               const uint8_t *mask, const double *q, double *phi, double tol,
                               uint32_t maxIter, void *work, uint32_t *iters, double *relres,
                               void *stream)
{
    if (!cx || !cy || !mask || !q || !phi || !work || np == 0 || nx < 3) return OECL_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    uint32_t nb = (np + kThreads - 1) / kThreads;
    double *r = (double *)work, *z = r + np, *pv = z + np, *ap = pv + np, *dinv = ap + np;
    double *partials = dinv + np, *s = partials + 3 * nb;
    hipLaunchKernelGGL(k_mask_phi, dim3(nb), dim3(kThreads), 0, st, np, mask, phi);
    hipLaunchKernelGGL(k_pcg_init, dim3(nb), dim3(kThreads), 0, st, nx, np, cx, cy, mask, q, phi, r,
                       z, pv, dinv, partials);
    hipLaunchKernelGGL(k_reduce_init, dim3(1), dim3(kThreads), 0, st, partials, nb, s);
    OECL_CHECK_LAUNCH("oecl_pcg_solve init");
    double hs[3];
    OECL_HIP(hipMemcpyAsync(hs, s, 3 * sizeof(double), hipMemcpyDeviceToHost, st), "pcg copy");
    OECL_HIP(hipStreamSynchronize(st), "pcg sync");
    double qn = sqrt(hs[2]);
    double target = tol * (qn > 0. ? qn : 1.);
    uint32_t it = 0;
    double rr = hs[1];
    const uint32_t check = 8;
    while (sqrt(rr) > target && it < maxIter) {
        for (uint32_t k = 0; k < check && it < maxIter; ++k, ++it) {
            hipLaunchKernelGGL(k_pcg_ap, dim3(nb), dim3(kThreads), 0, st, nx, np, cx, cy, mask, pv, ap,
                               partials);
            hipLaunchKernelGGL(k_reduce_alpha, dim3(1), dim3(kThreads), 0, st, partials, nb, s);
            hipLaunchKernelGGL(k_pcg_update, dim3(nb), dim3(kThreads), 0, st, np, s, phi, r, z, pv,
                               ap, dinv, partials);
            hipLaunchKernelGGL(k_reduce_beta, dim3(1), dim3(kThreads), 0, st, partials, nb, s);
            hipLaunchKernelGGL(k_pcg_dir, dim3(nb), dim3(kThreads), 0, st, np, s, z, pv);
        }
        OECL_CHECK_LAUNCH("oecl_pcg_solve iterate");
        OECL_HIP(hipMemcpyAsync(hs, s, 2 * sizeof(double), hipMemcpyDeviceToHost, st), "pcg copy");
        OECL_HIP(hipStreamSynchronize(st), "pcg sync");
        rr = hs[1];
        if (!(rr == rr)) break;  // NaN
    }
    if (iters) *iters = it;
    double rel = sqrt(rr) / (qn > 0. ? qn : 1.);
    if (relres) *relres = rel;
    if (!(rel <= tol)) return OECL_ERR_NOCONV;
    return OECL_OK;
}

OECL_EXPORT int oecl_fit_apply(uint32_t nx, uint32_t np, const double *cx, const double *cy, const uint8_t *mask,
                               const double *x, double *y, void *stream)
{
    if (!cx || !cy || !mask || !x || !y || nx < 3 || np < 3 * nx || np % nx) return OECL_ERR_ARG;
    hipLaunchKernelGGL(k_fit_apply, dim3(oecl_blocks(np, kThreads)), dim3(kThreads), 0, (hipStream_t)stream, nx,
                       np, cx, cy, mask, x, y);
    OECL_CHECK_LAUNCH("oecl_fit_apply");
    return OECL_OK;
}

OECL_EXPORT int oecl_csr_spmv(uint32_t nRows, const int32_t *rowPtr, const int32_t *col,
                              const double *val, const double *x, double *y, void *stream)
{
    if (!rowPtr || !col || !val || !x || !y) return OECL_ERR_ARG;
    if (nRows == 0) return OECL_OK;
    hipLaunchKernelGGL(k_csr_spmv, dim3(oecl_blocks(nRows, 256)), dim3(256), 0, (hipStream_t)stream,
                       nRows, rowPtr, col, val, x, y);
    OECL_CHECK_LAUNCH("oecl_csr_spmv");
    return OECL_OK;
}

OECL_EXPORT int oecl_axpy_rhs(uint32_t np, const double *rho, double scale, const double *profile,
                              double *out, void *stream)
{
    if (!rho || !out) return OECL_ERR_ARG;
    hipLaunchKernelGGL(k_axpy_rhs, dim3(oecl_blocks(np, 256)), dim3(256), 0, (hipStream_t)stream, np,
                       rho, scale, profile, out);
    OECL_CHECK_LAUNCH("oecl_axpy_rhs");
    return OECL_OK;
}
