// Direct FIT Poisson solve by the capacitance-matrix method.
//
// The operator A (poissonSolver.pyx:44-59) acts on the inside nodes, all of which lie in
// the box B of nodes 2 <= i < 2 + bx, 2 <= j <= nyExt-3 (bx x by nodes, column-major with
// x fastest; bx = 2^p - 1 >= nxExt - 4, the columns beyond nxExt - 3 are padding).  Let R be the box's uniform 5-point operator (x-edge coefficient cxu, y-edge
// cyu, zero Dirichlet data outside B) and M the box operator whose rows are A's rows on
// inside nodes and R's rows elsewhere.  M x = b with b = q on inside nodes reproduces
// A phi = q exactly on inside nodes (the inside rows never reference a non-inside node).  M
// and R differ only in the m rows S of inside nodes next to the wall: M = R + U W with U the
// unit columns of S.  By the Woodbury identity
//     x = R^-1 (b - U z),   z = C^-1 W R^-1 b,   C = I_m + W R^-1 U.
//
// R = cxu Tx (x) I + cyu I (x) Ty with T = tridiag(-1, 2, -1).  The type-I DST Sx
// (orthonormal, symmetric, Sx^2 = I) diagonalises Tx, so in the x-transformed space
// (hat b = Sx b) R splits into bx independent tridiagonal systems along y, one per x mode q:
//     (cxu lx_q I + cyu Ty) hat u[q, :] = hat b[q, :],   lx_q = 2 - 2 cos(pi (q+1)/(bx+1)),
// solved exactly by Thomas elimination with precomputed pivots.  The whole solve stays in
// the transformed space:
//     hat b = Sx b                                        (FFT of length bx + 1 per line)
//     hat y = T^-1 hat b                                  (Thomas, per x mode)
//     (W y)_a = sum_k W[a,k] (Sx hat y)[k]               (k near the wall: short dot products)
//     z = C^-1 W y ;  hat b -= Sx[:, i_a] z_a  for each a (column j_a of hat b)
//     hat x = T^-1 hat b ;  x = Sx hat x                  (transform fused with the scatter)
// Two transforms per solve.  C is built once on the device with the same kernels
// (batched unit right-hand sides), so that the correction matches the operator applied,
// and inverted with rocSOLVER.
#include <math.h>

#include <algorithm>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "common.h"
#include "fastpoisson_internal.h"

struct oecl_fastpoisson {
    uint32_t nx, ny, np;      // extended grid
    uint32_t bx, by, nb;      // box sizes, nb = bx * by; bx + 1 a power of two
    uint32_t bxg;             // interior grid columns nx - 4 (box columns beyond are padding)
    uint32_t m;               // irregular rows
    uint32_t ldc;             // m rounded up to even: row stride of C^-1 and of wy
    uint32_t nJ;              // distinct box columns j among the irregular nodes
    double cyu;
    rocblas_handle blas;
    double *Sx;               // bx * bx (symmetric)
    double2 *tab;             // 2 (bx + 1): e^{-i pi k / (bx + 1)}, the FFT's twiddles
    double *invw;             // nb: Thomas pivots 1 / w[q, j] of mode q's y system
    int32_t *sIdx;            // [m] box index of row a
    int32_t *wPtr, *wCol;     // CSR of W over box columns, m rows
    double *wVal;
    int32_t *jCol, *jStart;   // distinct columns j of the irregular nodes, their entry ranges
    int32_t *ePos;            // [m] entry of row a in the by-column order
    uint32_t *iPos;           // [m] box column i_a of entry e (by-column order)
    double *CinvT;            // m*ldc, C^-1 row-major (row a contiguous, 16-byte aligned, padding zero)
    double *b, *t1, *t2;      // nb each, per right-hand side
    double *wy, *zE;          // ldc (padding entry zero), m (zE in the by-column order), per right-hand side
    const uint8_t *mask;      // inside flags of the extended grid (caller owned)
    uint32_t rCap;            // right-hand sides the vectors above hold
    char *batchMem;           // their own allocation once rCap > 1
    uint32_t pipeGrid[5];     // blocks of k_dst_fft_pipe<EPI> that are resident at once (0: not used)
};

namespace {

// ------------------------------------------------------------------ per-solve kernels
// The orthonormal DST-I of order n = M - 1 (M = bx + 1, a power of two: the box is padded to
// that width), S_k = sqrt(2/M) sum_{j=1..n} v_j sin(pi j k / M), k = 1..n, through one complex
// FFT of length M per line:
//   y (length 2M) = odd extension of v: y_0 = y_M = 0, y_j = v_j, y_{2M-j} = -v_j, whose DFT
//   is Y_k = -2 i S_k;  z_m = y_{2m} + i y_{2m+1} (m < M), Z = FFT_M(z);
//   Y_k = (Z_k + conj Z_{M-k}) / 2 + w^k (Z_k - conj Z_{M-k}) / (2 i),  w = e^{-i pi / M},
// so that
//   S_k = -Im(Y_k) / 2 = -(1/4) [(Im Z_k - Im Z_{M-k}) - cos(pi k/M) (Re Z_k - Re Z_{M-k})
//                                - sin(pi k/M) (Im Z_k + Im Z_{M-k})].
// The FFT is a Stockham autosort in LDS (radix-4 passes, one radix-2 pass for odd log2 M), each
// pass reading all of a thread's inputs into registers before a barrier and writing after it,
// so one buffer of M complex values per line suffices.  Twiddles come from a table of
// e^{-i pi k / M}, k < 2M, built on the host in extended precision.  Error O(eps log M)
// against O(eps sqrt(M)) for the dense product.  T threads per line, LB lines per block of NT
// threads.
// EPI 0 stores the transform of line l at C + l * ldc; EPI 2 scatters it into the
// extended-grid potential: line l is row l % by of potential l / by (np nodes apart), only the
// box columns that are interior grid columns (k < bxg) are written, masked.
#ifndef OECL_FFT_BIG_T
#define OECL_FFT_BIG_T 256
#endif
template <int M>
struct FftShape {
    static constexpr int T = M >= 2048 ? OECL_FFT_BIG_T : (M >= 1024 ? 256 : (M >= 4 ? M / 4 : 1));
    static constexpr int LB = T >= 256 ? 1 : 256 / T;
    static constexpr int NT = T * LB;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// z value m of the packed odd extension of v (n = M - 1 values staged in LDS)
template <int M>
__device__ __forceinline__ double2 packed_odd(const double *v, uint32_t m)
{
    const uint32_t t0 = 2 * m, t1 = 2 * m + 1;
    const double y0 = t0 == 0 ? 0. : (t0 < (uint32_t)M ? v[t0 - 1] : (t0 == (uint32_t)M ? 0. : -v[2 * M - t0 - 1]));
    const double y1 = t1 < (uint32_t)M ? v[t1 - 1] : -v[2 * M - t1 - 1];
    return make_double2(y0, y1);
}

// LDS place of complex entry e of a line: the low 3 bits XOR-ed with bits 3-5, a permutation
// within each block of 64 entries that spreads the strided writes of a pass (entries R or
// 8 R apart in consecutive lanes) over all 8 16-byte bank groups, while the contiguous reads
// stay spread too.
#ifndef OECL_FFT_SWIZZLE
#define OECL_FFT_SWIZZLE 1
#endif
__device__ __forceinline__ uint32_t zsw(uint32_t e)
{
    return OECL_FFT_SWIZZLE ? e ^ ((e >> 3) & 7u) : e;
}

// 4-point DFT (e^{-2 pi i / 4} = -i) of a, b, c, d into X[0..3].
__device__ __forceinline__ void dft4(double2 a, double2 b, double2 c, double2 d, double2 *X)
{
    const double2 t0 = make_double2(a.x + c.x, a.y + c.y), t1 = make_double2(a.x - c.x, a.y - c.y);
    const double2 t2 = make_double2(b.x + d.x, b.y + d.y), t3 = make_double2(b.x - d.x, b.y - d.y);
    X[0] = make_double2(t0.x + t2.x, t0.y + t2.y);
    X[1] = make_double2(t1.x + t3.y, t1.y - t3.x);
    X[2] = make_double2(t0.x - t2.x, t0.y - t2.y);
    X[3] = make_double2(t1.x - t3.y, t1.y + t3.x);
}

// In-place R-point DFT (R = 2, 4, 8) of x[0..R-1].
template <int R>
__device__ __forceinline__ void bfly(double2 *x)
{
    if constexpr (R == 8) {
        // two 4-point DFTs of the even and odd inputs, combined with w8^k = e^{-i pi k / 4}
        double2 e[4], o[4];
        dft4(x[0], x[2], x[4], x[6], e);
        dft4(x[1], x[3], x[5], x[7], o);
        constexpr double h = 0.70710678118654752440;   // sqrt(1/2)
        const double2 o1 = make_double2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));   // (1 - i) h o1
        const double2 o2 = make_double2(o[2].y, -o[2].x);                                  // -i o2
        const double2 o3 = make_double2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));  // (-1 - i) h o3
        x[0] = make_double2(e[0].x + o[0].x, e[0].y + o[0].y);
        x[4] = make_double2(e[0].x - o[0].x, e[0].y - o[0].y);
        x[1] = make_double2(e[1].x + o1.x, e[1].y + o1.y);
        x[5] = make_double2(e[1].x - o1.x, e[1].y - o1.y);
        x[2] = make_double2(e[2].x + o2.x, e[2].y + o2.y);
        x[6] = make_double2(e[2].x - o2.x, e[2].y - o2.y);
        x[3] = make_double2(e[3].x + o3.x, e[3].y + o3.y);
        x[7] = make_double2(e[3].x - o3.x, e[3].y - o3.y);
    } else if constexpr (R == 4) {
        double2 X[4];
        dft4(x[0], x[1], x[2], x[3], X);
        x[0] = X[0];
        x[1] = X[1];
        x[2] = X[2];
        x[3] = X[3];
    } else {
        const double2 a = x[0], b = x[1];
        x[0] = make_double2(a.x + b.x, a.y + b.y);
        x[1] = make_double2(a.x - b.x, a.y - b.y);
    }
}

// One Stockham pass of radix R at stride NS: out[(j / NS) NS R + j % NS + r NS] =
// DFT_R(in[j + r M / R] w^{r (j % NS)}), w = e^{-2 pi i / (NS R)}; the first pass (NS = 1)
// reads the packed odd extension of v instead of z.
template <int M, int R, int NS>
__device__ __forceinline__ void fft_pass(double2 *z, const double *v, uint32_t t, const double2 *__restrict__ tab)
{
    constexpr int T = FftShape<M>::T, J = M / (R * T);
    static_assert(J >= 1 && J * R * T == M, "every thread takes whole butterflies of each pass");
    double2 x[J][R];
#pragma unroll
    for (int q = 0; q < J; ++q) {
        const uint32_t j = t + q * T;
#pragma unroll
        for (int r = 0; r < R; ++r) x[q][r] = NS == 1 ? packed_odd<M>(v, j + r * (M / R)) : z[zsw(j + r * (M / R))];
        if (NS > 1) {
            const uint32_t jm = j & (NS - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) x[q][r] = cmul(x[q][r], tab[2 * r * jm * (M / (NS * R))]);
        }
        bfly<R>(x[q]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < J; ++q) {
        const uint32_t j = t + q * T;
        const uint32_t o = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
        for (int r = 0; r < R; ++r) z[zsw(o + r * NS)] = x[q][r];
    }
    __syncthreads();
}

#ifndef OECL_FFT_RADIX8
#define OECL_FFT_RADIX8 0
#endif
// Radix 4 passes, then one radix 2 (with OECL_FFT_RADIX8, radix 8 first while 8 divides what is
// left and every thread takes whole 8-point butterflies: fewer passes, but measured 5 % slower
// for M = 2048 in this one-line-per-block form).
template <int M, int NS>
__device__ __forceinline__ void fft_passes(double2 *z, const double *v, uint32_t t, const double2 *__restrict__ tab)
{
    if constexpr (NS < M) {
        constexpr int T = FftShape<M>::T;
        constexpr int R = (OECL_FFT_RADIX8 && (M / NS) % 8 == 0 && M % (8 * T) == 0) ? 8
                          : ((M / NS) % 4 == 0 ? 4 : 2);
        fft_pass<M, R, NS>(z, v, t, tab);
        fft_passes<M, NS * R>(z, v, t, tab);
    }
}

struct FftArgs {
    uint32_t nLines;
    const double *X;            // EPI 0, 2: line l at X + l * ldx; EPI 3: zE (m per right-hand side)
    uint32_t ldx;
    double *C;                  // EPI 0, 3: line l at C + l * ldc (EPI 3: line jCol[c] of side r); EPI 2: phi
    uint32_t ldc;
    const double2 *tab;
    uint32_t nx, by, np, bxg;   // EPI 2: the extended grid and its interior columns
    const uint8_t *mask;
    const int32_t *jCol, *jStart;   // EPI 3: distinct box rows of the corrected nodes, entry ranges
    const uint32_t *iPos;           // EPI 3: box column of entry e
    uint32_t nJ, m;
    uint64_t nb;
    // EPI 4: the right-hand side gathered from the grid, q = rho + s(step) * prof on the inside
    // nodes (rho per run, np apart; s from the per-run table at the device step counter)
    double *rho;
    const double *prof, *table;
    uint32_t nScale;
    const uint32_t *stepDev;
};

// EPI 3 transforms the sparse lines u_c (z values of the entries of box row jCol[c] at their
// columns, zero elsewhere; line l = r nJ + c for right-hand side r) and subtracts the result
// from row jCol[c] of C: hat b -= Sx U z in the transformed space.  X and C may be the same
// buffer (EPI 0 in place): a line is read whole into LDS before any of it is written.
template <int M, int EPI>
__global__ void __launch_bounds__(FftShape<M>::NT) k_dst_fft(FftArgs g)
{
    constexpr int T = FftShape<M>::T, LB = FftShape<M>::LB, n = M - 1;
    __shared__ double2 sh[LB * M];
    const uint32_t t = threadIdx.x % T, sub = threadIdx.x / T;
    const uint32_t line = blockIdx.x * LB + sub;
    const bool live = line < g.nLines;
    double2 *z = sh + (size_t)sub * M;
    double *v = reinterpret_cast<double *>(z) + M;   // the upper half of the line's buffer
    uint32_t r = 0, c = 0;
    if (EPI == 3) {
        r = live ? line / g.nJ : 0;
        c = live ? line - r * g.nJ : 0;
        for (uint32_t k = t; k < (uint32_t)n; k += T) v[k] = 0.;
        __syncthreads();
        if (live) {
            const double *ze = g.X + (size_t)r * g.m;
            for (int32_t e = g.jStart[c] + (int32_t)t; e < g.jStart[c + 1]; e += T) v[g.iPos[e]] = ze[e];
        }
    } else if (EPI == 4) {
        // line = box row jj of run r: grid row jj + 2; box column k is grid column k + 2.  The
        // row's rho is read into the line (masked, plus the beam term) and zeroed for the
        // step's deposit, every node of the grid row (each thread zeroes what it read).
        const uint32_t rr = live ? line / g.by : 0, jj = live ? line - rr * g.by : 0;
        const uint32_t s = *g.stepDev;
        const double sc = (g.prof && g.table && s < g.nScale) ? g.table[(size_t)rr * g.nScale + s] : 0.;
        double *rrow = g.rho + (size_t)rr * g.np + (size_t)(jj + 2) * g.nx;
        const double *prow = g.prof ? g.prof + (size_t)(jj + 2) * g.nx : nullptr;
        const uint8_t *mrow = g.mask + (size_t)(jj + 2) * g.nx;
        for (uint32_t k = t; k < (uint32_t)n; k += T) {
            double q = 0.;
            if (live && k < g.bxg) {
                const double r0 = rrow[k + 2];
                q = mrow[k + 2] ? (sc != 0. ? r0 + sc * prow[k + 2] : r0) : 0.;
                rrow[k + 2] = 0.;
            }
            v[k] = q;
        }
        if (live)
            for (uint32_t c = t; c < 4; c += T) rrow[c < 2 ? c : g.nx - 4 + c] = 0.;   // grid columns 0, 1, nx-2, nx-1
    } else {
        const double *xl = g.X + (size_t)(live ? line : 0) * g.ldx;
        for (uint32_t k = t; k < (uint32_t)n; k += T) v[k] = live ? xl[k] : 0.;
    }
    __syncthreads();
    fft_passes<M, 1>(z, v, t, g.tab);
    if (!live) return;
    const double sc = -0.25 * sqrt(2. / M);
    double *out = g.C;
    uint32_t run = 0, jj = 0;
    if (EPI == 0 || EPI == 4) out += (size_t)line * g.ldc;
    if (EPI == 2) {
        run = line / g.by;
        jj = line - run * g.by;
        out += (size_t)run * g.np;
    }
    if (EPI == 3) out += (size_t)r * g.nb + (size_t)g.jCol[c] * g.ldc;
    for (uint32_t k = t; k < (uint32_t)n; k += T) {
        const double2 a = z[zsw(k + 1)], b = z[zsw(M - 1 - k)], w = g.tab[k + 1];
        // w = (cos(pi (k+1)/M), -sin(pi (k+1)/M))
        const double s = sc * (((a.y - b.y) - w.x * (a.x - b.x)) + w.y * (a.y + b.y));
        if (EPI == 0 || EPI == 4) {
            out[k] = s;
        } else if (EPI == 3) {
            out[k] -= s;
        } else if (k < g.bxg) {
            const uint32_t p = (k + 2) + (jj + 2) * g.nx;
            out[p] = g.mask[p] ? s : 0.;
        }
    }
}


// ---- the persistent, pipelined form for M = 2048 (256 threads a line, 8 entries a thread)
// Block b transforms lines b, b + G, b + 2G, ... (G = gridDim.x).  The next line's entries are
// loaded into registers while the current line is transformed, so that a line's HBM latency
// hides behind the previous line's passes.  For that the passes issue no global load: their
// twiddles are powers of one base per pass and butterfly (w^2 = w w, ...: a few ulp from the
// table values), and the epilogue's e^{-i pi (k+1)/M} for k = t + i T are e^{-i pi (t+1)/M}
// times i factors e^{-i pi T/M} = e^{-i pi / 8}, all formed from table entries loaded before the
// first line.  Radix 8, 8, 8, 4.
#ifndef OECL_FFT_PIPE
#define OECL_FFT_PIPE 1
#endif

template <int R>
__device__ __forceinline__ void apply_powers(double2 *x, double2 w)
{
    if constexpr (R == 2) {
        x[1] = cmul(x[1], w);
    } else if constexpr (R == 4) {
        const double2 w2 = cmul(w, w), w3 = cmul(w2, w);
        x[1] = cmul(x[1], w);
        x[2] = cmul(x[2], w2);
        x[3] = cmul(x[3], w3);
    } else {
        const double2 w2 = cmul(w, w), w3 = cmul(w2, w), w4 = cmul(w2, w2);
        x[1] = cmul(x[1], w);
        x[2] = cmul(x[2], w2);
        x[3] = cmul(x[3], w3);
        x[4] = cmul(x[4], w4);
        x[5] = cmul(x[5], cmul(w4, w));
        x[6] = cmul(x[6], cmul(w4, w2));
        x[7] = cmul(x[7], cmul(w4, w3));
    }
}

// A pass as fft_pass, its twiddle bases bw[q] (q < J) given.
template <int M, int T, int R, int NS>
__device__ __forceinline__ void fft_pass_reg(double2 *z, const double *v, uint32_t t, const double2 *bw)
{
    constexpr int J = M / (R * T);
    static_assert(J >= 1 && J * R * T == M, "every thread takes whole butterflies of each pass");
    double2 x[J][R];
#pragma unroll
    for (int q = 0; q < J; ++q) {
        const uint32_t j = t + q * T;
#pragma unroll
        for (int r = 0; r < R; ++r) x[q][r] = NS == 1 ? packed_odd<M>(v, j + r * (M / R)) : z[zsw(j + r * (M / R))];
        if (NS > 1) apply_powers<R>(x[q], bw[q]);
        bfly<R>(x[q]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < J; ++q) {
        const uint32_t j = t + q * T;
        const uint32_t o = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
        for (int r = 0; r < R; ++r) z[zsw(o + r * NS)] = x[q][r];
    }
    __syncthreads();
}

// EPI 0, 2, 3 and 4 as k_dst_fft.
template <int EPI>
__global__ void __launch_bounds__(256) k_dst_fft_pipe(FftArgs g)
{
    constexpr int M = 2048, T = 256, P = M / T, n = M - 1;
    __shared__ double2 z[M];
    double *v = reinterpret_cast<double *>(z) + M;
    const uint32_t t = threadIdx.x;
    const uint32_t step = EPI == 4 ? *g.stepDev : 0u;
    const double sc0 = -0.25 * sqrt(2. / M);
    // a line's inputs as loaded, one line ahead: pv the entries (EPI 0, 2: X; EPI 4: rho; EPI 3:
    // the row of C it updates), pp the beam profile (EPI 4), pm the inside flags (EPI 2, 4);
    // EPI 3: this thread's sparse entry (column pi, value pz) among the line's ne
    double pv[P], pp[P];
    uint32_t pm[P];
    double pscale = 0., pz = 0.;
    uint32_t pi = 0, ne = 0;
    auto fetch = [&](uint32_t line) {
        if (EPI == 4) {
            const uint32_t rr = line / g.by, jj = line - rr * g.by;
            pscale = (g.prof && g.table && step < g.nScale) ? g.table[(size_t)rr * g.nScale + step] : 0.;
            double *rrow = g.rho + (size_t)rr * g.np + (size_t)(jj + 2) * g.nx;
            const double *prow = g.prof ? g.prof + (size_t)(jj + 2) * g.nx : nullptr;
            const uint8_t *mrow = g.mask + (size_t)(jj + 2) * g.nx;
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const uint32_t k = t + i * T;
                pv[i] = 0.;
                pp[i] = 0.;
                pm[i] = 0u;
                if (k < g.bxg) {
                    pv[i] = rrow[k + 2];
                    pm[i] = mrow[k + 2];
                    if (prow && pscale != 0.) pp[i] = prow[k + 2];
                    rrow[k + 2] = 0.;
                }
            }
            if (t < 4) rrow[t < 2 ? t : g.nx - 4 + t] = 0.;   // grid columns 0, 1, nx-2, nx-1
        } else if (EPI == 3) {
            const uint32_t r = line / g.nJ, c = line - r * g.nJ;
            const int32_t e0 = g.jStart[c], e1 = g.jStart[c + 1];
            ne = (uint32_t)(e1 - e0);
            if (t < ne) {
                pi = g.iPos[e0 + t];
                pz = g.X[(size_t)r * g.m + e0 + t];
            }
            const double *orow = g.C + (size_t)r * g.nb + (size_t)g.jCol[c] * g.ldc;
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const uint32_t k = t + i * T;
                pv[i] = k < (uint32_t)n ? orow[k] : 0.;
            }
        } else {
            const double *xl = g.X + (size_t)line * g.ldx;
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const uint32_t k = t + i * T;
                pv[i] = k < (uint32_t)n ? xl[k] : 0.;
            }
            if (EPI == 2) {
                const uint32_t jj = line - (line / g.by) * g.by;
                const uint8_t *mrow = g.mask + (size_t)(jj + 2) * g.nx + 2;
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    const uint32_t k = t + i * T;
                    pm[i] = k < g.bxg ? mrow[k] : 0u;
                }
            }
        }
    };
    const uint32_t G = gridDim.x;
    if (blockIdx.x < g.nLines) fetch(blockIdx.x);
    for (uint32_t line = blockIdx.x; line < g.nLines; line += G) {
        // the thread index as an opaque value in each iteration: the LDS addresses derived from
        // it are formed again in each line rather than held in registers across the loop
        uint32_t t = threadIdx.x;
        asm volatile("" : "+v"(t));
        // the pass twiddle bases e^{-2 pi i jm / (NS R)} = tab[2 jm M / (NS R)] (jm = j % NS)
        // and the epilogue's e^{-i pi (t+1)/M}, loaded in each line before its prefetch (L2 hits;
        // the passes wait for them, not for the prefetch behind them)
        const double2 b8[1] = {g.tab[64 * (t & 7u)]}, b64[1] = {g.tab[8 * (t & 63u)]};
        const double2 b512[2] = {g.tab[2 * t], g.tab[2 * (t + 256u)]}, eb = g.tab[t + 1];
        // this line's values the epilogue needs (the prefetch below overwrites the registers)
        double co[EPI == 3 ? P : 1];
        uint32_t cm[EPI == 2 ? P : 1];
        if (EPI == 3) {
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const uint32_t k = t + i * T;
                co[i] = pv[i];
                if (k < (uint32_t)n) v[k] = 0.;
            }
            __syncthreads();
            if (t < ne) v[pi] = pz;
            if (ne > (uint32_t)T) {   // block-uniform: the entries beyond one per thread
                const uint32_t r = line / g.nJ, c = line - r * g.nJ;
                const double *ze = g.X + (size_t)r * g.m;
                for (int32_t e = g.jStart[c] + T + (int32_t)t; e < g.jStart[c + 1]; e += T) v[g.iPos[e]] = ze[e];
            }
        } else {
            const double scl = pscale;
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const uint32_t k = t + i * T;
                double q = pv[i];
                if (EPI == 4) q = pm[i] ? (scl != 0. ? pv[i] + scl * pp[i] : pv[i]) : 0.;
                if (k < (uint32_t)n) v[k] = q;
                if (EPI == 2) cm[i] = pm[i];
            }
        }
        __syncthreads();
        if (line + G < g.nLines) fetch(line + G);
        fft_pass_reg<M, T, 8, 1>(z, v, t, nullptr);
        fft_pass_reg<M, T, 8, 8>(z, v, t, b8);
        fft_pass_reg<M, T, 8, 64>(z, v, t, b64);
        fft_pass_reg<M, T, 4, 512>(z, v, t, b512);
        double *out = g.C;
        uint32_t jj = 0;
        if (EPI == 0 || EPI == 4) {
            out += (size_t)line * g.ldc;
        } else if (EPI == 3) {
            const uint32_t r = line / g.nJ, c = line - r * g.nJ;
            out += (size_t)r * g.nb + (size_t)g.jCol[c] * g.ldc;
        } else {
            const uint32_t run = line / g.by;
            jj = line - run * g.by;
            out += (size_t)run * g.np;
        }
        double2 w = eb;   // e^{-i pi (k+1)/M}, advanced by e^{-i pi / 8} from one k to the next
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const uint32_t k = t + i * T;
            if (i) w = cmul(w, make_double2(0.92387953251128675613, -0.38268343236508977173));
            if (k < (uint32_t)n) {
                const double2 a = z[zsw(k + 1)], b = z[zsw(M - 1 - k)];
                const double s = sc0 * (((a.y - b.y) - w.x * (a.x - b.x)) + w.y * (a.y + b.y));
                if (EPI == 0 || EPI == 4) {
                    out[k] = s;
                } else if (EPI == 3) {
                    out[k] = co[i] - s;
                } else if (k < g.bxg) {
                    out[(k + 2) + (jj + 2) * g.nx] = cm[i] ? s : 0.;
                }
            }
        }
        __syncthreads();   // the line's LDS reads are done before the next line is staged
    }
}

// b = masked charge on the box; the potential outside the box is zero.
__global__ void k_gather_box_zero(uint32_t nx, uint32_t ny, uint32_t bx, const uint8_t *__restrict__ mask,
                                  const double *__restrict__ q, double *__restrict__ b, double *__restrict__ phi)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nx * ny) return;
    uint32_t i = p % nx, j = p / nx;
    if (i < 2 || j < 2 || i >= nx - 2 || j >= ny - 2) phi[p] = 0.;
    else b[(i - 2) + (j - 2) * bx] = mask[p] ? q[p] : 0.;
}

// Thomas elimination along y for every x mode q, one wave per q:
//     g_j = (r_j + cyu g_{j-1}) / w_j   (g_{-1} = 0),
//     u_j = g_j + (cyu / w_j) u_{j+1}   (u_{by} = 0),   w_0 = d_q, w_j = d_q - cyu^2 / w_{j-1},
// with the pivots 1 / w precomputed.  Both sweeps are affine recurrences x_j = a_j x_{j-1} + b_j,
// evaluated as a prefix scan of affine maps: lane L holds the C consecutive rows
// [C L, C L + C); it composes its rows' maps, a shuffle scan over the 64 lanes gives every
// lane its incoming value, and the lane then runs its rows.  The order of operations is
// fixed, so results are reproducible.  UNIT: the right-hand side is Sx e_s for the irregular
// node s = sIdx[a0 + blockIdx.y] (row j_s = Sx[:, i_s], zero elsewhere), for building the
// capacitance matrix with exactly the operator the solve applies.  Batch element blockIdx.y
// (a unit column, or a right-hand side of a batched solve) is read from r + blockIdx.y * nb
// and written at u + blockIdx.y * nb.
template <int C, bool UNIT>
__global__ void __launch_bounds__(256)
    k_thomas_y(uint32_t bx, uint32_t by, double cyu, const double *__restrict__ invw,
               const double *__restrict__ r, double *__restrict__ u, const int32_t *__restrict__ sIdx,
               uint32_t a0, const double *__restrict__ Sx)
{
    const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (q >= bx) return;
    const size_t nb = (size_t)bx * by;
    double *ub = u + (size_t)blockIdx.y * nb;
    const double *rb = UNIT ? r : r + (size_t)blockIdx.y * nb;
    uint32_t js = 0;
    double rs = 0.;
    if (UNIT) {
        const uint32_t s = (uint32_t)sIdx[a0 + blockIdx.y];
        js = s / bx;
        rs = Sx[q + (size_t)(s % bx) * bx];
    }
    const uint32_t j0 = lane * C;
    double iw[C], g[C];
#pragma unroll
    for (int t = 0; t < C; ++t) {
        const uint32_t j = j0 + t;
        const size_t o = q + (size_t)j * bx;
        iw[t] = j < by ? invw[o] : 0.;
        double rv = 0.;
        if (j < by) rv = UNIT ? (j == js ? rs : 0.) : rb[o];
        g[t] = rv;
    }
    // forward: g_j = a_j g_{j-1} + b_j with a_j = cyu iw_j, b_j = r_j iw_j
    double A = 1., B = 0.;
#pragma unroll
    for (int t = 0; t < C; ++t) {
        const double a = cyu * iw[t], b = g[t] * iw[t];
        A *= a;
        B = fma(a, B, b);
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double Ao = __shfl_up(A, d, 64), Bo = __shfl_up(B, d, 64);
        if (lane >= (uint32_t)d) {
            B = fma(A, Bo, B);
            A *= Ao;
        }
    }
    double x = __shfl_up(B, 1, 64);
    if (lane == 0) x = 0.;
#pragma unroll
    for (int t = 0; t < C; ++t) {
        x = fma(cyu * iw[t], x, g[t] * iw[t]);
        g[t] = x;
    }
    // backward: u_j = e_j u_{j+1} + g_j with e_j = cyu iw_j (rows past by: e = g = 0)
    A = 1.;
    B = 0.;
#pragma unroll
    for (int t = C - 1; t >= 0; --t) {
        const double e = cyu * iw[t];
        B = fma(e, B, g[t]);
        A *= e;
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double Ao = __shfl_down(A, d, 64), Bo = __shfl_down(B, d, 64);
        if (lane + d < 64) {
            B = fma(A, Bo, B);
            A *= Ao;
        }
    }
    x = __shfl_down(B, 1, 64);
    if (lane == 63) x = 0.;
#pragma unroll
    for (int t = C - 1; t >= 0; --t) {
        x = fma(cyu * iw[t], x, g[t]);
        const uint32_t j = j0 + t;
        if (j < by) ub[q + (size_t)j * bx] = x;
    }
}

// The same Thomas solve for batched right-hand sides with coalesced loads: a block takes QB
// consecutive x modes (lanes run over the modes, so a row's QB entries are one load) and
// splits the by rows into S contiguous segments of L rows, one per thread.  Each thread
// composes its segment's affine recurrence, an LDS scan over the S segments of each mode
// gives every segment its entry value, and the segment is then applied from registers; the
// backward sweep is done the same way from the other end.  Block QB * S threads (QB, S
// powers of two, S * L >= by).  Deterministic (fixed segments and scan order).
template <int L>
__global__ void __launch_bounds__(1024)
    k_thomas_seg(uint32_t bx, uint32_t by, uint32_t lqb, uint32_t S, double cyu, const double *__restrict__ invw,
                 const double *__restrict__ r, double *__restrict__ u)
{
    __shared__ double sA[1024], sB[1024];
    const uint32_t t = threadIdx.x, QB = 1u << lqb, qq = t & (QB - 1), sg = t >> lqb;
    // blocks whose modes share a 128-byte row piece (16 modes) run on one XCD, so its L2 holds
    // the line for all of them
    const uint32_t q = xcd_block_chunk(blockIdx.x, gridDim.x, QB < 16 ? 16 / QB : 1) * QB + qq;
    const bool live = q < bx;
    const size_t nb = (size_t)bx * by;
    const double *rb = r + (size_t)blockIdx.y * nb;
    double *ub = u + (size_t)blockIdx.y * nb;
    const uint32_t j0 = sg * L;
    double iw[L], g[L];
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const uint32_t j = j0 + k;
        const size_t o = q + (size_t)j * bx;
        const bool in = live && j < by;
        iw[k] = in ? invw[o] : 0.;
        g[k] = in ? rb[o] : 0.;
    }
    // forward: g_j = a_j g_{j-1} + b_j with a_j = cyu iw_j, b_j = r_j iw_j; a segment maps its
    // entry value x to A x + B
    double A = 1., B = 0.;
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const double a = cyu * iw[k];
        A *= a;
        B = fma(a, B, g[k] * iw[k]);
    }
    sA[t] = A;
    sB[t] = B;
    __syncthreads();
    for (uint32_t d = 1; d < S; d <<= 1) {
        double Ao = 1., Bo = 0.;
        if (sg >= d) {
            Ao = sA[t - (d << lqb)];
            Bo = sB[t - (d << lqb)];
        }
        __syncthreads();
        if (sg >= d) {
            B = fma(A, Bo, B);
            A *= Ao;
            sA[t] = A;
            sB[t] = B;
        }
        __syncthreads();
    }
    double x = sg ? sB[t - QB] : 0.;   // g before the first row is zero
    __syncthreads();
#pragma unroll
    for (int k = 0; k < L; ++k) {
        x = fma(cyu * iw[k], x, g[k] * iw[k]);
        g[k] = x;
    }
    // backward: u_j = e_j u_{j+1} + g_j with e_j = cyu iw_j (rows past by: e = g = 0)
    A = 1.;
    B = 0.;
#pragma unroll
    for (int k = L - 1; k >= 0; --k) {
        const double e = cyu * iw[k];
        B = fma(e, B, g[k]);
        A *= e;
    }
    sA[t] = A;
    sB[t] = B;
    __syncthreads();
    for (uint32_t d = 1; d < S; d <<= 1) {
        double Ao = 1., Bo = 0.;
        if (sg + d < S) {
            Ao = sA[t + (d << lqb)];
            Bo = sB[t + (d << lqb)];
        }
        __syncthreads();
        if (sg + d < S) {
            B = fma(A, Bo, B);
            A *= Ao;
            sA[t] = A;
            sB[t] = B;
        }
        __syncthreads();
    }
    x = sg + 1 < S ? sB[t + QB] : 0.;   // u past the last row is zero
#pragma unroll
    for (int k = L - 1; k >= 0; --k) {
        x = fma(cyu * iw[k], x, g[k]);
        const uint32_t j = j0 + k;
        if (live && j < by) ub[q + (size_t)j * bx] = x;
    }
}

__device__ __forceinline__ double wave_sum(double s)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
    return s;
}

// wy[a] = sum_k W[a, k] y[k] with y = Sx hat y evaluated only where W needs it:
// y[i + j bx] = sum_q Sx[q, i] hat y[q, j] (Sx symmetric: column i).  One wave per row a;
// a row has at most WMAX entries (the node and its four neighbours), all of whose dot
// products are accumulated together so that every load of the row is in flight at once.
// CAP: batch element c = blockIdx.y holds R^-1 of the unit vector of row a0 + c, and the
// result is column a0 + c of C = I + W R^-1 U, stored row-major with row stride ld (C^T
// column-major).  Else blockIdx.y is the right-hand side of a batched solve: wy[blockIdx.y * ld + a].
constexpr int WMAX = 5;

template <int CQ, bool CAP>
__global__ void __launch_bounds__(256)
    k_cap_wy(uint32_t m, uint32_t ld, uint32_t bx, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
             const double *__restrict__ val, const double *__restrict__ Sx, const double *__restrict__ yh,
             double *__restrict__ wy, size_t nb, uint32_t a0)
{
    const uint32_t a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (a >= m) return;
    const double *yb = yh + (size_t)blockIdx.y * nb;
    const int32_t k0 = ptr[a], nk = ptr[a + 1] - k0;  // 1 <= nk <= WMAX (the node itself first)
    double acc[WMAX];
#pragma unroll
    for (int e = 0; e < WMAX; ++e) {
        // absent entries repeat the last one with weight zero, so that every load is issued
        // unconditionally and all of them are in flight together
        const int32_t k = k0 + (e < nk ? e : nk - 1);
        const uint32_t t = (uint32_t)col[k], i = t % bx, j = t / bx;
        const double *sc = Sx + (size_t)i * bx, *yc = yb + (size_t)j * bx;
        double d = 0.;
#pragma unroll
        for (int c = 0; c < CQ; ++c) {
            const uint32_t q = lane + 64 * c;
            const uint32_t qq = q < bx ? q : bx - 1;
            d = fma(sc[qq], q < bx ? yc[qq] : 0., d);
        }
        acc[e] = e < nk ? d * val[k] : 0.;
    }
    double s = 0.;
#pragma unroll
    for (int e = 0; e < WMAX; ++e) s += acc[e];
    s = wave_sum(s);
    if (lane == 0) {
        if (CAP) {
            const uint32_t c = a0 + blockIdx.y;
            wy[(size_t)a * ld + c] = s + (a == c ? 1. : 0.);
        } else {
            wy[(size_t)blockIdx.y * ld + a] = s;
        }
    }
}

// z = C^-1 wy, C^-1 row-major with rows padded to an even length ldc (16-byte aligned): RW rows
// per wave (a, a + 1, ...), 16-byte loads (two entries per lane and load, 8 / RW loads of each
// row in flight), each wy load serving all RW rows; z_a is stored at zE[ePos[a]], its place in
// the by-column order of the update.  blockIdx.y: the right-hand side (m apart in zE, ldc in
// wy; wy 16-byte aligned, its padding entry zero).  RW = 2: the 268 MB of C^-1 at 2048^2 stream
// at 4.5 TB/s (one row a wave: 4.2 TB/s, four: 4.1 TB/s).
template <int RW>
__global__ void __launch_bounds__(256)
    k_cap_zr(uint32_t m, uint32_t ldc, const double *__restrict__ CinvT, const double *__restrict__ wy,
             const int32_t *__restrict__ ePos, double *__restrict__ zE)
{
    constexpr int U = 8 / RW;   // loads of each row in flight
    const uint32_t a = RW * (blockIdx.x * 4 + (threadIdx.x >> 6)), lane = threadIdx.x & 63;
    if (a >= m) return;
    const double2 *w2 = reinterpret_cast<const double2 *>(wy + (size_t)blockIdx.y * ldc);
    zE += (size_t)blockIdx.y * m;
    const double2 *r[RW];
#pragma unroll
    for (int i = 0; i < RW; ++i) r[i] = reinterpret_cast<const double2 *>(CinvT + (size_t)(a + i < m ? a + i : a) * ldc);
    const uint32_t n2 = ldc / 2;
    double p[RW][U];
#pragma unroll
    for (int i = 0; i < RW; ++i)
#pragma unroll
        for (int u = 0; u < U; ++u) p[i][u] = 0.;
    uint32_t k = lane;
    for (; k + (U - 1) * 64 < n2; k += U * 64) {
        double2 c[RW][U];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < RW; ++i) c[i][u] = r[i][k + 64 * u];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double2 y = w2[k + 64 * u];
#pragma unroll
            for (int i = 0; i < RW; ++i) p[i][u] = fma(c[i][u].y, y.y, fma(c[i][u].x, y.x, p[i][u]));
        }
    }
    for (; k < n2; k += 64) {
        const double2 y = w2[k];
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const double2 c = r[i][k];
            p[i][0] = fma(c.y, y.y, fma(c.x, y.x, p[i][0]));
        }
    }
#pragma unroll
    for (int i = 0; i < RW; ++i) {
        double sum = p[i][0];
#pragma unroll
        for (int u = 1; u < U; ++u) sum += p[i][u];
        sum = wave_sum(sum);
        if (lane == 0 && a + i < m) zE[ePos[a + i]] = sum;
    }
}

// wy[a] = sum_k W[a, k] y[k] with y = R^-1 b in the box (physical) layout: one thread per row a,
// rows m..ld-1 (the padding) written zero; blockIdx.y: the right-hand side (ld apart in wy, nb in y).
__global__ void __launch_bounds__(256)
    k_wy(uint32_t m, uint32_t ld, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
         const double *__restrict__ val, const double *__restrict__ y, double *__restrict__ wy, size_t nb)
{
    const uint32_t a = blockIdx.x * 256 + threadIdx.x;
    if (a >= ld) return;
    y += (size_t)blockIdx.y * nb;
    double s = 0.;
    if (a < m)
        for (int32_t k = ptr[a]; k < ptr[a + 1]; ++k) s = fma(val[k], y[col[k]], s);
    wy[(size_t)blockIdx.y * ld + a] = s;
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

// Orthonormal DST-I matrix S[j, k] = sqrt(2/(n+1)) sin(pi (j+1)(k+1)/(n+1)), the argument
// reduced modulo 2(n+1) first so that sin sees at most 2 pi.
std::vector<double> dst1(uint32_t n)
{
    std::vector<double> S((size_t)n * n);
    const double f = sqrt(2. / (n + 1.)), w = M_PI / (n + 1.);
    for (uint32_t j = 0; j < n; ++j)
        for (uint32_t k = 0; k < n; ++k) {
            uint64_t t = ((uint64_t)(j + 1) * (k + 1)) % (2ull * (n + 1));
            S[j + (size_t)k * n] = f * sin(w * (double)t);
        }
    return S;
}

// Thomas scan over all x modes (unit: cnt batched unit right-hand sides from row a0 on; else
// cnt right-hand sides nb apart).
int launch_thomas(const oecl_fastpoisson *h, bool unit, const double *r, double *u, uint32_t a0, uint32_t cnt,
                  hipStream_t s)
{
    if (!unit && h->by <= 2048) {
        // segmented solve: S segments of L <= 16 rows, QB modes a block (4 to 16: 32- to 128-byte
        // row pieces), at most 1024 threads, and at least 256 blocks where the box allows
        uint32_t S = 1;
        while (S * 16 < h->by) S <<= 1;
        const uint32_t Lr = (h->by + S - 1) / S;
        uint32_t lqb = 2;
        while ((S << (lqb + 1)) <= 1024 && lqb < 4) ++lqb;
        while (lqb > 2 && (size_t)((h->bx + (1u << lqb) - 1) >> lqb) * cnt < 256) --lqb;
        const dim3 g((h->bx + (1u << lqb) - 1) >> lqb, cnt), b(S << lqb);
#define OECL_THOMAS_SEG(LL) \
    hipLaunchKernelGGL((k_thomas_seg<LL>), g, b, 0, s, h->bx, h->by, lqb, S, h->cyu, h->invw, r, u)
        if (Lr <= 4) OECL_THOMAS_SEG(4);
        else if (Lr <= 8) OECL_THOMAS_SEG(8);
        else if (Lr <= 12) OECL_THOMAS_SEG(12);
        else OECL_THOMAS_SEG(16);
#undef OECL_THOMAS_SEG
        return OECL_OK;
    }
    const dim3 g((h->bx + 3) / 4, cnt), b(256);
    const uint32_t c = (h->by + 63) / 64;
#define OECL_THOMAS(C)                                                                                   \
    do {                                                                                                 \
        if (unit)                                                                                        \
            hipLaunchKernelGGL((k_thomas_y<C, true>), g, b, 0, s, h->bx, h->by, h->cyu, h->invw, r, u, \
                               h->sIdx, a0, h->Sx);                                                      \
        else                                                                                             \
            hipLaunchKernelGGL((k_thomas_y<C, false>), g, b, 0, s, h->bx, h->by, h->cyu, h->invw, r, u, \
                               h->sIdx, a0, h->Sx);                                                      \
    } while (0)
    if (c <= 4) OECL_THOMAS(4);
    else if (c <= 8) OECL_THOMAS(8);
    else if (c <= 16) OECL_THOMAS(16);
    else if (c <= 32) OECL_THOMAS(32);
    else return OECL_ERR_ARG;
#undef OECL_THOMAS
    return OECL_OK;
}

// W Sx hat y over all rows (cap: cnt batched unit columns from row a0 on, into C; else cnt
// right-hand sides).
int launch_cap_wy(const oecl_fastpoisson *h, bool cap, const double *yh, double *out, uint32_t a0, uint32_t cnt,
                  hipStream_t s)
{
    const dim3 g((h->m + 3) / 4, cnt), b(256);
    const uint32_t c = (h->bx + 63) / 64;
#define OECL_WY(CQ)                                                                                        \
    do {                                                                                                   \
        if (cap)                                                                                           \
            hipLaunchKernelGGL((k_cap_wy<CQ, true>), g, b, 0, s, h->m, h->ldc, h->bx, h->wPtr, h->wCol,    \
                               h->wVal, h->Sx, yh, out, (size_t)h->nb, a0);                                \
        else                                                                                               \
            hipLaunchKernelGGL((k_cap_wy<CQ, false>), g, b, 0, s, h->m, h->ldc, h->bx, h->wPtr, h->wCol,   \
                               h->wVal, h->Sx, yh, out, (size_t)h->nb, a0);                                \
    } while (0)
    if (c <= 4) OECL_WY(4);
    else if (c <= 8) OECL_WY(8);
    else if (c <= 16) OECL_WY(16);
    else if (c <= 32) OECL_WY(32);
    else return OECL_ERR_ARG;
#undef OECL_WY
    return OECL_OK;
}

}  // namespace


uint32_t fastpoisson_box_width(uint32_t nx)
{
    if (nx < 5) return 0;
    uint32_t bx = 1;
    while (bx < nx - 4) bx = 2 * bx + 1;
    return bx;
}

OECL_EXPORT uint32_t oecl_fastpoisson_box_width(uint32_t nx)
{
    return fastpoisson_box_width(nx);
}

OECL_EXPORT int oecl_fastpoisson_create(uint32_t nx, uint32_t ny, double cxu, double cyu, uint32_t m,
                                        const int32_t *sIdxHost, const int32_t *wPtrHost,
                                        const int32_t *wColHost, const double *wValHost,
                                        const uint8_t *maskDev, oecl_fastpoisson **out, double *capDiag)
{
    if (!out || !maskDev || nx < 5 || ny < 5 || !(cxu > 0.) || !(cyu > 0.)) return OECL_ERR_ARG;
    if (ny - 4 > 64 * 32 || nx - 4 > 64 * 32 - 1) {
        oecl_set_error_msg("fastpoisson: nxExt - 4 must not exceed 2047 and nyExt - 4 not 2048");
        return OECL_ERR_ARG;
    }
    if (m > 0 && (!sIdxHost || !wPtrHost || !wColHost || !wValHost)) return OECL_ERR_ARG;
    // The box is padded in x to bx = 2^p - 1 >= nx - 4 columns (the transform is an FFT of
    // length bx + 1); the padding columns lie on the grid's boundary ring or beyond the grid and
    // hold no inside node.  The caller's rows S and W are in this box's indices
    // (i - 2) + (j - 2) bx, so an inside node next to the first padding column is an
    // irregular row (R couples it to that column, A does not).
    const uint32_t bxg = nx - 4, by = ny - 4, bx = fastpoisson_box_width(nx), nb = bx * by;
    for (uint32_t a = 0; a < m; ++a)
        if (sIdxHost[a] < 0 || (uint32_t)sIdxHost[a] >= nb || (uint32_t)sIdxHost[a] % bx >= bxg) return OECL_ERR_ARG;
    for (uint32_t a = 0; a < m; ++a)
        if (wPtrHost[a + 1] < wPtrHost[a] || wPtrHost[a + 1] - wPtrHost[a] > WMAX) return OECL_ERR_ARG;
    const uint32_t nnz = m ? (uint32_t)wPtrHost[m] : 0;
    for (uint32_t k = 0; k < nnz; ++k)
        if (wColHost[k] < 0 || (uint32_t)wColHost[k] >= nb) return OECL_ERR_ARG;
    oecl_fastpoisson *h = new oecl_fastpoisson();
    h->nx = nx;
    h->ny = ny;
    h->np = nx * ny;
    h->bx = bx;
    h->bxg = bxg;
    h->by = by;
    h->nb = nb;
    h->m = m;
    h->ldc = m + (m & 1u);
    h->cyu = cyu;
    h->mask = maskDev;
    h->rCap = 1;
    h->batchMem = nullptr;
    for (int e = 0; e < 5; ++e) h->pipeGrid[e] = 0;
    if (OECL_FFT_PIPE && bx + 1 == 2048) {
        int dev = 0, cus = 0, b0 = 0, b2 = 0, b3 = 0, b4 = 0;
        OECL_HIP(hipGetDevice(&dev), "device");
        OECL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
        OECL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b0, k_dst_fft_pipe<0>, 256, 0), "occupancy");
        OECL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b2, k_dst_fft_pipe<2>, 256, 0), "occupancy");
        OECL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b3, k_dst_fft_pipe<3>, 256, 0), "occupancy");
        OECL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b4, k_dst_fft_pipe<4>, 256, 0), "occupancy");
        h->pipeGrid[0] = (uint32_t)(cus * b0);
        h->pipeGrid[2] = (uint32_t)(cus * b2);
        h->pipeGrid[3] = (uint32_t)(cus * b3);
        h->pipeGrid[4] = (uint32_t)(cus * b4);
    }
    // rows a grouped by their box column j (stable), for the per-column update
    std::vector<int32_t> jA(m), jCol, jStart;
    for (uint32_t a = 0; a < m; ++a) jA[a] = (int32_t)a;
    std::stable_sort(jA.begin(), jA.end(), [&](int32_t u, int32_t v) {
        return sIdxHost[u] / (int32_t)bx < sIdxHost[v] / (int32_t)bx;
    });
    for (uint32_t e = 0; e < m; ++e) {
        int32_t j = sIdxHost[jA[e]] / (int32_t)bx;
        if (jCol.empty() || jCol.back() != j) {
            jCol.push_back(j);
            jStart.push_back((int32_t)e);
        }
    }
    jStart.push_back((int32_t)m);
    h->nJ = (uint32_t)jCol.size();

    if (rocblas_create_handle(&h->blas) != rocblas_status_success) {
        delete h;
        oecl_set_error_msg("rocblas_create_handle failed");
        return OECL_ERR_HIP;
    }
    const uint64_t M = bx + 1ull;
    const uint64_t ldc = h->ldc;
    const uint64_t bytes = 8ull * ((uint64_t)bx * bx + 4ull * M + 4ull * nb + ldc + m + (uint64_t)m * ldc + nnz) +
                           4ull * (m + (m + 1) + nnz + h->nJ + (h->nJ + 1) + 2ull * m) + 1024;
    char *mem = nullptr;
    OECL_HIP(hipMalloc(&mem, bytes), "fastpoisson alloc");
    char *cur = mem;
    h->Sx = (double *)cur; cur += 8ull * bx * bx;
    h->tab = (double2 *)cur; cur += 32ull * M;
    h->invw = (double *)cur; cur += 8ull * nb;
    h->b = (double *)cur; cur += 8ull * nb;
    h->t1 = (double *)cur; cur += 8ull * nb;
    h->t2 = (double *)cur; cur += 8ull * nb;
    cur = mem + ((cur - mem + 15) & ~(ptrdiff_t)15);  // wy and C^-1 rows 16-byte aligned
    h->wy = (double *)cur; cur += 8ull * ldc;
    h->CinvT = (double *)cur; cur += 8ull * m * ldc;
    h->zE = (double *)cur; cur += 8ull * m;
    h->wVal = (double *)cur; cur += 8ull * nnz;
    h->sIdx = (int32_t *)cur; cur += 4ull * m;
    h->wPtr = (int32_t *)cur; cur += 4ull * (m + 1);
    h->wCol = (int32_t *)cur; cur += 4ull * nnz;
    h->jCol = (int32_t *)cur; cur += 4ull * h->nJ;
    h->jStart = (int32_t *)cur; cur += 4ull * (h->nJ + 1);
    h->ePos = (int32_t *)cur; cur += 4ull * m;
    h->iPos = (uint32_t *)cur; cur += 4ull * m;

    const std::vector<double> Sx = dst1(bx);
    std::vector<double> lx(bx);
    for (uint32_t q = 0; q < bx; ++q) lx[q] = 2. - 2. * cos(M_PI * (q + 1.) / (bx + 1.));
    {
        // twiddles e^{-i pi k / M}, k < 2M, from the angle pi k / M < 2 pi in extended precision
        std::vector<double2> tw(2 * M);
        for (uint64_t k = 0; k < 2 * M; ++k) {
            const long double a = 3.141592653589793238462643383279502884L * (long double)k / (long double)M;
            tw[k] = make_double2((double)cosl(a), -(double)sinl(a));
        }
        // Thomas pivots of (cxu lx_q + 2 cyu) on the diagonal, -cyu off it
        std::vector<double> iw(nb);
        for (uint32_t q = 0; q < bx; ++q) {
            const double d = cxu * lx[q] + 2. * cyu;
            double w = d;
            for (uint32_t j = 0; j < by; ++j) {
                if (j) w = d - cyu * cyu / w;
                iw[q + (size_t)j * bx] = 1. / w;
            }
        }
        OECL_HIP(hipMemcpy(h->Sx, Sx.data(), 8ull * bx * bx, hipMemcpyHostToDevice), "copy Sx");
        OECL_HIP(hipMemcpy(h->tab, tw.data(), 32ull * M, hipMemcpyHostToDevice), "copy twiddles");
        // the padding columns of b are never written by the gathers: zero once
        OECL_HIP(hipMemset(h->b, 0, 8ull * nb), "zero box");
        OECL_HIP(hipMemcpy(h->invw, iw.data(), 8ull * nb, hipMemcpyHostToDevice), "copy pivots");
    }
    if (m) {
        OECL_HIP(hipMemcpy(h->sIdx, sIdxHost, 4ull * m, hipMemcpyHostToDevice), "copy S");
        OECL_HIP(hipMemcpy(h->wPtr, wPtrHost, 4ull * (m + 1), hipMemcpyHostToDevice), "copy W");
        OECL_HIP(hipMemcpy(h->wCol, wColHost, 4ull * nnz, hipMemcpyHostToDevice), "copy W");
        OECL_HIP(hipMemcpy(h->wVal, wValHost, 8ull * nnz, hipMemcpyHostToDevice), "copy W");
        OECL_HIP(hipMemcpy(h->jCol, jCol.data(), 4ull * h->nJ, hipMemcpyHostToDevice), "copy columns");
        OECL_HIP(hipMemcpy(h->jStart, jStart.data(), 4ull * (h->nJ + 1), hipMemcpyHostToDevice), "copy columns");
        std::vector<int32_t> ePos(m);
        std::vector<uint32_t> iPos(m);
        for (uint32_t e = 0; e < m; ++e) {
            ePos[jA[e]] = (int32_t)e;
            iPos[e] = (uint32_t)sIdxHost[jA[e]] % bx;
        }
        OECL_HIP(hipMemcpy(h->ePos, ePos.data(), 4ull * m, hipMemcpyHostToDevice), "copy columns");
        OECL_HIP(hipMemcpy(h->iPos, iPos.data(), 4ull * m, hipMemcpyHostToDevice), "copy columns");
        // Capacitance matrix C = I + W R^-1 U with exactly the operator the solve applies
        // (Sx, the Thomas scan, the W dot products), in chunks of `chunk` unit columns.
        uint32_t chunk = (uint32_t)((256ull << 20) / (8ull * nb));
        if (chunk < 1) chunk = 1;
        if (chunk > m) chunk = m;
        if (chunk > 65535) chunk = 65535;
        double *G = nullptr;
        OECL_HIP(hipMemset(h->CinvT, 0, 8ull * m * ldc), "zero C");  // the row padding stays zero
        OECL_HIP(hipMalloc(&G, 8ull * nb * chunk), "cap alloc");
        for (uint32_t a0 = 0; a0 < m; a0 += chunk) {
            const uint32_t cnt = m - a0 < chunk ? m - a0 : chunk;
            int rc = launch_thomas(h, true, nullptr, G, a0, cnt, 0);
            if (rc == OECL_OK) rc = launch_cap_wy(h, true, G, h->CinvT, a0, cnt, 0);
            if (rc != OECL_OK) {
                (void)hipFree(G);
                return rc;
            }
        }
        OECL_CHECK_LAUNCH("capacitance build");
        OECL_HIP(hipDeviceSynchronize(), "capacitance build");
        if (capDiag) {
            // report max |C_aa| for a scale check
            std::vector<double> hc(m);
            OECL_HIP(hipMemcpy2D(hc.data(), 8, h->CinvT, 8ull * (ldc + 1), 8, m, hipMemcpyDeviceToHost), "diag");
            double mx = 0.;
            for (uint32_t a = 0; a < m; ++a) mx = fabs(hc[a]) > mx ? fabs(hc[a]) : mx;
            *capDiag = mx;
        }
        (void)hipFree(G);
        // In-place inverse: getrf then getri.
        rocblas_int *ipiv = nullptr, *info = nullptr;
        OECL_HIP(hipMalloc(&ipiv, 4ull * m), "ipiv");
        OECL_HIP(hipMalloc(&info, 4), "info");
        RB(rocsolver_dgetrf(h->blas, m, m, h->CinvT, (rocblas_int)ldc, ipiv, info), "dgetrf");
        RB(rocsolver_dgetri(h->blas, m, h->CinvT, (rocblas_int)ldc, ipiv, info), "dgetri");
        int hinfo = 0;
        OECL_HIP(hipMemcpy(&hinfo, info, 4, hipMemcpyDeviceToHost), "info");
        (void)hipFree(ipiv);
        (void)hipFree(info);
        if (hinfo != 0) {
            oecl_set_error_msg("capacitance matrix is singular");
            return OECL_ERR_NOCONV;
        }
    } else if (capDiag) {
        *capDiag = 0.;
    }
    OECL_HIP(hipDeviceSynchronize(), "fastpoisson create");
    *out = h;
    return OECL_OK;
}

FpBox fastpoisson_box(const oecl_fastpoisson *h)
{
    return FpBox{h->b, h->mask, h->nx, h->ny, h->bx};
}

int fastpoisson_reserve(oecl_fastpoisson *h, uint32_t R)
{
    if (R <= h->rCap) return OECL_OK;
    const uint64_t nb = h->nb, m = h->m, ldc = h->ldc;
    char *mem = nullptr;
    OECL_HIP(hipMalloc(&mem, 8ull * R * (3 * nb + ldc + m) + 1024), "fastpoisson batch alloc");
    if (h->batchMem) (void)hipFree(h->batchMem);
    h->batchMem = mem;
    h->b = (double *)mem;
    OECL_HIP(hipMemset(h->b, 0, 8ull * R * nb), "zero box");
    h->t1 = h->b + R * nb;
    h->t2 = h->t1 + R * nb;
    h->wy = h->t2 + R * nb + ((R * nb) & 1u);  // 16-byte aligned (the allocation holds the slack)
    h->zE = h->wy + R * ldc;
    h->rCap = R;
    return OECL_OK;
}

OECL_EXPORT int oecl_fastpoisson_reserve(oecl_fastpoisson *h, uint32_t R)
{
    if (!h || R == 0) return OECL_ERR_ARG;
    return fastpoisson_reserve(h, R);
}

OECL_EXPORT int oecl_fastpoisson_solve(oecl_fastpoisson *h, const double *q, double *phi, void *stream)
{
    if (!h || !q || !phi) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_gather_box_zero, dim3(oecl_blocks(h->np, 256)), dim3(256), 0, s, h->nx, h->ny, h->bx,
                       h->mask, q, h->b, phi);
    return fastpoisson_solve_box(h, phi, 1, s);
}

// One DST transform of N lines: EPI 0 of X (ld bx) into C (ld ldc; C may be X), EPI 2 of X into
// the potential C = phi, EPI 3 of the sparse lines of zE (N = R nJ) subtracted from C (ld bx).
// The grid-side arguments of EPI 4.
struct FftGrid {
    double *rho;
    const double *prof, *table;
    uint32_t nScale;
    const uint32_t *stepDev;
};

template <int EPI>
static int dst_transform(oecl_fastpoisson *h, int N, const double *X, double *C, int ldc, hipStream_t s,
                         const FftGrid *gg = nullptr)
{
    FftArgs g;
    g.nLines = (uint32_t)N;
    g.X = X;
    g.ldx = h->bx;
    g.C = C;
    g.ldc = (uint32_t)ldc;
    g.tab = h->tab;
    g.nx = h->nx;
    g.by = h->by;
    g.np = h->np;
    g.bxg = h->bxg;
    g.mask = h->mask;
    g.jCol = h->jCol;
    g.jStart = h->jStart;
    g.iPos = h->iPos;
    g.nJ = h->nJ;
    g.m = h->m;
    g.nb = h->nb;
    g.rho = nullptr;
    g.prof = g.table = nullptr;
    g.nScale = 0;
    g.stepDev = nullptr;
    if (EPI == 4) {
        if (!gg) return OECL_ERR_ARG;
        g.rho = gg->rho;
        g.prof = gg->prof;
        g.table = gg->table;
        g.nScale = gg->nScale;
        g.stepDev = gg->stepDev;
    }
    if (h->bx + 1 == 2048 && h->pipeGrid[EPI] && N > 0) {
        // as many blocks as are resident at once, each taking every G-th line
        const uint32_t G = (uint32_t)N < h->pipeGrid[EPI] ? (uint32_t)N : h->pipeGrid[EPI];
        hipLaunchKernelGGL((k_dst_fft_pipe<EPI>), dim3(G), dim3(256), 0, s, g);
        return OECL_OK;
    }
#define OECL_FFT(MM)                                                                                    \
    case MM:                                                                                          \
        hipLaunchKernelGGL((k_dst_fft<MM, EPI>), dim3((N + FftShape<MM>::LB - 1) / FftShape<MM>::LB), \
                           dim3(FftShape<MM>::NT), 0, s, g);                                          \
        return OECL_OK;
    switch (h->bx + 1) {
        OECL_FFT(2) OECL_FFT(4) OECL_FFT(8) OECL_FFT(16) OECL_FFT(32) OECL_FFT(64) OECL_FFT(128) OECL_FFT(256)
        OECL_FFT(512) OECL_FFT(1024) OECL_FFT(2048) OECL_FFT(4096)
    default:
        oecl_set_error_msg("fastpoisson: padded box width out of range");
        return OECL_ERR_ARG;
    }
#undef OECL_FFT
}

static int solve_after_first(oecl_fastpoisson *h, double *phi, uint32_t R, hipStream_t s);

int fastpoisson_solve_box(oecl_fastpoisson *h, double *phi, uint32_t R, hipStream_t s)
{
    if (R < 1 || R > h->rCap || (uint64_t)h->by * R > 0x7FFFFFFFull) return OECL_ERR_ARG;
    // hat b = Sx b
    int rc = dst_transform<0>(h, (int)(h->by * R), h->b, h->t1, (int)h->bx, s);
    if (rc != OECL_OK) return rc;
    return solve_after_first(h, phi, R, s);
}

int fastpoisson_solve_grid(oecl_fastpoisson *h, double *rho, const double *prof, const double *table, uint32_t nScale,
                           const uint32_t *stepDev, double *phi, uint32_t R, hipStream_t s)
{
    if (R < 1 || R > h->rCap || (uint64_t)h->by * R > 0x7FFFFFFFull || !rho || !stepDev) return OECL_ERR_ARG;
    const FftGrid gg{rho, prof, table, nScale, stepDev};
    // hat b = Sx q, q gathered from the grid (rho zeroed on the way)
    int rc = dst_transform<4>(h, (int)(h->by * R), nullptr, h->t1, (int)h->bx, s, &gg);
    if (rc != OECL_OK) return rc;
    return solve_after_first(h, phi, R, s);
}

#ifndef OECL_CAPZ_ROWS
#define OECL_CAPZ_ROWS 2
#endif

static int solve_after_first(oecl_fastpoisson *h, double *phi, uint32_t R, hipStream_t s)
{
    const int bx = h->bx, by = h->by, N = by * (int)R;
    int rc = OECL_OK;
    if (h->m) {
        // hat y = T^-1 hat b ; y = Sx hat y (in place) ; wy = W y ; z = C^-1 wy ;
        // hat b -= Sx U z, line by line over the box rows holding corrected nodes
        rc = launch_thomas(h, false, h->t1, h->t2, 0, R, s);
        if (rc != OECL_OK) return rc;
        rc = dst_transform<0>(h, N, h->t2, h->t2, bx, s);
        if (rc != OECL_OK) return rc;
        hipLaunchKernelGGL(k_wy, dim3((h->ldc + 255) / 256, R), dim3(256), 0, s, h->m, h->ldc, h->wPtr, h->wCol,
                           h->wVal, h->t2, h->wy, (size_t)h->nb);
        hipLaunchKernelGGL(k_cap_zr<OECL_CAPZ_ROWS>, dim3((h->m + 4 * OECL_CAPZ_ROWS - 1) / (4 * OECL_CAPZ_ROWS), R),
                           dim3(256), 0, s, h->m, h->ldc, h->CinvT, h->wy, h->ePos, h->zE);
        rc = dst_transform<3>(h, (int)(h->nJ * R), h->zE, h->t1, bx, s);
        if (rc != OECL_OK) return rc;
    }
    // hat x = T^-1 hat b ; phi = Sx hat x scattered into the extended grid (np apart per side)
    rc = launch_thomas(h, false, h->t1, h->t2, 0, R, s);
    if (rc != OECL_OK) return rc;
    rc = dst_transform<2>(h, N, h->t2, phi, 0, s);
    if (rc != OECL_OK) return rc;
    OECL_CHECK_LAUNCH("oecl_fastpoisson_solve");
    return OECL_OK;
}

OECL_EXPORT int oecl_fastpoisson_destroy(oecl_fastpoisson *h)
{
    if (!h) return OECL_OK;
    (void)hipFree(h->Sx);
    if (h->batchMem) (void)hipFree(h->batchMem);
    rocblas_destroy_handle(h->blas);
    delete h;
    return OECL_OK;
}
