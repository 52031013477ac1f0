// Learned preconditioner for the FIT Poisson system (no counterpart in the reference, which
// back-substitutes SuperLU): a 4-level U-Net with 16 channels trained on the device, and a
// flexible conjugate-gradient solve that applies it.
//
// Network (C = 16, 3x3 convolutions with zero padding, ReLU, 2x2 average pooling, nearest
// upsampling with additive skips; 14 convolutions, all 16 -> 16 channels):
//     s0 = relu(c1(relu(c0(x))))            x: the input in channel 0
//     s1 = relu(c3(relu(c2(pool(s0)))))
//     s2 = relu(c5(relu(c4(pool(s1)))))
//     u3 = relu(c7(relu(c6(pool(s2)))))
//     u2 = relu(c9(relu(c8(up(u3) + s2))))
//     u1 = relu(c11(relu(c10(up(u2) + s1))))
//     y  = c13(relu(c12(up(u1) + s0)))[channel 0]
// Activations are NHWC bfloat16 images (16 channels = 32 bytes per pixel); weights are FP32
// masters with bfloat16 copies [conv][tap][16 out][16 in] (tap = 3 (dy + 1) + (dx + 1)); every
// product runs on bf16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32 accumulation.  One
// convolution wave computes 16 pixels of a row for all 16 output channels: 9 MFMAs, one per
// tap, with K = the 16 input channels.  The weight gradient is the same contraction with K =
// 16 pixels.  The solve keeps its vectors and reductions in FP64; the network sees the
// residual scaled to unit RMS, so the preconditioner is homogeneous of degree one.
#include <math.h>

#include <algorithm>
#include <vector>

#include "common.h"

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

struct oecl_unet {
    uint32_t H, W;            // level-0 image (multiples of 8)
    float *Wm, *bm;           // FP32 masters [14][9][16][16], [14][16]
    float *mW, *vW, *mb, *vb; // Adam moments
    float *gW, *gb;           // gradients
    float *wpart;             // per-block partial weight gradients of one convolution
    uint16_t *Wb, *WbT;       // bf16 weights, and flipped/transposed for the data gradient
    uint16_t *act[32];        // activations (see the layout in unet_layout)
    uint16_t *grad[8];        // gradient scratch images
    float *out;               // network output, H * W
    double *scal;             // device scalars (scale, reductions)
    double *part;             // per-block partial sums
    double *sv[4];            // grid vectors (H W doubles each) of the smoothed preconditioner
    char *mem;
    uint32_t adamT;
    int convRows;             // output rows per convolution wave: 0 = chosen per image size
    uint32_t batch;           // images the activation, grid-vector and reduction buffers hold
    char *bmem;               // their allocation once batch > 1
};

namespace {

constexpr int C = 16;
constexpr int NCONV = 14;
constexpr int WSZ = 9 * C * C;  // weights per convolution
constexpr int kThreads = 256;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// float -> bfloat16, round to nearest even with subnormals kept (gfx950's v_cvt_pk_bf16_f32;
// equal to the integer rounding u + 0x7FFF + ((u >> 16) & 1) for every non-NaN float, a NaN
// stays a NaN).  f2bf2 converts a pair into one 32-bit word (a in the low half).
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t f2bf2(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v2f{a, b}, v2bf));
}

__device__ __forceinline__ v4s load4(const uint16_t *p) { return *reinterpret_cast<const v4s *>(p); }

// out = conv3x3(in) + bias over an H x W image.  MODE 0: bf16 out with ReLU, 1: bf16 out
// without ReLU, 2: FP32 channel 0 only.  A block computes RW rows x 64 pixels (its 4 waves take
// 16 pixels each).  It first stages the RW + 2 input rows of its 66-pixel span (1 pixel of halo
// each side, zero outside the image) in LDS with coalesced 16-byte loads, all issued before the
// first use, so each wave keeps many rows in flight; a wave then walks the staged rows once,
// reads each row's three horizontally shifted operands from LDS, and feeds every output row
// that uses that input row (the taps of each output keep the order 0..8).  Tall tiles (RW 8)
// re-read fewer halo rows on large images; short ones (RW 1, 2) give small images more blocks.
// Blocks are numbered row group by row group and taken in XCD chunks (xcd_block), so a chunk's
// vertically adjacent tiles share their halo rows in one L2.
// MFMA operand layouts (16x16x16): lane l holds A[l % 16][4 (l / 16) + j], B[4 (l / 16) + j][l % 16]
// and D[4 (l / 16) + j][l % 16], j = 0..3; here A = weights (out x in), B = input (in x pixel).
// Grid vectors of the preconditioner (nx x ny, row-major; image pixel (y, x) is grid node
// y nx + x for x < nx, y < ny).  sc[0] = sum of r^2 over inside nodes, sc[1] = their count.
struct GridIO {
    const double *r;      // network input: channel 0 = r / rms on inside nodes, 0 elsewhere
    double *z;            // network output: z = rms * y on inside nodes, 0 elsewhere
    const uint8_t *mask;  // inside nodes
    const double *sc;
    uint32_t nx, ny;
    size_t gs;            // batch: image b's r and z at b gs (grid vectors), its sc at b ss
    uint32_t ss;
    const double *st;     // optional: image b is skipped when st[b ss] != 0 (a stopped run)
};

// Strides of a batch of independent problems (image or run b = blockIdx.y): grid vectors at
// b v, partial sums at b p, scalar blocks at b s; w is the stride between the slots of the
// flexible CG's direction rings (R v for a batch of R, v for one).
struct Bat {
    size_t v, w, p;
    uint32_t s;
    const double *st;  // optional: run b's kernels do nothing when st[b s] != 0 (it has stopped)
};

// Whether this block's run has stopped (uniform over the block: it depends on blockIdx.y only).
__device__ __forceinline__ bool stopped(const Bat &bt) { return bt.st && bt.st[blockIdx.y * bt.s] != 0.; }

__device__ __forceinline__ double grid_rms(const double *sc) { return sc[1] > 0. && sc[0] > 0. ? sqrt(sc[0] / sc[1]) : 1.; }

// SRC 1 (the first convolution of an inference pass) stages its input straight from the FP64
// grid vector g.r instead of an input image; MODE 3 (its last convolution) writes g.z.  SRC 2 (an
// up level's first convolution in an inference pass) stages up(up) + in, the nearest upsampling
// of the coarser level's output plus the skip image, instead of reading k_up_add's sum.  All give
// bit for bit what the unfused passes give.
constexpr int CONV_SPAN = 64 + 2;  // staged pixels per row
__device__ __forceinline__ uint32_t add_bf16x2(uint32_t a, uint32_t b)
{
    return f2bf2(__uint_as_float(a << 16) + __uint_as_float(b << 16),
                 __uint_as_float(a & 0xFFFF0000u) + __uint_as_float(b & 0xFFFF0000u));
}

template <int MODE, int RW, int SRC = 0>
__global__ void __launch_bounds__(kThreads)
    k_conv(int H, int W, const uint16_t *__restrict__ in, const uint16_t *__restrict__ wt,
           const float *__restrict__ bias, void *__restrict__ out, GridIO g, const uint16_t *__restrict__ up,
           uint16_t *__restrict__ pooled)
{
    __shared__ __align__(16) uint16_t tile[RW + 2][CONV_SPAN][C];
    // tile and image: the (tile, image) grid is flattened and ordered XCD-aware as a whole, so
    // that the tiles of one image sharing halo rows run on one XCD
    const uint32_t lin = xcd_block(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const uint32_t img = lin / gridDim.x;
    const int blk = (int)(lin - img * gridDim.x);
    if (g.st && g.st[(size_t)img * g.ss] != 0.) return;  // a stopped run: whole block
    if (img) {  // image img of a batch: its images and grid vectors
        const size_t b = img, is = (size_t)H * W * C, isc = (size_t)(H >> 1) * (W >> 1) * C;
        if (in) in += b * is;
        if (up) up += b * isc;
        if (pooled) pooled += b * isc;
        if (out) out = MODE == 2 ? (void *)(static_cast<float *>(out) + b * (size_t)H * W)
                                 : (void *)(static_cast<uint16_t *>(out) + b * is);
        if (g.r) g.r += b * g.gs;
        if (g.z) g.z += b * g.gs;
        if (g.sc) g.sc += b * g.ss;
    }
    const int lane = threadIdx.x & 63, px = lane & 15, cq = lane >> 4, wv = threadIdx.x >> 6;
    const int quads = (W + 63) >> 6;
    const int y0 = (blk / quads) * RW, xq = (blk % quads) << 6;  // grid.x = groups * quads exactly
    // stage rows y0 - 1 .. y0 + RW, pixels xq - 1 .. xq + 64 (two 16-byte halves per pixel):
    // every load of the thread is issued before the first LDS store
    constexpr int NE = (RW + 2) * CONV_SPAN * 2, NL = (NE + kThreads - 1) / kThreads;
    uint4 v[NL];
    double rv[SRC == 1 ? NL : 1];
    uint8_t mv[SRC == 1 ? NL : 1];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int e = threadIdx.x + k * kThreads;
        const int row = e / (CONV_SPAN * 2), rem = e - row * (CONV_SPAN * 2), sp = rem >> 1, half = rem & 1;
        const int gy = y0 - 1 + row, gx = xq - 1 + sp;
        v[k] = make_uint4(0u, 0u, 0u, 0u);
        if (SRC == 0) {
            if (e < NE && gy >= 0 && gy < H && gx >= 0 && gx < W)
                v[k] = *reinterpret_cast<const uint4 *>(in + ((size_t)gy * W + gx) * C + 8 * half);
        } else if (SRC == 2) {
            if (e < NE && gy >= 0 && gy < H && gx >= 0 && gx < W) {
                const uint4 sk = *reinterpret_cast<const uint4 *>(in + ((size_t)gy * W + gx) * C + 8 * half);
                const uint4 cu =
                    *reinterpret_cast<const uint4 *>(up + ((size_t)(gy >> 1) * (W >> 1) + (gx >> 1)) * C + 8 * half);
                v[k] = make_uint4(add_bf16x2(cu.x, sk.x), add_bf16x2(cu.y, sk.y), add_bf16x2(cu.z, sk.z),
                                  add_bf16x2(cu.w, sk.w));
            }
        } else {
            // both loads issued unconditionally for in-range nodes; the mask selects afterwards
            const bool ok = e < NE && half == 0 && gy >= 0 && gy < (int)g.ny && gx >= 0 && gx < (int)g.nx;
            const uint32_t i = ok ? (uint32_t)gy * g.nx + (uint32_t)gx : 0u;
            rv[k] = ok ? g.r[i] : 0.;
            mv[k] = ok ? g.mask[i] : (uint8_t)0;
        }
    }
    if constexpr (SRC == 1) {
        const double rms = grid_rms(g.sc);
#pragma unroll
        for (int k = 0; k < NL; ++k)
            if (mv[k]) v[k].x = f2bf((float)(rv[k] / rms));  // channel 0 in the low half
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int e = threadIdx.x + k * kThreads;
        const int row = e / (CONV_SPAN * 2), rem = e - row * (CONV_SPAN * 2), sp = rem >> 1, half = rem & 1;
        if (e < NE) *reinterpret_cast<uint4 *>(&tile[row][sp][8 * half]) = v[k];
    }
    __syncthreads();
    const int x = xq + (wv << 4) + px;
    if (xq + (wv << 4) >= W) return;  // whole wave, after the block's only barrier
    v4s a[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = load4(wt + (k * C + px) * C + 4 * cq);
    v4f acc[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) acc[r] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int iy = 0; iy < RW + 2; ++iy) {
        v4s b[3];
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) b[dx] = *reinterpret_cast<const v4s *>(&tile[iy][(wv << 4) + px + dx][4 * cq]);
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int ky = iy - r;  // tap row of output row y0 + r that reads input row y0 - 1 + iy
            if (ky >= 0 && ky <= 2)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx)
                    acc[r] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ky * 3 + dx], b[dx], acc[r], 0, 0, 0);
        }
    }
    if (x >= W) return;
    if (MODE == 3) {
        if (cq != 0 || x >= (int)g.nx) return;
        const double rms = grid_rms(g.sc);
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int y = y0 + r;
            if (y >= (int)g.ny) break;
            const uint32_t i = (uint32_t)y * g.nx + (uint32_t)x;
            const float yv = acc[r][0] + bias[0];
            g.z[i] = g.mask[i] ? rms * (double)yv : 0.;
        }
        return;
    }
    v4s pe{0, 0, 0, 0}, pen{0, 0, 0, 0};  // the even row's stored values at x and x + 1
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int y = y0 + r;
        if (y >= H) break;
        const size_t p = (size_t)y * W + x;
        if (MODE == 2) {
            if (cq == 0) static_cast<float *>(out)[p] = acc[r][0] + bias[0];
            continue;
        }
        float f[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            f[j] = acc[r][j] + bias[4 * cq + j];
            if (MODE == 0) f[j] = f[j] > 0.f ? f[j] : 0.f;
        }
        const uint32_t w01 = f2bf2(f[0], f[1]), w23 = f2bf2(f[2], f[3]);
        const v4s o{(short)(w01 & 0xFFFFu), (short)(w01 >> 16), (short)(w23 & 0xFFFFu), (short)(w23 >> 16)};
        *reinterpret_cast<v4s *>(static_cast<uint16_t *>(out) + p * C + 4 * cq) = o;
        if constexpr (RW >= 2) {
            // pooled != null: also the 2x2 average of the stored values into the next level's
            // image, as k_pool forms it ((a + b) + c) + d from rows y, y + 1 (this lane's r - 1,
            // r) and columns x, x + 1 (this lane and the next; W and H are even, y0 too)
            if (pooled) {
                v4s on;
#pragma unroll
                for (int j = 0; j < 4; ++j) on[j] = (short)__shfl((int)o[j], lane + 1, 64);
                if (r & 1) {
                    if (!(x & 1)) {
                        v4s po;
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            po[j] = (short)f2bf(0.25f * (((bf2f((uint16_t)pe[j]) + bf2f((uint16_t)pen[j])) +
                                                          bf2f((uint16_t)o[j])) + bf2f((uint16_t)on[j])));
                        *reinterpret_cast<v4s *>(pooled + ((size_t)(y >> 1) * (W >> 1) + (x >> 1)) * C + 4 * cq) = po;
                    }
                } else {
                    pe = o;
                    pen = on;
                }
            }
        }
    }
}

// Weight and bias gradients of one convolution: gW[k][o][i] += sum_p g[p][o] x[p + off_k][i],
// gb[o] += sum_p g[p][o].  MFMA with M = out channels, N = in channels, K = 16 pixels of a row
// segment; waves stride over the segments.  The sums are formed in a fixed order, so training is
// reproducible: each wave stores its sums in its own LDS row, the block adds its 4 rows in order
// and stores the block's partial sums (WG per block) in wpart; k_wgrad_sum adds the blocks.
constexpr int WG = WSZ + C;            // weight and bias entries of one convolution
constexpr uint32_t kWgradBlocks = 1024;
__global__ void __launch_bounds__(kThreads)
    k_conv_wgrad(int H, int W, const uint16_t *__restrict__ x, const uint16_t *__restrict__ g,
                 float *__restrict__ wpart)
{
    __shared__ float red[kThreads / 64][WSZ];
    __shared__ float bl[kThreads / 64][64];
    const int lane = threadIdx.x & 63, px = lane & 15, cq = lane >> 4;
    const int segs = (W + 15) >> 4, nSeg = H * segs;
    const int wave = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    const int nWaves = gridDim.x * (kThreads / 64);
    v4f acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = v4f{0.f, 0.f, 0.f, 0.f};
    float bs = 0.f;
    for (int s = wave; s < nSeg; s += nWaves) {
        const int y = s / segs, x0 = (s % segs) << 4;
        // A[o = px][p = 4 cq + j] = g[x0 + 4 cq + j][o]
        v4s a;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int xx = x0 + 4 * cq + j;
            const uint16_t v = xx < W ? g[((size_t)y * W + xx) * C + px] : (uint16_t)0;
            a[j] = (short)v;
            bs += bf2f(v);
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int yy = y + k / 3 - 1;
            v4s b;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int xo = x0 + 4 * cq + j, xx = xo + k % 3 - 1;
                const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W && xo < W;
                b[j] = ok ? (short)x[((size_t)yy * W + xx) * C + px] : (short)0;
            }
            acc[k] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[k], 0, 0, 0);
        }
    }
    // D[o = 4 cq + j][i = px]: every entry of the wave's row is written once
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[wv][(k * C + 4 * cq + j) * C + px] = acc[k][j];
    bl[wv][lane] = bs;  // channel px, pixels 4 cq .. 4 cq + 3 of each segment
    __syncthreads();
    float *out = wpart + (size_t)blockIdx.x * WG;
    for (int t = threadIdx.x; t < WSZ; t += kThreads) out[t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
    if (threadIdx.x < C) {
        float sb = 0.f;
        for (int w = 0; w < kThreads / 64; ++w)
            for (int q = 0; q < 4; ++q) sb += bl[w][threadIdx.x + 16 * q];
        out[WSZ + threadIdx.x] = sb;
    }
}

// gW[e] += sum over the nb blocks of wpart[b][e] (e < WSZ; gb for the last C entries), in a fixed
// order: 8 stripes each sum blocks s, s + 8, ... in turn, then the stripes are added in order.
__global__ void __launch_bounds__(256)
    k_wgrad_sum(uint32_t nb, const float *__restrict__ wpart, float *__restrict__ gW, float *__restrict__ gb)
{
    __shared__ float sh[8][32];
    const int el = threadIdx.x & 31, st = threadIdx.x >> 5;
    const int e = blockIdx.x * 32 + el;
    float s = 0.f;
    if (e < WG)
        for (uint32_t b = st; b < nb; b += 8) s += wpart[(size_t)b * WG + e];
    sh[st][el] = s;
    __syncthreads();
    if (threadIdx.x < 32 && e < WG) {
        float t = 0.f;
        for (int k = 0; k < 8; ++k) t += sh[k][el];
        if (e < WSZ) gW[e] += t;
        else gb[e - WSZ] += t;
    }
}

// g *= (act > 0): the ReLU of a layer applied to the gradient of its output (8-byte groups).
__global__ void k_relu_back(size_t n4, uint16_t *__restrict__ g, const uint16_t *__restrict__ act)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    v4s gv = load4(g + 4 * i), av = load4(act + 4 * i);
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (!(bf2f((uint16_t)av[j]) > 0.f)) gv[j] = 0;
    *reinterpret_cast<v4s *>(g + 4 * i) = gv;
}

// out (H x W) = 2x2 average of in (2H x 2W); one thread per 4 channels of an output pixel.
__global__ void k_pool(int H, int W, const uint16_t *__restrict__ in, uint16_t *__restrict__ out)
{
    in += (size_t)blockIdx.y * 4 * H * W * C;   // image blockIdx.y of a batch
    out += (size_t)blockIdx.y * H * W * C;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)H * W * 4) return;
    const int q = t & 3;
    const size_t p = t >> 2;
    const int y = (int)(p / W), x = (int)(p % W), W2 = 2 * W;
    const v4s a = load4(in + ((size_t)(2 * y) * W2 + 2 * x) * C + 4 * q);
    const v4s b = load4(in + ((size_t)(2 * y) * W2 + 2 * x + 1) * C + 4 * q);
    const v4s c = load4(in + ((size_t)(2 * y + 1) * W2 + 2 * x) * C + 4 * q);
    const v4s d = load4(in + ((size_t)(2 * y + 1) * W2 + 2 * x + 1) * C + 4 * q);
    v4s o;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        o[j] = (short)f2bf(0.25f * (((bf2f(a[j]) + bf2f(b[j])) + bf2f(c[j])) + bf2f(d[j])));
    *reinterpret_cast<v4s *>(out + p * C + 4 * q) = o;
}

// out (H x W) = up(u) + s, u (H/2 x W/2).  BACK: out (H x W) += 0.25 up(u) instead (the pooling's
// adjoint accumulated into a skip gradient).
template <bool BACK>
__global__ void k_up_add(int H, int W, const uint16_t *__restrict__ u, const uint16_t *__restrict__ s,
                         uint16_t *__restrict__ out)
{
    u += (size_t)blockIdx.y * (H >> 1) * (W >> 1) * C;   // image blockIdx.y of a batch
    s += (size_t)blockIdx.y * H * W * C;
    out += (size_t)blockIdx.y * H * W * C;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)H * W * 4) return;
    const int q = t & 3;
    const size_t p = t >> 2;
    const int y = (int)(p / W), x = (int)(p % W);
    const v4s uv = load4(u + ((size_t)(y >> 1) * (W >> 1) + (x >> 1)) * C + 4 * q);
    const v4s sv = load4(s + p * C + 4 * q);
    v4s o;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        o[j] = (short)f2bf(BACK ? bf2f(sv[j]) + 0.25f * bf2f(uv[j]) : bf2f(uv[j]) + bf2f(sv[j]));
    *reinterpret_cast<v4s *>(out + p * C + 4 * q) = o;
}

// out (H x W) = sum of the 2x2 block of g (2H x 2W): the adjoint of nearest upsampling.
__global__ void k_up_back(int H, int W, const uint16_t *__restrict__ g, uint16_t *__restrict__ out)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)H * W * 4) return;
    const int q = t & 3;
    const size_t p = t >> 2;
    const int y = (int)(p / W), x = (int)(p % W), W2 = 2 * W;
    const v4s a = load4(g + ((size_t)(2 * y) * W2 + 2 * x) * C + 4 * q);
    const v4s b = load4(g + ((size_t)(2 * y) * W2 + 2 * x + 1) * C + 4 * q);
    const v4s c = load4(g + ((size_t)(2 * y + 1) * W2 + 2 * x) * C + 4 * q);
    const v4s d = load4(g + ((size_t)(2 * y + 1) * W2 + 2 * x + 1) * C + 4 * q);
    v4s o;
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = (short)f2bf(((bf2f(a[j]) + bf2f(b[j])) + bf2f(c[j])) + bf2f(d[j]));
    *reinterpret_cast<v4s *>(out + p * C + 4 * q) = o;
}

// bf16 weights, and the flipped/transposed copy WT[k][i][o] = W[8 - k][o][i] that turns the
// data gradient into a forward convolution.
__global__ void k_weights_bf16(const float *__restrict__ Wm, uint16_t *__restrict__ Wb, uint16_t *__restrict__ WbT)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NCONV * WSZ) return;
    const int c = t / WSZ, r = t % WSZ, k = r / (C * C), o = (r / C) % C, i = r % C;
    const uint16_t v = f2bf(Wm[t]);
    Wb[t] = v;
    WbT[c * WSZ + ((8 - k) * C + i) * C + o] = v;
}

// Adam (Kingma & Ba, 2015) on the FP32 masters; bias corrections c1 = 1 - beta1^t, c2 = 1 - beta2^t.
__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    return v;
}

// Sum over the block (valid in thread 0) of K values.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K])
{
    __shared__ double lds[kThreads / 64][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.;
            for (int w = 0; w < kThreads / 64; ++w) s += lds[w][k];
            v[k] = s;
        }
}

// Final sum of K partials per block into s[off .. off + K), in a fixed order.
template <int K>
__global__ void __launch_bounds__(kThreads) k_finish(const double *__restrict__ part, uint32_t nb, double *__restrict__ s, int off,
                                                     Bat bt)
{
    if (stopped(bt)) return;
    part += blockIdx.y * bt.p;
    s += blockIdx.y * bt.s;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.;
    for (uint32_t b = threadIdx.x; b < nb; b += kThreads)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += part[K * b + k];
    block_sum<K>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) s[off + k] = v[k];
}

// The grid-vector reductions below run on at most kRedBlocks blocks (grid-stride loops), so
// that k_finish sums few partials.
constexpr uint32_t kRedBlocks = 1024;

// Partial sums of r^2 and of the inside-node count (for the RMS scale of the network input).
__global__ void __launch_bounds__(kThreads) k_rms_part(uint32_t np, const uint8_t *__restrict__ mask, const double *__restrict__ r, double *__restrict__ part,
                                                       Bat bt)
{
    if (stopped(bt)) return;
    r += blockIdx.y * bt.v;
    part += blockIdx.y * bt.p;
    double v[2] = {0., 0.};
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < np; i += gridDim.x * kThreads)
        if (mask[i]) {
            v[0] += r[i] * r[i];
            v[1] += 1.;
        }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

// Network input image (H x W x 16 bf16): channel 0 = r / rms on inside nodes of the grid (nx x
// ny, row-major), zero elsewhere and in the padding; sc[0] = sum r^2, sc[1] = inside count.
__global__ void k_net_input(uint32_t nx, uint32_t ny, int H, int W, const uint8_t *__restrict__ mask,
                            const double *__restrict__ r, const double *__restrict__ sc, uint16_t *__restrict__ img)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)H * W * 4) return;
    const int q = t & 3;
    const size_t p = t >> 2;
    const uint32_t y = (uint32_t)(p / W), x = (uint32_t)(p % W);
    v4s o = {0, 0, 0, 0};
    if (q == 0 && y < ny && x < nx) {
        const uint32_t i = y * nx + x;
        const double rms = sc[1] > 0. && sc[0] > 0. ? sqrt(sc[0] / sc[1]) : 1.;
        if (mask[i]) o[0] = (short)f2bf((float)(r[i] / rms));
    }
    *reinterpret_cast<v4s *>(img + p * C + 4 * q) = o;
}

__device__ __forceinline__ double apply_a(const double *__restrict__ cx, const double *__restrict__ cy,
                                          const double *__restrict__ x, uint32_t p, uint32_t nx)
{
    const double xp = x[p];
    return cx[p] * (xp - x[p + 1]) - cx[p - 1] * (x[p - 1] - xp) + cy[p] * (xp - x[p + nx]) -
           cy[p - nx] * (x[p - nx] - xp);
}

// Training loss: the A-norm error of the best multiple of the preconditioner's output z for
// the right-hand side q (exact solution phi = A^-1 q), which does not depend on the scale of
// z (the flexible CG does not either):
//     L = min_c (c z - phi)^T A (c z - phi) / (phi^T A phi) = 1 - (z^T q)^2 / ((z^T A z)(phi^T q)).
// With a = z^T q, b = z^T A z, c = phi^T q: dL/dz = (2 a / (b c)) ((a / b) A z - q), and
// z = rms * y on inside nodes (the forward pass writes z).  First pass: A z and the partials
// (b, c, a).  Second: dL/dy = rms dL/dz into channel 0 of the gradient image.

__global__ void __launch_bounds__(kThreads)
    k_loss_az(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
              const uint8_t *__restrict__ mask, const double *__restrict__ z, const double *__restrict__ phi,
              const double *__restrict__ q, double *__restrict__ az, double *__restrict__ part)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    double v[3] = {0., 0., 0.};
    if (i < np) {
        double a = 0.;
        if (mask[i]) {
            a = apply_a(cx, cy, z, i, nx);
            v[0] = z[i] * a;
            v[1] = phi[i] * q[i];
            v[2] = z[i] * q[i];
        }
        az[i] = a;
    }
    block_sum<3>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = v[k];
}

// sc[4] = z^T A z, sc[5] = phi^T q, sc[6] = z^T q
__global__ void k_loss_grad(uint32_t nx, uint32_t ny, int H, int W, const uint8_t *__restrict__ mask,
                            const double *__restrict__ az, const double *__restrict__ q, const double *__restrict__ sc,
                            uint16_t *__restrict__ gimg)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)H * W * 4) return;
    const int qd = t & 3;
    const size_t p = t >> 2;
    const uint32_t y = (uint32_t)(p / W), x = (uint32_t)(p % W);
    v4s o = {0, 0, 0, 0};
    if (qd == 0 && y < ny && x < nx && mask[y * nx + x]) {
        const double rms = sc[1] > 0. && sc[0] > 0. ? sqrt(sc[0] / sc[1]) : 1.;
        const double b = sc[4], c = sc[5], a = sc[6];
        const size_t i = (size_t)y * nx + x;
        const double g = (b != 0. && c != 0.) ? (2. * a / (b * c)) * ((a / b) * az[i] - q[i]) : 0.;
        o[0] = (short)f2bf((float)(rms * g));
    }
    *reinterpret_cast<v4s *>(gimg + p * C + 4 * qd) = o;
}

// ---- damped Jacobi around the network ----
// d = the FIT operator's diagonal at an inside node.
__device__ __forceinline__ double diag_a(const double *__restrict__ cx, const double *__restrict__ cy, uint32_t p,
                                         uint32_t nx)
{
    return cx[p] + cx[p - 1] + cy[p] + cy[p - nx];
}

// z = omega r / d on inside nodes, 0 elsewhere.
__global__ void k_jacobi(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
                         const uint8_t *__restrict__ mask, const double *__restrict__ r, double omega,
                         double *__restrict__ z)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) z[i] = mask[i] ? omega * r[i] / diag_a(cx, cy, i, nx) : 0.;
}

// out = r - A z on inside nodes, 0 elsewhere; e (optional) = phi - z.
__global__ void k_resid_az(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
                           const uint8_t *__restrict__ mask, const double *__restrict__ r, const double *__restrict__ z,
                           double *__restrict__ out, const double *__restrict__ phi, double *__restrict__ e)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const bool in = mask[i];
    out[i] = in ? r[i] - apply_a(cx, cy, z, i, nx) : 0.;
    if (e) e[i] = in ? phi[i] - z[i] : 0.;
}

// The pre-smoothing of the preconditioner in one pass: z1 = omega D^-1 r and r1 = r - A z1 on
// inside nodes (0 elsewhere; z1 at the neighbours is recomputed, giving k_jacobi and
// k_resid_az's values bit for bit); partials (r1.r1, inside count) for the network's input scale.
// Inside nodes lie off the grid's outer ring, so every neighbour index is in range, and a
// neighbour's own diagonal is only formed when it is inside.
__global__ void __launch_bounds__(kThreads)
    k_smooth1(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
              const uint8_t *__restrict__ mask, const double *__restrict__ r, double omega, double *__restrict__ z1,
              double *__restrict__ r1, double *__restrict__ part, Bat bt)
{
    if (stopped(bt)) return;
    r += blockIdx.y * bt.v;
    z1 += blockIdx.y * bt.v;
    r1 += blockIdx.y * bt.v;
    part += blockIdx.y * bt.p;
    double v[2] = {0., 0.};
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < np; i += gridDim.x * kThreads) {
        double zi = 0., ri = 0.;
        if (mask[i]) {
            auto zj = [&](uint32_t j) { return mask[j] ? omega * r[j] / diag_a(cx, cy, j, nx) : 0.; };
            zi = zj(i);
            const double ze = zj(i + 1), zw = zj(i - 1), zn = zj(i + nx), zs = zj(i - nx);
            ri = r[i] - (cx[i] * (zi - ze) - cx[i - 1] * (zw - zi) + cy[i] * (zi - zn) - cy[i - nx] * (zs - zi));
            v[0] += ri * ri;
            v[1] += 1.;
        }
        z1[i] = zi;
        r1[i] = ri;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

// an = A n, partials (n.r1, n.an)
__global__ void __launch_bounds__(kThreads)
    k_an_dots(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
              const uint8_t *__restrict__ mask, const double *__restrict__ n, const double *__restrict__ r1,
              double *__restrict__ an, double *__restrict__ part, Bat bt)
{
    if (stopped(bt)) return;
    n += blockIdx.y * bt.v;
    r1 += blockIdx.y * bt.v;
    an += blockIdx.y * bt.v;
    part += blockIdx.y * bt.p;
    double v[2] = {0., 0.};
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < np; i += gridDim.x * kThreads) {
        double a = 0.;
        if (mask[i]) {
            a = apply_a(cx, cy, n, i, nx);
            v[0] += n[i] * r1[i];
            v[1] += n[i] * a;
        }
        an[i] = a;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

// z = z1 + c n + omega (r1 - c A n) / d with c = (n.r1) / (n.A n) (sc[8], sc[9]), the A-norm
// optimal multiple of the network's output, of either sign: the training loss does not fix the
// output's sign, so a network may learn -A^-1 as well as A^-1.  c = 0 when n.A n is not
// positive (n = 0, or not finite).
__global__ void k_combine(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
                          const uint8_t *__restrict__ mask, const double *__restrict__ z1, const double *__restrict__ n,
                          const double *__restrict__ r1, const double *__restrict__ an, const double *__restrict__ sc,
                          double omega, double *__restrict__ z, Bat bt)
{
    if (stopped(bt)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const size_t o = blockIdx.y * bt.v;
    z1 += o;
    n += o;
    r1 += o;
    an += o;
    z += o;
    sc += blockIdx.y * bt.s;
    const double c = sc[9] > 0. ? sc[8] / sc[9] : 0.;
    z[i] = mask[i] ? z1[i] + c * n[i] + omega * (r1[i] - c * an[i]) / diag_a(cx, cy, i, nx) : 0.;
}

// ---- flexible conjugate gradients (Notay 2000: FCG with m_max kept directions) ----
// Direction k is the preconditioned residual z_k A-orthogonalised against the last
// min(k, m) directions, p_k = z_k - sum_j (z_k.Ap_j / p_j.Ap_j) p_j, held in a ring of m + 1
// slots (slot k mod (m + 1)); alpha_k = p_k.r_k / p_k.Ap_k.  m = 1 is the truncated form
// (beta = -z.Ap / p.Ap).  Keeping more directions restores what a preconditioner that varies
// from one application to the next (a ReLU network and its line search) takes away from the
// CG recurrence.
// scalars per run: s[3] alpha, s[6] p.Ap, s[7] p.r, s[8] r.r (after the update), s[9] q.q (init;
// then the update's node count), s[12] stopped, s[13] its iteration, s[14] q.q, s[16..23] z.Ap_j,
// s[24..32] p.Ap per ring slot
constexpr uint32_t FCG_MAXM = 8;

__global__ void __launch_bounds__(kThreads)
    k_fcg_init(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
               const uint8_t *__restrict__ mask, const double *__restrict__ q, double *__restrict__ phi,
               double *__restrict__ r, double *__restrict__ part, Bat bt)
{
    q += blockIdx.y * bt.v;
    phi += blockIdx.y * bt.v;
    r += blockIdx.y * bt.v;
    part += blockIdx.y * bt.p;
    double v[2] = {0., 0.};
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < np; i += gridDim.x * kThreads) {
        double ri = 0.;
        if (mask[i]) {
            ri = q[i] - apply_a(cx, cy, phi, i, nx);
            v[0] += ri * ri;
            v[1] += q[i] * q[i];
        }
        r[i] = ri;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

__global__ void k_mask_zero(uint32_t np, const uint8_t *__restrict__ mask, double *__restrict__ x, Bat bt)
{
    x += blockIdx.y * bt.v;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np && !mask[i]) x[i] = 0.;
}

// Partial z.Ap_j for the window's directions j = k - 1 - t, t < nUsed (FCG_MAXM per block).
__global__ void __launch_bounds__(kThreads)
    k_fcg_zdots(uint32_t np, const double *__restrict__ z, const double *__restrict__ AP, uint32_t k, uint32_t m1,
                uint32_t nUsed, double *__restrict__ part, Bat bt)
{
    if (stopped(bt)) return;
    z += blockIdx.y * bt.v;
    AP += blockIdx.y * bt.v;
    part += blockIdx.y * bt.p;
    double v[FCG_MAXM];
#pragma unroll
    for (uint32_t t = 0; t < FCG_MAXM; ++t) v[t] = 0.;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < np; i += gridDim.x * kThreads) {
        const double zi = z[i];
#pragma unroll
        for (uint32_t t = 0; t < FCG_MAXM; ++t)
            if (t < nUsed) v[t] += zi * AP[(size_t)((k - 1 - t) % m1) * bt.w + i];
    }
    block_sum<FCG_MAXM>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (uint32_t t = 0; t < FCG_MAXM; ++t) part[FCG_MAXM * blockIdx.x + t] = v[t];
}

// p_k = z - sum_t (d_t / pAp[slot]) p_slot into slot k mod m1 (d = z.Ap_j from k_fcg_zdots; a
// direction with p.Ap <= 0 is skipped).
__global__ void k_fcg_newdir(uint32_t np, const double *__restrict__ z, double *__restrict__ P,
                             const double *__restrict__ d, const double *__restrict__ pAp, uint32_t k, uint32_t m1,
                             uint32_t nUsed, Bat bt)
{
    if (stopped(bt)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    z += blockIdx.y * bt.v;
    P += blockIdx.y * bt.v;
    d += blockIdx.y * bt.s;
    pAp += blockIdx.y * bt.s;
    double v = z[i];
    for (uint32_t t = 0; t < nUsed; ++t) {
        const uint32_t sl = (k - 1 - t) % m1;
        const double den = pAp[sl];
        if (den > 0.) v -= (d[t] / den) * P[(size_t)sl * bt.w + i];
    }
    P[(size_t)(k % m1) * bt.w + i] = v;
}

// Ap; partial (p.Ap, p.r)
__global__ void __launch_bounds__(kThreads)
    k_fcg_ap(uint32_t nx, uint32_t np, const double *__restrict__ cx, const double *__restrict__ cy,
             const uint8_t *__restrict__ mask, const double *__restrict__ p, const double *__restrict__ r,
             double *__restrict__ ap, double *__restrict__ part, Bat bt)
{
    if (stopped(bt)) return;
    p += blockIdx.y * bt.v;
    r += blockIdx.y * bt.v;
    ap += blockIdx.y * bt.v;
    part += blockIdx.y * bt.p;
    double v[2] = {0., 0.};
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < np; i += gridDim.x * kThreads) {
        double a = 0.;
        if (mask[i]) {
            a = apply_a(cx, cy, p, i, nx);
            v[0] += p[i] * a;
            v[1] += p[i] * r[i];
        }
        ap[i] = a;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

// alpha = p.r / p.Ap (0 when p.Ap is not positive); p.Ap kept for the direction's slot.  One
// thread per run; a run that has stopped (s[12] != 0) is left as it is (its other kernels skip
// it too, so its phi, r and s[8] keep their values bit for bit).
__global__ void k_alpha(double *s, double *pAp, uint32_t slot, uint32_t R, Bat bt)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= R) return;
    s += b * bt.s;
    pAp += b * bt.s;
    if (s[12] != 0.) return;
    s[3] = s[6] > 0. ? s[7] / s[6] : 0.;
    pAp[slot] = s[6];
}

// The convergence check every `check` iterations, per run (one thread each): a run stops (s[12]
// = 1, s[13] = it) once sqrt(r.r) <= tol |q| (|q|^2 in s[14], taken as 1 when q = 0) or r.r is
// not finite
// -- the test the single-problem loop applies at the same iterations.
__global__ void k_fcg_check(double *s, uint32_t R, Bat bt, double tol, uint32_t it)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= R) return;
    s += b * bt.s;
    if (s[12] != 0.) return;
    const double rr = s[8], qn = sqrt(s[14]) > 0. ? sqrt(s[14]) : 1.;
    if (!(sqrt(rr) > tol * qn) || !(rr == rr)) {
        s[12] = 1.;
        s[13] = (double)it;
    }
}

// Clears the stop flags (s[12], s[13]) of R runs and keeps q.q (s[9] after k_fcg_init's
// reduction; later reductions reuse s[9]) in s[14].
__global__ void k_fcg_start(double *s, uint32_t R, Bat bt)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= R) return;
    s += b * bt.s;
    s[12] = 0.;
    s[13] = 0.;
    s[14] = s[9];
}

// phi += alpha p, r -= alpha Ap; partial r.r.
__global__ void __launch_bounds__(kThreads)
    k_fcg_update(uint32_t np, const double *__restrict__ s, double *__restrict__ phi, double *__restrict__ r,
                 const double *__restrict__ p, const double *__restrict__ ap, double *__restrict__ part, Bat bt)
{
    if (stopped(bt)) return;
    s += blockIdx.y * bt.s;
    phi += blockIdx.y * bt.v;
    r += blockIdx.y * bt.v;
    p += blockIdx.y * bt.v;
    ap += blockIdx.y * bt.v;
    part += blockIdx.y * bt.p;
    double v[2] = {0., 0.};
    const double a = s[3];
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < np; i += gridDim.x * kThreads) {
        const double ri = r[i] - a * ap[i];
        phi[i] += a * p[i];
        r[i] = ri;
        v[0] += ri * ri;
        v[1] += 1.;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

uint32_t nblk(size_t n, int b = kThreads) { return (uint32_t)((n + b - 1) / b); }

// Per-image strides of the preconditioner's own buffers (sv: grid vectors np apart, part:
// kPartImg partial sums, scal: kScalImg scalars) for a batch of problems on one nx x ny grid.
constexpr size_t kPartImg = 4 * (size_t)kRedBlocks + 64;
constexpr uint32_t kScalImg = 64;
Bat unet_bat(uint32_t np) { return Bat{np, np, kPartImg, kScalImg}; }


// Activation images per level l (H_l = H >> l): x0 (input), a (first conv), s (second conv /
// skip), p (pooled input of the level), t (up + skip input of an up level), u (up level's
// first conv), v (up level's output).  Index: 7 l + slot.
enum { IX = 0, IA = 1, IS = 2, IP = 3, IT = 4, IU = 5, IV = 6 };

size_t img_elems(const oecl_unet *u, int l) { return (size_t)(u->H >> l) * (u->W >> l) * C; }

// Rows per convolution block: the tallest tile that still leaves >= 1024 blocks (4 per CU) over
// the nimg images of a launch.
int conv_rows(const oecl_unet *u, int H, int W, uint32_t nimg = 1)
{
    if (u->convRows) return u->convRows;
    const long quads = (W + 63) / 64;
    for (int rw = 8; rw > 1; rw >>= 1)
        if ((H + rw - 1) / rw * quads * (long)nimg >= 1024) return rw;
    return 1;
}

template <int MODE, int SRC = 0>
void launch_conv_mode(int rw, dim3 g, hipStream_t s, int H, int W, const uint16_t *in, const uint16_t *w,
                      const float *b, void *out, const GridIO &io, const uint16_t *up, uint16_t *pooled)
{
    const dim3 bl(kThreads);
    if (rw == 8) hipLaunchKernelGGL((k_conv<MODE, 8, SRC>), g, bl, 0, s, H, W, in, w, b, out, io, up, pooled);
    else if (rw == 4) hipLaunchKernelGGL((k_conv<MODE, 4, SRC>), g, bl, 0, s, H, W, in, w, b, out, io, up, pooled);
    else if (rw == 2) hipLaunchKernelGGL((k_conv<MODE, 2, SRC>), g, bl, 0, s, H, W, in, w, b, out, io, up, pooled);
    else hipLaunchKernelGGL((k_conv<MODE, 1, SRC>), g, bl, 0, s, H, W, in, w, b, out, io, up, nullptr);
}

// mode 0: ReLU, 1: linear, 2: FP32 channel 0; 3: the same into io.z; 4: ReLU with the input
// staged from io.r (both on level 0 only); 5: ReLU with the input staged as up(up) + in.
// pooled (modes 0, 1): also write the 2x2 average of the output there when the tile has two or
// more rows; returns 1 when it did (the caller pools otherwise), 0 when not, < 0 on error.
int launch_conv(const oecl_unet *u, int l, int mode, const uint16_t *in, const uint16_t *w, const float *b, void *out,
                hipStream_t s, const GridIO &io = GridIO{}, const uint16_t *up = nullptr, uint16_t *pooled = nullptr,
                uint32_t nimg = 1)
{
    const int H = u->H >> l, W = u->W >> l;
    const int rw = conv_rows(u, H, W, nimg);
    const dim3 g((uint32_t)((H + rw - 1) / rw) * ((W + 63) / 64), nimg);  // one block per RW x 64 tile and image
    uint16_t *pl = (mode == 0 || mode == 1) && rw >= 2 ? pooled : nullptr;
    if (mode == 0) launch_conv_mode<0>(rw, g, s, H, W, in, w, b, out, io, up, pl);
    else if (mode == 1) launch_conv_mode<1>(rw, g, s, H, W, in, w, b, out, io, up, pl);
    else if (mode == 2) launch_conv_mode<2>(rw, g, s, H, W, in, w, b, out, io, up, nullptr);
    else if (mode == 3) launch_conv_mode<3>(rw, g, s, H, W, in, w, b, out, io, up, nullptr);
    else if (mode == 4) launch_conv_mode<0, 1>(rw, g, s, H, W, in, w, b, out, io, up, nullptr);
    else launch_conv_mode<0, 2>(rw, g, s, H, W, in, w, b, out, io, up, nullptr);
    return pl ? 1 : 0;
}

const uint16_t *Wc(const oecl_unet *u, int c) { return u->Wb + (size_t)c * WSZ; }
const uint16_t *WTc(const oecl_unet *u, int c) { return u->WbT + (size_t)c * WSZ; }
const float *bc(const oecl_unet *u, int c) { return u->bm + (size_t)c * C; }

// Forward pass from the input image act[IX] (level 0) to u->out.
// io != null: a pass from the grid vector io->r to io->z (no FP32 output image).  Without tape
// it stages the input and the up levels' sums inside the convolutions (inference); with tape
// (training) it writes the input image and those sums, which the backward pass reads.
void unet_forward(oecl_unet *u, hipStream_t s, const GridIO *io = nullptr, bool tape = false, uint32_t nimg = 1)
{
    uint16_t **A = u->act;
    auto pool = [&](int l, const uint16_t *in, uint16_t *out) {  // level l-1 -> l
        const int H = u->H >> l, W = u->W >> l;
        hipLaunchKernelGGL(k_pool, dim3(nblk((size_t)H * W * 4), nimg), dim3(kThreads), 0, s, H, W, in, out);
    };
    auto upadd = [&](int l, const uint16_t *uu, const uint16_t *sk, uint16_t *out) {  // level l+1 -> l
        const int H = u->H >> l, W = u->W >> l;
        hipLaunchKernelGGL(k_up_add<false>, dim3(nblk((size_t)H * W * 4), nimg), dim3(kThreads), 0, s, H, W, uu, sk,
                           out);
    };
    GridIO none{};  // the convolutions without grid input or output: only the stop flags
    if (io) {
        none.st = io->st;
        none.ss = io->ss;
    }
    auto conv = [&](int l, int mode, const uint16_t *in, int c, void *out, const GridIO &g,
                    const uint16_t *up = nullptr, uint16_t *pooled = nullptr) {
        return launch_conv(u, l, mode, in, Wc(u, c), bc(u, c), out, s, g, up, pooled, nimg);
    };
    const bool fused = io && !tape;
    if (fused) {
        conv(0, 4, nullptr, 0, A[IA], *io);
    } else {
        if (io)
            hipLaunchKernelGGL(k_net_input, dim3(nblk((size_t)u->H * u->W * 4)), dim3(kThreads), 0, s, io->nx, io->ny,
                               (int)u->H, (int)u->W, io->mask, io->r, io->sc, A[IX]);
        conv(0, 0, A[IX], 0, A[IA], none);
    }
    // each level's second convolution also pools its output for the next level when it can
    int pooledNext = conv(0, 0, A[IA], 1, A[IS], none, nullptr, A[7 + IP]);
    for (int l = 1; l <= 3; ++l) {
        if (!pooledNext) pool(l, A[7 * (l - 1) + IS], A[7 * l + IP]);
        conv(l, 0, A[7 * l + IP], 2 * l, A[7 * l + IA], none);
        pooledNext = conv(l, 0, A[7 * l + IA], 2 * l + 1, A[7 * l + IS], none, nullptr,
                          l < 3 ? A[7 * (l + 1) + IP] : nullptr);
    }
    // level 3's output is u3 = A[7*3 + IS]
    for (int l = 2; l >= 0; --l) {
        const uint16_t *below = l == 2 ? A[7 * 3 + IS] : A[7 * (l + 1) + IV];
        const int c0 = 8 + 2 * (2 - l);
        if (fused) {  // inference: the up level's sum is staged by its first convolution
            conv(l, 5, A[7 * l + IS], c0, A[7 * l + IU], *io, below);
        } else {
            upadd(l, below, A[7 * l + IS], A[7 * l + IT]);
            conv(l, 0, A[7 * l + IT], c0, A[7 * l + IU], none);
        }
        if (l > 0) conv(l, 0, A[7 * l + IU], c0 + 1, A[7 * l + IV], none);
        else if (io) conv(0, 3, A[IU], 13, nullptr, *io);
        else conv(0, 2, A[IU], 13, u->out, none);
    }
}

// Backward pass: gout (level-0 image, channel 0 = dL/dy) -> gW, gb (accumulated) and, when gin
// is not null, the gradient of the input image.  Scratch: P_l = grad[2 l], Q_l = grad[2 l + 1]
// (level-l size).  The skip gradients (grad of t_l, which feeds s_l directly) stay in Q_0, P_1,
// P_2 until the down path adds the pooling adjoint of the level below and consumes them.
void unet_backward(oecl_unet *u, const uint16_t *gout, uint16_t *gin, hipStream_t s)
{
    uint16_t **A = u->act;
    uint16_t *P[4], *Q[4];
    for (int l = 0; l < 4; ++l) {
        P[l] = u->grad[2 * l];
        Q[l] = u->grad[2 * l + 1];
    }
    auto relu_back = [&](int l, uint16_t *g, const uint16_t *act) {
        const size_t n4 = img_elems(u, l) / 4;
        hipLaunchKernelGGL(k_relu_back, dim3(nblk(n4)), dim3(kThreads), 0, s, n4, g, act);
    };
    // conv c at level l: input x, gradient g of its (pre-activation) output -> wgrad, and the
    // input gradient into gx (when not null)
    auto convb = [&](int l, int c, const uint16_t *x, const uint16_t *g, uint16_t *gx) {
        const int H = u->H >> l, W = u->W >> l;
        const uint32_t segs = (uint32_t)H * ((W + 15) / 16);
        const uint32_t blocks = std::min(nblk(segs, kThreads / 64), kWgradBlocks);
        hipLaunchKernelGGL(k_conv_wgrad, dim3(blocks), dim3(kThreads), 0, s, H, W, x, g, u->wpart);
        hipLaunchKernelGGL(k_wgrad_sum, dim3((WG + 31) / 32), dim3(256), 0, s, blocks, u->wpart,
                           u->gW + (size_t)c * WSZ, u->gb + (size_t)c * C);
        if (gx) launch_conv(u, l, 1, g, WTc(u, c), u->bm + (size_t)NCONV * C /* zero bias row */, gx, s);
    };
    auto up_back = [&](int l, const uint16_t *big, uint16_t *small) {  // level l-1 -> level l
        const int H = u->H >> l, W = u->W >> l;
        hipLaunchKernelGGL(k_up_back, dim3(nblk((size_t)H * W * 4)), dim3(kThreads), 0, s, H, W, big, small);
    };
    auto pool_back_add = [&](int l, const uint16_t *small, uint16_t *skip) {  // level l+1 -> level l
        const int H = u->H >> l, W = u->W >> l;
        hipLaunchKernelGGL(k_up_add<true>, dim3(nblk((size_t)H * W * 4)), dim3(kThreads), 0, s, H, W, small, skip,
                           skip);
    };
    // up level 0: y = c13(relu(c12(t0)))
    convb(0, 13, A[IU], gout, P[0]);
    relu_back(0, P[0], A[IU]);
    convb(0, 12, A[IT], P[0], Q[0]);              // Q0 = grad t0 (skip of s0)
    // up level 1: u1 = relu(c11(relu(c10(t1))))
    up_back(1, Q[0], P[1]);                        // grad u1
    relu_back(1, P[1], A[7 + IV]);
    convb(1, 11, A[7 + IU], P[1], Q[1]);
    relu_back(1, Q[1], A[7 + IU]);
    convb(1, 10, A[7 + IT], Q[1], P[1]);           // P1 = grad t1 (skip of s1)
    // up level 2: u2 = relu(c9(relu(c8(t2))))
    up_back(2, P[1], P[2]);                        // grad u2
    relu_back(2, P[2], A[14 + IV]);
    convb(2, 9, A[14 + IU], P[2], Q[2]);
    relu_back(2, Q[2], A[14 + IU]);
    convb(2, 8, A[14 + IT], Q[2], P[2]);           // P2 = grad t2 (skip of s2)
    // bottom, level 3: u3 = relu(c7(relu(c6(pool(s2)))))
    up_back(3, P[2], P[3]);                        // grad u3
    relu_back(3, P[3], A[21 + IS]);
    convb(3, 7, A[21 + IA], P[3], Q[3]);
    relu_back(3, Q[3], A[21 + IA]);
    convb(3, 6, A[21 + IP], Q[3], P[3]);           // P3 = grad pool(s2)
    pool_back_add(2, P[3], P[2]);                  // P2 = grad s2
    // down level 2: s2 = relu(c5(relu(c4(pool(s1)))))
    relu_back(2, P[2], A[14 + IS]);
    convb(2, 5, A[14 + IA], P[2], Q[2]);
    relu_back(2, Q[2], A[14 + IA]);
    convb(2, 4, A[14 + IP], Q[2], P[2]);           // P2 = grad pool(s1)
    pool_back_add(1, P[2], P[1]);                  // P1 = grad s1
    // down level 1
    relu_back(1, P[1], A[7 + IS]);
    convb(1, 3, A[7 + IA], P[1], Q[1]);
    relu_back(1, Q[1], A[7 + IA]);
    convb(1, 2, A[7 + IP], Q[1], P[1]);            // P1 = grad pool(s0)
    pool_back_add(0, P[1], Q[0]);                  // Q0 = grad s0
    // down level 0: s0 = relu(c1(relu(c0(x))))
    relu_back(0, Q[0], A[IS]);
    convb(0, 1, A[IA], Q[0], P[0]);
    relu_back(0, P[0], A[IA]);
    convb(0, 0, A[IX], P[0], gin);
}

}  // namespace

OECL_EXPORT int oecl_unet_create(uint32_t H, uint32_t W, uint32_t seed, oecl_unet **out)
{
    if (!out || H < 8 || W < 8 || (H & 7) || (W & 7) || H > 16384 || W > 16384) return OECL_ERR_ARG;
    oecl_unet *u = new oecl_unet();
    u->H = H;
    u->W = W;
    u->batch = 1;
    const size_t nW = (size_t)NCONV * WSZ, nb = (size_t)(NCONV + 1) * C;  // + one zero bias row
    // carve the buffers out of one allocation: a first pass sizes it, the second assigns
    char *base = nullptr;
    size_t used = 0;
    auto take = [&](size_t b) {
        char *r = base ? base + used : nullptr;
        used += (b + 255) / 256 * 256;
        return r;
    };
    for (int pass = 0; pass < 2; ++pass) {
        used = 0;
        u->Wm = (float *)take(4 * nW);
        u->bm = (float *)take(4 * nb);
        u->mW = (float *)take(4 * nW);
        u->vW = (float *)take(4 * nW);
        u->mb = (float *)take(4 * nb);
        u->vb = (float *)take(4 * nb);
        u->gW = (float *)take(4 * nW);
        u->gb = (float *)take(4 * nb);
        u->wpart = (float *)take(4ull * kWgradBlocks * WG);
        u->Wb = (uint16_t *)take(2 * nW);
        u->WbT = (uint16_t *)take(2 * nW);
        for (int l = 0; l < 4; ++l)
            for (int k = 0; k < 7; ++k) u->act[7 * l + k] = (uint16_t *)take(2 * img_elems(u, l));
        for (int l = 0; l < 4; ++l) {
            u->grad[2 * l] = (uint16_t *)take(2 * img_elems(u, l));
            u->grad[2 * l + 1] = (uint16_t *)take(2 * img_elems(u, l));
        }
        u->out = (float *)take(4ull * H * W);
        for (int k = 0; k < 4; ++k) u->sv[k] = (double *)take(8ull * H * W);
        u->scal = (double *)take(8 * kScalImg);
        u->part = (double *)take(8ull * std::max(4 * ((size_t)nblk((size_t)H * W) + 16), kPartImg));
        if (pass == 0) {
            if (hipMalloc(&base, used) != hipSuccess) {
                delete u;
                oecl_set_error_msg("unet alloc failed");
                return OECL_ERR_HIP;
            }
            u->mem = base;
        }
    }
    OECL_HIP(hipMemset(u->mem, 0, used), "unet clear");
    // He initialisation from a fixed-seed host generator (xorshift + Box-Muller)
    std::vector<float> w(nW);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ ((uint64_t)seed << 17 | seed);
    auto next = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    const double sd = sqrt(2. / (9. * C));
    for (size_t i = 0; i < nW; ++i) {
        const double a = next(), b = next();
        w[i] = (float)(sd * sqrt(-2. * log(1. - a)) * cos(6.283185307179586 * b));
    }
    OECL_HIP(hipMemcpy(u->Wm, w.data(), 4 * nW, hipMemcpyHostToDevice), "unet weights");
    hipLaunchKernelGGL(k_weights_bf16, dim3(nblk(nW)), dim3(kThreads), 0, 0, u->Wm, u->Wb, u->WbT);
    OECL_CHECK_LAUNCH("oecl_unet_create");
    OECL_HIP(hipDeviceSynchronize(), "unet create");
    *out = u;
    return OECL_OK;
}

OECL_EXPORT int oecl_unet_destroy(oecl_unet *u)
{
    if (!u) return OECL_OK;
    (void)hipFree(u->mem);
    if (u->bmem) (void)hipFree(u->bmem);
    delete u;
    return OECL_OK;
}

// Room for batches of up to R inference images (oecl_fcg_solve_batch): the activation images,
// the smoothed preconditioner's grid vectors and the reduction buffers are reallocated R times
// over (contents are not kept; the weights and the training state are untouched).
OECL_EXPORT int oecl_unet_reserve(oecl_unet *u, uint32_t R)
{
    if (!u || R < 1 || R > 4096) return OECL_ERR_ARG;
    if (R <= u->batch) return OECL_OK;
    const size_t hw = (size_t)u->H * u->W;
    char *base = nullptr;
    size_t used = 0;
    auto take = [&](size_t b) {
        char *r = base ? base + used : nullptr;
        used += (b + 255) / 256 * 256;
        return r;
    };
    uint16_t *act[32];
    double *sv[4], *scal = nullptr, *part = nullptr;
    for (int pass = 0; pass < 2; ++pass) {
        used = 0;
        for (int l = 0; l < 4; ++l)
            for (int k = 0; k < 7; ++k) act[7 * l + k] = (uint16_t *)take(2 * img_elems(u, l) * R);
        for (int k = 0; k < 4; ++k) sv[k] = (double *)take(8 * hw * R);
        scal = (double *)take(8ull * kScalImg * R);
        part = (double *)take(8 * (std::max(4 * ((size_t)nblk(hw) + 16), kPartImg) + kPartImg * (R - 1)));
        if (pass == 0 && hipMalloc(&base, used) != hipSuccess) {
            oecl_set_error_msg("unet batch alloc failed");
            return OECL_ERR_HIP;
        }
    }
    OECL_HIP(hipMemset(base, 0, used), "unet batch clear");
    OECL_HIP(hipDeviceSynchronize(), "unet batch");
    if (u->bmem) (void)hipFree(u->bmem);
    u->bmem = base;
    for (int k = 0; k < 28; ++k) u->act[k] = act[k];
    for (int k = 0; k < 4; ++k) u->sv[k] = sv[k];
    u->scal = scal;
    u->part = part;
    u->batch = R;
    return OECL_OK;
}

OECL_EXPORT int oecl_unet_set_conv_rows(oecl_unet *u, int rows)
{
    if (!u || !(rows == 0 || rows == 1 || rows == 2 || rows == 4 || rows == 8)) return OECL_ERR_ARG;
    u->convRows = rows;
    return OECL_OK;
}

// Host copies of the FP32 parameters: W [14][9][16][16], b [14][16].
OECL_EXPORT int oecl_unet_set_params(oecl_unet *u, const float *W, const float *b)
{
    if (!u || !W || !b) return OECL_ERR_ARG;
    OECL_HIP(hipMemcpy(u->Wm, W, 4ull * NCONV * WSZ, hipMemcpyHostToDevice), "unet set W");
    OECL_HIP(hipMemcpy(u->bm, b, 4ull * NCONV * C, hipMemcpyHostToDevice), "unet set b");
    hipLaunchKernelGGL(k_weights_bf16, dim3(nblk((size_t)NCONV * WSZ)), dim3(kThreads), 0, 0, u->Wm, u->Wb, u->WbT);
    OECL_CHECK_LAUNCH("oecl_unet_set_params");
    OECL_HIP(hipDeviceSynchronize(), "unet set");
    return OECL_OK;
}

OECL_EXPORT int oecl_unet_get_params(const oecl_unet *u, float *W, float *b, float *gW, float *gb)
{
    if (!u) return OECL_ERR_ARG;
    OECL_HIP(hipDeviceSynchronize(), "unet get");
    if (W) OECL_HIP(hipMemcpy(W, u->Wm, 4ull * NCONV * WSZ, hipMemcpyDeviceToHost), "unet get W");
    if (b) OECL_HIP(hipMemcpy(b, u->bm, 4ull * NCONV * C, hipMemcpyDeviceToHost), "unet get b");
    if (gW) OECL_HIP(hipMemcpy(gW, u->gW, 4ull * NCONV * WSZ, hipMemcpyDeviceToHost), "unet get gW");
    if (gb) OECL_HIP(hipMemcpy(gb, u->gb, 4ull * NCONV * C, hipMemcpyDeviceToHost), "unet get gb");
    return OECL_OK;
}

// Forward pass of an input image img (H x W x 16 bf16, channel 0 used) into y (H x W FP32).
OECL_EXPORT int oecl_unet_forward(oecl_unet *u, const uint16_t *img, float *y, void *stream)
{
    if (!u || !img) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    OECL_HIP(hipMemcpyAsync(u->act[IX], img, img_elems(u, 0) * 2, hipMemcpyDeviceToDevice, s), "unet input");
    unet_forward(u, s);
    if (y) OECL_HIP(hipMemcpyAsync(y, u->out, 4ull * u->H * u->W, hipMemcpyDeviceToDevice, s), "unet output");
    OECL_CHECK_LAUNCH("oecl_unet_forward");
    return OECL_OK;
}

// Backward pass after oecl_unet_forward: gout (H x W x 16 bf16, channel 0 = dL/dy).  The weight
// and bias gradients are zeroed first when zero != 0, else accumulated; gin (optional) receives
// the gradient of the input image.
OECL_EXPORT int oecl_unet_backward(oecl_unet *u, const uint16_t *gout, uint16_t *gin, int zero, void *stream)
{
    if (!u || !gout) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (zero) {
        OECL_HIP(hipMemsetAsync(u->gW, 0, 4ull * NCONV * WSZ, s), "unet zero");
        OECL_HIP(hipMemsetAsync(u->gb, 0, 4ull * NCONV * C, s), "unet zero");
    }
    unet_backward(u, gout, gin, s);
    OECL_CHECK_LAUNCH("oecl_unet_backward");
    return OECL_OK;
}

// One Adam step on the accumulated gradients, then the bf16 weight copies.
OECL_EXPORT int oecl_unet_adam(oecl_unet *u, float lr, float beta1, float beta2, float eps, void *stream)
{
    if (!u) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    u->adamT += 1;
    const float c1 = 1.f - powf(beta1, (float)u->adamT), c2 = 1.f - powf(beta2, (float)u->adamT);
    const int nW = NCONV * WSZ, nb = NCONV * C;
    hipLaunchKernelGGL(k_adam, dim3(nblk(nW)), dim3(kThreads), 0, s, nW, u->Wm, u->gW, u->mW, u->vW, lr, beta1, beta2,
                       eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(nblk(nb)), dim3(kThreads), 0, s, nb, u->bm, u->gb, u->mb, u->vb, lr, beta1, beta2,
                       eps, c1, c2);
    hipLaunchKernelGGL(k_weights_bf16, dim3(nblk((size_t)nW)), dim3(kThreads), 0, s, u->Wm, u->Wb, u->WbT);
    OECL_CHECK_LAUNCH("oecl_unet_adam");
    return OECL_OK;
}

// z = rms(r) * Net(r / rms(r)) on the inside nodes of an nx x ny grid (row-major, mask = inside).
// z = rms(r) * Net(r / rms(r)) on the inside nodes of an nx x ny grid (row-major, mask = inside);
// nimg problems at once (r and z np apart, np = nx ny).
static int precondition(oecl_unet *u, uint32_t nx, uint32_t ny, const uint8_t *mask, const double *r, double *z,
                        hipStream_t s, bool tape = false, uint32_t nimg = 1, const double *st = nullptr)
{
    const uint32_t np = nx * ny, nbr = std::min(nblk(np), kRedBlocks);
    Bat bt = unet_bat(np);
    bt.st = st;
    hipLaunchKernelGGL(k_rms_part, dim3(nbr, nimg), dim3(kThreads), 0, s, np, mask, r, u->part, bt);
    hipLaunchKernelGGL(k_finish<2>, dim3(1, nimg), dim3(kThreads), 0, s, u->part, nbr, u->scal, 0, bt);
    const GridIO io{r, z, mask, u->scal, nx, ny, np, kScalImg, st};
    unet_forward(u, s, &io, tape, nimg);
    return OECL_OK;
}

// The network between damped Jacobi smoothings (omega > 0): z1 = omega D^-1 r; the network's
// n for r1 = r - A z1, scaled by the A-norm line search c = n.r1 / n.A n; z = z1 + c n +
// omega D^-1 (r1 - c A n).  omega = 0: the network alone.
static int precondition_smoothed(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                                 const uint8_t *mask, const double *r, double *z, double omega, hipStream_t s,
                                 uint32_t nimg = 1, const double *st = nullptr)
{
    if (nimg > u->batch) return OECL_ERR_ARG;
    if (!(omega > 0.)) return precondition(u, nx, ny, mask, r, z, s, false, nimg, st);
    const uint32_t np = nx * ny, nb = nblk(np), nbr = std::min(nb, kRedBlocks);
    Bat bt = unet_bat(np);
    bt.st = st;
    double *z1 = u->sv[0], *r1 = u->sv[1], *n = u->sv[2], *an = u->sv[3];
    hipLaunchKernelGGL(k_smooth1, dim3(nbr, nimg), dim3(kThreads), 0, s, nx, np, cx, cy, mask, r, omega, z1, r1, u->part,
                       bt);
    hipLaunchKernelGGL(k_finish<2>, dim3(1, nimg), dim3(kThreads), 0, s, u->part, nbr, u->scal, 0, bt);
    const GridIO io{r1, n, mask, u->scal, nx, ny, np, kScalImg, st};
    unet_forward(u, s, &io, false, nimg);
    hipLaunchKernelGGL(k_an_dots, dim3(nbr, nimg), dim3(kThreads), 0, s, nx, np, cx, cy, mask, n, r1, an, u->part, bt);
    hipLaunchKernelGGL(k_finish<2>, dim3(1, nimg), dim3(kThreads), 0, s, u->part, nbr, u->scal, 8, bt);
    hipLaunchKernelGGL(k_combine, dim3(nb, nimg), dim3(kThreads), 0, s, nx, np, cx, cy, mask, z1, n, r1, an, u->scal,
                       omega, z, bt);
    return OECL_OK;
}

OECL_EXPORT int oecl_unet_precondition(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                                       const uint8_t *mask, const double *r, double *z, double omega, void *stream)
{
    if (!u || !mask || !r || !z || ny > u->H || nx > u->W) return OECL_ERR_ARG;
    if (omega > 0. && (!cx || !cy)) return OECL_ERR_ARG;
    int rc = precondition_smoothed(u, nx, ny, cx, cy, mask, r, z, omega, (hipStream_t)stream);
    OECL_CHECK_LAUNCH("oecl_unet_precondition");
    return rc;
}

// The pair the network meets inside the smoothed preconditioner: for (q, phi = A^-1 q),
// r1 = q - A (omega D^-1 q) and e1 = phi - omega D^-1 q, so that A e1 = r1.
OECL_EXPORT int oecl_unet_smooth_pair(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                                      const uint8_t *mask, const double *q, const double *phi, double omega, double *r1,
                                      double *e1, void *stream)
{
    if (!u || !cx || !cy || !mask || !q || !phi || !r1 || !e1 || ny > u->H || nx > u->W) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t np = nx * ny, nb = nblk(np);
    hipLaunchKernelGGL(k_jacobi, dim3(nb), dim3(kThreads), 0, s, nx, np, cx, cy, mask, q, omega, u->sv[0]);
    hipLaunchKernelGGL(k_resid_az, dim3(nb), dim3(kThreads), 0, s, nx, np, cx, cy, mask, q, u->sv[0], r1, phi, e1);
    OECL_CHECK_LAUNCH("oecl_unet_smooth_pair");
    return OECL_OK;
}

// One training sample: forward on q (the right-hand side, scaled to unit RMS), the
// scale-free A-norm loss against the exact solution phi, backward (gradients accumulated;
// zeroed when zero != 0).  work: 2 np doubles.  *loss (optional, synchronises) receives L.
OECL_EXPORT int oecl_unet_train_sample(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                                       const uint8_t *mask, const double *q, const double *phi, double *work,
                                       int zero, double *loss, void *stream)
{
    if (!u || !cx || !cy || !mask || !q || !phi || !work || ny > u->H || nx > u->W) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t np = nx * ny, nb = nblk(np);
    double *z = work, *az = work + np;
    int rc = precondition(u, nx, ny, mask, q, z, s, true);  // z, input RMS in scal[0..1], the tape
    if (rc) return rc;
    hipLaunchKernelGGL(k_loss_az, dim3(nb), dim3(kThreads), 0, s, nx, np, cx, cy, mask, z, phi, q, az, u->part);
    hipLaunchKernelGGL(k_finish<3>, dim3(1), dim3(kThreads), 0, s, u->part, nb, u->scal, 4, unet_bat(np));
    // the output gradient goes into Q_0, which the backward first reads and only later reuses
    hipLaunchKernelGGL(k_loss_grad, dim3(nblk((size_t)u->H * u->W * 4)), dim3(kThreads), 0, s, nx, ny, (int)u->H,
                       (int)u->W, mask, az, q, u->scal, u->grad[1]);
    if (zero) {
        OECL_HIP(hipMemsetAsync(u->gW, 0, 4ull * NCONV * WSZ, s), "unet zero");
        OECL_HIP(hipMemsetAsync(u->gb, 0, 4ull * NCONV * C, s), "unet zero");
    }
    unet_backward(u, u->grad[1], nullptr, s);
    OECL_CHECK_LAUNCH("oecl_unet_train_sample");
    if (loss) {
        double h[3];
        OECL_HIP(hipMemcpyAsync(h, u->scal + 4, 24, hipMemcpyDeviceToHost, s), "loss");
        OECL_HIP(hipStreamSynchronize(s), "loss");
        *loss = (h[0] != 0. && h[1] != 0.) ? 1. - h[2] * h[2] / (h[0] * h[1]) : 1.;
    }
    return OECL_OK;
}

OECL_EXPORT uint64_t oecl_fcg_work_bytes(uint32_t np, uint32_t mDirs)
{
    const uint64_t m1 = (mDirs < 1 ? 1 : mDirs > FCG_MAXM ? FCG_MAXM : mDirs) + 1;
    return 8ull * ((2 + 2 * m1) * (uint64_t)np + FCG_MAXM * (nblk(np) + 1) + 2 * FCG_MAXM + 64) + 256;
}

namespace {

// One block of `check` flexible-CG iterations for R problems on one grid (R = 1: a single
// solve).  Vectors of run b at b np; the direction rings' slots R np apart; the scalar block of
// run b at S + b kScalImg (D = S + 16: z.Ap_j, pAp = S + 24: p.Ap per slot).
int fcg_iterations(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy, const uint8_t *mask,
                   uint32_t R, double *phi, double omega, uint32_t mDirs, double *r, double *z, double *P, double *AP,
                   double *part, size_t pst, double *S, uint32_t &it, uint32_t check, uint32_t maxIter,
                   hipStream_t st)
{
    const uint32_t np = nx * ny, nbr = std::min(nblk(np), kRedBlocks), m1 = mDirs + 1;
    const Bat bt{np, (size_t)R * np, pst, kScalImg, S + 12};  // stopped runs skip every kernel
    double *D = S + 16, *pAp = S + 24;
    const dim3 gr(nbr, R), g1(1, R);
    for (uint32_t k = 0; k < check && it < maxIter; ++k, ++it) {
        int rc = precondition_smoothed(u, nx, ny, cx, cy, mask, r, z, omega, st, R, S + 12);
        if (rc) return rc;
        const uint32_t nUsed = it < mDirs ? it : mDirs;
        if (nUsed) {
            hipLaunchKernelGGL(k_fcg_zdots, gr, dim3(kThreads), 0, st, np, z, AP, it, m1, nUsed, part, bt);
            hipLaunchKernelGGL(k_finish<FCG_MAXM>, g1, dim3(kThreads), 0, st, part, nbr, D, 0, bt);
        }
        hipLaunchKernelGGL(k_fcg_newdir, dim3(nblk(np, 256), R), dim3(256), 0, st, np, z, P, D, pAp, it, m1, nUsed,
                           bt);
        double *p = P + (size_t)(it % m1) * bt.w, *ap = AP + (size_t)(it % m1) * bt.w;
        hipLaunchKernelGGL(k_fcg_ap, gr, dim3(kThreads), 0, st, nx, np, cx, cy, mask, p, r, ap, part, bt);
        hipLaunchKernelGGL(k_finish<2>, g1, dim3(kThreads), 0, st, part, nbr, S, 6, bt);
        hipLaunchKernelGGL(k_alpha, dim3((R + 63) / 64), dim3(64), 0, st, S, pAp, it % m1, R, bt);
        hipLaunchKernelGGL(k_fcg_update, gr, dim3(kThreads), 0, st, np, S, phi, r, p, ap, part, bt);
        hipLaunchKernelGGL(k_finish<2>, g1, dim3(kThreads), 0, st, part, nbr, S, 8, bt);
    }
    return OECL_OK;
}

// r = q - A phi (phi zeroed off the inside nodes), s[8] = r.r, s[9] = q.q, stop flags cleared.
void fcg_begin(uint32_t nx, uint32_t ny, const double *cx, const double *cy, const uint8_t *mask, uint32_t R,
               const double *q, double *phi, double *r, double *part, size_t pst, double *S, hipStream_t st)
{
    const uint32_t np = nx * ny, nb = nblk(np), nbr = std::min(nb, kRedBlocks);
    const Bat bt{np, (size_t)R * np, pst, kScalImg};
    hipLaunchKernelGGL(k_mask_zero, dim3(nb, R), dim3(kThreads), 0, st, np, mask, phi, bt);
    hipLaunchKernelGGL(k_fcg_init, dim3(nbr, R), dim3(kThreads), 0, st, nx, np, cx, cy, mask, q, phi, r, part, bt);
    hipLaunchKernelGGL(k_finish<2>, dim3(1, R), dim3(kThreads), 0, st, part, nbr, S, 8, bt);
    hipLaunchKernelGGL(k_fcg_start, dim3((R + 63) / 64), dim3(64), 0, st, S, R, bt);
}

size_t fcg_part_stride(uint32_t np) { return FCG_MAXM * ((size_t)std::min(nblk(np), kRedBlocks) + 1); }

}  // namespace

// Flexible CG on the FIT system with the network as preconditioner, keeping mDirs directions
// (1 .. 8).  phi: initial guess in, solution out (zero off the inside nodes).  Converged when
// ||q - A phi|| <= tol ||q||; the residual norm is read on the host every `check` iterations.
OECL_EXPORT int oecl_fcg_solve(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                               const uint8_t *mask, const double *q, double *phi, double tol, uint32_t maxIter,
                               uint32_t check, double omega, uint32_t mDirs, void *work, uint32_t *iters,
                               double *relres, void *stream)
{
    if (!u || !cx || !cy || !mask || !q || !phi || !work || ny > u->H || nx > u->W || nx < 3 || ny < 3 ||
        mDirs < 1 || mDirs > FCG_MAXM)
        return OECL_ERR_ARG;
    if (check == 0) check = 1;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t np = nx * ny, m1 = mDirs + 1;
    const size_t pst = fcg_part_stride(np);
    double *r = (double *)work, *z = r + np, *P = z + np, *AP = P + (size_t)m1 * np;
    double *part = AP + (size_t)m1 * np, *S = part + pst;
    fcg_begin(nx, ny, cx, cy, mask, 1, q, phi, r, part, pst, S, st);
    double h[2];
    OECL_HIP(hipMemcpyAsync(h, S + 8, 16, hipMemcpyDeviceToHost, st), "fcg copy");
    OECL_HIP(hipStreamSynchronize(st), "fcg sync");
    const double qn = sqrt(h[1]) > 0. ? sqrt(h[1]) : 1.;
    double rr = h[0];
    uint32_t it = 0;
    while (sqrt(rr) > tol * qn && it < maxIter) {
        int rc = fcg_iterations(u, nx, ny, cx, cy, mask, 1, phi, omega, mDirs, r, z, P, AP, part, pst, S, it, check,
                                maxIter, st);
        if (rc) return rc;
        OECL_CHECK_LAUNCH("oecl_fcg_solve");
        OECL_HIP(hipMemcpyAsync(h, S + 8, 8, hipMemcpyDeviceToHost, st), "fcg copy");
        OECL_HIP(hipStreamSynchronize(st), "fcg sync");
        rr = h[0];
        if (!(rr == rr)) break;
    }
    if (iters) *iters = it;
    const double rel = sqrt(rr) / qn;
    if (relres) *relres = rel;
    return rel <= tol ? OECL_OK : OECL_ERR_NOCONV;
}

OECL_EXPORT uint64_t oecl_fcg_batch_work_bytes(uint32_t np, uint32_t mDirs, uint32_t R)
{
    const uint64_t m1 = (mDirs < 1 ? 1 : mDirs > FCG_MAXM ? FCG_MAXM : mDirs) + 1;
    return 8ull * R * ((2 + 2 * m1) * (uint64_t)np + fcg_part_stride(np) + kScalImg) + 256;
}

// R flexible CGs on one grid in the same launches (oecl_unet_reserve(u, R) first): q and phi of
// run b at b np; iters[b], relres[b] as oecl_fcg_solve gives them.  Each run stops at the first
// check point where it has converged (its phi and r are then left as they are), so every run's
// result is bit for bit that of oecl_fcg_solve on its own problem.
OECL_EXPORT int oecl_fcg_solve_batch(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                                     const uint8_t *mask, uint32_t R, const double *q, double *phi, double tol,
                                     uint32_t maxIter, uint32_t check, double omega, uint32_t mDirs, void *work,
                                     uint32_t *iters, double *relres, void *stream)
{
    if (!u || !cx || !cy || !mask || !q || !phi || !work || ny > u->H || nx > u->W || nx < 3 || ny < 3 ||
        mDirs < 1 || mDirs > FCG_MAXM || R < 1 || R > u->batch)
        return OECL_ERR_ARG;
    if (check == 0) check = 1;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t np = nx * ny, m1 = mDirs + 1;
    const size_t pst = fcg_part_stride(np), Rn = (size_t)R * np;
    double *r = (double *)work, *z = r + Rn, *P = z + Rn, *AP = P + m1 * Rn;
    double *part = AP + m1 * Rn, *S = part + pst * R;
    const Bat bt{np, Rn, pst, kScalImg};
    fcg_begin(nx, ny, cx, cy, mask, R, q, phi, r, part, pst, S, st);
    std::vector<double> h((size_t)R * kScalImg);
    uint32_t it = 0;
    for (;;) {
        hipLaunchKernelGGL(k_fcg_check, dim3((R + 63) / 64), dim3(64), 0, st, S, R, bt, tol, it);
        OECL_CHECK_LAUNCH("oecl_fcg_solve_batch");
        OECL_HIP(hipMemcpyAsync(h.data(), S, 8 * h.size(), hipMemcpyDeviceToHost, st), "fcg copy");
        OECL_HIP(hipStreamSynchronize(st), "fcg sync");
        bool live = false;
        for (uint32_t b = 0; b < R; ++b) live |= h[(size_t)b * kScalImg + 12] == 0.;
        if (!live || it >= maxIter) break;
        int rc = fcg_iterations(u, nx, ny, cx, cy, mask, R, phi, omega, mDirs, r, z, P, AP, part, pst, S, it, check,
                                maxIter, st);
        if (rc) return rc;
    }
    bool ok = true;
    for (uint32_t b = 0; b < R; ++b) {
        const double *sb = h.data() + (size_t)b * kScalImg;
        const double qn = sqrt(sb[14]) > 0. ? sqrt(sb[14]) : 1., rel = sqrt(sb[8]) / qn;
        if (iters) iters[b] = sb[12] != 0. ? (uint32_t)sb[13] : it;
        if (relres) relres[b] = rel;
        ok &= rel <= tol;
    }
    return ok ? OECL_OK : OECL_ERR_NOCONV;
}
