// Shared device/host helpers of liboecl (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/oecl.h"

#define OECL_EXPORT extern "C" __attribute__((visibility("default")))

// Records the last HIP error text for oecl_last_error().
void oecl_set_error(const char *where, hipError_t e);
void oecl_set_error_msg(const char *msg);

#define OECL_CHECK_LAUNCH(where)                          \
    do {                                                  \
        hipError_t _e = hipGetLastError();                \
        if (_e != hipSuccess) {                           \
            oecl_set_error(where, _e);                    \
            return OECL_ERR_HIP;                          \
        }                                                 \
    } while (0)

#define OECL_HIP(call, where)                             \
    do {                                                  \
        hipError_t _e = (call);                           \
        if (_e != hipSuccess) {                           \
            oecl_set_error(where, _e);                    \
            return OECL_ERR_HIP;                          \
        }                                                 \
    } while (0)

static inline unsigned oecl_blocks(uint64_t n, unsigned per_block)
{
    uint64_t b = (n + per_block - 1) / per_block;
    if (b == 0) b = 1;
    return (unsigned)b;
}

// ----------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Bit-identical to the oracle's CPU version.
// ----------------------------------------------------------------------------------------
struct Philox4 {
    uint32_t v[4];
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                  uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c0;
        uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    Philox4 r;
    r.v[0] = c0;
    r.v[1] = c1;
    r.v[2] = c2;
    r.v[3] = c3;
    return r;
}

// Uniform in [0, 1) from the 53 high bits of (hi:lo).
__device__ __forceinline__ double u01_from_words(uint32_t lo, uint32_t hi)
{
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return (double)(x >> 11) * 0x1.0p-53;
}

// Uniform number k of the stream addressed by (item, purpose, step) under key (k0, k1).
// Block index k/2 of item `item` holds uniforms 2*(k/2) and 2*(k/2)+1.
__device__ __forceinline__ double philox_uniform(uint32_t item, uint32_t k, uint32_t step,
                                                 uint32_t purpose, uint32_t k0, uint32_t k1)
{
    Philox4 r = philox4x32_10(item, k >> 1, step, purpose, k0, k1);
    return (k & 1) ? u01_from_words(r.v[2], r.v[3]) : u01_from_words(r.v[0], r.v[1]);
}

// Particle component accessor for the (6, stride) SoA block: x, y, vx, vy, vz, w.
template <typename T>
struct Soa {
    T *base;
    uint64_t stride;
    __device__ __forceinline__ T &operator()(int c, uint64_t i) const
    {
        return base[(uint64_t)c * stride + i];
    }
};

// Workgroups are dispatched to the 8 XCDs round-robin (block b to XCD b % 8), each XCD with its
// own L2.  xcd_block maps hardware block b to a logical block so that every XCD works through
// whole chunks of XCD_CHUNK consecutive logical blocks (chunk c on XCD c % 8): neighbouring
// logical blocks then share their XCD's L2, while the chunks of every XCD still spread over the
// whole grid (any live prefix keeps all XCDs busy).  Blocks past the last whole group of 8
// chunks keep their index.  A bijection on [0, n).
constexpr uint32_t XCD_CHUNK = 32;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n)
{
    const uint32_t full = n / (8 * XCD_CHUNK) * (8 * XCD_CHUNK);
    if (b >= full) return b;
    const uint32_t x = b & 7u, k = b >> 3;
    return ((k / XCD_CHUNK) * 8 + x) * XCD_CHUNK + k % XCD_CHUNK;
}

// The same mapping with chunks of ch blocks (ch >= 1), for kernels whose neighbouring blocks
// share cache lines in small groups.
__device__ __forceinline__ uint32_t xcd_block_chunk(uint32_t b, uint32_t n, uint32_t ch)
{
    const uint32_t full = n / (8 * ch) * (8 * ch);
    if (b >= full) return b;
    const uint32_t x = b & 7u, k = b >> 3;
    return ((k / ch) * 8 + x) * ch + k % ch;
}

// Particles per push block.  Batched runs occupy segments that are multiples of it, and the
// per-block run table (blockRun) has one entry per kPushBlock particle slots: every kernel that
// reads that table divides slot indices by this constant.
constexpr uint32_t kPushBlock = 256;

// ---- raw buffer access: a 128-bit buffer resource in SGPRs plus a 32-bit byte offset per lane
// (one VALU offset serves several loads: the constant part goes to the instruction's immediate
// offset, a uniform part to the scalar offset).  Loads past `bytes` read 0 and stores past it are
// dropped by the hardware.  The resource's third dword is the gfx9 raw-buffer format word.
typedef __amdgpu_buffer_rsrc_t BufRes;

__device__ __forceinline__ BufRes buf_res(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

// CPOL: the instruction's cache-policy bits (bit 1: non-temporal, for data streamed once).
template <int CPOL = 0>
__device__ __forceinline__ float buf_ld_f32(BufRes r, uint32_t off, uint32_t soff = 0)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, CPOL));
}

// Two adjacent floats: two dword loads, which the backend merges into one dwordx2 load.  (The b64
// form of the load builtin is not used: with this compiler it loads the low dword only.)
__device__ __forceinline__ float2 buf_ld_f32x2(BufRes r, uint32_t off, uint32_t soff = 0)
{
    return make_float2(buf_ld_f32(r, off, soff), buf_ld_f32(r, off + 4u, soff));
}

template <int CPOL = 0>
__device__ __forceinline__ void buf_st_f32(float v, BufRes r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)off, 0, CPOL);
}

__device__ __forceinline__ void buf_st_u8(uint8_t v, BufRes r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b8(v, r, (int)off, 0, 0);
}
