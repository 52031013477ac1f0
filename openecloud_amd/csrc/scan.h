// Tile scan helpers: 256-thread blocks, 16 items per thread (4096-item tiles), wave64.
#pragma once
#include "common.h"

#define OECL_SCAN_THREADS 256
#define OECL_SCAN_ITEMS 16
#define OECL_SCAN_TILE (OECL_SCAN_THREADS * OECL_SCAN_ITEMS)

static inline uint32_t oecl_scan_tiles(uint32_t n) { return (n + OECL_SCAN_TILE - 1) / OECL_SCAN_TILE; }

// Exclusive scan over a 256-thread block; returns the thread's prefix, *total the block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *lds4, uint32_t *total)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds4[wave] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < OECL_SCAN_THREADS / 64; ++w) {
        uint32_t t = lds4[w];
        if (w < wave) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// Visits every element of the block's tile with its exclusive hole rank (holes before it):
// F(i, isHole, holesBefore).  Elements at or beyond n are skipped.
template <typename F>
__device__ __forceinline__ void for_each_ranked(const uint8_t *__restrict__ flags, uint32_t n,
                                                const uint32_t *__restrict__ tileOff, F f)
{
    __shared__ uint32_t lds[4];
    uint64_t base = (uint64_t)blockIdx.x * OECL_SCAN_TILE + threadIdx.x * OECL_SCAN_ITEMS;
    uint8_t fl[OECL_SCAN_ITEMS];
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < OECL_SCAN_ITEMS; ++k) {
        uint64_t i = base + k;
        fl[k] = i < n ? flags[i] : 1;
        c += fl[k] == 0;
    }
    uint32_t tot;
    uint32_t run = tileOff[blockIdx.x] + block_excl_scan(c, lds, &tot);
#pragma unroll
    for (int k = 0; k < OECL_SCAN_ITEMS; ++k) {
        uint64_t i = base + k;
        if (i < n) f((uint32_t)i, fl[k] == 0, run);
        run += fl[k] == 0;
    }
}

// Host launchers (defined in boundary.hip): tile counts of zero flags and their exclusive
// scan.  counts[nTiles], offsets[nTiles + 1]; offsets[nTiles] is the total.
int oecl_launch_hole_scan(const uint8_t *flags, uint32_t n, uint32_t *counts,
                          uint32_t *offsets, hipStream_t s);
// Exclusive scan of a u32 array whose length is n (host value) into offsets[n + 1];
// work holds 2*nTiles + 2 u32.
int oecl_launch_scan_u32(const uint32_t *in, uint32_t n, uint32_t *offsets, uint32_t *work,
                         hipStream_t s);
