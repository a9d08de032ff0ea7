/*
 * common.h - device helpers shared by the em2d HIP kernels (gfx950, wave64).
 */
#ifndef ZK_COMMON_H
#define ZK_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zk.h"

#define ZK_WAVE 64

/* Element at byte offset o of an array.  The particle arrays of the push
 * and of the re-sort are addressed as a uniform base plus one shared 32-bit
 * offset (the global saddr form): one offset register instead of a 64-bit
 * address per array, which leaves registers for more loads in flight.
 * Offsets fit 32 bits because a species holds at most ZK_MAX_PARTS slots. */
template <class T>
__device__ inline T &zk_at(T *base, unsigned o)
{
    return *(T *)((char *)base + o);
}

/* Blocks are dealt to the 8 XCDs round-robin.  Map block b to a work item
 * so that each XCD receives one contiguous run of items: neighbouring tiles
 * then share the halo lines in that XCD's L2. */
__device__ inline int zk_xcd_item(int b, int n)
{
    const int nx = 8;
    int x = b % nx, k = b / nx;
    int base = n / nx, rem = n % nx;
    return x * base + (x < rem ? x : rem) + k;
}

/* Float to 64-bit fixed point: v * scale rounded to the nearest integer,
 * for a power-of-two scale and |v * scale| < 2^51.  The product is exact in
 * FP64; adding 1.5 * 2^52 in the same fused operation rounds it to an
 * integer held in the low mantissa bits, and subtracting the bit pattern
 * of 1.5 * 2^52 (whose low word is zero) leaves its two's complement. */
__device__ inline unsigned long long zk_fx(float v, float scale)
{
    const double r = __builtin_fma((double)v, (double)scale, 6755399441055744.0);
    return (unsigned long long)__double_as_longlong(r) - 0x4338000000000000ull;
}

/* The same conversion without the subtraction, for sums that are read
 * back with zk_fx_read: the raw bit pattern of 1.5 * 2^52 + k is
 * 2151 * 2^51 + k, so a sum of N such patterns is N * 2151 * 2^51 + S
 * (mod 2^64), whose low 51 bits are those of the sum S of the integers k.
 * One instruction fewer per contribution; the read-back is exact while
 * |S| < 2^50. */
__device__ inline unsigned long long zk_fx_raw(float v, float scale)
{
    return (unsigned long long)__double_as_longlong(__builtin_fma((double)v, (double)scale, 6755399441055744.0));
}

/* S from a sum of zk_fx_raw patterns: its low 51 bits, sign-extended. */
__device__ inline long long zk_fx_read(unsigned long long r) { return (long long)(r << 13) >> 13; }

/* Packed cell index of the particle arrays (zk_parts): ix | iy << cbits,
 * -1 for a removed particle, whose ix reads as ZK_DEAD. */
__device__ inline int zk_cell_ix(int c, int cbits) { return c < 0 ? ZK_DEAD : (c & ((1 << cbits) - 1)); }
__device__ inline int zk_cell_iy(int c, int cbits) { return c >> cbits; }
__device__ inline int zk_cell(int ix, int iy, int cbits) { return ix == ZK_DEAD ? -1 : (ix | (iy << cbits)); }

/* Floor modulo for periodic wrap. */
__device__ inline int zk_wrap(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

/*
 * Reserve one slot per active lane in counters[key], aggregating runs of
 * equal keys held by consecutive lanes into a single atomic.  Lanes with
 * key < 0 reserve nothing.  Must be called by all 64 lanes of the wave.
 * Returns the reserved slot (counter value before the run + rank in run).
 */
__device__ inline int zk_run_reserve(int key, int *counters)
{
    const int lane = __lane_id();
    int prev = __shfl_up(key, 1, ZK_WAVE);
    bool head = (lane == 0) || (prev != key);
    unsigned long long heads = __ballot(head);
    unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    int myhead = 63 - __clzll(heads & upto);
    unsigned long long after = heads & ~upto;
    int next = after ? (__ffsll((long long)after) - 1) : ZK_WAVE;
    int base = 0;
    if (head && key >= 0) base = atomicAdd(&counters[key], next - lane);
    base = __shfl(base, myhead, ZK_WAVE);
    return base + lane - myhead;
}

/* Sum of a double over the block (blockDim.x a multiple of 64, <= 1024).
 * Result valid in thread 0. */
__device__ inline double zk_block_sum(double v, double *lds32)
{
    for (int o = ZK_WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, ZK_WAVE);
    const int w = threadIdx.x / ZK_WAVE, nw = blockDim.x / ZK_WAVE;
    if ((threadIdx.x & (ZK_WAVE - 1)) == 0) lds32[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < nw; k++) s += lds32[k];
    __syncthreads();
    return s;
}

#define ZK_LAUNCH_CHECK() return (int)hipGetLastError()

#endif
