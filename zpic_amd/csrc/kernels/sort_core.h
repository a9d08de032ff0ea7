/*
 * sort_core.h - device code of the tile counting sort, shared by sort.hip
 * and the batched kernels of batch.hip.  See sort.hip.
 */
#ifndef ZK_SORT_CORE_H
#define ZK_SORT_CORE_H

#include "common.h"

namespace zk {


__device__ inline int tile_key(int ix, int iy, const zk_tiles &t)
{
    if (ix == ZK_DEAD) return -1;
    return ix / t.TX + (iy / t.TY) * t.ntx;
}

__device__ inline void count_body(const int *__restrict__ ix, const int *__restrict__ iy, const zk_spec_ctr *ctr,
                                  int *cnt, const zk_tiles &t, const int block, const int nblocks)
{
    const int n = ctr->np_total;
    const int stride = nblocks * blockDim.x;
    for (int base = block * blockDim.x; base < n; base += stride) {
        const int i = base + threadIdx.x;
        const int key = (i < n) ? tile_key(ix[i], iy[i], t) : -1;
        zk_run_reserve(key, cnt);
    }
}

/* Exclusive scan of the tile counts into off[0..ntiles], cursor = off;
 * counts are cleared for the next sort.  One workgroup of 1024 threads. */
__device__ inline void scan_body(int *cnt, int *off, int *cursor, zk_spec_ctr *ctr, int ntiles)
{
    __shared__ int part[1024];
    const int per = (ntiles + blockDim.x - 1) / blockDim.x;
    const int b = threadIdx.x * per, e = min(ntiles, b + per);
    int s = 0;
    for (int i = b; i < e; i++) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
        int v = (threadIdx.x >= (unsigned)o) ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - s;
    for (int i = b; i < e; i++) {
        const int c = cnt[i];
        off[i] = run;
        cursor[i] = run;
        cnt[i] = 0;
        run += c;
    }
    if (threadIdx.x == blockDim.x - 1) {
        off[ntiles] = part[threadIdx.x];
        ctr->np_total = part[threadIdx.x];
        ctr->n_sorted = part[threadIdx.x];
        ctr->n_dead = 0;
    }
}

/* Slot of the rank-th particle of a tile: consecutive ranks (which come
 * from one cell, as the input is cell-ordered within a tile) are spread by
 * a prime stride over the tile's range, so the 64 lanes of a wave push
 * particles of different cells and their LDS atomics do not collide. */
__device__ inline int interleave(int rank, int count)
{
    const int primes[6] = {97, 101, 103, 107, 109, 113};
    int p = primes[0];
    for (int k = 0; k < 6; k++)
        if (count % primes[k]) {
            p = primes[k];
            break;
        }
    return (int)(((long long)rank * p) % count);
}

__device__ inline void scatter_body(const zk_parts &src, const zk_parts &dst, const int *n_old, int *cursor,
                                    const int *off, const zk_tiles &t, int spread, const int block, const int nblocks)
{
    const int n = *n_old;
    const int stride = nblocks * blockDim.x;
    for (int base = block * blockDim.x; base < n; base += stride) {
        const int i = base + threadIdx.x;
        int ix = 0, iy = 0, key = -1;
        if (i < n) {
            ix = src.ix[i];
            iy = src.iy[i];
            key = tile_key(ix, iy, t);
        }
        int d = zk_run_reserve(key, cursor);
        if (key >= 0) {
            if (spread) {
                const int base = off[key];
                d = base + interleave(d - base, off[key + 1] - base);
            }
            dst.ix[d] = ix;
            dst.iy[d] = iy;
            dst.x[d] = src.x[i];
            dst.y[d] = src.y[i];
            dst.ux[d] = src.ux[i];
            dst.uy[d] = src.uy[i];
            dst.uz[d] = src.uz[i];
            if (src.id) dst.id[d] = src.id[i];
        }
    }
}

/* ---- sort of a tile-sorted input: particles that stay in their tile keep
 * their order at the front of the tile's new range, the others are
 * appended after the stayers of their new tile ---- */

/* Exclusive prefix of `flag` over the block in thread order; returns the
 * rank and sets *total.  lds needs blockDim.x / 64 ints. */
__device__ inline int block_rank(bool flag, int *lds, int *total)
{
    const int lane = __lane_id(), w = threadIdx.x / ZK_WAVE, nw = blockDim.x / ZK_WAVE;
    const unsigned long long m = __ballot(flag);
    const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) lds[w] = __popcll(m);
    __syncthreads();
    int before = 0, all = 0;
    for (int k = 0; k < nw; k++) {
        const int c = lds[k];
        before += (k < w) ? c : 0;
        all += c;
    }
    __syncthreads();
    *total = all;
    return before + in_wave;
}

/* Block b < ntiles counts the stayers of source tile b (stay[b]) and
 * reserves the movers in move[]; blocks beyond take chunks of the tail.
 * old_off[ntiles] is the end of the tile ranges, old_off[ntiles + 1] the
 * slot count in use. */
__device__ inline void count_stay_body(const int *__restrict__ ix, const int *__restrict__ iy,
                                       const int *__restrict__ old_off, int *stay, int *move, const zk_tiles &t,
                                       const int block, const int nblocks)
{
    __shared__ int red[16];
    if (block < t.ntiles) {
        const int lo = old_off[block], hi = old_off[block + 1];
        int nstay = 0;
        for (int base = lo; base < hi; base += blockDim.x) {
            const int i = base + threadIdx.x;
            const int key = (i < hi) ? tile_key(ix[i], iy[i], t) : -1;
            const bool st = key == block;
            nstay += st;
            zk_run_reserve(st ? -1 : key, move);
        }
        for (int o = ZK_WAVE / 2; o > 0; o >>= 1) nstay += __shfl_down(nstay, o, ZK_WAVE);
        if ((threadIdx.x & (ZK_WAVE - 1)) == 0) red[threadIdx.x / ZK_WAVE] = nstay;
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = 0;
            for (int k = 0; k < (int)(blockDim.x / ZK_WAVE); k++) s += red[k];
            stay[block] = s;
        }
    } else {
        const int n0 = old_off[t.ntiles], n1 = old_off[t.ntiles + 1];
        const int stride = (nblocks - t.ntiles) * blockDim.x;
        for (int base = n0 + (block - t.ntiles) * blockDim.x; base < n1; base += stride) {
            const int i = base + threadIdx.x;
            const int key = (i < n1) ? tile_key(ix[i], iy[i], t) : -1;
            zk_run_reserve(key, move);
        }
    }
}

/* new_off = exclusive scan of stay + move; cursor = new_off + stay (where
 * the movers go); move is cleared.  One workgroup of 1024 threads. */
__device__ inline void scan_stay_body(const int *stay, int *move, int *new_off, int *cursor, zk_spec_ctr *ctr,
                                      int ntiles)
{
    __shared__ int part[1024];
    const int per = (ntiles + blockDim.x - 1) / blockDim.x;
    const int b = threadIdx.x * per, e = min(ntiles, b + per);
    int s = 0;
    for (int i = b; i < e; i++) s += stay[i] + move[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
        int v = (threadIdx.x >= (unsigned)o) ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - s;
    for (int i = b; i < e; i++) {
        const int c = stay[i] + move[i];
        new_off[i] = run;
        cursor[i] = run + stay[i];
        move[i] = 0;
        run += c;
    }
    if (threadIdx.x == blockDim.x - 1) {
        new_off[ntiles] = part[threadIdx.x];
        ctr->np_total = part[threadIdx.x];
        ctr->n_sorted = part[threadIdx.x];
        ctr->n_dead = 0;
    }
}

__device__ inline void copy_part(const zk_parts &src, const zk_parts &dst, int i, int d, int ix, int iy)
{
    dst.ix[d] = ix;
    dst.iy[d] = iy;
    dst.x[d] = src.x[i];
    dst.y[d] = src.y[i];
    dst.ux[d] = src.ux[i];
    dst.uy[d] = src.uy[i];
    dst.uz[d] = src.uz[i];
    if (src.id) dst.id[d] = src.id[i];
}

/* Block b < ntiles moves source tile b: stayers to new_off[b] + their
 * rank among the tile's stayers (contiguous, in order), movers to the
 * cursor of their new tile; blocks beyond move chunks of the tail. */
__device__ inline void scatter_stay_body(const zk_parts &src, const zk_parts &dst, const int *__restrict__ old_off,
                                         const int *__restrict__ new_off, int *cursor, const zk_tiles &t,
                                         const int block, const int nblocks)
{
    __shared__ int red[16];
    if (block < t.ntiles) {
        const int lo = old_off[block], hi = old_off[block + 1];
        const int dst0 = new_off[block];
        int placed = 0;
        for (int base = lo; base < hi; base += blockDim.x) {
            const int i = base + threadIdx.x;
            int ix = 0, iy = 0, key = -1;
            if (i < hi) {
                ix = src.ix[i];
                iy = src.iy[i];
                key = tile_key(ix, iy, t);
            }
            const bool st = key == block;
            int total;
            const int r = block_rank(st, red, &total);
            const int dm = zk_run_reserve(st ? -1 : key, cursor);
            if (st) copy_part(src, dst, i, dst0 + placed + r, ix, iy);
            else if (key >= 0) copy_part(src, dst, i, dm, ix, iy);
            placed += total;
        }
    } else {
        const int n0 = old_off[t.ntiles], n1 = old_off[t.ntiles + 1];
        const int stride = (nblocks - t.ntiles) * blockDim.x;
        for (int base = n0 + (block - t.ntiles) * blockDim.x; base < n1; base += stride) {
            const int i = base + threadIdx.x;
            int ix = 0, iy = 0, key = -1;
            if (i < n1) {
                ix = src.ix[i];
                iy = src.iy[i];
                key = tile_key(ix, iy, t);
            }
            const int dm = zk_run_reserve(key, cursor);
            if (key >= 0) copy_part(src, dst, i, dm, ix, iy);
        }
    }
}

} // namespace zk

#endif
