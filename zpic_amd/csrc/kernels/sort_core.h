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

} // namespace zk

#endif
