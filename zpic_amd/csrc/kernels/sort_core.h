/*
 * sort_core.h - device code of the tile counting sort, shared by sort.hip
 * and the batched kernels of batch.hip.  See sort.hip.
 */
#ifndef ZK_SORT_CORE_H
#define ZK_SORT_CORE_H

#include "common.h"

namespace zk {


/* floor(a / d) for 0 <= a < 2^22, with inv = 1.0f / d (correctly rounded):
 * a + 0.5 lies at least 0.5 from a multiple of d, farther than the product
 * can be off (relative 2^-23, i.e. below 0.5 / d for a < 2^22) */
__device__ inline int zk_div(int a, float inv) { return (int)(((float)a + 0.5f) * inv); }

__device__ inline int tile_key(int ix, int iy, const zk_tiles &t)
{
    if (ix == ZK_DEAD) return -1;
    return zk_div(ix, 1.0f / t.TX) + zk_div(iy, 1.0f / t.TY) * t.ntx;
}

/* tile of a packed cell (zk_parts), -1 for a removed particle */
__device__ inline int tile_key_cell(int c, int cbits, const zk_tiles &t)
{
    return tile_key(zk_cell_ix(c, cbits), zk_cell_iy(c, cbits), t);
}

/* Slots of a tile range for `count` particles that were `before` at the
 * previous sort (-1: unknown): with slack 1, room for the arrivals of the
 * staged re-sorts that follow, the largest of half as many more, twice
 * the last growth (tiles of a forming filament fill up fast) and `floor`
 * (the mean count: sparse tiles next to dense ones), plus 32; slack 2
 * leaves one free slot (tests of full ranges); 0 none. */
__device__ inline int tile_cap(int count, int slack, int before = -1, int floor = 0)
{
    if (slack == 2) return count + 1;
    if (slack != 1) return count;
    const int grew = before >= 0 ? count - before : 0;
    return count + max(max(count >> 1, 2 * grew), floor) + 32;
}

/* Staging region of tile t in a set with ranges off[]: behind the ranges,
 * half the size of the tile's range. */
__device__ inline int stage_lo(const int *off, int ntiles, int t) { return off[ntiles] + (off[t] >> 1); }
__device__ inline int stage_room(const int *off, int t) { return (off[t + 1] >> 1) - (off[t] >> 1); }

/* Whether the ranges get slack: the set of cap_slots slots must hold the
 * ranges and the staging area behind them (half the ranges).  The counts
 * of tile i are a[i] (+ b[i] + c[i]); *floor receives tile_cap's floor.
 * Called by every thread of one workgroup; returns the same values in all
 * of them. */
__device__ inline int decide_slack(const int *a, const int *b, const int *c, int ntiles, int slack,
                                   long long cap_slots, const int *before, int *floor)
{
    __shared__ unsigned long long tot, cnt;
    *floor = 0;
    if (!slack) return 0;
    if (threadIdx.x == 0) tot = cnt = 0ull;
    __syncthreads();
    unsigned long long n = 0ull;
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) n += (unsigned long long)(a[i] + (b ? b[i] : 0) + (c ? c[i] : 0));
    atomicAdd(&cnt, n);
    __syncthreads();
    const int fl = slack == 1 && ntiles > 0 ? (int)(cnt / (unsigned long long)ntiles) : 0;
    unsigned long long s = 0ull;
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x)
        s += (unsigned long long)tile_cap(a[i] + (b ? b[i] : 0) + (c ? c[i] : 0), slack, before ? before[i] : -1, fl);
    atomicAdd(&tot, s);
    __syncthreads();
    const long long need = (long long)tot + ((long long)tot >> 1) + 64;
    __syncthreads();
    *floor = fl;
    return need <= cap_slots ? slack : 0;
}

__device__ inline void count_body(const int *__restrict__ cell, int cbits, const zk_spec_ctr *ctr, int *cnt,
                                  const zk_tiles &t, const int block, const int nblocks)
{
    const int n = ctr->np_total;
    const int stride = nblocks * blockDim.x;
    for (int base = block * blockDim.x; base < n; base += stride) {
        const int i = base + threadIdx.x;
        const int key = (i < n) ? tile_key_cell(cell[i], cbits, t) : -1;
        zk_run_reserve(key, cnt);
    }
}

/* Exclusive scan of the tile capacities (tile_cap of the counts) into
 * off[0..ntiles], cursor = off, tile_n = counts (when given); counts are
 * cleared for the next sort.  One workgroup of 1024 threads. */
__device__ inline void scan_body(int *cnt, int *off, int *cursor, zk_spec_ctr *ctr, int ntiles,
                                 int *tile_n = nullptr, int slack = 0, long long cap_slots = 0)
{
    __shared__ int part[1024];
    int floor;
    slack = decide_slack(cnt, nullptr, nullptr, ntiles, slack, cap_slots, nullptr, &floor);
    const int per = (ntiles + blockDim.x - 1) / blockDim.x;
    const int b = threadIdx.x * per, e = min(ntiles, b + per);
    int s = 0, live = 0;
    for (int i = b; i < e; i++) {
        s += tile_cap(cnt[i], slack, -1, floor);
        live += cnt[i];
    }
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
        if (tile_n) tile_n[i] = c;
        cnt[i] = 0;
        run += tile_cap(c, slack, -1, floor);
    }
    __syncthreads();
    /* the free slots: total capacity less the particles */
    part[threadIdx.x] = live;
    __syncthreads();
    for (int o = (int)blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == blockDim.x - 1) {
        off[ntiles] = run;
        ctr->np_total = run;
        ctr->n_sorted = run;
        ctr->n_dead = 0;
        ctr->n_slack = run - part[0];
        ctr->n_unplaced = 0;
        ctr->n_overfull = 0;
        ctr->ovf_sorted = ctr->n_overflow;
    }
}

/* Marks the free slots of the tile ranges removed (cell -1): block t
 * covers [off[t] + tile_n[t], off[t + 1]). */
__device__ inline void mark_free_body(int *cell, const int *off, const int *tile_n, int tile)
{
    const int a = off[tile] + tile_n[tile], b = off[tile + 1];
    for (int i = a + (int)threadIdx.x; i < b; i += blockDim.x) cell[i] = -1;
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
                                    const int *off, const zk_tiles &t, int spread, const int block, const int nblocks,
                                    const int *tile_n = nullptr)
{
    const int n = *n_old;
    const int stride = nblocks * blockDim.x;
    for (int base = block * blockDim.x; base < n; base += stride) {
        const int i = base + threadIdx.x;
        int c = -1, key = -1;
        if (i < n) {
            c = src.cell[i];
            key = tile_key_cell(c, src.cbits, t);
        }
        int d = zk_run_reserve(key, cursor);
        if (key >= 0) {
            if (spread) {
                const int base = off[key];
                d = base + interleave(d - base, tile_n ? tile_n[key] : off[key + 1] - base);
            }
            dst.cell[d] = c;
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
 * appended after the stayers of their new tile ----
 *
 * Particles move at most a few cells between sorts, so nearly every mover
 * lands in one of the 8 tiles around its source tile.  The count pass
 * counts, per source tile, its stayers and its movers per neighbour slot
 * (LDS counters, one ballot per distinct slot in a wave) and reserves each
 * slot's run in the destination's mover region with one global atomic
 * (nbr[ZK_NBR b + slot] = start of the run).  Each of the block's waves
 * owns a contiguous piece of the tile and records its stayer count
 * (nbr[ZK_NBR b + 9 + wave]).  The scatter pass writes stayers in order,
 * each wave from its offset, and movers at their run start plus an LDS
 * rank.  The
 * rare particle that moved farther, and the unsorted tail, reserve slots
 * after the movers with wave-aggregated global atomics (tailn, cursor). */

/* per tile in nbr: 9 run starts and the stayer counts of the 4 waves of a
 * 256-thread block */
#define ZK_NBR 13
#define ZK_SORT_THREADS 256
/* The count and scatter passes start each wave's chunks at the multiple of
 * 32 at or below its piece, with the slots before the piece masked, so
 * that the loads of a chunk cover whole 128-byte lines (scatter 1.35 ->
 * 1.32 ms per 2048^2 species). */

/* 64-particle chunks a wave loads before processing them (loads in flight;
 * 2048^2 Weibel re-sort, unaligned chunks: 1.45 ms with 4, 1.35 ms with
 * 8, 1.31 ms with 12 at occupancy 3, 1.36 ms with 16; aligned chunks:
 * 1.314 ms with 8, 1.278 ms with 12) */
#define ZK_SORT_UNROLL 12

/* piece [*lo, *hi) of the tile range [a, b) owned by wave w of nw */
__device__ inline void wave_piece(int a, int b, int w, int nw, int *lo, int *hi)
{
    const int piece = (b - a + nw - 1) / nw;
    *lo = min(b, a + w * piece);
    *hi = min(b, *lo + piece);
}

/* neighbour slot 0..8 of tile `key` seen from tile `b` (periodic), 4 for
 * b itself, -1 when not adjacent */
__device__ inline int nbr_slot(int key, int b, const zk_tiles &t)
{
    const float inv = 1.0f / t.ntx;
    const int ky = zk_div(key, inv), by = zk_div(b, inv);
    int dx = (key - ky * t.ntx) - (b - by * t.ntx), dy = ky - by;
    if (dx > 1) dx -= t.ntx;
    else if (dx < -1) dx += t.ntx;
    if (dy > 1) dy -= t.nty;
    else if (dy < -1) dy += t.nty;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1) return -1;
    return (dx + 1) + 3 * (dy + 1);
}

/* tile in neighbour slot j of tile b */
__device__ inline int nbr_tile(int j, int b, const zk_tiles &t)
{
    const int tx = (b % t.ntx + j % 3 - 1 + t.ntx) % t.ntx;
    const int ty = (b / t.ntx + j / 3 - 1 + t.nty) % t.nty;
    return tx + ty * t.ntx;
}

/* Adds each lane with slot >= 0 to lcnt[slot] (LDS), one atomic per
 * distinct slot in the wave; returns the lane's rank: the counter before
 * the wave's addition plus the lanes of the same slot below it.  Must be
 * called by all 64 lanes of the wave. */
__device__ inline int wave_slot_add(int slot, int *lcnt)
{
    const int lane = __lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(slot >= 0);
    int mine = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int sl = __shfl(slot, leader, ZK_WAVE);
        const unsigned long long m = __ballot(slot == sl);
        int base = 0;
        if (lane == leader) base = atomicAdd(&lcnt[sl], __popcll(m));
        base = __shfl(base, leader, ZK_WAVE);
        if (slot == sl) mine = base + __popcll(m & lt);
        todo &= ~m;
    }
    return mine;
}

/* end of the particles of range t: off[t] + n[t], or off[t + 1] without counts */
__device__ inline int range_end(const int *off, const int *n, int t) { return n ? off[t] + n[t] : off[t + 1]; }

/* Block b < ntiles (ZK_SORT_THREADS threads) counts the stayers of source
 * tile b (stay[b], and per wave in nbr) and its movers per neighbour slot,
 * reserving their runs (nbr, move); farther movers count in tailn.  Blocks
 * beyond count chunks of the tail into tailn.  old_off[ntiles] is the end
 * of the tile ranges, old_off[ntiles + 1] the slot count in use. */
__device__ inline void count_stay_body(const int *__restrict__ cell, int cbits, const int *__restrict__ old_off,
                                       const int *__restrict__ old_n, int *stay, int *move, int *nbr, int *tailn,
                                       const zk_tiles &t, const int block, const int nblocks)
{
    __shared__ int lc[9], wst[ZK_SORT_THREADS / ZK_WAVE];
    const int lane = __lane_id();
    if (block < t.ntiles) {
        if (threadIdx.x < 9) lc[threadIdx.x] = 0;
        __syncthreads();
        const int nw = ZK_SORT_THREADS / ZK_WAVE, w = threadIdx.x / ZK_WAVE;
        int wlo, whi;
        wave_piece(old_off[block], range_end(old_off, old_n, block), w, nw, &wlo, &whi);
        int nstay = 0;
        for (int base = wlo & ~31; base < whi; base += ZK_SORT_UNROLL * ZK_WAVE) {
            int cc[ZK_SORT_UNROLL];
#pragma unroll
            for (int u = 0; u < ZK_SORT_UNROLL; u++) {
                const int i = base + u * ZK_WAVE + lane;
                cc[u] = (i >= wlo && i < whi) ? cell[i] : -1;
            }
#pragma unroll
            for (int u = 0; u < ZK_SORT_UNROLL; u++) {
                const int key = tile_key_cell(cc[u], cbits, t);
                const bool st = key == block;
                nstay += __popcll(__ballot(st));
                const int slot = (key >= 0 && !st) ? nbr_slot(key, block, t) : -1;
                if (slot >= 0) atomicAdd(&lc[slot], 1); /* totals only: no ranks needed here */
                zk_run_reserve((key >= 0 && !st && slot < 0) ? key : -1, tailn);
            }
        }
        if (lane == 0) {
            wst[w] = nstay;
            nbr[ZK_NBR * block + 9 + w] = nstay;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = 0;
            for (int k = 0; k < nw; k++) s += wst[k];
            stay[block] = s;
        }
        if (threadIdx.x < 9) {
            const int c = lc[threadIdx.x];
            nbr[ZK_NBR * block + threadIdx.x] = c ? atomicAdd(&move[nbr_tile(threadIdx.x, block, t)], c) : 0;
        }
    } else {
        const int n0 = old_off[t.ntiles], n1 = old_off[t.ntiles + 1];
        const int stride = (nblocks - t.ntiles) * blockDim.x;
        for (int base = n0 + (block - t.ntiles) * blockDim.x; base < n1; base += stride) {
            const int i = base + threadIdx.x;
            const int key = (i < n1) ? tile_key_cell(cell[i], cbits, t) : -1;
            zk_run_reserve(key, tailn);
        }
    }
}

/* new_off = exclusive scan of stay + move + tailn; cursor = new_off +
 * stay + move (where the farther movers and the tail go); move and tailn
 * are cleared.  One workgroup of 1024 threads. */
__device__ inline void scan_stay_body(const int *stay, int *move, int *tailn, int *new_off, int *new_n,
                                      int *cursor, zk_spec_ctr *ctr, int ntiles, int slack, long long cap_slots,
                                      const int *old_n = nullptr)
{
    /* chunks of blockDim.x consecutive tiles, read and written coalesced:
     * a wave scan by shuffles, a scan of the wave totals, and a running
     * carry from chunk to chunk */
    __shared__ int wsum[1024 / ZK_WAVE];
    __shared__ int carry_s, live_s;
    int floor;
    slack = decide_slack(stay, move, tailn, ntiles, slack, cap_slots, old_n, &floor);
    const int lane = __lane_id(), w = threadIdx.x / ZK_WAVE, nw = blockDim.x / ZK_WAVE;
    if (threadIdx.x == 0) carry_s = live_s = 0;
    __syncthreads();
    int live = 0;
    for (int base = 0; base < ntiles; base += blockDim.x) {
        const int i = base + threadIdx.x;
        const int st = i < ntiles ? stay[i] : 0, mv = i < ntiles ? move[i] : 0, tl = i < ntiles ? tailn[i] : 0;
        const int cnt = st + mv + tl, c = i < ntiles ? tile_cap(cnt, slack, old_n ? old_n[i] : -1, floor) : 0;
        live += cnt;
        int x = c; /* inclusive scan within the wave */
        for (int o = 1; o < ZK_WAVE; o <<= 1) {
            const int y = __shfl_up(x, o, ZK_WAVE);
            if (lane >= o) x += y;
        }
        if (lane == ZK_WAVE - 1) wsum[w] = x;
        __syncthreads();
        if (w == 0) {
            int v = lane < nw ? wsum[lane] : 0;
            for (int o = 1; o < ZK_WAVE; o <<= 1) {
                const int y = __shfl_up(v, o, ZK_WAVE);
                if (lane >= o) v += y;
            }
            if (lane < nw) wsum[lane] = v;
        }
        __syncthreads();
        const int carry = carry_s;
        const int run = carry + (w > 0 ? wsum[w - 1] : 0) + x - c;
        if (i < ntiles) {
            new_off[i] = run;
            if (new_n) new_n[i] = cnt;
            cursor[i] = run + st + mv;
            move[i] = 0;
            tailn[i] = 0;
        }
        __syncthreads();
        if (threadIdx.x == 0) carry_s = carry + wsum[nw - 1];
        __syncthreads();
    }
    atomicAdd(&live_s, live);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int total = carry_s;
        new_off[ntiles] = total;
        ctr->np_total = total;
        ctr->n_sorted = total;
        ctr->n_dead = 0;
        ctr->n_slack = total - live_s;
        ctr->n_unplaced = 0;
        ctr->n_overfull = 0;
        ctr->ovf_sorted = ctr->n_overflow;
    }
}

__device__ inline void copy_part(const zk_parts &src, const zk_parts &dst, int i, int d, int c)
{
    dst.cell[d] = c;
    dst.x[d] = src.x[i];
    dst.y[d] = src.y[i];
    dst.ux[d] = src.ux[i];
    dst.uy[d] = src.uy[i];
    dst.uz[d] = src.uz[i];
    if (src.id) dst.id[d] = src.id[i];
}

/* Block b < ntiles (ZK_SORT_THREADS threads) moves source tile b:
 * stayers to new_off[b] + their rank among the tile's stayers (contiguous,
 * in order), movers to their slot's run; blocks beyond move chunks of the
 * tail.  Every lane loads its particle before the slot bookkeeping, so the
 * loads are in flight during it. */
__device__ inline void scatter_stay_body(const zk_parts &src, const zk_parts &dst, const int *__restrict__ old_off,
                                         const int *__restrict__ old_n, const int *__restrict__ new_off,
                                         const int *__restrict__ stay,
                                         const int *__restrict__ nbr, int *cursor, const zk_tiles &t,
                                         const int block, const int nblocks)
{
    __shared__ int lc[9], run0[9];
    const int lane = __lane_id();
    if (block < t.ntiles) {
        if (threadIdx.x < 9) {
            const int d = nbr_tile(threadIdx.x, block, t);
            lc[threadIdx.x] = 0;
            run0[threadIdx.x] = new_off[d] + stay[d] + nbr[ZK_NBR * block + threadIdx.x];
        }
        const int nw = ZK_SORT_THREADS / ZK_WAVE, w = threadIdx.x / ZK_WAVE;
        int wlo, whi;
        wave_piece(old_off[block], range_end(old_off, old_n, block), w, nw, &wlo, &whi);
        int placed = new_off[block];
        for (int k = 0; k < w; k++) placed += nbr[ZK_NBR * block + 9 + k];
        __syncthreads();
        const unsigned long long lt = (1ull << lane) - 1ull;
        for (int base = wlo & ~31; base < whi; base += ZK_SORT_UNROLL * ZK_WAVE) {
            int cc[ZK_SORT_UNROLL];
            float x[ZK_SORT_UNROLL], y[ZK_SORT_UNROLL], ux[ZK_SORT_UNROLL], uy[ZK_SORT_UNROLL], uz[ZK_SORT_UNROLL];
            int64_t id[ZK_SORT_UNROLL];
#pragma unroll
            for (int u = 0; u < ZK_SORT_UNROLL; u++) {
                const int i = base + u * ZK_WAVE + lane;
                const bool in = i >= wlo && i < whi;
                const unsigned o = (unsigned)i * 4u;
                cc[u] = in ? zk_at(src.cell, o) : -1;
                x[u] = in ? zk_at(src.x, o) : 0.f;
                y[u] = in ? zk_at(src.y, o) : 0.f;
                ux[u] = in ? zk_at(src.ux, o) : 0.f;
                uy[u] = in ? zk_at(src.uy, o) : 0.f;
                uz[u] = in ? zk_at(src.uz, o) : 0.f;
                id[u] = (in && src.id) ? src.id[i] : 0;
            }
#pragma unroll
            for (int u = 0; u < ZK_SORT_UNROLL; u++) {
                const int key = tile_key_cell(cc[u], src.cbits, t);
                const bool st = key == block;
                const unsigned long long ms = __ballot(st);
                int d = placed + __popcll(ms & lt);
                placed += __popcll(ms);
                const int slot = (key >= 0 && !st) ? nbr_slot(key, block, t) : -1;
                const int r = wave_slot_add(slot, lc);
                if (slot >= 0) d = run0[slot] + r;
                const bool far = key >= 0 && !st && slot < 0;
                const int dm = zk_run_reserve(far ? key : -1, cursor);
                if (far) d = dm;
                if (key >= 0) {
                    const unsigned od = (unsigned)d * 4u;
                    zk_at(dst.cell, od) = cc[u];
                    zk_at(dst.x, od) = x[u];
                    zk_at(dst.y, od) = y[u];
                    zk_at(dst.ux, od) = ux[u];
                    zk_at(dst.uy, od) = uy[u];
                    zk_at(dst.uz, od) = uz[u];
                    if (src.id) dst.id[d] = id[u];
                }
            }
        }
    } else {
        const int n0 = old_off[t.ntiles], n1 = old_off[t.ntiles + 1];
        const int stride = (nblocks - t.ntiles) * blockDim.x;
        for (int base = n0 + (block - t.ntiles) * blockDim.x; base < n1; base += stride) {
            const int i = base + threadIdx.x;
            int c = -1, key = -1;
            if (i < n1) {
                c = src.cell[i];
                key = tile_key_cell(c, src.cbits, t);
            }
            const int dm = zk_run_reserve(key, cursor);
            if (key >= 0) copy_part(src, dst, i, dm, c);
        }
    }
}

/* ---- staged re-sort ----
 *
 * A push with `sorting` set has written the particles that stayed in
 * their tile to the front of the tile's range of the other buffer set
 * (sstay[t] of them) and the others to the tile's staging region behind
 * the ranges (sleave[t], stage_lo).  Three launches move only the staged
 * particles: count (runs per neighbour slot, as in the re-sort above),
 * scan (arrival starts and new counts per tile) and scatter.  Arrivals for
 * a range without free slots stay where they are, and the staging area
 * becomes the unsorted tail (n_overfull); those and the leavers that found
 * no staging room (n_unplaced, which stayed with their old tile) make the
 * host follow with a full re-sort, so that after every sort each particle
 * sits in the range of its own tile (which tile deposits a particle's
 * current decides how the windows round, and runs must not depend on the
 * order of atomics). */

__device__ inline void staged_count_body(const int *__restrict__ cell, int cbits, const int *__restrict__ off,
                                         const int *__restrict__ sleave, int *arr, int *far, int *nbr,
                                         const zk_tiles &t, const int tile)
{
    __shared__ int lc[9];
    if (threadIdx.x < 9) lc[threadIdx.x] = 0;
    __syncthreads();
    const int b = stage_lo(off, t.ntiles, tile), m = sleave[tile];
    for (int base = 0; base < m; base += blockDim.x) {
        const int j = base + threadIdx.x;
        const int key = j < m ? tile_key_cell(cell[b + j], cbits, t) : -1;
        const int slot = key >= 0 ? nbr_slot(key, tile, t) : -1;
        if (slot >= 0) atomicAdd(&lc[slot], 1);
        zk_run_reserve((key >= 0 && slot < 0) ? key : -1, far);
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        const int c = lc[threadIdx.x];
        nbr[ZK_NBR * tile + threadIdx.x] = c ? atomicAdd(&arr[nbr_tile(threadIdx.x, tile, t)], c) : 0;
    }
}

/* Tile k keeps min(stayers + arrivals, capacity) particles: the stayers,
 * then the arrivals from its neighbours from arr0[k], then the farther
 * ones from farcur[k].  arr and far are cleared; sstay[k] receives the
 * tile's count before this sort (the growth a following full sort sizes
 * the free slots by).  One workgroup. */
__device__ inline void staged_scan_body(const int *off, int *tile_n, int *sstay, int *arr, int *far, int *arr0,
                                        int *farcur, zk_spec_ctr *ctr, int ntiles)
{
    __shared__ int free_s, over_s;
    if (threadIdx.x == 0) free_s = over_s = 0;
    __syncthreads();
    int fr = 0, ov = 0;
    for (int k = threadIdx.x; k < ntiles; k += blockDim.x) {
        const int o = off[k], cap = off[k + 1] - o;
        const int s = sstay[k], a = arr[k], f = far[k];
        const int placed = min(s + a + f, cap);
        sstay[k] = tile_n[k];
        tile_n[k] = placed;
        arr0[k] = o + s;
        farcur[k] = o + s + a;
        arr[k] = 0;
        far[k] = 0;
        fr += cap - placed;
        ov += s + a + f - placed;
    }
    atomicAdd(&free_s, fr);
    atomicAdd(&over_s, ov);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int end = off[ntiles];
        ctr->n_sorted = end;
        /* arrivals left in the staging area: it is the unsorted tail, its
         * other slots removed particles */
        ctr->np_total = over_s > 0 ? end + (end >> 1) : end;
        ctr->n_dead = over_s > 0 ? (end >> 1) - over_s : 0;
        ctr->n_slack = free_s;
        ctr->n_overfull = over_s;
        ctr->ovf_sorted = ctr->n_overflow;
    }
}

/* Block `tile` moves its staged particles to their new tiles and marks its
 * range's free slots removed.  When some range was full (n_overfull), the
 * staging area is the tail that the following full re-sort reads: the
 * particles moved out of it, and its unused slots, are marked removed. */
__device__ inline void staged_scatter_body(const zk_parts &dst, const int *__restrict__ off,
                                           const int *__restrict__ tile_n, const int *__restrict__ sleave,
                                           const int *__restrict__ arr0, int *farcur, const int *__restrict__ nbr,
                                           const zk_spec_ctr *ctr, const zk_tiles &t, const int tile)
{
    __shared__ int lc[9], run0[9], end0[9];
    if (threadIdx.x < 9) {
        const int d = nbr_tile(threadIdx.x, tile, t);
        lc[threadIdx.x] = 0;
        run0[threadIdx.x] = arr0[d] + nbr[ZK_NBR * tile + threadIdx.x];
        end0[threadIdx.x] = off[d + 1];
    }
    __syncthreads();
    const bool tail = ctr->n_overfull > 0;
    const int b = stage_lo(off, t.ntiles, tile), m = sleave[tile];
    for (int base = 0; base < m; base += blockDim.x) {
        const int j = base + threadIdx.x;
        const bool in = j < m;
        const int i = b + (in ? j : 0);
        const int c = in ? dst.cell[i] : -1;
        const float x = dst.x[i], y = dst.y[i], ux = dst.ux[i], uy = dst.uy[i], uz = dst.uz[i];
        const int64_t id = dst.id ? dst.id[i] : 0;
        const int key = tile_key_cell(c, dst.cbits, t);
        const int slot = key >= 0 ? nbr_slot(key, tile, t) : -1;
        const int r = wave_slot_add(slot, lc);
        int d = 0, end = 0;
        if (slot >= 0) {
            d = run0[slot] + r;
            end = end0[slot];
        }
        const bool farp = key >= 0 && slot < 0;
        const int dfar = zk_run_reserve(farp ? key : -1, farcur);
        if (farp) {
            d = dfar;
            end = off[key + 1];
        }
        if (key >= 0 && d < end) {
            dst.cell[d] = c;
            dst.x[d] = x;
            dst.y[d] = y;
            dst.ux[d] = ux;
            dst.uy[d] = uy;
            dst.uz[d] = uz;
            if (dst.id) dst.id[d] = id;
            if (tail) dst.cell[i] = -1;
        }
    }
    if (tail) {
        const int room = stage_room(off, tile);
        for (int j = m + (int)threadIdx.x; j < room; j += blockDim.x) dst.cell[b + j] = -1;
    }
    mark_free_body(dst.cell, off, tile_n, tile);
}

} // namespace zk

#endif
