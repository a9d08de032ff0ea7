/*
 * sort.hip - counting sort of the particle arrays by tile (gfx950).
 *
 * Plays the role of the reference spec_sort (em2d/particles.c:942-1007),
 * which orders particles by cell to keep memory access local.  Here the key
 * is the tile (TX x TY cells) that the push kernel assigns one workgroup
 * to, so after the sort every tile owns one contiguous range of slots.
 * Removed particles (ZK_DEAD) are dropped.  With `spread` (used when a
 * cell-ordered host buffer is uploaded) the particles of a tile are
 * interleaved across cells (see interleave()); later sorts keep runs
 * contiguous, which preserves that interleaving for particles that stay
 * and keeps the writes coalesced.  The order is not stable and is not part
 * of the em2d contract (the reference itself reorders on sort and on
 * removal).
 *
 * Three launches: count (per-tile histogram), scan (tile offsets, one
 * workgroup), scatter (write every live particle to its tile range).  Count
 * and scatter reserve slots with one atomic per run of equal keys in a
 * wave, so a tile-ordered input costs about one atomic per wave.
 * zk_sort_tiles is the variant for inputs that are already tile-sorted
 * (every regular re-sort): particles that stay in their tile are written
 * in order, contiguously, and only those that changed tile use atomics.
 */
#include "sort_core.h"

using namespace zk;

namespace {

__global__ void count_kernel(const int *__restrict__ cell, int cbits, const zk_spec_ctr *ctr, int *cnt, zk_tiles t)
{
    count_body(cell, cbits, ctr, cnt, t, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(1024) scan_kernel(int *cnt, int *off, int *cursor, zk_spec_ctr *ctr, int ntiles,
                                                    int *tile_n, int slack, int cap_slots)
{
    scan_body(cnt, off, cursor, ctr, ntiles, tile_n, slack, cap_slots);
}

__global__ void __launch_bounds__(256) mark_free_kernel(int *cell, const int *off, const int *tile_n)
{
    mark_free_body(cell, off, tile_n, blockIdx.x);
}

__global__ void snapshot_kernel(const zk_spec_ctr *ctr, int *n_old)
{
    *n_old = ctr->np_total;
}

__global__ void scatter_indirect(zk_parts src, zk_parts dst, const int *n_old, int *cursor, const int *off,
                                 zk_tiles t, int spread, const int *tile_n)
{
    scatter_body(src, dst, n_old, cursor, off, t, spread, blockIdx.x, gridDim.x, tile_n);
}

__global__ void __launch_bounds__(ZK_SORT_THREADS)
    count_stay_kernel(const int *__restrict__ cell, int cbits, const int *old_off, const int *old_n, int *stay,
                      int *move, int *nbr, int *tailn, zk_tiles t)
{
    count_stay_body(cell, cbits, old_off, old_n, stay, move, nbr, tailn, t, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(1024) scan_stay_kernel(const int *stay, int *move, int *tailn, int *new_off,
                                                         int *new_n, int *cursor, zk_spec_ctr *ctr, int ntiles,
                                                         int slack, int cap_slots, const int *old_n)
{
    scan_stay_body(stay, move, tailn, new_off, new_n, cursor, ctr, ntiles, slack, cap_slots, old_n);
}

__global__ void __launch_bounds__(ZK_SORT_THREADS)
    scatter_stay_kernel(zk_parts src, zk_parts dst, const int *old_off, const int *old_n, const int *new_off,
                        const int *stay, const int *nbr, int *cursor, zk_tiles t)
{
    scatter_stay_body(src, dst, old_off, old_n, new_off, stay, nbr, cursor, t, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(ZK_SORT_THREADS)
    staged_count_kernel(const int *__restrict__ cell, int cbits, const int *off, const int *sleave, int *arr,
                        int *far, int *nbr, zk_tiles t)
{
    staged_count_body(cell, cbits, off, sleave, arr, far, nbr, t, blockIdx.x);
}

__global__ void __launch_bounds__(1024) staged_scan_kernel(const int *off, int *tile_n, int *sstay, int *arr,
                                                           int *far, int *arr0, int *farcur, zk_spec_ctr *ctr,
                                                           int ntiles)
{
    staged_scan_body(off, tile_n, sstay, arr, far, arr0, farcur, ctr, ntiles);
}

__global__ void __launch_bounds__(ZK_SORT_THREADS)
    staged_scatter_kernel(zk_parts dst, const int *off, const int *tile_n, const int *sleave, const int *arr0,
                          int *farcur, const int *nbr, const zk_spec_ctr *ctr, zk_tiles t)
{
    staged_scatter_body(dst, off, tile_n, sleave, arr0, farcur, nbr, ctr, t, blockIdx.x);
}

/* t_part AoS (28 bytes: ix, iy, x, y, ux, uy, uz) <-> SoA with the cell
 * index pair packed */
__global__ void unpack_kernel(const int *__restrict__ aos, int n, zk_parts dst, int64_t id0, zk_spec_ctr *ctr,
                              int append)
{
    const int base = append ? ctr->np_total : 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int *r = aos + 7 * (size_t)i;
        const int d = base + i;
        dst.cell[d] = zk_cell(r[0], r[1], dst.cbits);
        dst.x[d] = __int_as_float(r[2]);
        dst.y[d] = __int_as_float(r[3]);
        dst.ux[d] = __int_as_float(r[4]);
        dst.uy[d] = __int_as_float(r[5]);
        dst.uz[d] = __int_as_float(r[6]);
        if (dst.id) dst.id[d] = id0 + i;
    }
}

/* An injected column appended in one block: the block reads the tail's
 * start, converts the column (read straight from page-locked host memory:
 * no copy on the stream) and commits the new count itself, so one launch
 * replaces a copy, the unpack and the commit. */
__global__ void __launch_bounds__(1024) unpack_append_kernel(const int *__restrict__ aos, int n, zk_parts dst,
                                                             int64_t id0, zk_spec_ctr *ctr)
{
    __shared__ int base_s;
    if (threadIdx.x == 0) base_s = ctr->np_total;
    __syncthreads();
    const int base = base_s;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int *r = aos + 7 * (size_t)i;
        int v[7];
#pragma unroll
        for (int c = 0; c < 7; c++) v[c] = r[c];
        const int d = base + i;
        dst.cell[d] = zk_cell(v[0], v[1], dst.cbits);
        dst.x[d] = __int_as_float(v[2]);
        dst.y[d] = __int_as_float(v[3]);
        dst.ux[d] = __int_as_float(v[4]);
        dst.uy[d] = __int_as_float(v[5]);
        dst.uz[d] = __int_as_float(v[6]);
        if (dst.id) dst.id[d] = id0 + i;
    }
    __syncthreads();
    if (threadIdx.x == 0) ctr->np_total = base + n;
}

__global__ void unpack_commit(zk_spec_ctr *ctr, int n, int append)
{
    if (append) {
        ctr->np_total += n;
    } else {
        ctr->np_total = n;
        ctr->n_sorted = 0;
        ctr->n_dead = 0;
    }
}

__global__ void pack_kernel(zk_parts src, int n, int *__restrict__ aos)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int *r = aos + 7 * (size_t)i;
        const int c = src.cell[i];
        r[0] = zk_cell_ix(c, src.cbits);
        r[1] = zk_cell_iy(c, src.cbits);
        r[2] = __float_as_int(src.x[i]);
        r[3] = __float_as_int(src.y[i]);
        r[4] = __float_as_int(src.ux[i]);
        r[5] = __float_as_int(src.uy[i]);
        r[6] = __float_as_int(src.uz[i]);
    }
}

unsigned grid_for(long long n, int bs, unsigned cap)
{
    long long b = (n + bs - 1) / bs;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

} // namespace

extern "C" int zk_sort(const zk_parts *src, const zk_parts *dst, zk_spec_ctr *ctr, int *cnt, int *off,
                              int *cursor, int *tile_n, const zk_tiles *t, int np_upper, int spread, int slack,
                              int cap_slots, zk_stream_t s)
{
    const unsigned g = grid_for(np_upper, 256, 8192);
    /* the slot count in use before the scan is kept in off[ntiles + 1] */
    int *n_old = off + t->ntiles + 1;
    hipLaunchKernelGGL(snapshot_kernel, dim3(1), dim3(1), 0, s, ctr, n_old);
    hipLaunchKernelGGL(count_kernel, dim3(g), dim3(256), 0, s, src->cell, src->cbits, ctr, cnt, *t);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, cnt, off, cursor, ctr, t->ntiles, tile_n, slack,
                       cap_slots);
    hipLaunchKernelGGL(scatter_indirect, dim3(g), dim3(256), 0, s, *src, *dst, n_old, cursor, off, *t, spread, tile_n);
    if (t->ntiles > 0 && tile_n && slack)
        hipLaunchKernelGGL(mark_free_kernel, dim3(t->ntiles), dim3(256), 0, s, dst->cell, off, tile_n);
    ZK_LAUNCH_CHECK();
}

/* Sort of a tile-sorted input (ranges old_off, plus an unsorted tail of
 * at most tail_upper slots): one workgroup per source tile keeps the
 * particles that stay in their tile in order at the front of its new
 * range with coalesced writes; particles that changed tile (and the tail)
 * are appended to their new tile. */
extern "C" int zk_sort_tiles(const zk_parts *src, const zk_parts *dst, zk_spec_ctr *ctr, int *move, int *stay,
                             int *nbr, int *tailn, int *old_off, const int *old_n, const int *before,
                             int *new_off, int *new_n, int *cursor, const zk_tiles *t, int tail_upper, int slack,
                             int cap_slots, zk_stream_t s)
{
    const unsigned tb = tail_upper > 0 ? grid_for(tail_upper, 1024, 4096) : 0;
    const unsigned g = (unsigned)t->ntiles + tb;
    hipLaunchKernelGGL(snapshot_kernel, dim3(1), dim3(1), 0, s, ctr, old_off + t->ntiles + 1);
    hipLaunchKernelGGL(count_stay_kernel, dim3(g), dim3(ZK_SORT_THREADS), 0, s, src->cell, src->cbits, old_off, old_n,
                       stay, move, nbr, tailn, *t);
    hipLaunchKernelGGL(scan_stay_kernel, dim3(1), dim3(1024), 0, s, stay, move, tailn, new_off, new_n, cursor, ctr,
                       t->ntiles, slack, cap_slots, before);
    hipLaunchKernelGGL(scatter_stay_kernel, dim3(g), dim3(ZK_SORT_THREADS), 0, s, *src, *dst, old_off, old_n,
                       new_off, stay, nbr, cursor, *t);
    if (t->ntiles > 0 && new_n && slack)
        hipLaunchKernelGGL(mark_free_kernel, dim3(t->ntiles), dim3(256), 0, s, dst->cell, new_off, new_n);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_resort_staged(const zk_parts *dst, zk_spec_ctr *ctr, const int *tile_off, int *tile_n,
                                int *sstay, const int *sleave, int *arr, int *far, int *arr0, int *farcur,
                                int *nbr, const zk_tiles *t, zk_stream_t s)
{
    if (t->ntiles <= 0) return 0;
    hipLaunchKernelGGL(staged_count_kernel, dim3(t->ntiles), dim3(ZK_SORT_THREADS), 0, s, dst->cell, dst->cbits,
                       tile_off, sleave, arr, far, nbr, *t);
    hipLaunchKernelGGL(staged_scan_kernel, dim3(1), dim3(1024), 0, s, tile_off, tile_n, sstay, arr, far, arr0, farcur,
                       ctr, t->ntiles);
    hipLaunchKernelGGL(staged_scatter_kernel, dim3(t->ntiles), dim3(ZK_SORT_THREADS), 0, s, *dst, tile_off, tile_n,
                       sleave, arr0, farcur, nbr, ctr, *t);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_unpack(const void *aos, int n, const zk_parts *dst, int64_t id0, zk_spec_ctr *ctr,
                                int append, zk_stream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(unpack_kernel, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, (const int *)aos, n, *dst,
                           id0, ctr, append);
    hipLaunchKernelGGL(unpack_commit, dim3(1), dim3(1), 0, s, ctr, n, append);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_unpack_append(const void *aos, int n, const zk_parts *dst, int64_t id0, zk_spec_ctr *ctr,
                                 zk_stream_t s)
{
    if (n <= 0) return 0;
    if (n > ZK_APPEND_MAX) return 1;
    hipLaunchKernelGGL(unpack_append_kernel, dim3(1), dim3(1024), 0, s, (const int *)aos, n, *dst, id0, ctr);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_pack(const zk_parts *src, int n, void *aos, zk_stream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(pack_kernel, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, *src, n, (int *)aos);
    ZK_LAUNCH_CHECK();
}
