/*
 * batch.hip - one launch per phase for many independent em2d simulations.
 *
 * A batch runs the reference step sequence (em2d/simulation.c:45-56) of
 * every simulation, but each phase of the step (current reset, push of the
 * k-th species, out-of-window push, energy, current reduction, filter
 * passes, Yee update, guard cells) is a single kernel launch over an item
 * table of (simulation or species, block) pairs, so simulations of uneven
 * size share the GPU and a whole step is a fixed sequence of launches that
 * a HIP graph captures once and replays.
 *
 * The buffers a phase reads or writes are selected through indices kept in
 * the device descriptors (ecur, jcur, pcur); batch_end_step advances them,
 * so the same captured graph is valid for every step.  The per-block work
 * is the device code of the one-simulation kernels (push_core.h,
 * current_core.h, fields_core.h, sort_core.h), so results are identical to
 * running each simulation on its own.
 */
#include "current_core.h"
#include "fields_core.h"
#include "push_core.h"
#include "sort_core.h"

using namespace zk;

/* waves per SIMD the fixed-point tile push is compiled for: 5 (96 VGPRs,
 * 24 bytes per lane spilled) against 4: 512-member scan 146.5 -> 148.3
 * sims/s */
#ifndef ZK_PUSH_OCC
#define ZK_PUSH_OCC 5
#endif

namespace {

__device__ inline zk_push_args spec_args(const zk_bspec &S, const zk_bsim &M)
{
    zk_push_args a = S.a;
    a.p = S.set[S.pcur];
    a.tile_off = S.offs[S.ocur];
    a.E = M.E[M.ecur];
    a.B = M.B[M.ecur];
    return a;
}

/* the sort of this step runs for species whose schedule asks for it */
__device__ inline bool sorts_now(const zk_bspec &S) { return S.n_sort > 0 && (S.iter + 1) % S.n_sort == 0; }

__device__ inline float *j_after(const zk_bsim &M, int passes) { return M.J[(M.jcur + passes) & 1]; }

__global__ void fill_jov_kernel(const zk_bsim *sims, const int *items)
{
    const int s = items[2 * blockIdx.x], b = items[2 * blockIdx.x + 1];
    const zk_bsim &M = sims[s];
    const size_t n = 3 * (size_t)(M.nx + 3) * (M.ny + 3);
    const size_t i = (size_t)b * blockDim.x + threadIdx.x;
    if (i < n) M.Jov[i] = 0;
}

template <int MIN_WAVES, bool FX>
__global__ void __launch_bounds__(256, MIN_WAVES) push_tile_batch(const zk_bspec *specs, const zk_bsim *sims,
                                                                  const int *items)
{
    const zk_bspec &S = specs[items[2 * blockIdx.x]];
    const zk_push_args a = spec_args(S, sims[S.sim]);
    push_tile_body<1, FX>(OneSpecies{a}, items[2 * blockIdx.x + 1]);
}

__global__ void __launch_bounds__(256) push_global_batch(const zk_bspec *specs, const zk_bsim *sims,
                                                         const int *items)
{
    const zk_bspec &S = specs[items[2 * blockIdx.x]];
    const zk_push_args a = spec_args(S, sims[S.sim]);
    push_global_body(a, items[2 * blockIdx.x + 1], a.global_blocks);
}

__global__ void finalize_batch(const zk_bspec *specs)
{
    const zk_bspec &S = specs[blockIdx.x];
    finalize_body(S.a.ctr, S.a.epart, S.nparts, S.qm, S.dx0, S.dx1);
}

__global__ void reduce_batch(const zk_bsim *sims, const int *items)
{
    const zk_bsim &M = sims[items[2 * blockIdx.x]];
    zk_reduce_args r;
    r.region = M.deposited ? M.region : nullptr;
    r.Jov = M.Jov;
    r.jov_unscale = M.jov_unscale;
    r.J = j_after(M, 0);
    r.nrow = M.nrow;
    r.t = M.t;
    r.periodic_x = 1;
    r.open_y = 0;
    reduce_node(r, items[2 * blockIdx.x + 1] * blockDim.x + threadIdx.x);
}

__global__ void smooth_batch(const zk_bsim *sims, const int *items, int pass)
{
    const zk_bsim &M = sims[items[2 * blockIdx.x]];
    if (pass >= M.npass) return;
    smooth_node(j_after(M, pass), j_after(M, pass + 1), M.nx, M.ny, M.nrow, M.pass_axis[pass], M.pass_sa[pass],
                M.pass_sb[pass], 1, items[2 * blockIdx.x + 1] * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) yee_batch(const zk_bsim *sims, const int *items)
{
    const int s = items[2 * blockIdx.x], t = items[2 * blockIdx.x + 1];
    const zk_bsim &M = sims[s];
    const int ntx = (M.nx + 3 + BX - 1) / BX;
    yee_tile(M.E[M.ecur], M.B[M.ecur], j_after(M, M.npass), M.E[1 - M.ecur], M.B[1 - M.ecur], M.nx, M.ny, M.nrow,
             M.yk, t % ntx, t / ntx);
}

__global__ void guards_batch(const zk_bsim *sims, const int *items)
{
    const zk_bsim &M = sims[items[2 * blockIdx.x]];
    guard_node(M.E[1 - M.ecur], M.B[1 - M.ecur], M.nx, M.ny, M.nrow, 1,
               items[2 * blockIdx.x + 1] * blockDim.x + threadIdx.x);
}

/* The re-sort of the single-simulation path (zk_sort_tiles: stayers in
 * order, movers by neighbour slot), one workgroup per (species, tile)
 * item; batch members have no unsorted tail (no moving window). */
__global__ void snapshot_batch(const zk_bspec *specs)
{
    const zk_bspec &S = specs[blockIdx.x];
    if (sorts_now(S)) S.offs[S.ocur][S.a.t.ntiles + 1] = S.a.ctr->np_total;
}

__global__ void __launch_bounds__(ZK_SORT_THREADS) count_batch(const zk_bspec *specs, const int *items)
{
    const zk_bspec &S = specs[items[2 * blockIdx.x]];
    if (!sorts_now(S)) return;
    const zk_parts &p = S.set[S.pcur];
    count_stay_body(p.cell, p.cbits, S.offs[S.ocur], nullptr, S.stay, S.cnt, S.nbr, S.tailn, S.a.t, items[2 * blockIdx.x + 1],
                    S.a.t.ntiles);
}

__global__ void __launch_bounds__(1024) scan_batch(const zk_bspec *specs)
{
    const zk_bspec &S = specs[blockIdx.x];
    if (!sorts_now(S)) return;
    scan_stay_body(S.stay, S.cnt, S.tailn, S.offs[1 - S.ocur], nullptr, S.cursor, S.a.ctr, S.a.t.ntiles, 0, 0);
}

__global__ void __launch_bounds__(ZK_SORT_THREADS) scatter_batch(const zk_bspec *specs, const int *items)
{
    const zk_bspec &S = specs[items[2 * blockIdx.x]];
    if (!sorts_now(S)) return;
    scatter_stay_body(S.set[S.pcur], S.set[1 - S.pcur], S.offs[S.ocur], nullptr, S.offs[1 - S.ocur], S.stay, S.nbr, S.cursor,
                      S.a.t, items[2 * blockIdx.x + 1], S.a.t.ntiles);
}

/* Advance the buffer indices of every simulation and species so the next
 * replay reads this step's results: new fields, filtered current, and the
 * sorted particle set of species sorted this step. */
__global__ void end_step_batch(zk_bsim *sims, int nsims, zk_bspec *specs, int nspecs, int sorted)
{
    for (int s = threadIdx.x; s < nsims; s += blockDim.x) {
        sims[s].ecur ^= 1;
        sims[s].jcur = (sims[s].jcur + sims[s].npass) & 1;
    }
    for (int k = threadIdx.x; k < nspecs; k += blockDim.x) {
        if (sorted && sorts_now(specs[k])) {
            specs[k].pcur ^= 1;
            specs[k].ocur ^= 1;
        }
        specs[k].iter++;
    }
}

} // namespace

#define LAUNCH_ITEMS(kernel, n, ...)                                                                                   \
    do {                                                                                                               \
        if ((n) > 0) hipLaunchKernelGGL(kernel, dim3(n), dim3(256), 0, s, __VA_ARGS__);                                \
    } while (0)

extern "C" int zk_batch_fill_jov(const zk_bsim *sims, const int *items, int n, zk_stream_t s)
{
    LAUNCH_ITEMS(fill_jov_kernel, n, sims, items);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_batch_push(const zk_bspec *specs, const zk_bsim *sims, const int *tile_items, int ntile,
                             const int *global_items, int nglobal, size_t lds, int waves, int fixed, zk_stream_t s)
{
    if (ntile > 0) {
        if (fixed)
            hipLaunchKernelGGL((push_tile_batch<ZK_PUSH_OCC, true>), dim3(ntile), dim3(256), lds, s, specs, sims, tile_items);
        else if (waves == 8)
            hipLaunchKernelGGL((push_tile_batch<8, false>), dim3(ntile), dim3(256), lds, s, specs, sims, tile_items);
        else
            hipLaunchKernelGGL((push_tile_batch<1, false>), dim3(ntile), dim3(256), lds, s, specs, sims, tile_items);
    }
    LAUNCH_ITEMS(push_global_batch, nglobal, specs, sims, global_items);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_batch_prepare_push(size_t lds)
{
    if (lds <= 65536) return 0;
    const void *fns[3] = {(const void *)push_tile_batch<1, false>, (const void *)push_tile_batch<8, false>,
                          (const void *)push_tile_batch<ZK_PUSH_OCC, true>};
    for (int k = 0; k < 3; k++) {
        const int e = (int)hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e) return e;
    }
    return 0;
}

extern "C" int zk_batch_finalize(const zk_bspec *specs, int nspecs, zk_stream_t s)
{
    if (nspecs > 0) hipLaunchKernelGGL(finalize_batch, dim3(nspecs), dim3(256), 0, s, specs);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_batch_current(const zk_bsim *sims, const int *node_items, int n, int max_passes, zk_stream_t s)
{
    LAUNCH_ITEMS(reduce_batch, n, sims, node_items);
    for (int p = 0; p < max_passes; p++) LAUNCH_ITEMS(smooth_batch, n, sims, node_items, p);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_batch_fields(const zk_bsim *sims, const int *yee_items, int nyee, const int *guard_items,
                               int nguard, zk_stream_t s)
{
    LAUNCH_ITEMS(yee_batch, nyee, sims, yee_items);
    LAUNCH_ITEMS(guards_batch, nguard, sims, guard_items);
    ZK_LAUNCH_CHECK();
}

/* The sort of the species whose schedule asks for it this step; a graph
 * holding these launches must end its step with zk_batch_end_step_sorted. */
extern "C" int zk_batch_sort(const zk_bspec *specs, int nspecs, const int *items, int n, zk_stream_t s)
{
    if (nspecs <= 0) return 0;
    hipLaunchKernelGGL(snapshot_batch, dim3(nspecs), dim3(1), 0, s, specs);
    LAUNCH_ITEMS(count_batch, n, specs, items);
    hipLaunchKernelGGL(scan_batch, dim3(nspecs), dim3(1024), 0, s, specs);
    LAUNCH_ITEMS(scatter_batch, n, specs, items);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_batch_end_step(zk_bsim *sims, int nsims, zk_bspec *specs, int nspecs, zk_stream_t s)
{
    hipLaunchKernelGGL(end_step_batch, dim3(1), dim3(256), 0, s, sims, nsims, specs, nspecs, 0);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_batch_end_step_sorted(zk_bsim *sims, int nsims, zk_bspec *specs, int nspecs, zk_stream_t s)
{
    hipLaunchKernelGGL(end_step_batch, dim3(1), dim3(256), 0, s, sims, nsims, specs, nspecs, 1);
    ZK_LAUNCH_CHECK();
}
