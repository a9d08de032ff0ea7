/*
 * push.hip - particle push and charge-conserving current deposit (gfx950).
 *
 * Replaces the particle loop of the reference spec_advance
 * (em2d/particles.c:1125-1259) with interpolate_fld (:1029-1071), the
 * Boris push (:1137-1192), ltrim (:1081-1084) and dep_current_zamb
 * (:773-924).  The per-particle arithmetic follows the reference operation
 * by operation; with -ffp-contract=off and IEEE division/square root the
 * new momenta, positions and cell indices are bit-identical to the
 * reference given the same fields.
 *
 * Layout of one launch: one workgroup per tile of TX x TY cells.  The
 * workgroup stages the E and B values of its window (the tile plus an H
 * cell drift margin plus the 1-below/2-above stencil) into LDS, pushes the
 * particles of its tile range with gathers from LDS, accumulates the
 * current of its window in LDS (64-bit fixed point with ds_add_u64 by
 * default, float ds_add_f32 on request), and finally stores the window to
 * its own slot of `region` with plain coalesced stores.  No two
 * workgroups write the same address, so no global atomics are needed; the
 * windows are summed into J afterwards by current.hip in a fixed order.
 *
 * A particle that has drifted more than H cells out of its tile since the
 * last sort is listed by the tile pass and pushed by a second kernel on the
 * global path: gathers from the guard-inclusive field buffers exactly as
 * the reference indexes them, and deposits into the raw overflow grid
 * `Jov` with 64-bit fixed-point atomics (exact, order-free sums).  The
 * same kernel pushes the unsorted tail (particles injected by the moving
 * window since the last sort).
 */
#include "push_core.h"

using namespace zk;

/* waves per SIMD of the single-species tile push: 5 fits 96 VGPRs, with a
 * few values spilled outside the loop (LWFA 4000 x 256: 0.2537 -> 0.2505
 * ms per step against 4 waves) */
#ifndef ZK_PUSH_OCC
#define ZK_PUSH_OCC 5
#endif
/* waves per SIMD of the fused multi-species push: 5 fits 96 VGPRs, with
 * 16 bytes per lane spilled before the loop and reloaded after it; the
 * extra wave hides more latency (2048^2 Weibel: 5 waves 1.727 ms per
 * launch against 1.750 with 4, when the 5-wave build still reloaded two
 * spilled address terms in the loop; 1.697 ms without them) */
#ifndef ZK_SPECIES_OCC
#define ZK_SPECIES_OCC 5
#endif

namespace {

/* WC: the window is WC x WC nodes (compile-time LDS offsets), 0 any */
constexpr int ZK_WIN_DEFAULT = 25; /* 16 x 16 tiles, 3-cell margin */

/* Blocks past the tiles push the unsorted tail (the global pass) in the
 * same launch: the two touch disjoint particles, deposit into disjoint
 * windows or with exact atomics into the overflow grid, and write their own
 * energy partials, so they may run side by side (one launch and its gap
 * less per step while a moving window's injected columns wait for a sort). */
template <int MIN_WAVES, int PAIR, bool FX, int WC = 0>
__global__ void __launch_bounds__(256, MIN_WAVES) push_tile_kernel(zk_push_args a)
{
    if ((int)blockIdx.x >= a.t.ntiles) {
        push_global_body(a, (int)blockIdx.x - a.t.ntiles, a.global_blocks);
        return;
    }
    push_tile_body<PAIR, FX, OneSpecies, WC>(OneSpecies{a}, zk_xcd_item(blockIdx.x, a.t.ntiles));
}

/* every species of a simulation in one launch (shared staging and window) */
template <int WC>
__global__ void __launch_bounds__(256, ZK_SPECIES_OCC) push_species_kernel(zk_push_multi m)
{
    push_tile_body<1, true, SpeciesSet, WC>(SpeciesSet{m}, zk_xcd_item(blockIdx.x, m.s[0].t.ntiles));
}

bool default_window(const zk_tiles &t) { return t.WX == ZK_WIN_DEFAULT && t.WY == ZK_WIN_DEFAULT; }

__global__ void __launch_bounds__(256) push_global_kernel(zk_push_args a)
{
    push_global_body(a, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(1024) finalize_kernel(zk_spec_ctr *ctr, const double *epart, int nparts, double qm, double dx0,
                                double dx1)
{
    finalize_body(ctr, epart, nparts, qm, dx0, dx1);
}

/* one block per species */
__global__ void __launch_bounds__(1024) finalize_set_kernel(zk_finalize_set f)
{
    const int k = blockIdx.x;
    finalize_body(f.ctr[k], f.epart[k], f.nparts[k], f.qm[k], f.dx0[k], f.dx1[k]);
}

__global__ void rescale_fx_kernel(long long *p, size_t n, int shift)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] >>= shift;
}

__global__ void fill_kernel(float *p, size_t n, float v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

} // namespace

extern "C" int zk_push(const zk_push_args *a, zk_stream_t s)
{
    const int n = a->t.WX * a->t.WY;
    /* E, B window (6 n floats) and the current window: 3 n floats, or 3 n
     * 64-bit integers with the fixed-point deposit */
    const size_t lds = (size_t)6 * n * sizeof(float) + (size_t)3 * n * (a->fixed ? 8 : 4);
    /* variants: 0 lets the compiler keep every value in registers (about
     * 6 waves per SIMD); 8 asks for 8 waves per SIMD at the price of a few
     * register spills; 2 pushes two particles per thread and iteration */
    const bool dw = default_window(a->t);
    const void *fn = a->fixed ? (dw ? (const void *)push_tile_kernel<ZK_PUSH_OCC, 1, true, ZK_WIN_DEFAULT>
                                    : (const void *)push_tile_kernel<ZK_PUSH_OCC, 1, true>)
                     : a->waves == 8 ? (const void *)push_tile_kernel<8, 1, false>
                     : a->waves == 2 ? (const void *)push_tile_kernel<1, 2, false>
                                     : (const void *)push_tile_kernel<1, 1, false>;
    if (lds > 65536) {
        /* windows of large tiles need more than the default 64 KiB of LDS */
        const int e = (int)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e) return e;
    }
    if (a->t.ntiles > 0) {
        /* the tiles, then the tail's blocks (global_blocks) */
        const dim3 grid(a->t.ntiles + (a->global_blocks > 0 ? a->global_blocks : 0));
        if (a->fixed && dw)
            hipLaunchKernelGGL((push_tile_kernel<ZK_PUSH_OCC, 1, true, ZK_WIN_DEFAULT>), grid, dim3(256), lds, s, *a);
        else if (a->fixed)
            hipLaunchKernelGGL((push_tile_kernel<ZK_PUSH_OCC, 1, true>), grid, dim3(256), lds, s, *a);
        else if (a->waves == 8)
            hipLaunchKernelGGL((push_tile_kernel<8, 1, false>), grid, dim3(256), lds, s, *a);
        else if (a->waves == 2)
            hipLaunchKernelGGL((push_tile_kernel<1, 2, false>), grid, dim3(256), lds, s, *a);
        else
            hipLaunchKernelGGL((push_tile_kernel<1, 1, false>), grid, dim3(256), lds, s, *a);
    } else if (a->global_blocks > 0) {
        hipLaunchKernelGGL(push_global_kernel, dim3(a->global_blocks), dim3(256), 0, s, *a);
    }
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_push_species(const zk_push_multi *m, zk_stream_t s)
{
    const zk_push_args *a = &m->s[0];
    if (m->n < 1 || m->n > ZK_FUSE_MAX || !a->fixed) return (int)hipErrorInvalidValue;
    const int n = a->t.WX * a->t.WY;
    const size_t lds = (size_t)6 * n * sizeof(float) + (size_t)3 * n * 8;
    const bool dw = default_window(a->t);
    const void *fn = dw ? (const void *)push_species_kernel<ZK_WIN_DEFAULT> : (const void *)push_species_kernel<0>;
    if (lds > 65536) {
        const int e = (int)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e) return e;
    }
    if (a->t.ntiles > 0) {
        if (dw)
            hipLaunchKernelGGL(push_species_kernel<ZK_WIN_DEFAULT>, dim3(a->t.ntiles), dim3(256), lds, s, *m);
        else
            hipLaunchKernelGGL(push_species_kernel<0>, dim3(a->t.ntiles), dim3(256), lds, s, *m);
    }
    for (int k = 0; k < m->n; k++)
        if (m->s[k].global_blocks > 0)
            hipLaunchKernelGGL(push_global_kernel, dim3(m->s[k].global_blocks), dim3(256), 0, s, m->s[k]);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_spec_finalize(zk_spec_ctr *ctr, const double *epart, int nparts, double qm,
                                       double dx0, double dx1, zk_stream_t s)
{
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(1024), 0, s, ctr, epart, nparts, qm, dx0, dx1);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_spec_finalize_set(const zk_finalize_set *f, zk_stream_t s)
{
    if (f->n < 1 || f->n > ZK_FUSE_MAX) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(finalize_set_kernel, dim3(f->n), dim3(1024), 0, s, *f);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_fill_f32(float *p, size_t n, float v, zk_stream_t s)
{
    if (n == 0) return 0;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_rescale_fx(long long *p, size_t n, int shift, zk_stream_t s)
{
    if (n == 0 || shift <= 0) return 0;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(rescale_fx_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, shift);
    ZK_LAUNCH_CHECK();
}
