/*
 * init.hip - particle initialisation of spec_new on the device (gfx950).
 *
 * Produces the particles and momenta that the host initialisation
 * (em2d/particles.c:85-354, em2d/random.c:48-101) produces, in the same
 * order, for the uniform, step and slab densities:
 *
 * - positions: the lattice of ppc[0] x ppc[1] points per cell, emitted row
 *   by row; a row's (cell, lattice point) entries come from a table built
 *   on the host (the step and slab profiles drop points column by column);
 * - thermal momenta: u = uth * rand_norm() in the order ux, uy, uz of
 *   particle 0, 1, ..., from the reference's generator: the two 16-bit
 *   multiply-with-carry halves z <- 36969 (z & 0xffff) + (z >> 16) and
 *   w <- 18000 (w & 0xffff) + (w >> 16) are, after their first step, the
 *   multiplicative congruential generators z <- 36969 z mod (36969 * 2^16 - 1)
 *   and w <- 18000 w mod (18000 * 2^16 - 1), so a thread jumps to its place
 *   in the stream by modular exponentiation.  The polar method draws
 *   attempts of two uniforms and accepts those inside the unit circle; a
 *   count pass, a prefix sum over the blocks and an emit pass give every
 *   accepted attempt its pair index, and pair p supplies deviates 2p and
 *   2p + 1 (after a deviate kept from before, if any);
 * - the per-cell mean of the thermal momenta (summed in particle order,
 *   cell key ix + nx[1] iy as in the reference) is removed and the fluid
 *   drift added.
 *
 * log(s) of the polar method is evaluated in double-double arithmetic and
 * rounded once (Dekker's exact products and sums; log(m) = 2 atanh(f),
 * f = (m - 1)/(m + 1), by its series to relative 2^-106), which is the
 * correctly rounded double unless log(s) lies within about 2^-100 of a
 * rounding boundary; such cases are counted (Ziv's rounding test).  The
 * square root and divisions are IEEE double operations.
 */
#include <stdint.h>

#include "common.h"
#include "init_core.h"

using namespace zk_init;

namespace {

/* the attempt's two uniforms in [-1, 1] (em2d/random.c:82-83) */
__device__ inline bool attempt(uint32_t &z, uint32_t &w, double &u, double &v, double &s)
{
    u = (mwc_next(z, w) + 0.5) / 2147483649.0 - 1.0;
    v = (mwc_next(z, w) + 0.5) / 2147483649.0 - 1.0;
    s = u * u + v * v;
    return !(s == 0.0 || s >= 1.0);
}

/* Count pass: thread g handles attempts [g per, (g + 1) per) of the
 * `nat` attempts; per-thread counts within the block (exclusive prefix)
 * and the block's total. */
__global__ void __launch_bounds__(256) rng_count_kernel(zk_init_args a, long long nat, int per, int *thr_off,
                                                        int *blk_cnt)
{
    __shared__ int sc[256];
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long t0 = g * per;
    int c = 0;
    if (t0 < nat) {
        /* attempt t uses calls 2t + 1 and 2t + 2: start from the state
         * after call 2 t0 (the initial state for t0 = 0) */
        uint32_t zz, ww;
        if (t0 == 0) {
            zz = a.z0;
            ww = a.w0;
        } else {
            mwc_at(a.z1, a.w1, (uint64_t)(2 * t0), zz, ww);
        }
        const long long t1 = t0 + per < nat ? t0 + per : nat;
        for (long long t = t0; t < t1; t++) {
            double u, v, s;
            c += attempt(zz, ww, u, v, s);
        }
    }
    sc[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = 0; k < (int)blockDim.x; k++) {
            const int v = sc[k];
            sc[k] = run;
            run += v;
        }
        blk_cnt[blockIdx.x] = run;
    }
    __syncthreads();
    thr_off[g] = sc[threadIdx.x];
}

/* Emit pass: the accepted attempts' deviates become thermal momenta;
 * blk_off is the exclusive prefix of the block counts. */
__global__ void __launch_bounds__(256) rng_emit_kernel(zk_init_args a, long long nat, int per, const int *thr_off,
                                                       const long long *blk_off, zk_init_out *out)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long t0 = g * per;
    if (t0 >= nat) return;
    const long long D = 3LL * a.np;
    long long p = blk_off[blockIdx.x] + thr_off[g];
    if (p >= a.pairs) return;
    uint32_t zz, ww;
    if (t0 == 0) {
        zz = a.z0;
        ww = a.w0;
    } else {
        mwc_at(a.z1, a.w1, (uint64_t)(2 * t0), zz, ww);
    }
    const long long t1 = t0 + per < nat ? t0 + per : nat;
    int amb = 0;
    for (long long t = t0; t < t1 && p < a.pairs; t++) {
        double u, v, s;
        if (!attempt(zz, ww, u, v, s)) continue;
        const double fac = __builtin_sqrt(-2.0 * log_cr(s, &amb) / s);
        const double dv[2] = {v * fac, u * fac};
        for (int h = 0; h < 2; h++) {
            const long long q = a.kept0 + 2 * p + h;
            if (q < D) {
                const long long i = q / 3;
                const int c = (int)(q - 3 * i);
                float *dst = c == 0 ? a.ux : (c == 1 ? a.uy : a.uz);
                dst[i] = (float)((double)a.uth[c] * dv[h]);
            } else {
                out->kept = dv[h];
                out->have_kept = 1;
            }
        }
        if (p == a.pairs - 1) out->last_attempt = t;
        p++;
    }
    if (amb) atomicAdd(&out->ambiguous, 1);
}

/* the kept deviate from before the species is particle 0's u_x */
__global__ void kept_kernel(zk_init_args a)
{
    if (a.kept0 && a.np > 0) a.ux[0] = (float)((double)a.uth[0] * a.kept_in);
}

/* positions: row j = p / perrow, (cell, lattice point) of the row from the
 * table */
__global__ void positions_kernel(zk_init_args a)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < a.np;
         p += (long long)gridDim.x * blockDim.x) {
        const long long j = p / a.perrow;
        const int e = (int)(p - j * a.perrow);
        const int i = a.tab_i[e], k = a.tab_k[e];
        a.cell[p] = zk_cell(i, a.row0 + (int)j, a.cbits);
        a.x[p] = a.lat[2 * k];
        a.y[p] = a.lat[2 * k + 1];
    }
}

/* Mean thermal momentum of each key idx = ix + ny iy: the cells with that
 * key (several when nx > ny) in particle order, each cell's particles in
 * order (em2d/particles.c:115-129). */
__global__ void cell_mean_kernel(zk_init_args a, float *mean, int nkeys)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nkeys) return;
    const int ny = a.stride;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int n = 0;
    /* rows j with i = idx - ny j in the injected columns, increasing j:
     * col0 <= idx - ny j <= col0 + ncols - 1 */
    int jlo = a.row0, jhi = a.row0 + a.rows - 1;
    const int top = idx - a.col0;                   /* ny j <= top */
    const int bot = idx - (a.col0 + a.ncols - 1);   /* ny j >= bot */
    if (top < 0) return;
    if (top / ny < jhi) jhi = top / ny;
    if (bot > 0 && (bot + ny - 1) / ny > jlo) jlo = (bot + ny - 1) / ny;
    for (int j = jlo; j <= jhi; j++) {
        const int i = idx - ny * j;
        const int c = i - a.col0;
        const long long p0 = (long long)(j - a.row0) * a.perrow + a.col_start[c];
        for (int m = 0; m < a.col_count[c]; m++) {
            sx += a.ux[p0 + m];
            sy += a.uy[p0 + m];
            sz += a.uz[p0 + m];
        }
        n += a.col_count[c];
    }
    if (n > 0) {
        const float w = 1.0f / n;
        mean[3 * idx + 0] = sx * w;
        mean[3 * idx + 1] = sy * w;
        mean[3 * idx + 2] = sz * w;
    }
}

__global__ void apply_mean_kernel(zk_init_args a, const float *mean)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < a.np;
         p += (long long)gridDim.x * blockDim.x) {
        const int idx = zk_cell_ix(a.cell[p], a.cbits) + a.stride * zk_cell_iy(a.cell[p], a.cbits);
        a.ux[p] += a.ufl[0] - mean[3 * idx + 0];
        a.uy[p] += a.ufl[1] - mean[3 * idx + 1];
        a.uz[p] += a.ufl[2] - mean[3 * idx + 2];
    }
}

unsigned blocks_for(long long n, unsigned cap)
{
    long long b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

} // namespace

extern "C" int zk_init_positions(const zk_init_args *a, zk_stream_t s)
{
    if (a->np > 0) hipLaunchKernelGGL(positions_kernel, dim3(blocks_for(a->np, 16384)), dim3(256), 0, s, *a);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_init_rng_count(const zk_init_args *a, long long nat, int per, int *thr_off, int *blk_cnt,
                                 int nblocks, zk_stream_t s)
{
    if (nblocks > 0) hipLaunchKernelGGL(rng_count_kernel, dim3(nblocks), dim3(256), 0, s, *a, nat, per, thr_off, blk_cnt);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_init_rng_emit(const zk_init_args *a, long long nat, int per, const int *thr_off,
                                const long long *blk_off, int nblocks, zk_init_out *out, zk_stream_t s)
{
    hipLaunchKernelGGL(kept_kernel, dim3(1), dim3(1), 0, s, *a);
    if (nblocks > 0)
        hipLaunchKernelGGL(rng_emit_kernel, dim3(nblocks), dim3(256), 0, s, *a, nat, per, thr_off, blk_off, out);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_init_mean(const zk_init_args *a, float *mean, int nkeys, zk_stream_t s)
{
    if (nkeys > 0) hipLaunchKernelGGL(cell_mean_kernel, dim3((nkeys + 255) / 256), dim3(256), 0, s, *a, mean, nkeys);
    if (a->np > 0) hipLaunchKernelGGL(apply_mean_kernel, dim3(blocks_for(a->np, 16384)), dim3(256), 0, s, *a, mean);
    ZK_LAUNCH_CHECK();
}
