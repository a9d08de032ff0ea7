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

/* One injected column in one block of 1024 threads (zk_inject_column):
 * rounds of 4 attempts per thread, each round counted, scanned and then
 * emitted, until the pairs are drawn; the positions, the momenta, the
 * per-cell mean (one thread per row: a column's cells are its rows, each
 * row's particles contiguous and summed in order, em2d/particles.c:115-129)
 * and the generator state after the last attempt used. */
constexpr int INJ_THREADS = 1024, INJ_PER = 4;

__global__ void __launch_bounds__(1024) inject_column_kernel(zk_inject_args a)
{
    extern __shared__ int tab[]; /* tab_k, then the lattice (floats) */
    __shared__ int scan[INJ_THREADS / ZK_WAVE];
    __shared__ int base_s, last_s, total_s;
    __shared__ double kept_s;
    __shared__ int have_s;
    float *lat = (float *)(tab + a.perrow);
    const int t = threadIdx.x;
    for (int e = t; e < a.perrow; e += INJ_THREADS) tab[e] = a.tab_k[e];
    for (int e = t; e < 2 * a.nlat; e += INJ_THREADS) lat[e] = a.lat[e];
    if (t == 0) {
        base_s = a.ctr->np_total;
        last_s = -1;
        have_s = 0;
        kept_s = 0.0;
    }
    __syncthreads();
    const int base = base_s;
    const long long n = (long long)a.perrow * a.rows;
    const zk_parts &P = a.dst;
    /* positions */
    for (long long p = t; p < n; p += INJ_THREADS) {
        const int j = (int)(p / a.perrow), k = tab[p - (long long)j * a.perrow];
        const long long d = base + p;
        P.cell[d] = zk_cell(a.col, a.row0 + j, P.cbits);
        P.x[d] = lat[2 * k];
        P.y[d] = lat[2 * k + 1];
        if (P.id) P.id[d] = a.id0 + p;
    }
    /* thermal momenta: deviate q (after a kept one) is particle q / 3,
     * component q mod 3 */
    const uint32_t z0 = a.rng->z, w0 = a.rng->w;
    const int kept0 = a.rng->have_kept;
    const double kept_in = a.rng->kept;
    const uint32_t z1 = 36969u * (z0 & 0xffffu) + (z0 >> 16), w1 = 18000u * (w0 & 0xffffu) + (w0 >> 16);
    const long long D = 3 * n, pairs = (D - kept0 + 1) / 2;
    if (kept0 && t == 0) P.ux[base] = (float)((double)a.uth[0] * kept_in);
    int amb = 0;
    long long pbase = 0, ab = 0;
    /* 64 rounds hold 262144 attempts; a column needs at most 24577 pairs
     * (ZK_APPEND_MAX particles), which take 31300 attempts on average
     * (acceptance pi / 4): the bound only guarantees that the loop ends */
    for (int round = 0; round < 64 && pbase < pairs; round++) {
        const long long t0 = ab + (long long)t * INJ_PER;
        uint32_t zz, ww;
        if (t0 == 0) {
            zz = z0;
            ww = w0;
        } else {
            mwc_at(z1, w1, (uint64_t)(2 * t0), zz, ww);
        }
        const uint32_t zs = zz, ws = ww;
        int c = 0;
        for (int m = 0; m < INJ_PER; m++) {
            double u, v, sq;
            c += attempt(zz, ww, u, v, sq);
        }
        /* exclusive prefix of the counts over the block: within each wave
         * by shuffles, then over the 16 wave totals */
        const int lane = t & (ZK_WAVE - 1), wv = t / ZK_WAVE;
        int inc = c;
        for (int o = 1; o < ZK_WAVE; o <<= 1) {
            const int y = __shfl_up(inc, o, ZK_WAVE);
            if (lane >= o) inc += y;
        }
        if (lane == ZK_WAVE - 1) scan[wv] = inc;
        __syncthreads();
        if (t == 0) {
            int run = 0;
            for (int k = 0; k < INJ_THREADS / ZK_WAVE; k++) {
                const int v = scan[k];
                scan[k] = run;
                run += v;
            }
            total_s = run;
        }
        __syncthreads();
        long long p = pbase + scan[wv] + (inc - c);
        zz = zs;
        ww = ws;
        for (int m = 0; m < INJ_PER && p < pairs; m++) {
            double u, v, sq;
            if (!attempt(zz, ww, u, v, sq)) continue;
            const double fac = __builtin_sqrt(-2.0 * log_cr(sq, &amb) / sq);
            const double dv[2] = {v * fac, u * fac};
            for (int h = 0; h < 2; h++) {
                const long long q = kept0 + 2 * p + h;
                if (q < D) {
                    const long long i = q / 3;
                    const int cmp = (int)(q - 3 * i);
                    float *dst = cmp == 0 ? P.ux : (cmp == 1 ? P.uy : P.uz);
                    dst[base + i] = (float)((double)a.uth[cmp] * dv[h]);
                } else {
                    kept_s = dv[h];
                    have_s = 1;
                }
            }
            if (p == pairs - 1) last_s = (int)(t0 + m);
            p++;
        }
        pbase += total_s;
        ab += (long long)INJ_THREADS * INJ_PER;
        __syncthreads();
    }
    if (amb) atomicAdd(&a.rng->ambiguous, amb);
    /* per-cell mean (the column's cells are its rows), then the drift */
    for (int j = t; j < a.rows; j += INJ_THREADS) {
        const long long p0 = base + (long long)j * a.perrow;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (int m = 0; m < a.perrow; m++) {
            sx += P.ux[p0 + m];
            sy += P.uy[p0 + m];
            sz += P.uz[p0 + m];
        }
        if (a.perrow > 0) {
            const float wgt = 1.0f / a.perrow;
            sx *= wgt;
            sy *= wgt;
            sz *= wgt;
        }
        for (int m = 0; m < a.perrow; m++) {
            P.ux[p0 + m] += a.ufl[0] - sx;
            P.uy[p0 + m] += a.ufl[1] - sy;
            P.uz[p0 + m] += a.ufl[2] - sz;
        }
    }
    __syncthreads();
    if (t == 0) {
        /* the generator after the calls of the last attempt used */
        if (pairs > 0) {
            uint32_t zz, ww;
            mwc_at(z1, w1, (uint64_t)(2 * (long long)last_s + 2), zz, ww);
            a.rng->z = zz;
            a.rng->w = ww;
            a.rng->have_kept = have_s;
            a.rng->kept = have_s ? kept_s : 0.0;
        }
        a.ctr->np_total = base + (int)n;
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

extern "C" int zk_inject_column(const zk_inject_args *a, zk_stream_t s)
{
    const long long n = (long long)a->perrow * a->rows;
    const size_t lds = sizeof(int) * (size_t)a->perrow + sizeof(float) * 2 * (size_t)a->nlat;
    if (n <= 0 || n > ZK_APPEND_MAX || lds > ZK_INJECT_TABLE_MAX) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(inject_column_kernel, dim3(1), dim3(INJ_THREADS), lds, s, *a);
    ZK_LAUNCH_CHECK();
}
