/*
 * current.hip - current density after deposition (gfx950).
 *
 * zk_current_reduce turns the per-tile current windows written by the
 * push kernel (plus the raw current of out-of-window pushes) into the final
 * guard-inclusive J.  It reproduces the reference's deposit-then-fold
 * semantics (em2d/current.c:118-159): deposits are accumulated at raw
 * node positions in [-1, n+1], and every node of a periodic axis receives
 * the sum of all its periodic images, guard copies included.  Along x with
 * a moving window, and along y with open boundaries, there is no fold and
 * the raw guard columns (rows) are kept.
 * One thread per output node; contributions are summed in a fixed order,
 * so the result does not depend on scheduling.
 *
 * zk_smooth applies one [sa, sb, sa] pass along x or y with the guard
 * handling of kernel_x / kernel_y (em2d/current.c:316-413).
 */
#include "current_core.h"

using namespace zk;

namespace {

__global__ void reduce_kernel(zk_reduce_args a) { reduce_node(a, blockIdx.x * blockDim.x + threadIdx.x); }

__global__ void smooth_kernel(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nrow,
                              int axis, float sa, float sb, int periodic_x, int open_y)
{
    smooth_node(in, out, nx, ny, nrow, axis, sa, sb, periodic_x, blockIdx.x * blockDim.x + threadIdx.x, open_y);
}

__global__ void smooth_xy_kernel(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nrow,
                                 float sax, float sbx, float say, float sby, int periodic_x)
{
    smooth_xy_node(in, out, nx, ny, nrow, sax, sbx, say, sby, periodic_x, blockIdx.x * blockDim.x + threadIdx.x);
}

/* Consecutive x passes in one launch: a block stages SX_R buffer rows of
 * SX_W output columns plus npass columns on each side in LDS and applies
 * every pass there (the valid region shrinks by one column per pass), with
 * the arithmetic and guard handling of kernel_x (em2d/current.c:316-354)
 * pass after pass: interior rows only; with periodic x every column is the
 * image of an interior one (the reference refreshes the guards from their
 * images after each pass); with a moving window the guard columns keep
 * their values.  Bit-identical to the single-pass launches. */
constexpr int SX_W = 256, SX_R = 4, SX_MAXP = ZK_MAX_PASSES;

struct SmoothPasses {
    float sa[SX_MAXP], sb[SX_MAXP];
    int n;
};

__global__ void __launch_bounds__(256) smooth_x_passes_kernel(const float *__restrict__ in, float *__restrict__ out,
                                                              int nx, int ny, int nrow, SmoothPasses ps,
                                                              int periodic_x)
{
    extern __shared__ float sx[];
    const int P = ps.n, W = SX_W + 2 * P;
    float *bufA = sx, *bufB = sx + 3 * SX_R * W;
    const int bw = nx + 3, bh = ny + 3;
    const int c0 = blockIdx.x * SX_W, r0 = blockIdx.y * SX_R;
    /* stage: local column m <-> buffer column c0 - P + m, raw i = that - 1 */
    for (int e = threadIdx.x; e < SX_R * W; e += blockDim.x) {
        const int r = e / W, m = e % W;
        const int bj = r0 + r, b = c0 - P + m;
        float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
        if (bj < bh) {
            const int j = bj - 1;
            const bool interior_row = j >= 0 && j < ny;
            int src = b;
            if (interior_row && periodic_x) src = zk_wrap(b - 1, nx) + 1;
            if (src >= 0 && src < bw) {
                const float *q = in + 3 * ((size_t)src + (size_t)bj * nrow);
                v0 = q[0];
                v1 = q[1];
                v2 = q[2];
            }
        }
        float *d = bufA + 3 * e;
        d[0] = v0;
        d[1] = v1;
        d[2] = v2;
    }
    __syncthreads();
    float *cur = bufA, *nxt = bufB;
    for (int p = 0; p < P; p++) {
        const float sa = ps.sa[p], sb = ps.sb[p];
        for (int e = threadIdx.x; e < SX_R * W; e += blockDim.x) {
            const int r = e / W, m = e % W;
            const int bj = r0 + r, b = c0 - P + m;
            const int j = bj - 1, i = b - 1;
            const bool interior_row = j >= 0 && j < ny;
            const bool live = interior_row && (periodic_x || (i >= 0 && i < nx)) && m > 0 && m < W - 1;
            const float *c = cur + 3 * e;
            float *o = nxt + 3 * e;
            if (live) {
                stencil(sa, sb, c - 3, c, c + 3, o);
            } else {
                o[0] = c[0];
                o[1] = c[1];
                o[2] = c[2];
            }
        }
        __syncthreads();
        float *t = cur;
        cur = nxt;
        nxt = t;
    }
    for (int e = threadIdx.x; e < SX_R * SX_W; e += blockDim.x) {
        const int r = e / SX_W, m = e % SX_W;
        const int bj = r0 + r, b = c0 + m;
        if (bj >= bh || b >= bw) continue;
        const float *c = cur + 3 * (r * W + m + P);
        float *o = out + 3 * ((size_t)b + (size_t)bj * nrow);
        o[0] = c[0];
        o[1] = c[1];
        o[2] = c[2];
    }
}

} // namespace

extern "C" int zk_smooth_x_passes(const float *in, float *out, int nx, int ny, int nrow, int npass, const float *sa,
                                  const float *sb, int periodic_x, zk_stream_t s)
{
    if (npass < 1 || npass > SX_MAXP) return (int)hipErrorInvalidValue;
    SmoothPasses ps;
    ps.n = npass;
    for (int k = 0; k < npass; k++) {
        ps.sa[k] = sa[k];
        ps.sb[k] = sb[k];
    }
    const size_t lds = (size_t)2 * 3 * SX_R * (SX_W + 2 * npass) * sizeof(float);
    dim3 grid((nx + 3 + SX_W - 1) / SX_W, (ny + 3 + SX_R - 1) / SX_R);
    hipLaunchKernelGGL(smooth_x_passes_kernel, grid, dim3(256), lds, s, in, out, nx, ny, nrow, ps, periodic_x);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_current_reduce(const zk_reduce_args *a, zk_stream_t s)
{
    const int nodes = (a->t.nx + 3) * (a->t.ny + 3);
    hipLaunchKernelGGL(reduce_kernel, dim3((nodes + 255) / 256), dim3(256), 0, s, *a);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_smooth(const float *in, float *out, int nx, int ny, int nrow, int axis, float sa, float sb,
                                int periodic_x, int open_y, zk_stream_t s)
{
    const int nodes = (nx + 3) * (ny + 3);
    hipLaunchKernelGGL(smooth_kernel, dim3((nodes + 255) / 256), dim3(256), 0, s, in, out, nx, ny, nrow, axis, sa, sb,
                       periodic_x, open_y);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_smooth_xy(const float *in, float *out, int nx, int ny, int nrow, float sax, float sbx, float say,
                            float sby, int periodic_x, zk_stream_t s)
{
    const int nodes = (nx + 3) * (ny + 3);
    hipLaunchKernelGGL(smooth_xy_kernel, dim3((nodes + 255) / 256), dim3(256), 0, s, in, out, nx, ny, nrow, sax, sbx,
                       say, sby, periodic_x);
    ZK_LAUNCH_CHECK();
}
