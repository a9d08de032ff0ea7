/*
 * current_core.h - device code of the current reduction and filters, shared
 * by current.hip and the batched kernels of batch.hip.
 */
#ifndef ZK_CURRENT_CORE_H
#define ZK_CURRENT_CORE_H

#include "common.h"

namespace zk {


__device__ inline int floor_div(int a, int b)
{
    int q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

__device__ inline void reduce_node(const zk_reduce_args &a, const int node)
{
    const zk_tiles &t = a.t;
    const int bw = t.nx + 3, bh = t.ny + 3;
    if (node >= bw * bh) return;
    const int i = node % bw - 1, j = node / bw - 1;
    const int ci = a.periodic_x ? zk_wrap(i, t.nx) : i;
    const int cj = a.open_y ? j : zk_wrap(j, t.ny);
    const int n = t.WX * t.WY;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int ky = -1; ky <= 1; ky++) {
        if (a.open_y && ky != 0) continue;
        const int ry = cj + ky * t.ny;
        if (ry < -1 || ry > t.ny + 1) continue;
        const int ty_lo = max(0, floor_div(ry + 1 + t.H - t.WY, t.TY) + 1);
        const int ty_hi = min(t.nty - 1, floor_div(ry + 1 + t.H, t.TY));
        for (int kx = -1; kx <= 1; kx++) {
            if (!a.periodic_x && kx != 0) continue;
            const int rx = ci + kx * t.nx;
            if (rx < -1 || rx > t.nx + 1) continue;
            const int tx_lo = max(0, floor_div(rx + 1 + t.H - t.WX, t.TX) + 1);
            const int tx_hi = min(t.ntx - 1, floor_div(rx + 1 + t.H, t.TX));
            for (int ty = ty_lo; a.region && ty <= ty_hi; ty++) {
                const int v = ry - (ty * t.TY - 1 - t.H);
                for (int tx = tx_lo; tx <= tx_hi; tx++) {
                    const int u = rx - (tx * t.TX - 1 - t.H);
                    const float *w = a.region + (size_t)(tx + ty * t.ntx) * 3 * n + u + v * t.WX;
                    sx += w[0];
                    sy += w[n];
                    sz += w[2 * n];
                }
            }
            const long long *o = a.Jov + 3 * ((rx + 1) + (ry + 1) * a.nrow);
            sx += (float)o[0] * a.jov_unscale;
            sy += (float)o[1] * a.jov_unscale;
            sz += (float)o[2] * a.jov_unscale;
        }
    }
    float *out = a.J + 3 * (size_t)node;
    out[0] = sx;
    out[1] = sy;
    out[2] = sz;
}

__device__ inline void stencil(float sa, float sb, const float *l, const float *m, const float *u, float *o)
{
    o[0] = sa * l[0] + sb * m[0] + sa * u[0];
    o[1] = sa * l[1] + sb * m[1] + sa * u[1];
    o[2] = sa * l[2] + sb * m[2] + sa * u[2];
}

__device__ inline void smooth_node(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                   int nrow, int axis, float sa, float sb, int periodic_x, const int node,
                                   const int open_y = 0)
{
    const int bw = nx + 3, bh = ny + 3;
    if (node >= bw * bh) return;
    const int i = node % bw - 1, j = node / bw - 1;
#define AT(a, b) (in + 3 * ((size_t)((a) + 1) + (size_t)((b) + 1) * nrow))
    float *o = out + 3 * (size_t)node;
    const bool ix_in = (i >= 0 && i < nx), jy_in = (j >= 0 && j < ny);
    if (axis == 0) {
        /* kernel_x: interior rows only; x guards of those rows are copies
         * unless the window moves */
        if (jy_in && (ix_in || periodic_x)) {
            const int c = ix_in ? i : zk_wrap(i, nx);
            stencil(sa, sb, AT(c - 1, j), AT(c, j), AT(c + 1, j), o);
            return;
        }
    } else if (jy_in || !open_y) {
        /* kernel_y: interior nodes; then y guards of every column copy the
         * final value of their image row (with open y boundaries the guard
         * rows keep their values, below) */
        const int c = jy_in ? j : zk_wrap(j, ny);
        if (ix_in) {
            stencil(sa, sb, AT(i, c - 1), AT(i, c), AT(i, c + 1), o);
            return;
        }
        const float *s = AT(i, c);
        o[0] = s[0];
        o[1] = s[1];
        o[2] = s[2];
        return;
    }
    const float *s = AT(i, j);
    o[0] = s[0];
    o[1] = s[1];
    o[2] = s[2];
#undef AT
}

/* One x pass and then one y pass (kernel_x, kernel_y, em2d/current.c:316-413)
 * in one launch, periodic y: node (i, j) of the y pass reads the x pass's
 * values at (i, c - 1 .. c + 1), c the image row of j, and each of those is
 * formed here from `in` with smooth_node's x rules (guard rows copied,
 * guard columns from their images unless the window moves) — the same
 * operations in the same order as the two launches, without writing and
 * re-reading the intermediate grid. */
__device__ inline void x_pass_value(const float *__restrict__ in, int nx, int ny, int nrow, float sa, float sb,
                                    int periodic_x, int i, int r, float *o)
{
#define AT(a, b) (in + 3 * ((size_t)((a) + 1) + (size_t)((b) + 1) * nrow))
    if (r >= 0 && r < ny && ((i >= 0 && i < nx) || periodic_x)) {
        const int c = (i >= 0 && i < nx) ? i : zk_wrap(i, nx);
        stencil(sa, sb, AT(c - 1, r), AT(c, r), AT(c + 1, r), o);
    } else {
        const float *s = AT(i, r);
        o[0] = s[0];
        o[1] = s[1];
        o[2] = s[2];
    }
#undef AT
}

__device__ inline void smooth_xy_node(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                      int nrow, float sax, float sbx, float say, float sby, int periodic_x,
                                      const int node)
{
    const int bw = nx + 3, bh = ny + 3;
    if (node >= bw * bh) return;
    const int i = node % bw - 1, j = node / bw - 1;
    const int c = (j >= 0 && j < ny) ? j : zk_wrap(j, ny);
    float *o = out + 3 * (size_t)node;
    float m[3];
    x_pass_value(in, nx, ny, nrow, sax, sbx, periodic_x, i, c, m);
    if (i >= 0 && i < nx) {
        float l[3], u[3];
        x_pass_value(in, nx, ny, nrow, sax, sbx, periodic_x, i, c - 1, l);
        x_pass_value(in, nx, ny, nrow, sax, sbx, periodic_x, i, c + 1, u);
        stencil(say, sby, l, m, u, o);
    } else {
        o[0] = m[0];
        o[1] = m[1];
        o[2] = m[2];
    }
}

} // namespace zk

#endif
