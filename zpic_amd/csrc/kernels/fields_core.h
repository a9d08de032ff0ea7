/*
 * fields_core.h - device code of the field kernels, shared by fields.hip
 * and the batched kernels of batch.hip.
 */
#ifndef ZK_FIELDS_CORE_H
#define ZK_FIELDS_CORE_H

#include "common.h"

namespace zk {


constexpr int BX = ZK_YEE_BX, BY = ZK_YEE_BY; /* output nodes per tile */
constexpr int LX = BX + 3, LY = BY + 3;

typedef zk_yee_k YeeK; /* (dt/2)/dx, (dt/2)/dy, dt/dx, dt/dy, dt */

/* One BX x BY tile of output nodes, tile (tx, ty) of the buffer.  The
 * half and whole steps update the staged B and E in place: a node's new
 * value depends on its own old value and on neighbours of the other field,
 * which the previous step finished before the barrier (three LDS arrays,
 * 24 KB, instead of five). */
__device__ inline void yee_tile(const float *__restrict__ E, const float *__restrict__ B,
                                const float *__restrict__ J, float *__restrict__ Eo, float *__restrict__ Bo, int nx,
                                int ny, int nrow, const YeeK &k, const int tx, const int ty)
{
    __shared__ float e[3][LY][LX], b[3][LY][LX], jj[3][LY][LX];
    const int bw = nx + 3, bh = ny + 3;
    const int bi0 = tx * BX, bj0 = ty * BY;
    const int tid = threadIdx.x;
    /* local (u, v) <-> buffer node (bi0 - 1 + u, bj0 - 1 + v) */
    for (int m = tid; m < LX * LY; m += blockDim.x) {
        const int u = m % LX, v = m / LX;
        const int bi = bi0 - 1 + u, bj = bj0 - 1 + v;
        const bool in = bi >= 0 && bi < bw && bj >= 0 && bj < bh;
        const size_t g = 3 * ((size_t)bi + (size_t)bj * nrow);
        for (int c = 0; c < 3; c++) {
            e[c][v][u] = in ? E[g + c] : 0.0f;
            b[c][v][u] = in ? B[g + c] : 0.0f;
            jj[c][v][u] = in ? J[g + c] : 0.0f;
        }
    }
    __syncthreads();
    /* B1 = B + dt/2 * (-curl E) on u in [0, LX-1), v in [0, LY-1) */
    for (int m = tid; m < (LX - 1) * (LY - 1); m += blockDim.x) {
        const int u = m % (LX - 1), v = m / (LX - 1);
        const int i = bi0 - 2 + u, j = bj0 - 2 + v; /* raw indices */
        if (i >= -1 && i <= nx && j >= -1 && j <= ny) {
            b[0][v][u] += (-k.dty2 * (e[2][v + 1][u] - e[2][v][u]));
            b[1][v][u] += (k.dtx2 * (e[2][v][u + 1] - e[2][v][u]));
            b[2][v][u] += (-k.dtx2 * (e[1][v][u + 1] - e[1][v][u]) + k.dty2 * (e[0][v + 1][u] - e[0][v][u]));
        }
    }
    __syncthreads();
    /* E1 = E + dt * (curl B1 - J) on u in [1, LX-1), v in [1, LY-1) */
    for (int m = tid; m < (LX - 2) * (LY - 2); m += blockDim.x) {
        const int u = 1 + m % (LX - 2), v = 1 + m / (LX - 2);
        const int i = bi0 - 2 + u, j = bj0 - 2 + v;
        if (i >= 0 && i <= nx + 1 && j >= 0 && j <= ny + 1) {
            e[0][v][u] += (+k.dty * (b[2][v][u] - b[2][v - 1][u])) - k.dt * jj[0][v][u];
            e[1][v][u] += (-k.dtx * (b[2][v][u] - b[2][v][u - 1])) - k.dt * jj[1][v][u];
            e[2][v][u] += (+k.dtx * (b[1][v][u] - b[1][v][u - 1]) - k.dty * (b[0][v][u] - b[0][v - 1][u])) -
                          k.dt * jj[2][v][u];
        }
    }
    __syncthreads();
    /* B2 = B1 + dt/2 * (-curl E1) and outputs on u in [1, BX+1), v in [1, BY+1) */
    for (int m = tid; m < BX * BY; m += blockDim.x) {
        const int u = 1 + m % BX, v = 1 + m / BX;
        const int bi = bi0 - 1 + u, bj = bj0 - 1 + v;
        if (bi >= bw || bj >= bh) continue;
        const int i = bi - 1, j = bj - 1;
        float bx = b[0][v][u], by = b[1][v][u], bz = b[2][v][u];
        if (i >= -1 && i <= nx && j >= -1 && j <= ny) {
            bx += (-k.dty2 * (e[2][v + 1][u] - e[2][v][u]));
            by += (k.dtx2 * (e[2][v][u + 1] - e[2][v][u]));
            bz += (-k.dtx2 * (e[1][v][u + 1] - e[1][v][u]) + k.dty2 * (e[0][v + 1][u] - e[0][v][u]));
        }
        const size_t g = 3 * ((size_t)bi + (size_t)bj * nrow);
        Bo[g + 0] = bx;
        Bo[g + 1] = by;
        Bo[g + 2] = bz;
        Eo[g + 0] = e[0][v][u];
        Eo[g + 1] = e[1][v][u];
        Eo[g + 2] = e[2][v][u];
    }
}

/* Guard nodes take the value of their image: x (unless the window moves)
 * and y.  One thread per guard node. */
__device__ inline void guard_node(float *E, float *B, int nx, int ny, int nrow, int periodic_x, const int k)
{
    const int bw = nx + 3;
    const int nrows_g = 3 * bw;       /* guard rows: j = -1, ny, ny+1 */
    const int ncols_g = 3 * ny;       /* guard columns of interior rows */
    int i, j;
    if (k < nrows_g) {
        i = k % bw - 1;
        const int r = k / bw;
        j = (r == 0) ? -1 : ny + r - 1;
    } else if (k < nrows_g + ncols_g) {
        const int q = k - nrows_g;
        const int c = q % 3;
        j = q / 3;
        i = (c == 0) ? -1 : nx + c - 1;
    } else {
        return;
    }
    const int ci = periodic_x ? zk_wrap(i, nx) : i;
    const int cj = zk_wrap(j, ny);
    if (ci == i && cj == j) return;
    const size_t d = 3 * ((size_t)(i + 1) + (size_t)(j + 1) * nrow);
    const size_t s = 3 * ((size_t)(ci + 1) + (size_t)(cj + 1) * nrow);
    for (int c = 0; c < 3; c++) {
        E[d + c] = E[s + c];
        B[d + c] = B[s + c];
    }
}

} // namespace zk

#endif
