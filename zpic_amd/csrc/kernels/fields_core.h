/*
 * fields_core.h - device code of the field kernels, shared by fields.hip
 * and the batched kernels of batch.hip.
 */
#ifndef ZK_FIELDS_CORE_H
#define ZK_FIELDS_CORE_H

#include "common.h"

namespace zk {


constexpr int BX = ZK_YEE_BX, BY = ZK_YEE_BY; /* output nodes per tile */
/* rows staged below the first output row: two, so that the whole step of E
 * is known one row below the outputs (the upper Mur boundary needs it) */
constexpr int OY = 2;
constexpr int LX = BX + 3, LY = BY + OY + 2;

typedef zk_yee_k YeeK; /* (dt/2)/dx, (dt/2)/dy, dt/dx, dt/dy, dt, Mur coefficient along y */

/* One BX x BY tile of output nodes, tile (tx, ty) of the buffer.  The
 * half and whole steps update the staged B and E in place: a node's new
 * value depends on its own old value and on neighbours of the other field,
 * which the previous step finished before the barrier (three LDS arrays
 * instead of five).
 *
 * open_y: first-order Mur absorbing boundaries at the y edges, after the
 * whole step of E and before the second half step of B, as the reference's
 * one-dimensional code applies them along x (em1d/emf.c:379-406, 548-561):
 * the tangential components Ex, Ez of the guard rows j = -1 and j = ny
 * become E_in(n) + S (E_in(n+1) - E_guard(n)) with E_in the adjacent
 * interior row (0, ny - 1) and S = (dt - dy) / (dt + dy); the old values
 * are read from the input buffer E, which the kernel does not write. */
/* shift: the step ends in a window move (em2d/emf.c:646-676): every node
 * is written one column to the left, the leftmost guard column's values
 * are dropped and raw columns nx - 1 .. nx + 1 are zeroed, as the
 * reference's shift after the step leaves them (the y guards, copied after
 * this kernel, commute with the shift). */
__device__ inline void yee_tile(const float *__restrict__ E, const float *__restrict__ B,
                                const float *__restrict__ J, float *__restrict__ Eo, float *__restrict__ Bo, int nx,
                                int ny, int nrow, const YeeK &k, const int tx, const int ty, const int open_y = 0,
                                const int shift = 0)
{
    __shared__ float e[3][LY][LX], b[3][LY][LX], jj[3][LY][LX];
    const int bw = nx + 3, bh = ny + 3;
    const int bi0 = tx * BX, bj0 = ty * BY;
    const int tid = threadIdx.x;
    /* local (u, v) <-> buffer node (bi0 - 1 + u, bj0 - OY + v), raw node
     * (bi0 - 2 + u, bj0 - OY - 1 + v) */
    for (int m = tid; m < LX * LY; m += blockDim.x) {
        const int u = m % LX, v = m / LX;
        const int bi = bi0 - 1 + u, bj = bj0 - OY + v;
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
        const int i = bi0 - 2 + u, j = bj0 - OY - 1 + v; /* raw indices */
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
        const int i = bi0 - 2 + u, j = bj0 - OY - 1 + v;
        if (i >= 0 && i <= nx + 1 && j >= 0 && j <= ny + 1) {
            e[0][v][u] += (+k.dty * (b[2][v][u] - b[2][v - 1][u])) - k.dt * jj[0][v][u];
            e[1][v][u] += (-k.dtx * (b[2][v][u] - b[2][v][u - 1])) - k.dt * jj[1][v][u];
            e[2][v][u] += (+k.dtx * (b[1][v][u] - b[1][v][u - 1]) - k.dty * (b[0][v][u] - b[0][v - 1][u])) -
                          k.dt * jj[2][v][u];
        }
    }
    __syncthreads();
    if (open_y) {
        /* Mur rows among this tile's outputs (raw rows bj0 - 1 .. bj0 + BY - 2) */
        for (int s = 0; s < 2; s++) {
            const int jb = s ? ny : -1, jin = s ? ny - 1 : 0;
            if (jb < bj0 - 1 || jb >= bj0 - 1 + BY) continue;
            const int v = jb - (bj0 - OY - 1), vin = v + (jin - jb);
            for (int u = 1 + tid; u < BX + 1; u += blockDim.x) {
                const int bi = bi0 - 1 + u;
                if (bi >= bw) continue;
                const size_t gb = 3 * ((size_t)bi + (size_t)(jb + 1) * nrow);
                const size_t gi = 3 * ((size_t)bi + (size_t)(jin + 1) * nrow);
                e[0][v][u] = E[gi + 0] + k.mur_y * (e[0][vin][u] - E[gb + 0]);
                e[2][v][u] = E[gi + 2] + k.mur_y * (e[2][vin][u] - E[gb + 2]);
            }
        }
        __syncthreads();
    }
    /* B2 = B1 + dt/2 * (-curl E1) and outputs on u in [1, BX+1), v in [OY, BY+OY) */
    for (int m = tid; m < BX * BY; m += blockDim.x) {
        const int u = 1 + m % BX, v = OY + m / BX;
        const int bi = bi0 - 1 + u, bj = bj0 - OY + v;
        if (bi >= bw || bj >= bh) continue;
        const int i = bi - 1, j = bj - 1;
        float bx = b[0][v][u], by = b[1][v][u], bz = b[2][v][u];
        if (i >= -1 && i <= nx && j >= -1 && j <= ny) {
            bx += (-k.dty2 * (e[2][v + 1][u] - e[2][v][u]));
            by += (k.dtx2 * (e[2][v][u + 1] - e[2][v][u]));
            bz += (-k.dtx2 * (e[1][v][u + 1] - e[1][v][u]) + k.dty2 * (e[0][v + 1][u] - e[0][v][u]));
        }
        if (shift) {
            /* buffer column bi goes to bi - 1 (raw i - 1); raw columns from
             * nx - 1 on are zero, the last one written by the last node */
            if (bi == 0) continue;
            const size_t g = 3 * ((size_t)(bi - 1) + (size_t)bj * nrow);
            const bool zero = bi - 1 >= nx;
            Bo[g + 0] = zero ? 0.0f : bx;
            Bo[g + 1] = zero ? 0.0f : by;
            Bo[g + 2] = zero ? 0.0f : bz;
            Eo[g + 0] = zero ? 0.0f : e[0][v][u];
            Eo[g + 1] = zero ? 0.0f : e[1][v][u];
            Eo[g + 2] = zero ? 0.0f : e[2][v][u];
            if (bi == bw - 1)
                for (int c = 0; c < 3; c++) Bo[g + 3 + c] = Eo[g + 3 + c] = 0.0f;
            continue;
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
 * and y (unless the y boundaries are open).  One thread per guard node. */
__device__ inline void guard_node(float *E, float *B, int nx, int ny, int nrow, int periodic_x, const int k,
                                  const int open_y = 0)
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
    const int cj = open_y ? j : zk_wrap(j, ny);
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
