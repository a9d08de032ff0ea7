/*
 * adjoint.hip - reverse-mode (adjoint) kernels of the em2d time step.
 *
 * The reference has no differentiation; these kernels are the exact
 * transposes of the forward step of push.hip, current.hip and fields.hip
 * for periodic grids without current filter or external fields:
 *
 *   forward step n:  particles (X^n, u^n), fields (E^n, B^n)
 *     gather E,B at X^n -> Boris -> u^{n+1} -> displacement and v_z
 *     -> zigzag deposit -> raw current -> periodic fold -> J^n
 *     -> Yee (B half, E, B half) -> guard copies -> (E^{n+1}, B^{n+1})
 *
 * Given the adjoints of (X^{n+1}, u^{n+1}, E^{n+1}, B^{n+1}) the reverse
 * step produces those of (X^n, u^n, E^n, B^n).  Positions are continuous
 * (cell index + offset), so cell crossings and the periodic wrap pass
 * adjoints through unchanged; the discrete choices of the forward step
 * (interpolation stagger, which cell edges are crossed) are taken from a
 * recomputation of the forward step from the stored state.
 *
 * Field adjoints are gathered per node (no atomics); particle adjoints use
 * the tile layout of the forward push: each tile stages its E,B window and
 * its window of the current adjoint in LDS, accumulates the adjoint of the
 * gathered fields in LDS windows, and stores them per tile; a node-parallel
 * pass adds the windows into the field adjoints.  The windows are summed in
 * 64-bit fixed point (integer LDS atomics; a float LDS atomic costs about
 * three integer ones here) with a power-of-two scale set from the running
 * maximum of the summed quantities: 2^8 above it still fits, and a
 * workgroup that meets a larger one sums its share again in float.  A
 * tile's particles may be
 * split over several workgroups (each with its own windows) so that small
 * grids still fill the GPU.  Particles outside their tile window are
 * listed and reversed by the same workgroup after the others, through the
 * global grids with atomics.
 */
#include "common.h"
#include "zk_adjoint.h"

namespace {

/* ------------------------------------------------------------------------
 * field adjoints
 * ---------------------------------------------------------------------- */

struct Grid {
    int nx, ny, nrow;
    __device__ bool in_buf(int i, int j) const { return i >= -1 && i <= nx + 1 && j >= -1 && j <= ny + 1; }
    __device__ size_t at(int i, int j) const { return 3 * ((size_t)(i + 1) + (size_t)(j + 1) * nrow); }
};

__device__ inline int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* adjoint of the guard copies: guard adjoints move to their images */
__global__ void adj_guards_kernel(const float *__restrict__ lEo, const float *__restrict__ lBo, float *__restrict__ aE,
                                  float *__restrict__ aB, Grid g)
{
    const int bw = g.nx + 3, bh = g.ny + 3;
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= bw * bh) return;
    const int i = node % bw - 1, j = node / bw - 1;
    float e[3] = {0.f, 0.f, 0.f}, b[3] = {0.f, 0.f, 0.f};
    if (i >= 0 && i < g.nx && j >= 0 && j < g.ny) {
        /* every buffer node whose image is (i, j): offsets 0, +-n */
        for (int dj = -1; dj <= 1; dj++) {
            const int jj = j + dj * g.ny;
            if (jj < -1 || jj > g.ny + 1) continue;
            for (int di = -1; di <= 1; di++) {
                const int ii = i + di * g.nx;
                if (ii < -1 || ii > g.nx + 1) continue;
                const size_t s = g.at(ii, jj);
                for (int c = 0; c < 3; c++) {
                    e[c] += lEo[s + c];
                    b[c] += lBo[s + c];
                }
            }
        }
    }
    const size_t d = g.at(i, j);
    for (int c = 0; c < 3; c++) {
        aE[d + c] = e[c];
        aB[d + c] = b[c];
    }
}

/* B node (i, j) is updated by the half steps when i, j in [-1, n] */
__device__ inline bool in_rb(const Grid &g, int i, int j) { return i >= -1 && i <= g.nx && j >= -1 && j <= g.ny; }
/* E node (i, j) is updated by the full step when i, j in [0, n + 1] */
__device__ inline bool in_re(const Grid &g, int i, int j) { return i >= 0 && i <= g.nx + 1 && j >= 0 && j <= g.ny + 1; }

/* E adjoint += transpose of the B half step (B += dt/2 (-curl E)) applied
 * to the B adjoint aB, at E node (i, j). */
__device__ inline void curl_e_t(const Grid &g, const float *aB, int i, int j, float dtx2, float dty2, float out[3])
{
    float x = 0.f, y = 0.f, z = 0.f;
    if (in_rb(g, i, j)) {
        const size_t s = g.at(i, j);
        /* Bx(i,j) += -dty2 (Ez(i,j+1) - Ez(i,j)); By(i,j) += dtx2 (Ez(i+1,j) - Ez(i,j));
         * Bz(i,j) += -dtx2 (Ey(i+1,j) - Ey(i,j)) + dty2 (Ex(i,j+1) - Ex(i,j)) */
        z += dty2 * aB[s + 0] - dtx2 * aB[s + 1];
        y += dtx2 * aB[s + 2];
        x += -dty2 * aB[s + 2];
    }
    if (in_rb(g, i, j - 1)) {
        const size_t s = g.at(i, j - 1);
        z += -dty2 * aB[s + 0];
        x += dty2 * aB[s + 2];
    }
    if (in_rb(g, i - 1, j)) {
        const size_t s = g.at(i - 1, j);
        z += dtx2 * aB[s + 1];
        y += -dtx2 * aB[s + 2];
    }
    out[0] = x;
    out[1] = y;
    out[2] = z;
}

/* second half step: aE += (dB/dE)^T aB (aB is the adjoint of the output B) */
__global__ void adj_bhalf2_kernel(float *__restrict__ aE, const float *__restrict__ aB, Grid g, float dtx2, float dty2)
{
    const int bw = g.nx + 3, bh = g.ny + 3;
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= bw * bh) return;
    const int i = node % bw - 1, j = node / bw - 1;
    float t[3];
    curl_e_t(g, aB, i, j, dtx2, dty2, t);
    const size_t d = g.at(i, j);
    aE[d + 0] += t[0];
    aE[d + 1] += t[1];
    aE[d + 2] += t[2];
}

/* full E step: lE = aE; aB += (dE/dB)^T aE; lJ = -dt aE on its range */
__global__ void adj_efull_kernel(const float *__restrict__ aE, float *__restrict__ aB, float *__restrict__ lE,
                                 float *__restrict__ lJ, Grid g, float dtx, float dty, float dt)
{
    const int bw = g.nx + 3, bh = g.ny + 3;
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= bw * bh) return;
    const int i = node % bw - 1, j = node / bw - 1;
    const size_t d = g.at(i, j);
    /* Ex(i,j) += dty (Bz(i,j) - Bz(i,j-1)); Ey(i,j) += -dtx (Bz(i,j) - Bz(i-1,j));
     * Ez(i,j) += dtx (By(i,j) - By(i-1,j)) - dty (Bx(i,j) - Bx(i,j-1)); E -= dt J */
    float bx = 0.f, by = 0.f, bz = 0.f;
    if (in_re(g, i, j)) {
        bz += dty * aE[d + 0] - dtx * aE[d + 1];
        by += dtx * aE[d + 2];
        bx += -dty * aE[d + 2];
    }
    if (in_re(g, i, j + 1) && j + 1 <= g.ny + 1) {
        const size_t s = g.at(i, j + 1);
        bz += -dty * aE[s + 0];
        bx += dty * aE[s + 2];
    }
    if (in_re(g, i + 1, j) && i + 1 <= g.nx + 1) {
        const size_t s = g.at(i + 1, j);
        bz += dtx * aE[s + 1];
        by += -dtx * aE[s + 2];
    }
    aB[d + 0] += bx;
    aB[d + 1] += by;
    aB[d + 2] += bz;
    const bool r = in_re(g, i, j);
    for (int c = 0; c < 3; c++) {
        lE[d + c] = aE[d + c];
        lJ[d + c] = r ? -dt * aE[d + c] : 0.0f;
    }
}

/* first half step: lB = aB; lE += (dB/dE)^T aB */
__global__ void adj_bhalf1_kernel(float *__restrict__ lE, const float *__restrict__ aB, float *__restrict__ lB, Grid g,
                                  float dtx2, float dty2)
{
    const int bw = g.nx + 3, bh = g.ny + 3;
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= bw * bh) return;
    const int i = node % bw - 1, j = node / bw - 1;
    float t[3];
    curl_e_t(g, aB, i, j, dtx2, dty2, t);
    const size_t d = g.at(i, j);
    for (int c = 0; c < 3; c++) {
        lE[d + c] += t[c];
        lB[d + c] = aB[d + c];
    }
}

/* adjoint of the periodic fold of the raw current: every raw node takes
 * the summed adjoint of all buffer nodes sharing its image */
__global__ void adj_images_kernel(const float *__restrict__ lJ, float *__restrict__ lam, Grid g)
{
    const int bw = g.nx + 3, bh = g.ny + 3;
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= bw * bh) return;
    const int i = node % bw - 1, j = node / bw - 1;
    const int ci = wrapi(i, g.nx), cj = wrapi(j, g.ny);
    float s[3] = {0.f, 0.f, 0.f};
    for (int dj = -1; dj <= 1; dj++) {
        const int jj = cj + dj * g.ny;
        if (jj < -1 || jj > g.ny + 1) continue;
        for (int di = -1; di <= 1; di++) {
            const int ii = ci + di * g.nx;
            if (ii < -1 || ii > g.nx + 1) continue;
            const size_t q = g.at(ii, jj);
            s[0] += lJ[q + 0];
            s[1] += lJ[q + 1];
            s[2] += lJ[q + 2];
        }
    }
    const size_t d = g.at(i, j);
    lam[d + 0] = s[0];
    lam[d + 1] = s[1];
    lam[d + 2] = s[2];
}

/* loss term: lB(interior) += scale * B(interior) */
__global__ void adj_loss_kernel(float *__restrict__ lB, const float *__restrict__ B, Grid g, float scale)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= g.nx * g.ny) return;
    const int i = node % g.nx, j = node / g.nx;
    const size_t d = g.at(i, j);
    lB[d + 0] += scale * B[d + 0];
    lB[d + 1] += scale * B[d + 1];
    lB[d + 2] += scale * B[d + 2];
}

/* ------------------------------------------------------------------------
 * particle adjoint
 * ---------------------------------------------------------------------- */

/* Six field components and three current-adjoint components of one
 * index space (LDS window or guard-inclusive global grid). */
struct LdsSrc {
    const float *f; /* 6 n: Ex Ey Ez Bx By Bz */
    const float *l; /* 3 n: adjoint of the raw current */
    float *acc;     /* 6 n: adjoint of the gathered fields */
    int w, n;
    __device__ float fld(int c, int a, int b) const { return f[c * n + a + b * w]; }
    __device__ float lam(int c, int a, int b) const { return l[c * n + a + b * w]; }
    __device__ void add(int c, int a, int b, float v) const { atomicAdd(&acc[c * n + a + b * w], v); }
};

/* the same windows with the gathered-field adjoints in fixed point */
struct LdsSrcFx {
    const float *f;
    const float *l;
    unsigned long long *acc;
    int w, n;
    float scale;
    __device__ float fld(int c, int a, int b) const { return f[c * n + a + b * w]; }
    __device__ float lam(int c, int a, int b) const { return l[c * n + a + b * w]; }
    __device__ void add(int c, int a, int b, float v) const { atomicAdd(&acc[c * n + a + b * w], zk_fx(v, scale)); }
};

struct GlobalSrc {
    const float *E, *B, *L;
    float *aE, *aB;
    int nrow;
    __device__ size_t at(int a, int b) const { return 3 * ((size_t)a + (size_t)b * nrow); }
    __device__ float fld(int c, int a, int b) const { return c < 3 ? E[at(a, b) + c] : B[at(a, b) + c - 3]; }
    __device__ float lam(int c, int a, int b) const { return L[at(a, b) + c]; }
    __device__ void add(int c, int a, int b, float v) const
    {
        if (c < 3) unsafeAtomicAdd(&aE[at(a, b) + c], v);
        else unsafeAtomicAdd(&aB[at(a, b) + c - 3], v);
    }
};

/* Interpolation of component c from base node (ia, jb) with weights
 * (wa, wb) along x and y; returns the value and its derivatives. */
template <class S>
__device__ inline float interp(const S &src, int c, int ia, int jb, float wa, float wb, float &da, float &db)
{
    const float f00 = src.fld(c, ia, jb), f10 = src.fld(c, ia + 1, jb);
    const float f01 = src.fld(c, ia, jb + 1), f11 = src.fld(c, ia + 1, jb + 1);
    const float lo = f00 * (1.0f - wa) + f10 * wa, hi = f01 * (1.0f - wa) + f11 * wa;
    da = (f10 - f00) * (1.0f - wb) + (f11 - f01) * wb;
    db = hi - lo;
    return lo * (1.0f - wb) + hi * wb;
}

template <class S>
__device__ inline void interp_scatter(const S &src, int c, int ia, int jb, float wa, float wb, float lam)
{
    if (lam == 0.0f) return;
    src.add(c, ia, jb, lam * (1.0f - wa) * (1.0f - wb));
    src.add(c, ia + 1, jb, lam * wa * (1.0f - wb));
    src.add(c, ia, jb + 1, lam * (1.0f - wa) * wb);
    src.add(c, ia + 1, jb + 1, lam * wa * wb);
}

/* staggering of the six components (em2d/particles.c:1029-1071):
 * x uses (i, wx) or (ih, wxh), y uses (j, wy) or (jh, wyh) */
struct Stagger {
    int i, j, ih, jh;
    float wx, wy, wxh, wyh;
    __device__ Stagger(int ci, int cj, float x, float y)
    {
        i = ci;
        j = cj;
        wx = x;
        wy = y;
        ih = ci + ((x < 0.5f) ? -1 : 0);
        jh = cj + ((y < 0.5f) ? -1 : 0);
        wxh = x + ((x < 0.5f) ? 0.5f : -0.5f);
        wyh = y + ((y < 0.5f) ? 0.5f : -0.5f);
    }
    /* component c: Ex (ih,j), Ey (i,jh), Ez (i,j), Bx (i,jh), By (ih,j), Bz (ih,jh) */
    __device__ void of(int c, int &ia, int &jb, float &wa, float &wb) const
    {
        const bool hx = (c == 0 || c == 4 || c == 5);
        const bool hy = (c == 1 || c == 3 || c == 5);
        ia = hx ? ih : i;
        wa = hx ? wxh : wx;
        jb = hy ? jh : j;
        wb = hy ? wyh : wy;
    }
};

struct SegF {
    float x0, x1, y0, y1, dx, dy, q;
    int ix, iy;
};

/* adjoint of one segment's deposit (deposit_seg of push_core.h) with
 * respect to its seven parameters */
template <class S>
__device__ inline SegF seg_adj(const S &src, const SegF &g, float qnx, float qny)
{
    const float ax0 = 1.0f - g.x0, bx0 = g.x0, ax1 = 1.0f - g.x1, bx1 = g.x1;
    const float ay0 = 1.0f - g.y0, by0 = g.y0, ay1 = 1.0f - g.y1, by1 = g.y1;
    const float jx = qnx * g.dx, jy = qny * g.dy;
    const float wy0 = 0.5f * (ay0 + ay1), wy1 = 0.5f * (by0 + by1);
    const float wx0 = 0.5f * (ax0 + ax1), wx1 = 0.5f * (bx0 + bx1);
    const float lx00 = src.lam(0, g.ix, g.iy), lx01 = src.lam(0, g.ix, g.iy + 1);
    const float ly00 = src.lam(1, g.ix, g.iy), ly10 = src.lam(1, g.ix + 1, g.iy);
    const float lz00 = src.lam(2, g.ix, g.iy), lz10 = src.lam(2, g.ix + 1, g.iy);
    const float lz01 = src.lam(2, g.ix, g.iy + 1), lz11 = src.lam(2, g.ix + 1, g.iy + 1);
    const float s00 = ax0 * ay0 + ax1 * ay1 + (ax0 * ay1 - ax1 * ay0) * 0.5f;
    const float s10 = bx0 * ay0 + bx1 * ay1 + (bx0 * ay1 - bx1 * ay0) * 0.5f;
    const float s01 = ax0 * by0 + ax1 * by1 + (ax0 * by1 - ax1 * by0) * 0.5f;
    const float s11 = bx0 * by0 + bx1 * by1 + (bx0 * by1 - bx1 * by0) * 0.5f;
    /* derivatives of sum(lz * s) with respect to the a/b weights */
    const float d_ax0 = lz00 * (ay0 + 0.5f * ay1) + lz01 * (by0 + 0.5f * by1);
    const float d_bx0 = lz10 * (ay0 + 0.5f * ay1) + lz11 * (by0 + 0.5f * by1);
    const float d_ax1 = lz00 * (ay1 - 0.5f * ay0) + lz01 * (by1 - 0.5f * by0);
    const float d_bx1 = lz10 * (ay1 - 0.5f * ay0) + lz11 * (by1 - 0.5f * by0);
    const float d_ay0 = lz00 * (ax0 - 0.5f * ax1) + lz10 * (bx0 - 0.5f * bx1);
    const float d_by0 = lz01 * (ax0 - 0.5f * ax1) + lz11 * (bx0 - 0.5f * bx1);
    const float d_ay1 = lz00 * (ax1 + 0.5f * ax0) + lz10 * (bx1 + 0.5f * bx0);
    const float d_by1 = lz01 * (ax1 + 0.5f * ax0) + lz11 * (bx1 + 0.5f * bx0);
    SegF a;
    a.dx = qnx * (lx00 * wy0 + lx01 * wy1);
    a.dy = qny * (ly00 * wx0 + ly10 * wx1);
    a.q = lz00 * s00 + lz10 * s10 + lz01 * s01 + lz11 * s11;
    const float tx = 0.5f * jy * (ly10 - ly00), ty = 0.5f * jx * (lx01 - lx00);
    a.x0 = tx + g.q * (d_bx0 - d_ax0);
    a.x1 = tx + g.q * (d_bx1 - d_ax1);
    a.y0 = ty + g.q * (d_by0 - d_ay0);
    a.y1 = ty + g.q * (d_by1 - d_ay1);
    a.ix = a.iy = 0;
    return a;
}

/* forward y cut of push_core.h (ysplit) */
__device__ inline void ysplit_f(SegF &s, SegF &t, int dj, float &f)
{
    const int edge = (dj == 1);
    f = (s.y1 - edge) / s.dy;
    t.y0 = 1 - edge;
    t.y1 = s.y1 - dj;
    t.dy = s.dy * f;
    t.iy = s.iy + dj;
    const float xc = s.x0 + s.dx * (1.0f - f);
    t.x0 = xc;
    t.x1 = s.x1;
    t.dx = s.dx * f;
    t.ix = s.ix;
    t.q = s.q * f;
    s.y1 = edge;
    s.dy *= (1.0f - f);
    s.dx *= (1.0f - f);
    s.x1 = xc;
    s.q *= (1.0f - f);
}

/* adjoint of the y cut: pre is the segment before the cut, as (after the
 * cut) the segment a_s and the new segment a_t */
__device__ inline SegF ysplit_adj(const SegF &pre, float f, const SegF &a_s, const SegF &a_t)
{
    const float lxc = a_t.x0 + a_s.x1;
    const float lf = (a_t.dy - a_s.dy) * pre.dy + (a_t.dx - a_s.dx) * pre.dx + (a_t.q - a_s.q) * pre.q - lxc * pre.dx;
    SegF a;
    a.x0 = a_s.x0 + lxc;
    a.y0 = a_s.y0;
    a.dx = a_t.dx * f + (a_s.dx + lxc) * (1.0f - f);
    a.dy = a_t.dy * f + a_s.dy * (1.0f - f) - lf * f / pre.dy;
    a.x1 = a_t.x1;
    a.y1 = a_t.y1 + lf / pre.dy;
    a.q = a_t.q * f + a_s.q * (1.0f - f);
    a.ix = a.iy = 0;
    return a;
}

/* Adjoint of the zigzag deposit (deposit of push_core.h) with respect to
 * the start offsets (x0, y0), the displacement (mx, my) and qvz. */
template <class S>
__device__ inline void deposit_adj(const S &src, int ix, int iy, int di, int dj, float x0, float y0, float mx,
                                   float my, float qnx, float qny, float qvz, float &gx, float &gy, float &gmx,
                                   float &gmy, float &gq)
{
    SegF s0, s1, s2, pre;
    s0.x0 = x0;
    s0.y0 = y0;
    s0.dx = mx;
    s0.dy = my;
    s0.x1 = x0 + mx;
    s0.y1 = y0 + my;
    s0.q = qvz * 0.5f;
    s0.ix = ix;
    s0.iy = iy;
    s1 = s0;
    s2 = s0;
    float fx = 0.0f, fy = 0.0f;
    int n = 1, ycase = 0;
    if (di != 0) {
        const int edge = (di == 1);
        fx = (x0 + mx - edge) / mx;
        s1.x0 = 1 - edge;
        s1.x1 = (x0 + mx) - di;
        s1.dx = mx * fx;
        s1.ix = ix + di;
        const float yc = y0 + my * (1.0f - fx);
        s1.y0 = yc;
        s1.y1 = s0.y1;
        s1.dy = my * fx;
        s1.iy = iy;
        s1.q = s0.q * fx;
        s0.x1 = edge;
        s0.dx *= (1.0f - fx);
        s0.dy *= (1.0f - fx);
        s0.y1 = yc;
        s0.q *= (1.0f - fx);
        n = 2;
    }
    if (dj != 0) {
        if (n == 1) {
            pre = s0;
            ysplit_f(s0, s1, dj, fy);
            ycase = 1;
        } else if (s0.y1 < 0.0f || s0.y1 >= 1.0f) {
            pre = s0;
            ysplit_f(s0, s2, dj, fy);
            s1.y0 -= dj;
            s1.y1 -= dj;
            s1.iy += dj;
            ycase = 2;
        } else {
            pre = s1;
            ysplit_f(s1, s2, dj, fy);
            ycase = 3;
        }
        n++;
    }
    SegF a0 = seg_adj(src, s0, qnx, qny);
    SegF a1, a2;
    if (n > 1) a1 = seg_adj(src, s1, qnx, qny);
    if (n > 2) a2 = seg_adj(src, s2, qnx, qny);
    if (ycase == 1) {
        a0 = ysplit_adj(pre, fy, a0, a1);
    } else if (ycase == 2) {
        a0 = ysplit_adj(pre, fy, a0, a2);
    } else if (ycase == 3) {
        a1 = ysplit_adj(pre, fy, a1, a2);
    }
    SegF A = a0;
    if (di != 0) {
        /* adjoint of the x cut of the initial segment (x0, y0, mx, my,
         * x0 + mx, y0 + my, qvz / 2) into (s0, s1) */
        const float q0 = qvz * 0.5f;
        const float lyc = a1.y0 + a0.y1;
        const float lf = (a1.dx - a0.dx) * mx + (a1.dy - a0.dy) * my + (a1.q - a0.q) * q0 - lyc * my;
        A.x0 = a0.x0;
        A.y0 = a0.y0 + lyc;
        A.dx = a1.dx * fx + a0.dx * (1.0f - fx) - lf * fx / mx;
        A.dy = a1.dy * fx + (a0.dy + lyc) * (1.0f - fx);
        A.x1 = a1.x1 + lf / mx;
        A.y1 = a1.y1;
        A.q = a1.q * fx + a0.q * (1.0f - fx);
    }
    /* init: x1 = x0 + mx, y1 = y0 + my, q = qvz / 2 */
    gx = A.x0 + A.x1;
    gy = A.y0 + A.y1;
    gmx = A.dx + A.x1;
    gmy = A.dy + A.y1;
    gq = 0.5f * A.q;
}

struct AdjConsts {
    float tem, dt_dx, dt_dy, qnx, qny, q;
};

/* Reverse step of one particle: state (x, y, ux, uy, uz) at the start of
 * the step in cell (ci, cj) of the source's index space; on entry l[5]
 * holds the adjoints of (x, y, ux, uy, uz) after the step, on exit those
 * before it.  Field adjoints are added through src.add; the largest
 * magnitude of the six gathered-field adjoints (the largest contribution
 * to any window node) is returned. */
template <class S>
__device__ inline float adjoint_one(const S &src, int ci, int cj, float x, float y, float ux, float uy, float uz,
                                    float l[5], const AdjConsts &k)
{
    const Stagger st(ci, cj, x, y);
    float f[6];
#pragma unroll
    for (int c = 0; c < 6; c++) {
        int ia, jb;
        float wa, wb, da, db;
        st.of(c, ia, jb, wa, wb);
        f[c] = interp(src, c, ia, jb, wa, wb, da, db);
    }
    /* forward Boris (push_one of push_core.h) */
    const float e0 = f[0] * k.tem, e1 = f[1] * k.tem, e2 = f[2] * k.tem;
    const float ax = ux + e0, ay = uy + e1, az = uz + e2;
    const float usq = ax * ax + ay * ay + az * az;
    const float g = sqrtf(1.0f + usq);
    const float tg = k.tem / g;
    const float tx = f[3] * tg, ty = f[4] * tg, tz = f[5] * tg;
    const float s2 = 2.0f / (1.0f + tx * tx + ty * ty + tz * tz);
    const float vx = ax + ay * tz - az * ty;
    const float vy = ay + az * tx - ax * tz;
    const float vz = az + ax * ty - ay * tx;
    const float sx = tx * s2, sy = ty * s2, sz = tz * s2;
    const float px = ax + vy * sz - vz * sy;
    const float py = ay + vz * sx - vx * sz;
    const float pz = az + vx * sy - vy * sx;
    const float nux = px + e0, nuy = py + e1, nuz = pz + e2;
    const float rg = 1.0f / sqrtf(1.0f + nux * nux + nuy * nuy + nuz * nuz);
    const float mx = k.dt_dx * rg * nux;
    const float my = k.dt_dy * rg * nuy;
    const float x1 = x + mx, y1 = y + my;
    const int di = (x1 >= 1.0f) - (x1 < 0.0f);
    const int dj = (y1 >= 1.0f) - (y1 < 0.0f);
    const float qvz = k.q * nuz * rg;

    /* reverse: position update and deposit */
    float gx, gy, gmx, gmy, gq;
    deposit_adj(src, ci, cj, di, dj, x, y, mx, my, k.qnx, k.qny, qvz, gx, gy, gmx, gmy, gq);
    const float lmx = l[0] + gmx, lmy = l[1] + gmy;
    /* mx = dt_dx rg u'x, my = dt_dy rg u'y, qvz = q u'z rg */
    const float cx = lmx * k.dt_dx, cy = lmy * k.dt_dy, cz = gq * k.q;
    const float sr = (cx * nux + cy * nuy + cz * nuz) * rg * rg * rg;
    float lpx = l[2] + cx * rg - sr * nux;
    float lpy = l[3] + cy * rg - sr * nuy;
    float lpz = l[4] + cz * rg - sr * nuz;
    /* u' = p + e */
    float le0 = lpx, le1 = lpy, le2 = lpz;
    /* p = a + v x s */
    float lax = lpx, lay = lpy, laz = lpz;
    const float lvx = sy * lpz - sz * lpy, lvy = sz * lpx - sx * lpz, lvz = sx * lpy - sy * lpx;
    const float lsx = lpy * vz - lpz * vy, lsy = lpz * vx - lpx * vz, lsz = lpx * vy - lpy * vx;
    /* s = s2 t, s2 = 2 / (1 + |t|^2) */
    const float ls2 = lsx * tx + lsy * ty + lsz * tz;
    float ltx = s2 * lsx - s2 * s2 * tx * ls2;
    float lty = s2 * lsy - s2 * s2 * ty * ls2;
    float ltz = s2 * lsz - s2 * s2 * tz * ls2;
    /* v = a + a x t */
    lax += lvx + (ty * lvz - tz * lvy);
    lay += lvy + (tz * lvx - tx * lvz);
    laz += lvz + (tx * lvy - ty * lvx);
    ltx += lvy * az - lvz * ay;
    lty += lvz * ax - lvx * az;
    ltz += lvx * ay - lvy * ax;
    /* t = b tg, tg = tem / g, g = sqrt(1 + |a|^2) */
    const float lb0 = ltx * tg, lb1 = lty * tg, lb2 = ltz * tg;
    const float ltg = ltx * f[3] + lty * f[4] + ltz * f[5];
    const float lg = -ltg * tg / g;
    lax += lg * ax / g;
    lay += lg * ay / g;
    laz += lg * az / g;
    /* a = u + e, e = tem f */
    le0 += lax;
    le1 += lay;
    le2 += laz;
    const float lf[6] = {le0 * k.tem, le1 * k.tem, le2 * k.tem, lb0, lb1, lb2};
    const float lmax = fmaxf(fmaxf(fmaxf(fabsf(lf[0]), fabsf(lf[1])), fmaxf(fabsf(lf[2]), fabsf(lf[3]))),
                             fmaxf(fabsf(lf[4]), fabsf(lf[5])));
    float gxs = 0.0f, gys = 0.0f;
    /* the derivatives of the gather are formed again here, which keeps
     * them out of the registers live across the deposit adjoint */
#pragma unroll
    for (int c = 0; c < 6; c++) {
        int ia, jb;
        float wa, wb, da, db;
        st.of(c, ia, jb, wa, wb);
        interp(src, c, ia, jb, wa, wb, da, db);
        interp_scatter(src, c, ia, jb, wa, wb, lf[c]);
        gxs += lf[c] * da;
        gys += lf[c] * db;
    }
    l[0] += gx + gxs;
    l[1] += gy + gys;
    l[2] = lax;
    l[3] = lay;
    l[4] = laz;
    /* NaN when any is NaN, so that a fixed-point pass rejects it */
    return (lf[0] == lf[0] && lf[1] == lf[1] && lf[2] == lf[2] && lf[3] == lf[3] && lf[4] == lf[4] && lf[5] == lf[5])
               ? lmax
               : __builtin_nanf("");
}

__device__ inline void load_adj(const zk_adj_push_args &a, int i, float l[5])
{
    l[0] = a.lin.lx[i];
    l[1] = a.lin.ly[i];
    l[2] = a.lin.lux[i];
    l[3] = a.lin.luy[i];
    l[4] = a.lin.luz[i];
}

__device__ inline void store_adj(const zk_adj_push_args &a, int i, const float l[5])
{
    a.lout.lx[i] = l[0];
    a.lout.ly[i] = l[1];
    a.lout.lux[i] = l[2];
    a.lout.luy[i] = l[3];
    a.lout.luz[i] = l[4];
}

/* contributions up to 2^46 per atomic: sums of up to 2^17 of them fit */
#define ZK_ADJ_FX_LIMIT 70368744177664.0f /* 2^46 */
#define ZK_ADJ_FX_MARGIN 8                /* 2^8 above the running maximum */
#define ZK_ADJ_FX_MAX_SHARE 131072

/* The windowed particles of [i0, i1): particles outside the window are
 * listed (at ovf_list[i0...]) instead.  Returns the largest gathered-field
 * adjoint met (NaN if a NaN was met). */
template <class S>
__device__ inline float adj_share(const zk_adj_push_args &a, const S &src, int i0, int i1, int tx0, int ty0, int ox,
                                  int oy, const AdjConsts &k, int *nlist)
{
    const zk_tiles &t = a.t;
    const int lane = __lane_id();
    float mx = 0.0f;
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const int ix = zk_cell_ix(a.p.cell[i], a.p.cbits), iy = zk_cell_iy(a.p.cell[i], a.p.cbits);
        const int lx = ix - tx0, ly = iy - ty0;
        const bool inside = lx >= -t.H && lx < t.TX + t.H && ly >= -t.H && ly < t.TY + t.H;
        const unsigned long long out = __ballot(!inside);
        if (out) {
            int base = 0;
            const int leader = __ffsll((long long)out) - 1;
            if (lane == leader) base = atomicAdd(nlist, __popcll(out));
            base = __shfl(base, leader, ZK_WAVE);
            if (!inside) a.ovf_list[i0 + base + __popcll(out & ((1ull << lane) - 1ull))] = i;
        }
        if (!inside) continue;
        float l[5];
        load_adj(a, i, l);
        const float m = adjoint_one(src, ix - ox, iy - oy, a.p.x[i], a.p.y[i], a.p.ux[i], a.p.uy[i], a.p.uz[i], l, k);
        mx = (m == m && mx == mx) ? fmaxf(mx, m) : __builtin_nanf("");
        store_adj(a, i, l);
    }
    return mx;
}

__global__ void __launch_bounds__(256) adj_tile_kernel(zk_adj_push_args a)
{
    /* declared as 64-bit words so that the dynamic LDS starts 8-byte
     * aligned after the static variables (ds_add_u64 needs it) */
    extern __shared__ unsigned long long smx[];
    float *sm = (float *)smx;
    const zk_tiles &t = a.t;
    /* workgroup `part` of the tile's `splits` takes one contiguous share of
     * the tile's particles and owns window slot blockIdx.x */
    const int tile = blockIdx.x / a.splits, part = blockIdx.x % a.splits;
    const AdjConsts k = {a.tem, a.dt_dx, a.dt_dy, a.qnx, a.qny, a.q};
    const int tx0 = (tile % t.ntx) * t.TX, ty0 = (tile / t.ntx) * t.TY;
    const int n = t.WX * t.WY;
    const int ox = tx0 - 1 - t.H, oy = ty0 - 1 - t.H;
    float *fl = sm, *ll = sm + 6 * n;
    unsigned long long *accx = smx + (9 * n + 1) / 2; /* after the 9 n floats */
    float *accf = (float *)accx;
    __shared__ int nlist, bad;
    __shared__ unsigned bmax;
    const int r0 = a.tile_off[tile], r1 = a.tile_off[tile + 1];
    const int share = (r1 - r0 + a.splits - 1) / a.splits;
    const int i0 = min(r1, r0 + part * share), i1 = min(r1, i0 + share);
    /* the scale: a power of two with 2^8 times the running maximum at
     * 2^46; none (float sums) before any maximum is known or when the
     * scale would not be a normal float */
    const float amax = __uint_as_float(__atomic_load_n(a.amax, __ATOMIC_RELAXED));
    int e = 0;
    frexpf(amax, &e); /* amax < 2^e */
    const int se = 46 - ZK_ADJ_FX_MARGIN - e;
    const bool fx = amax > 0.0f && amax <= 3.0e38f && se > -120 && se < 120 && i1 - i0 <= ZK_ADJ_FX_MAX_SHARE;
    const float scale = fx ? ldexpf(1.0f, se) : 0.0f;
    if (threadIdx.x == 0) {
        nlist = 0;
        bad = 0;
        bmax = 0u;
    }
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        const int gx = ox + m % t.WX, gy = oy + m / t.WX;
        const bool in = gx >= -1 && gx <= t.nx + 1 && gy >= -1 && gy <= t.ny + 1;
        const size_t g = 3 * ((size_t)(gx + 1) + (size_t)(gy + 1) * a.nrow);
        for (int c = 0; c < 3; c++) {
            fl[c * n + m] = in ? a.E[g + c] : 0.0f;
            fl[(3 + c) * n + m] = in ? a.B[g + c] : 0.0f;
            ll[c * n + m] = in ? a.lam[g + c] : 0.0f;
        }
    }
    for (int m = threadIdx.x; m < 6 * n; m += blockDim.x) accx[m] = 0ull;
    __syncthreads();
    float mx;
    bool fixed = fx;
    if (fx) {
        const LdsSrcFx src = {fl, ll, accx, t.WX, n, scale};
        mx = adj_share(a, src, i0, i1, tx0, ty0, ox, oy, k, &nlist);
        if (!(mx * scale < ZK_ADJ_FX_LIMIT)) bad = 1;
        __syncthreads();
        if (bad) {
            /* a contribution beyond the fixed-point range: this share again
             * with float sums (the particle adjoints are read from lin and
             * written to lout, so the pass can be repeated) */
            fixed = false;
            for (int m = threadIdx.x; m < 6 * n; m += blockDim.x) accx[m] = 0ull;
            if (threadIdx.x == 0) nlist = 0;
            __syncthreads();
        }
    }
    if (!fixed) {
        const LdsSrc src = {fl, ll, accf, t.WX, n};
        mx = adj_share(a, src, i0, i1, tx0, ty0, ox, oy, k, &nlist);
    }
    /* running maximum of this species: block maximum, one global atomic */
    for (int o = ZK_WAVE / 2; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, ZK_WAVE));
    if (__lane_id() == 0 && mx > 0.0f) atomicMax(&bmax, __float_as_uint(mx));
    __syncthreads();
    if (threadIdx.x == 0 && bmax) atomicMax(a.amax, bmax);
    float *dst = a.region + (size_t)blockIdx.x * 6 * n;
    const float unscale = fixed ? 1.0f / scale : 0.0f;
    for (int m = threadIdx.x; m < 6 * n; m += blockDim.x)
        dst[m] = fixed ? (float)(long long)accx[m] * unscale : accf[m];
    /* the listed particles, through the guard-inclusive global grids: their
     * field adjoints go straight into lE, lB with atomics (the windows are
     * added by the next kernel) */
    const int nl = nlist;
    if (nl == 0) return;
    const GlobalSrc gsrc = {a.E, a.B, a.lam, a.lE, a.lB, a.nrow};
    for (int q = threadIdx.x; q < nl; q += blockDim.x) {
        const int i = a.ovf_list[i0 + q];
        float l[5];
        load_adj(a, i, l);
        const int c = a.p.cell[i];
        const float m = adjoint_one(gsrc, zk_cell_ix(c, a.p.cbits) + 1, zk_cell_iy(c, a.p.cbits) + 1, a.p.x[i], a.p.y[i],
                                    a.p.ux[i], a.p.uy[i], a.p.uz[i], l, k);
        if (m > 0.0f) atomicMax(a.amax, __float_as_uint(m));
        store_adj(a, i, l);
    }
}

/* field adjoints += the tile windows (splits slots per tile) covering each
 * buffer node */
__global__ void adj_reduce_kernel(const float *__restrict__ region, float *__restrict__ lE, float *__restrict__ lB,
                                  zk_tiles t, int nrow, int splits)
{
    const int bw = t.nx + 3, bh = t.ny + 3;
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= bw * bh) return;
    const int gx = node % bw - 1, gy = node / bw - 1;
    const int n = t.WX * t.WY;
    float s[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    /* window of tile (tx, ty) covers x in [tx TX - 1 - H, tx TX - 1 - H + WX) */
    const int txlo = max(0, (gx + 1 + t.H - t.WX + 1 + t.TX - 1) / t.TX);
    const int txhi = min(t.ntx - 1, (gx + 1 + t.H) / t.TX);
    const int tylo = max(0, (gy + 1 + t.H - t.WY + 1 + t.TY - 1) / t.TY);
    const int tyhi = min(t.nty - 1, (gy + 1 + t.H) / t.TY);
    for (int ty = tylo; ty <= tyhi; ty++)
        for (int tx = txlo; tx <= txhi; tx++) {
            const int u = gx - (tx * t.TX - 1 - t.H), v = gy - (ty * t.TY - 1 - t.H);
            if (u < 0 || u >= t.WX || v < 0 || v >= t.WY) continue;
            const float *w = region + (size_t)(tx + ty * t.ntx) * splits * 6 * n + u + v * t.WX;
            for (int p = 0; p < splits; p++, w += 6 * n)
                for (int c = 0; c < 6; c++) s[c] += w[c * n];
        }
    const size_t d = 3 * ((size_t)(gx + 1) + (size_t)(gy + 1) * nrow);
    for (int c = 0; c < 3; c++) {
        lE[d + c] += s[c];
        lB[d + c] += s[3 + c];
    }
}

/* pos[ids[i]] = i */
__global__ void adj_positions_kernel(const int64_t *__restrict__ ids, int n, int *__restrict__ pos)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pos[ids[i]] = i;
}

/* dst[c][i] = src[c][pos[ids[i]]] for the five adjoint components */
__global__ void adj_permute_kernel(zk_adj_vec src, zk_adj_vec dst, const int64_t *__restrict__ ids,
                                   const int *__restrict__ pos, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = pos[ids[i]];
    dst.lx[i] = src.lx[s];
    dst.ly[i] = src.ly[s];
    dst.lux[i] = src.lux[s];
    dst.luy[i] = src.luy[s];
    dst.luz[i] = src.luz[s];
}

/* out[5 id + c] = component c of slot i, id = ids[i] */
__global__ void adj_by_id_kernel(zk_adj_vec v, const int64_t *__restrict__ ids, int n, float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *o = out + 5 * (size_t)ids[i];
    o[0] = v.lx[i];
    o[1] = v.ly[i];
    o[2] = v.lux[i];
    o[3] = v.luy[i];
    o[4] = v.luz[i];
}

/* copies of particle sets and grid pairs for checkpoints and records:
 * one launch instead of one runtime copy per array */
__global__ void copy_parts_kernel(zk_parts dst, zk_parts src, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        dst.cell[i] = src.cell[i];
        dst.x[i] = src.x[i];
        dst.y[i] = src.y[i];
        dst.ux[i] = src.ux[i];
        dst.uy[i] = src.uy[i];
        dst.uz[i] = src.uz[i];
        if (dst.id && src.id) dst.id[i] = src.id[i];
    }
}

__global__ void copy_pair_kernel(float *__restrict__ a, const float *__restrict__ sa, float *__restrict__ b,
                                 const float *__restrict__ sb, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        a[i] = sa[i];
        b[i] = sb[i];
    }
}

unsigned grid_of(size_t n) { return (unsigned)((n + 255) / 256); }

} // namespace

extern "C" int zk_copy_parts(const zk_parts *dst, const zk_parts *src, int n, zk_stream_t s)
{
    if (n <= 0) return 0;
    const unsigned g = grid_of((size_t)n) < 4096u ? grid_of((size_t)n) : 4096u;
    hipLaunchKernelGGL(copy_parts_kernel, dim3(g), dim3(256), 0, s, *dst, *src, n);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_copy_pair(float *a, const float *sa, float *b, const float *sb, size_t n, zk_stream_t s)
{
    if (n == 0) return 0;
    const unsigned g = grid_of(n) < 4096u ? grid_of(n) : 4096u;
    hipLaunchKernelGGL(copy_pair_kernel, dim3(g), dim3(256), 0, s, a, sa, b, sb, n);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_adj_loss(float *lB, const float *B, int nx, int ny, int nrow, float scale, zk_stream_t s)
{
    const Grid g = {nx, ny, nrow};
    hipLaunchKernelGGL(adj_loss_kernel, dim3(grid_of((size_t)nx * ny)), dim3(256), 0, s, lB, B, g, scale);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_adj_fields(float *lE, float *lB, float *lJ, float *aE, float *aB, int nx, int ny, int nrow, float dt,
                             float dx, float dy, zk_stream_t s)
{
    const Grid g = {nx, ny, nrow};
    const unsigned nb = grid_of((size_t)(nx + 3) * (ny + 3));
    /* constants of the forward step (zk_yee) */
    const float dt2 = dt / 2.0f;
    const float dtx2 = dt2 / dx, dty2 = dt2 / dy, dtx = dt / dx, dty = dt / dy;
    hipLaunchKernelGGL(adj_guards_kernel, dim3(nb), dim3(256), 0, s, lE, lB, aE, aB, g);
    hipLaunchKernelGGL(adj_bhalf2_kernel, dim3(nb), dim3(256), 0, s, aE, aB, g, dtx2, dty2);
    hipLaunchKernelGGL(adj_efull_kernel, dim3(nb), dim3(256), 0, s, aE, aB, lE, lJ, g, dtx, dty, dt);
    hipLaunchKernelGGL(adj_bhalf1_kernel, dim3(nb), dim3(256), 0, s, lE, aB, lB, g, dtx2, dty2);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_adj_images(const float *lJ, float *lam, int nx, int ny, int nrow, zk_stream_t s)
{
    const Grid g = {nx, ny, nrow};
    hipLaunchKernelGGL(adj_images_kernel, dim3(grid_of((size_t)(nx + 3) * (ny + 3))), dim3(256), 0, s, lJ, lam, g);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_adj_push(const zk_adj_push_args *a, zk_stream_t s)
{
    const int n = a->t.WX * a->t.WY;
    const size_t lds = ((size_t)(9 * n + 1) / 2 + (size_t)6 * n) * sizeof(unsigned long long);
    if (lds > 65536) {
        const int e = (int)hipFuncSetAttribute((const void *)adj_tile_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e) return e;
    }
    if (a->splits < 1) return (int)hipErrorInvalidValue;
    if (a->t.ntiles > 0)
        hipLaunchKernelGGL(adj_tile_kernel, dim3(a->t.ntiles * a->splits), dim3(256), lds, s, *a);
    const int nodes = (a->t.nx + 3) * (a->t.ny + 3);
    hipLaunchKernelGGL(adj_reduce_kernel, dim3(grid_of((size_t)nodes)), dim3(256), 0, s, a->region, a->lE, a->lB,
                       a->t, a->nrow, a->splits);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_adj_permute(const zk_adj_vec *src, const zk_adj_vec *dst, const int64_t *ids_src,
                              const int64_t *ids_dst, int *pos, int n, zk_stream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(adj_positions_kernel, dim3(grid_of((size_t)n)), dim3(256), 0, s, ids_src, n, pos);
    hipLaunchKernelGGL(adj_permute_kernel, dim3(grid_of((size_t)n)), dim3(256), 0, s, *src, *dst, ids_dst, pos, n);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_adj_by_id(const zk_adj_vec *v, const int64_t *ids, int n, float *out, zk_stream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(adj_by_id_kernel, dim3(grid_of((size_t)n)), dim3(256), 0, s, *v, ids, n, out);
    ZK_LAUNCH_CHECK();
}
