/*
 * push_core.h - device code of the particle push shared by the
 * one-simulation kernels (push.hip) and the batched kernels (batch.hip).
 * See push.hip for the description of the algorithm.
 */
#ifndef ZK_PUSH_CORE_H
#define ZK_PUSH_CORE_H

#include <type_traits>

#include "common.h"


/* 1: the fixed-point push defers the deposits of cell-crossing particles
 * to a per-wave queue (DepQueue); 0: every lane runs deposit() in place */
#ifndef ZK_DEFER_CROSSING
#define ZK_DEFER_CROSSING 1
#endif

namespace zk {

struct Seg {
    float x0, x1, y0, y1, dx, dy, qvz;
    int ix, iy;
};

/* Window geometry of the LDS sources and sinks: w nodes per row, n per
 * plane; with WC > 0 a WC x WC window known at compile time, so that the
 * neighbour nodes and the planes become immediate LDS offsets of one
 * address (the default 16 x 16 tiles with a 3-cell margin have 25 x 25;
 * 2048^2 Weibel push 1.758 -> 1.745 ms). */
template <int WC>
struct Win {
    int w_, n_;
    __device__ int w() const { return WC > 0 ? WC : w_; }
    __device__ int n() const { return WC > 0 ? WC * WC : n_; }
};

/* Field sources: node (a, b) in the source's index space. */
template <int WC = 0>
struct LdsFieldsT : Win<WC> {
    const float *f;
    __device__ float at(int p, int a, int b) const { return f[p * this->n() + a + b * this->w()]; }
    __device__ float ex(int a, int b) const { return at(0, a, b); }
    __device__ float ey(int a, int b) const { return at(1, a, b); }
    __device__ float ez(int a, int b) const { return at(2, a, b); }
    __device__ float bx(int a, int b) const { return at(3, a, b); }
    __device__ float by(int a, int b) const { return at(4, a, b); }
    __device__ float bz(int a, int b) const { return at(5, a, b); }
};

struct GlobalFields {
    const float *E, *B;
    int nrow;
    __device__ float ex(int a, int b) const { return E[3 * (a + b * nrow) + 0]; }
    __device__ float ey(int a, int b) const { return E[3 * (a + b * nrow) + 1]; }
    __device__ float ez(int a, int b) const { return E[3 * (a + b * nrow) + 2]; }
    __device__ float bx(int a, int b) const { return B[3 * (a + b * nrow) + 0]; }
    __device__ float by(int a, int b) const { return B[3 * (a + b * nrow) + 1]; }
    __device__ float bz(int a, int b) const { return B[3 * (a + b * nrow) + 2]; }
};

/* Current sinks. */
template <int WC = 0>
struct LdsCurrentT : Win<WC> {
    float *j;
    __device__ void add(int a, int b, int c, float v) const
    {
        atomicAdd(&j[c * this->n() + a + b * this->w()], v);
    }
};

/* Fixed-point current sink: contributions are divided by Q (the power of
 * two at or above the particle charge |q|; a particle adds at most |q| to a
 * node), scaled by 2^32 and rounded to integers, which integer LDS atomics
 * sum exactly, independent of the order of the atomics.  The sums are of
 * zk_fx_raw patterns, read back with zk_fx_read: exact while a node's sum
 * stays below 2^18 Q, which a tile of at most ZK_FX_TILE_MAX particles
 * guarantees (larger tiles are counted in n_fx_range). */
template <int WC = 0>
struct LdsCurrentFxT : Win<WC> {
    unsigned long long *j;
    float scale;
    __device__ void add(int a, int b, int c, float v) const
    {
        atomicAdd(&j[c * this->n() + a + b * this->w()], zk_fx_raw(v, scale));
    }
};

/* Raw overflow grid of out-of-window pushes: 64-bit fixed point at the
 * grid's own scale (2^40 / Q_J), summed with global integer atomics, so
 * every node holds an exact sum that does not depend on the order of the
 * atomics and J is the same in every run. */
struct GlobalCurrent {
    long long *J;
    int nrow;
    float scale;
    __device__ void add(int a, int b, int c, float v) const
    {
        atomicAdd((unsigned long long *)&J[3 * (a + b * nrow) + c], zk_fx(v, scale));
    }
};

/* Bilinear gather on the staggered mesh (em2d/particles.c:1029-1071);
 * (i, j) is the particle's cell in the source index space. */
template <class F>
__device__ inline void gather(const F &S, int i, int j, float wx, float wy, float e[3], float b[3])
{
    const int ih = i + ((wx < 0.5f) ? -1 : 0);
    const int jh = j + ((wy < 0.5f) ? -1 : 0);
    const float wxh = wx + ((wx < 0.5f) ? 0.5f : -0.5f);
    const float wyh = wy + ((wy < 0.5f) ? 0.5f : -0.5f);
    e[0] = (S.ex(ih, j) * (1.0f - wxh) + S.ex(ih + 1, j) * wxh) * (1.0f - wy) +
           (S.ex(ih, j + 1) * (1.0f - wxh) + S.ex(ih + 1, j + 1) * wxh) * wy;
    e[1] = (S.ey(i, jh) * (1.0f - wx) + S.ey(i + 1, jh) * wx) * (1.0f - wyh) +
           (S.ey(i, jh + 1) * (1.0f - wx) + S.ey(i + 1, jh + 1) * wx) * wyh;
    e[2] = (S.ez(i, j) * (1.0f - wx) + S.ez(i + 1, j) * wx) * (1.0f - wy) +
           (S.ez(i, j + 1) * (1.0f - wx) + S.ez(i + 1, j + 1) * wx) * wy;
    b[0] = (S.bx(i, jh) * (1.0f - wx) + S.bx(i + 1, jh) * wx) * (1.0f - wyh) +
           (S.bx(i, jh + 1) * (1.0f - wx) + S.bx(i + 1, jh + 1) * wx) * wyh;
    b[1] = (S.by(ih, j) * (1.0f - wxh) + S.by(ih + 1, j) * wxh) * (1.0f - wy) +
           (S.by(ih, j + 1) * (1.0f - wxh) + S.by(ih + 1, j + 1) * wxh) * wy;
    b[2] = (S.bz(ih, jh) * (1.0f - wxh) + S.bz(ih + 1, jh) * wxh) * (1.0f - wyh) +
           (S.bz(ih, jh + 1) * (1.0f - wxh) + S.bz(ih + 1, jh + 1) * wxh) * wyh;
}

/* Cut segment s at the y cell edge crossed (dj = +-1) into s and t
 * (em2d/particles.c:838-879). */
__device__ inline void ysplit(Seg &s, Seg &t, int dj)
{
    const int edge = (dj == 1);
    const float f = (s.y1 - edge) / s.dy;
    t.y0 = 1 - edge;
    t.y1 = s.y1 - dj;
    t.dy = s.dy * f;
    t.iy = s.iy + dj;
    const float xc = s.x0 + s.dx * (1.0f - f);
    t.x0 = xc;
    t.x1 = s.x1;
    t.dx = s.dx * f;
    t.ix = s.ix;
    t.qvz = s.qvz * f;
    s.y1 = edge;
    s.dy *= (1.0f - f);
    s.dx *= (1.0f - f);
    s.x1 = xc;
    s.qvz *= (1.0f - f);
}

/* Current of one straight segment inside one cell (em2d/particles.c:884-921).
 *
 * The same weights as the reference, refactored: with h = (y0 + y1)/2 the
 * Jx weights are 1 - h and h (likewise in x for Jy), and the four Jz
 * weights ax0 ay0 + ax1 ay1 + (ax0 ay1 - ax1 ay0)/2 (a = 1 - b, b the
 * coordinate) are ax0 A + ax1 B, bx0 A + bx1 B, ax0 C + ax1 D,
 * bx0 C + bx1 D with C = y0 + y1/2, D = y1 - y0/2, A = 3/2 - C, B = 1/2 - D,
 * evaluated with fused multiply-adds (26 fewer instructions per segment
 * than the reference's expressions).  The current is only required to
 * agree with the reference at FP32 rounding (its sum order differs
 * anyway); the particles do not depend on it within a step. */
template <class D>
__device__ inline void deposit_seg(const D &J, const Seg &g, float qnx, float qny)
{
    const float ax0 = 1.0f - g.x0, ax1 = 1.0f - g.x1;
    const float jx = qnx * g.dx, jy = qny * g.dy;
    const float jxh = jx * (0.5f * (g.y0 + g.y1));
    const float jyh = jy * (0.5f * (g.x0 + g.x1));
    J.add(g.ix, g.iy, 0, jx - jxh);
    J.add(g.ix, g.iy + 1, 0, jxh);
    J.add(g.ix, g.iy, 1, jy - jyh);
    J.add(g.ix + 1, g.iy, 1, jyh);
    const float c = __builtin_fmaf(0.5f, g.y1, g.y0), d = __builtin_fmaf(-0.5f, g.y0, g.y1);
    const float qa = g.qvz * (1.5f - c), qb = g.qvz * (0.5f - d);
    const float qc = g.qvz * c, qd = g.qvz * d;
    J.add(g.ix, g.iy, 2, __builtin_fmaf(ax0, qa, ax1 * qb));
    J.add(g.ix + 1, g.iy, 2, __builtin_fmaf(g.x0, qa, g.x1 * qb));
    J.add(g.ix, g.iy + 1, 2, __builtin_fmaf(ax0, qc, ax1 * qd));
    J.add(g.ix + 1, g.iy + 1, 2, __builtin_fmaf(g.x0, qc, g.x1 * qd));
}

/* Split the motion at cell edges into up to three segments and deposit
 * them (em2d/particles.c:773-924).  Each segment is computed as in the
 * reference; the deposit order differs (the fixed-point windows do not
 * depend on it): the segment cut in y, the other x-cut segment, then the
 * piece beyond the y edge. */
template <class D>
__device__ inline void deposit(const D &J, int ix, int iy, int di, int dj, float x0, float y0,
                               float dx, float dy, float qnx, float qny, float qvz)
{
    Seg a;
    a.x0 = x0;
    a.y0 = y0;
    a.dx = dx;
    a.dy = dy;
    a.x1 = x0 + dx;
    a.y1 = y0 + dy;
    a.qvz = qvz * 0.5f; /* equals (float)(qvz / 2.0): halving is exact */
    a.ix = ix;
    a.iy = iy;
    const bool xcut = di != 0;
    Seg b = a; /* used only when xcut */
    if (xcut) {
        const int edge = (di == 1);
        const float f = (x0 + dx - edge) / dx;
        b.x0 = 1 - edge;
        b.x1 = (x0 + dx) - di;
        b.dx = dx * f;
        b.ix = ix + di;
        const float yc = y0 + dy * (1.0f - f);
        b.y0 = yc;
        b.y1 = a.y1;
        b.dy = dy * f;
        b.iy = iy;
        b.qvz = a.qvz * f;
        a.x1 = edge;
        a.dx *= (1.0f - f);
        a.dy *= (1.0f - f);
        a.y1 = yc;
        a.qvz *= (1.0f - f);
    }
    /* p: the segment that crosses the y edge (a when it is the only one or
     * when it crosses, else b); q: the other one, when there are two.  When
     * a crosses, b lies beyond the y edge: it moves one cell in y. */
    const bool bcuts = xcut && !(a.y1 < 0.0f || a.y1 >= 1.0f);
    Seg p = bcuts ? b : a, q = bcuts ? a : b, t;
    if (dj != 0) {
        if (xcut && !bcuts) {
            q.y0 -= dj;
            q.y1 -= dj;
            q.iy += dj;
        }
        ysplit(p, t, dj);
    }
    deposit_seg(J, p, qnx, qny);
    if (xcut) deposit_seg(J, q, qnx, qny);
    if (dj != 0) deposit_seg(J, t, qnx, qny);
}

struct Consts {
    float tem, dt_dx, dt_dy, qnx, qny, q;
};

/* The motion of a particle that stays in its cell is one segment: the
 * deposit of deposit() with di = dj = 0, without its split arithmetic. */
template <class D>
__device__ inline void deposit_in_cell(const D &J, int ix, int iy, float x0, float y0, float dx, float dy,
                                       float qnx, float qny, float qvz)
{
    Seg a;
    a.x0 = x0;
    a.y0 = y0;
    a.dx = dx;
    a.dy = dy;
    a.x1 = x0 + dx;
    a.y1 = y0 + dy;
    a.qvz = qvz * 0.5f;
    a.ix = ix;
    a.iy = iy;
    deposit_seg(J, a, qnx, qny);
}

/* Deferred deposits of the particles that cross a cell edge.
 *
 * A wave executes each branch of deposit() for all its lanes as soon as
 * one lane takes it, and in nearly every wave of 64 some particle crosses
 * a cell edge: without deferral every wave pays for both splits and three
 * segment deposits.  Instead, the lanes whose particle stays in its cell
 * deposit its one segment at once, and the crossing lanes hand their
 * motion to a per-wave queue held one entry per lane in registers (entries
 * are moved between lanes with ds_permute).  When the queue cannot take
 * the next batch it is deposited with deposit() by the occupied lanes, so
 * the split arithmetic runs on (nearly) full waves.  The fixed-point sums
 * of the window do not depend on the order of the contributions, so J is
 * unchanged.  Window-local cells fit 8 bits each (windows are at most
 * 73 nodes per side), di + 1 and dj + 1 2 bits each. */
struct DepQueue {
    float x0, y0, dx, dy, qvz;
    int code;
    int n; /* entries held (wave-uniform), in lanes 0 .. n - 1 */

    __device__ void init() { n = 0; }

    template <class D>
    __device__ void flush(const D &J, const Consts &k)
    {
        if (n == 0) return;
        if ((int)__lane_id() < n) {
            const int ci = code & 0xff, cj = (code >> 8) & 0xff;
            const int di = ((code >> 16) & 3) - 1, dj = ((code >> 18) & 3) - 1;
            deposit(J, ci, cj, di, dj, x0, y0, dx, dy, k.qnx, k.qny, qvz);
        }
        n = 0;
    }

    /* called by every lane of the wave; lanes with `cross` append */
    template <class D>
    __device__ void push(bool cross, int ci, int cj, int di, int dj, float px, float py, float mx, float my,
                         float pq, const D &J, const Consts &k)
    {
        const unsigned long long m = __ballot(cross);
        const int add = __popcll(m);
        if (add == 0) return;
        if (n + add > ZK_WAVE) flush(J, k);
        const int lane = __lane_id();
        const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        /* lanes that do not append send to a lane outside [n, n + add) */
        const int dst = ((cross ? n + rank : n + add) & (ZK_WAVE - 1)) * 4;
        const int c = ci | (cj << 8) | ((di + 1) << 16) | ((dj + 1) << 18);
        const bool take = lane >= n && lane < n + add;
        const float rx0 = __int_as_float(__builtin_amdgcn_ds_permute(dst, __float_as_int(px)));
        const float ry0 = __int_as_float(__builtin_amdgcn_ds_permute(dst, __float_as_int(py)));
        const float rdx = __int_as_float(__builtin_amdgcn_ds_permute(dst, __float_as_int(mx)));
        const float rdy = __int_as_float(__builtin_amdgcn_ds_permute(dst, __float_as_int(my)));
        const float rq = __int_as_float(__builtin_amdgcn_ds_permute(dst, __float_as_int(pq)));
        const int rc = __builtin_amdgcn_ds_permute(dst, c);
        x0 = take ? rx0 : x0;
        y0 = take ? ry0 : y0;
        dx = take ? rdx : dx;
        dy = take ? rdy : dy;
        qvz = take ? rq : qvz;
        code = take ? rc : code;
        n += add;
    }
};

struct Part {
    int ix, iy;
    float x, y, ux, uy, uz;
};

/* Particle streams are read and written once per step with ordinary
 * loads and stores: nontemporal hints on them made the 2048^2 Weibel push
 * 14 % slower (1.35 against 1.16 ms per launch). */
__device__ inline Part load_part(const zk_parts &P, int i)
{
    const unsigned o = (unsigned)i * 4u;
    Part c;
    const int cell = zk_at(P.cell, o);
    c.ix = zk_cell_ix(cell, P.cbits);
    c.iy = zk_cell_iy(cell, P.cbits);
    c.x = zk_at(P.x, o);
    c.y = zk_at(P.y, o);
    c.ux = zk_at(P.ux, o);
    c.uy = zk_at(P.uy, o);
    c.uz = zk_at(P.uz, o);
    return c;
}

__device__ inline void store_part(const zk_parts &P, int i, const Part &c)
{
    const unsigned o = (unsigned)i * 4u;
    zk_at(P.cell, o) = zk_cell(c.ix, c.iy, P.cbits);
    zk_at(P.x, o) = c.x;
    zk_at(P.y, o) = c.y;
    zk_at(P.ux, o) = c.ux;
    zk_at(P.uy, o) = c.uy;
    zk_at(P.uz, o) = c.uz;
}

/* IEEE division a / b and square root of x for the Boris push, where the
 * divisors are >= 1 and the square-root arguments are >= 1: the sequences
 * of the correctly rounded hardware expansions (reciprocal and square-root
 * estimates refined by fused multiply-adds), without the operand scaling
 * and special-value fix-ups that only denormal, infinite, zero or
 * extreme-exponent operands need.  For finite operands in those ranges
 * the results are the correctly rounded ones, bit for bit. */
__device__ inline float div_ge1(float a, float b)
{
    float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = a * y;
    float r = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}

__device__ inline float sqrt_ge1(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __int_as_float(__float_as_int(s) - 1), sp = __int_as_float(__float_as_int(s) + 1);
    float r = (__builtin_fmaf(-sm, s, x) <= 0.0f) ? sm : s;
    r = (__builtin_fmaf(-sp, s, x) > 0.0f) ? sp : r;
    return r;
}

/* Motion of one particle whose cell is (ci, cj) in the sources' index
 * space: gather, Boris push, displacement (mx, my), cells crossed (di, dj)
 * and q v_z; returns the time-centred kinetic energy term
 * (em2d/particles.c:1125-1228). */
template <class F>
__device__ inline double move_one(const F &S, int ci, int cj, float x, float y, float &ux, float &uy, float &uz,
                                  float &mx, float &my, float &qvz, int &di, int &dj, const Consts &k)
{
    float e[3], b[3];
    gather(S, ci, cj, x, y, e, b);
    e[0] *= k.tem;
    e[1] *= k.tem;
    e[2] *= k.tem;
    float ax = ux + e[0], ay = uy + e[1], az = uz + e[2];
    const float usq = ax * ax + ay * ay + az * az;
    const float g = sqrt_ge1(1.0f + usq);
    const double kin = (double)div_ge1(usq, g + 1.0f);
    const float tg = div_ge1(k.tem, g);
    b[0] *= tg;
    b[1] *= tg;
    b[2] *= tg;
    const float s2 = div_ge1(2.0f, 1.0f + b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    const float vx = ax + ay * b[2] - az * b[1];
    const float vy = ay + az * b[0] - ax * b[2];
    const float vz = az + ax * b[1] - ay * b[0];
    b[0] *= s2;
    b[1] *= s2;
    b[2] *= s2;
    ax += vy * b[2] - vz * b[1];
    ay += vz * b[0] - vx * b[2];
    az += vx * b[1] - vy * b[0];
    ux = ax + e[0];
    uy = ay + e[1];
    uz = az + e[2];

    const float rg = div_ge1(1.0f, sqrt_ge1(1.0f + ux * ux + uy * uy + uz * uz));
    mx = k.dt_dx * rg * ux;
    my = k.dt_dy * rg * uy;
    const float x1 = x + mx, y1 = y + my;
    di = (x1 >= 1.0f) - (x1 < 0.0f);
    dj = (y1 >= 1.0f) - (y1 < 0.0f);
    qvz = k.q * uz * rg;
    return kin;
}

/* Push one particle: move_one, the deposit of its motion, and its new cell
 * offsets (ltrim, em2d/particles.c:1081-1084, 1200-1212). */
template <class F, class D>
__device__ inline double push_one(const F &S, const D &J, int ci, int cj, float &x, float &y,
                                  float &ux, float &uy, float &uz, int &di, int &dj, const Consts &k)
{
    float mx, my, qvz;
    const double kin = move_one(S, ci, cj, x, y, ux, uy, uz, mx, my, qvz, di, dj, k);
    float x1 = x + mx, y1 = y + my;
    x1 -= di;
    y1 -= dj;
    deposit(J, ci, cj, di, dj, x, y, mx, my, k.qnx, k.qny, qvz);
    x = x1;
    y = y1;
    return kin;
}

/* Boundary conditions after the push (em2d/particles.c:1236-1259); the
 * window shift of spec_move_window (:619-640) is applied first.  With open
 * y boundaries a particle that left through y is removed, as the moving
 * window removes those leaving through x (the open particle boundary of
 * the reference's 1D code, em1d/particles.h:77-80). */
__device__ inline bool boundaries(int &ix, int &iy, int nx, int ny, int mw, int shift, int open_y)
{
    if (mw) {
        ix -= shift;
        if (ix < 0 || ix >= nx) {
            ix = ZK_DEAD;
            return true;
        }
    } else {
        ix += ((ix < 0) ? nx : 0) - ((ix >= nx) ? nx : 0);
    }
    if (open_y) {
        if (iy < 0 || iy >= ny) {
            ix = ZK_DEAD;
            return true;
        }
    } else {
        iy += ((iy < 0) ? ny : 0) - ((iy >= ny) ? ny : 0);
    }
    return false;
}

/* List the live particles outside the window in the tile's own range of
 * ovf_list (one LDS atomic per wave). */
__device__ inline void list_out(int *ovf_list, int i0, int i, bool out_here, int *nlist)
{
    const unsigned long long out = __ballot(out_here);
    if (out) {
        const int lane = __lane_id();
        int base = 0;
        const int leader = __ffsll((long long)out) - 1;
        if (lane == leader) base = atomicAdd(nlist, __popcll(out));
        base = __shfl(base, leader, ZK_WAVE);
        if (out_here) ovf_list[i0 + base + __popcll(out & ((1ull << lane) - 1ull))] = i;
    }
}

/* Stores of a push with `sorting` set (staged re-sort).  A particle that
 * survived the step and is still in its tile goes to the front of the
 * tile's range of the other set, one that moved to another tile to the
 * tile's staging region behind the ranges of that set while it has room (a
 * leaver that finds none stays with the tile; the re-sort counts it as
 * unplaced).  sort_reserve takes the slots of a wave with one LDS atomic
 * on a pair of counters and returns the lane's slot in sdst, or -1 for
 * none; one set of stores then writes both kinds.  It is called by all
 * lanes of the wave.
 *
 * SortSlots: the slots a sorting push writes to for one tile, the range's
 * start and the staging region's start and size, read once per tile (the
 * loop's stores may alias them for the compiler, whose reloads would wait
 * for every load in flight: 2.71 -> 2.22 ms for a sorting 2048^2 push). */
struct SortSlots {
    int dst0, stage0, room;
    __device__ static SortSlots of(const zk_push_args &a, int tile)
    {
        SortSlots r = {0, 0, 0};
        if (a.sorting) {
            const int lo = a.tile_off[tile] >> 1;
            r.dst0 = a.tile_off[tile];
            r.stage0 = a.tile_off[a.t.ntiles] + lo;
            r.room = (a.tile_off[tile + 1] >> 1) - lo;
        }
        return r;
    }
};

__device__ inline int sort_reserve(const SortSlots &ss, bool keep, bool in_tile, unsigned long long *nsl)
{
    const bool lv = keep && !in_tile, st = keep && in_tile;
    const unsigned lane = __lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const unsigned long long ml = __ballot(lv), ms = __ballot(st);
    if (!(ml | ms)) return -1;
    /* one atomic for both counters: stayers in the low word, leavers in
     * the high word, taken by the wave's first active lane */
    const unsigned first = __builtin_amdgcn_readfirstlane(lane);
    unsigned long long old = 0ull;
    if (lane == first)
        old = atomicAdd(nsl, (unsigned long long)__popcll(ms) | ((unsigned long long)__popcll(ml) << 32));
    const unsigned s0 = __builtin_amdgcn_readfirstlane((unsigned)old);
    const unsigned l0 = __builtin_amdgcn_readfirstlane((unsigned)(old >> 32));
    int d = -1;
    bool over = false;
    if (st) d = ss.dst0 + (int)s0 + __popcll(ms & lt);
    if (lv) {
        const int r = (int)l0 + __popcll(ml & lt);
        if (r < ss.room) d = ss.stage0 + r;
        else over = true;
    }
    /* rare: leavers without staging room stay with the tile */
    const unsigned long long mo = __ballot(over);
    if (mo) {
        unsigned long long o2 = 0ull;
        if (lane == first) o2 = atomicAdd(nsl, (unsigned long long)__popcll(mo));
        const unsigned s1 = __builtin_amdgcn_readfirstlane((unsigned)o2);
        if (over) d = ss.dst0 + (int)s1 + __popcll(mo & lt);
    }
    return d;
}

__device__ inline void sort_write(const zk_push_args &a, int d, int i, const Part &c)
{
    if (d >= 0) {
        store_part(a.sdst, d, c);
        if (a.p.id) {
            /* the id addresses are formed here from the slot numbers: left
             * to itself the compiler keeps 64-bit per-lane address terms
             * alive across the push loop, which spills registers */
            int src = i, dst = d;
            asm volatile("" : "+v"(src), "+v"(dst));
            a.sdst.id[dst] = a.p.id[src];
        }
    }
}

/* Cells [x0, x1) x [y0, y1) of a tile (the last row and column of tiles
 * may be narrower). */
struct TileBox {
    int x0, x1, y0, y1;
    __device__ bool has(int ix, int iy) const { return ix >= x0 && ix < x1 && iy >= y0 && iy < y1; }
};

/* The species a tile workgroup pushes: one (the plain and batched
 * launches), or several species of one simulation that share the tiling,
 * the fields, the current and the fixed-point scale (fused launch). */
struct OneSpecies {
    const zk_push_args &a;
    __device__ int n() const { return 1; }
    __device__ const zk_push_args &operator[](int) const { return a; }
};

struct SpeciesSet {
    const zk_push_multi &m;
    __device__ int n() const { return m.n; }
    __device__ const zk_push_args &operator[](int k) const { return m.s[k]; }
};

/* Tile pass.  The workgroup stages the fields of its window and clears
 * the current window once, pushes the tile's particles of each species in
 * turn, and stores the window once: with several species the staging and
 * the window store (0.12 ms of a 2048^2 launch) are not repeated per
 * species.  PAIR = 2 pushes two particles per thread and iteration: the
 * gathers and Boris pushes of both are independent and interleave, which
 * hides LDS and arithmetic latency at the price of registers.  FX selects
 * the fixed-point window (LdsCurrentFxT) instead of float LDS atomics. */
template <int PAIR = 1, bool FX = false, class Set, int WC = 0>
__device__ __forceinline__ void push_tile_body(const Set &S, const int tile)
{
    extern __shared__ float sm[];
    __shared__ double red[16];
    __shared__ int ndead_blk, nlist;
    __shared__ int nls[ZK_FUSE_MAX];    /* out-of-window list lengths per species */
    __shared__ double ekin[ZK_FUSE_MAX]; /* energy partials per species */
    __shared__ unsigned long long nsl[ZK_FUSE_MAX]; /* sorting push: stayers | leavers << 32 */
    const zk_push_args &a0 = S[0];
    const int ns = S.n();
    const zk_tiles &t = a0.t;
    if (threadIdx.x == 0) ndead_blk = nlist = 0;
    if (threadIdx.x < ZK_FUSE_MAX) nsl[threadIdx.x] = 0ull;

    const int tx0 = (tile % t.ntx) * t.TX, ty0 = (tile / t.ntx) * t.TY;
    const TileBox box = {tx0, min(tx0 + t.TX, t.nx), ty0, min(ty0 + t.TY, t.ny)};
    const int n = t.WX * t.WY;
    const int ox = tx0 - 1 - t.H, oy = ty0 - 1 - t.H; /* cell index of window node 0 */
    float *fl = sm;
    float *jl = sm + 6 * n;
    unsigned long long *jx = (unsigned long long *)(sm + 6 * n); /* FX: 3 n integers */
    /* stage the window of E and B; nodes outside the guard-inclusive
     * buffer are never read by an in-window particle */
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        const int gx = ox + m % t.WX, gy = oy + m / t.WX;
        const bool in = gx >= -1 && gx <= t.nx + 1 && gy >= -1 && gy <= t.ny + 1;
        const int g = 3 * ((gx + 1) + (gy + 1) * a0.nrow);
        fl[0 * n + m] = in ? a0.E[g + 0] : 0.0f;
        fl[1 * n + m] = in ? a0.E[g + 1] : 0.0f;
        fl[2 * n + m] = in ? a0.E[g + 2] : 0.0f;
        fl[3 * n + m] = in ? a0.B[g + 0] : 0.0f;
        fl[4 * n + m] = in ? a0.B[g + 1] : 0.0f;
        fl[5 * n + m] = in ? a0.B[g + 2] : 0.0f;
        if (FX) {
            jx[0 * n + m] = 0ull;
            jx[1 * n + m] = 0ull;
            jx[2 * n + m] = 0ull;
        } else {
            jl[0 * n + m] = 0.0f;
            jl[1 * n + m] = 0.0f;
            jl[2 * n + m] = 0.0f;
        }
    }
    __syncthreads();
    LdsFieldsT<WC> lf;
    lf.f = fl;
    lf.w_ = t.WX;
    lf.n_ = n;
    typedef typename std::conditional<FX, LdsCurrentFxT<WC>, LdsCurrentT<WC>>::type Sink;
    Sink lj;
    lj.w_ = t.WX;
    lj.n_ = n;
    if constexpr (FX) {
        lj.j = jx;
        lj.scale = a0.fx_scale;
    } else {
        lj.j = jl;
    }
    int fx_count = 0; /* slots of the tile over all species of the launch */
    for (int sp = 0; sp < S.n(); sp++) {
        const zk_push_args &a = S[sp];
        const Consts k = {a.tem, a.dt_dx, a.dt_dy, a.qnx, a.qny, a.q};
        zk_parts P = a.p;
        double kin = 0.0;
        int ndead = 0;
        const int i0 = a.tile_off[tile], i1 = a.tile_n ? i0 + a.tile_n[tile] : a.tile_off[tile + 1];
        fx_count += i1 - i0;
        const SortSlots ss = SortSlots::of(a, tile);
        if (PAIR == 1) {
            /* Each thread pushes particles i, i + bs, i + 2 bs, ... with the
             * loads running two particles ahead in three register sets that
             * take turns (A, B, C), so that the loads' latency is hidden behind
             * two pushes.  The sets are unrolled rather than copied at the
             * loop's back edge: such a copy waits for every outstanding memory
             * operation, the previous particle's stores included.  Loads past
             * the range re-read its last particle and are not pushed.  (2048^2
             * Weibel push: 1.16 ms per launch with no load ahead, 1.08 ms with
             * one particle ahead, 1.03-1.05 ms with two; three ahead spills.) */
            DepQueue dq;
            dq.init();
            auto step = [&](Part &c, int i) {
                const int lx = c.ix - tx0, ly = c.iy - ty0;
                const bool inside = lx >= -t.H && lx < t.TX + t.H && ly >= -t.H && ly < t.TY + t.H;
                const bool live = i >= i0 && i < i1 && c.ix != ZK_DEAD;
                list_out(a.ovf_list, i0, i, live && !inside, &nlist);
                if constexpr (FX && ZK_DEFER_CROSSING) {
                    /* deposits of cell-crossing particles are deferred (DepQueue) */
                    const bool go = live && inside;
                    const int ci = c.ix - ox, cj = c.iy - oy;
                    float mx = 0.0f, my = 0.0f, qvz = 0.0f;
                    int di = 0, dj = 0;
                    if (go) kin += move_one(lf, ci, cj, c.x, c.y, c.ux, c.uy, c.uz, mx, my, qvz, di, dj, k);
                    const bool cross = go && (di != 0 || dj != 0);
                    /* the particle after the step */
                    Part nc = c;
                    if (go) {
                        float x1 = c.x + mx, y1 = c.y + my;
                        x1 -= di;
                        y1 -= dj;
                        nc.x = x1;
                        nc.y = y1;
                        nc.ix += di;
                        nc.iy += dj;
                        if (boundaries(nc.ix, nc.iy, t.nx, t.ny, a.moving_window, a.shift, a.open_y)) ndead++;
                    }
                    int sd = -1;
                    if (a.sorting)
                        sd = sort_reserve(ss, go && nc.ix != ZK_DEAD, box.has(nc.ix, nc.iy), &nsl[sp]);
                    if (go && !cross) deposit_in_cell(lj, ci, cj, c.x, c.y, mx, my, k.qnx, k.qny, qvz);
                    dq.push(cross, ci, cj, di, dj, c.x, c.y, mx, my, qvz, lj, k);
                    if (a.sorting) sort_write(a, sd, i, nc);
                    else if (go) store_part(P, i, nc);
                } else {
                    const bool go = live && inside;
                    if (go) {
                        int di, dj;
                        kin += push_one(lf, lj, c.ix - ox, c.iy - oy, c.x, c.y, c.ux, c.uy, c.uz, di, dj, k);
                        c.ix += di;
                        c.iy += dj;
                        if (boundaries(c.ix, c.iy, t.nx, t.ny, a.moving_window, a.shift, a.open_y)) ndead++;
                        if (!a.sorting) store_part(P, i, c);
                    }
                    if (a.sorting) {
                        const int sd = sort_reserve(ss, go && c.ix != ZK_DEAD, box.has(c.ix, c.iy), &nsl[sp]);
                        sort_write(a, sd, i, c);
                    }
                }
            };
            /* The loop starts at the multiple of 32 at or below i0, so each
             * wave's 64 consecutive slots cover exactly two 128-byte lines
             * of every array (up to 31 slots of the previous tile are
             * loaded and left alone): 1.98 -> 1.94 ms per 2048^2 launch. */
            const int bs = blockDim.x, last = i1 - 1;
            int i = (i0 & ~31) + threadIdx.x;
            if (i1 > i0) {
                /* The loop runs while any lane of the wave has a slot: the
                 * deferred deposits move data between the wave's lanes, so
                 * no lane may leave the loop early.  Lanes past the range
                 * load its last particle and leave it alone. */
                Part ca, cb, cc;
                ca = load_part(P, min(i, last));
                cb = load_part(P, min(i + bs, last));
                int wave0 = __builtin_amdgcn_readfirstlane(i - (int)__lane_id());
                for (; wave0 < i1; i += 3 * bs, wave0 += 3 * bs) {
                    cc = load_part(P, min(i + 2 * bs, last));
                    step(ca, i);
                    ca = load_part(P, min(i + 3 * bs, last));
                    step(cb, i + bs);
                    cb = load_part(P, min(i + 4 * bs, last));
                    step(cc, i + 2 * bs);
                }
            }
            if constexpr (FX && ZK_DEFER_CROSSING) dq.flush(lj, k);
        } else {
            for (int i = i0 + threadIdx.x; i < i1; i += 2 * blockDim.x) {
                const int j = i + blockDim.x;
                const bool hasb = j < i1;
                Part ca = load_part(P, i);
                Part cb = hasb ? load_part(P, j) : ca;
                const int lxa = ca.ix - tx0, lya = ca.iy - ty0, lxb = cb.ix - tx0, lyb = cb.iy - ty0;
                const bool ina = lxa >= -t.H && lxa < t.TX + t.H && lya >= -t.H && lya < t.TY + t.H;
                const bool inb = lxb >= -t.H && lxb < t.TX + t.H && lyb >= -t.H && lyb < t.TY + t.H;
                const bool livea = ca.ix != ZK_DEAD, liveb = hasb && cb.ix != ZK_DEAD;
                list_out(a.ovf_list, i0, i, livea && !ina, &nlist);
                list_out(a.ovf_list, i0, j, liveb && !inb, &nlist);
                const bool pa = livea && ina, pb = liveb && inb;
                /* a particle that is not pushed here moves at a safe cell */
                const int cia = pa ? ca.ix - ox : t.H + 1, cja = pa ? ca.iy - oy : t.H + 1;
                const int cib = pb ? cb.ix - ox : t.H + 1, cjb = pb ? cb.iy - oy : t.H + 1;
                float mxa, mya, qa, mxb, myb, qb;
                int dia, dja, dib, djb;
                const double ka = move_one(lf, cia, cja, ca.x, ca.y, ca.ux, ca.uy, ca.uz, mxa, mya, qa, dia, dja, k);
                const double kb = move_one(lf, cib, cjb, cb.x, cb.y, cb.ux, cb.uy, cb.uz, mxb, myb, qb, dib, djb, k);
                if (pa) {
                    float x1 = ca.x + mxa, y1 = ca.y + mya;
                    x1 -= dia;
                    y1 -= dja;
                    deposit(lj, cia, cja, dia, dja, ca.x, ca.y, mxa, mya, k.qnx, k.qny, qa);
                    ca.x = x1;
                    ca.y = y1;
                    kin += ka;
                    ca.ix += dia;
                    ca.iy += dja;
                    if (boundaries(ca.ix, ca.iy, t.nx, t.ny, a.moving_window, a.shift, a.open_y)) ndead++;
                    store_part(P, i, ca);
                }
                if (pb) {
                    float x1 = cb.x + mxb, y1 = cb.y + myb;
                    x1 -= dib;
                    y1 -= djb;
                    deposit(lj, cib, cjb, dib, djb, cb.x, cb.y, mxb, myb, k.qnx, k.qny, qb);
                    cb.x = x1;
                    cb.y = y1;
                    kin += kb;
                    cb.ix += dib;
                    cb.iy += djb;
                    if (boundaries(cb.ix, cb.iy, t.nx, t.ny, a.moving_window, a.shift, a.open_y)) ndead++;
                    store_part(P, j, cb);
                }
            }
        }
        __syncthreads();
        /* the tile's particles outside its window: global path, by this block
         * (gathers from the guard-inclusive grids, atomics into Jov); with
         * several species after all main loops, below */
        const int nl = nlist;
        if (ns > 1) {
            if (threadIdx.x == 0) nls[sp] = nl;
        } else if (nl > 0) {
            const GlobalFields gf = {a.E, a.B, a.nrow};
            const GlobalCurrent gj = {a.Jov, a.nrow, a.jov_scale};
            /* whole waves per round: the sorting store needs every lane */
            for (int qb = 0; qb < nl; qb += blockDim.x) {
                const int q = qb + (int)threadIdx.x;
                int i = 0;
                Part c = {ZK_DEAD, 0, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                if (q < nl) {
                    i = a.ovf_list[i0 + q];
                    c = load_part(P, i);
                    int di, dj;
                    kin += push_one(gf, gj, c.ix + 1, c.iy + 1, c.x, c.y, c.ux, c.uy, c.uz, di, dj, k);
                    c.ix += di;
                    c.iy += dj;
                    if (boundaries(c.ix, c.iy, t.nx, t.ny, a.moving_window, a.shift, a.open_y)) ndead++;
                    if (!a.sorting) store_part(P, i, c);
                }
                if (a.sorting) {
                    const int sd = sort_reserve(ss, c.ix != ZK_DEAD, box.has(c.ix, c.iy), &nsl[sp]);
                    sort_write(a, sd, i, c);
                }
            }
            if (threadIdx.x == 0) atomicAdd(&a.ctr->n_overflow, nl);
        }
        /* per-tile energy partial (summed in fixed order by zk_spec_finalize) */
        const double e = zk_block_sum(kin, red);
        if (threadIdx.x == 0) {
            if (ns > 1) ekin[sp] = e;
            else a.epart[tile] = e;
        }
        if (ndead) atomicAdd(&ndead_blk, ndead);
        __syncthreads();
        if (threadIdx.x == 0) {
            if (ndead_blk) atomicAdd(&a.ctr->n_dead, ndead_blk);
            ndead_blk = nlist = 0;
        }
        __syncthreads();
    }
    if (ns > 1) {
        /* each species' out-of-window particles on its own wave, so that
         * their chains of dependent global accesses run side by side
         * (one species after the other, they took 3 % of the launch) */
        const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / ZK_WAVE));
        if (w < ns) {
            const zk_push_args &a = S[w];
            const int nl = nls[w];
            double kin = 0.0;
            int ndead = 0;
            if (nl > 0) {
                const Consts k = {a.tem, a.dt_dx, a.dt_dy, a.qnx, a.qny, a.q};
                const zk_parts P = a.p;
                const GlobalFields gf = {a.E, a.B, a.nrow};
                const GlobalCurrent gj = {a.Jov, a.nrow, a.jov_scale};
                const int i0 = a.tile_off[tile];
                for (int qb = 0; qb < nl; qb += ZK_WAVE) {
                    const int q = qb + (int)__lane_id();
                    int i = 0;
                    Part c = {ZK_DEAD, 0, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                    if (q < nl) {
                        i = a.ovf_list[i0 + q];
                        c = load_part(P, i);
                        int di, dj;
                        kin += push_one(gf, gj, c.ix + 1, c.iy + 1, c.x, c.y, c.ux, c.uy, c.uz, di, dj, k);
                        c.ix += di;
                        c.iy += dj;
                        if (boundaries(c.ix, c.iy, t.nx, t.ny, a.moving_window, a.shift, a.open_y)) ndead++;
                        if (!a.sorting) store_part(P, i, c);
                    }
                    if (a.sorting) {
                        const SortSlots sw = SortSlots::of(a, tile);
                        const int sd = sort_reserve(sw, c.ix != ZK_DEAD, box.has(c.ix, c.iy), &nsl[w]);
                        sort_write(a, sd, i, c);
                    }
                }
                for (int o = ZK_WAVE / 2; o > 0; o >>= 1) {
                    kin += __shfl_down(kin, o, ZK_WAVE);
                    ndead += __shfl_down(ndead, o, ZK_WAVE);
                }
            }
            if (__lane_id() == 0) {
                a.epart[tile] = ekin[w] + kin;
                if (nl > 0) atomicAdd(&a.ctr->n_overflow, nl);
                if (ndead) atomicAdd(&a.ctr->n_dead, ndead);
            }
        }
    }
    /* sorting push: the slots taken per species */
    __syncthreads();
    if ((int)threadIdx.x < ns) {
        const zk_push_args &a = S[threadIdx.x];
        if (a.sorting) {
            const int cap = (a.tile_off[tile + 1] >> 1) - (a.tile_off[tile] >> 1);
            const int nl = (int)(unsigned)(nsl[threadIdx.x] >> 32);
            a.sstay[tile] = (int)(unsigned)nsl[threadIdx.x];
            a.sleave[tile] = min(nl, cap);
            if (nl > cap) atomicAdd(&a.ctr->n_unplaced, nl - cap);
        }
    }
    /* a tile too large for the windows' exact range is counted (the host
     * stops the run when it reads the counter) */
    if (FX && threadIdx.x == 0 && fx_count > ZK_FX_TILE_MAX) atomicAdd(&a0.ctr->n_fx_range, 1);
    /* store the current window of this tile */
    float *dst = a0.region + (size_t)tile * 3 * n;
    for (int m = threadIdx.x; m < 3 * n; m += blockDim.x) {
        const float v = FX ? (float)zk_fx_read(jx[m]) * a0.fx_unscale : jl[m];
        dst[m] = a0.fresh ? v : dst[m] + v;
    }
}

/* Global pass: the unsorted tail (particles injected since the last sort),
 * with gathers from the guard-inclusive field grids as the reference
 * indexes them and atomics into the raw overflow current. */
__device__ inline void push_global_body(const zk_push_args &a, const int block, const int nblocks)
{
    __shared__ double red[16];
    __shared__ int ndead_blk, nov_blk;
    const zk_tiles &t = a.t;
    const Consts k = {a.tem, a.dt_dx, a.dt_dy, a.qnx, a.qny, a.q};
    const GlobalFields gf = {a.E, a.B, a.nrow};
    const GlobalCurrent gj = {a.Jov, a.nrow, a.jov_scale};
    zk_parts P = a.p;
    double kin = 0.0;
    int ndead = 0, nov = 0;
    if (threadIdx.x == 0) ndead_blk = nov_blk = 0;
    /* the unsorted tail */
    const int tail0 = a.ctr->n_sorted;
    const int ntail = a.ctr->np_total - tail0;
    for (int q = block * blockDim.x + threadIdx.x; q < ntail; q += nblocks * blockDim.x) {
        const int i = tail0 + q;
        Part c = load_part(P, i);
        if (c.ix == ZK_DEAD) continue;
        int di, dj;
        nov++;
        kin += push_one(gf, gj, c.ix + 1, c.iy + 1, c.x, c.y, c.ux, c.uy, c.uz, di, dj, k);
        c.ix += di;
        c.iy += dj;
        if (boundaries(c.ix, c.iy, t.nx, t.ny, a.moving_window, a.shift, a.open_y)) ndead++;
        store_part(P, i, c);
    }
    const double s = zk_block_sum(kin, red);
    if (threadIdx.x == 0) a.epart[t.ntiles + block] = s;
    if (ndead) atomicAdd(&ndead_blk, ndead);
    if (nov) atomicAdd(&nov_blk, nov);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (ndead_blk) atomicAdd(&a.ctr->n_dead, ndead_blk);
        if (nov_blk) atomicAdd(&a.ctr->n_overflow, nov_blk);
    }
}

/* Sum the energy partials and count the pushes (em2d/particles.c:1231,
 * 1267).  The order of the sum is fixed (thread t adds partials t, t + B,
 * t + 2B, ... with coalesced loads, then a tree over the B threads; B a
 * power of two up to 1024), so every run sums alike. */
__device__ inline void finalize_body(zk_spec_ctr *ctr, const double *epart, int nparts, double qm, double dx0,
                                     double dx1)
{
    __shared__ double part[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += epart[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = (int)blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double tot = part[0];
        ctr->energy = qm * tot * dx0 * dx1;
        ctr->npush += (unsigned long long)(ctr->np_total - ctr->n_dead - ctr->n_slack);
    }
}

} // namespace zk

#endif
