/*
 * em2d_oracle.c - CPU restatement of the ZPIC em2d time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU library
 * is checked against; nothing in the product links, loads or calls it.
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * use it, and only as a checker or a baseline.
 *
 * It exports the same C API as the reference em2d (types from the headers in include/),
 * written from the reference's documented behaviour with the floating-point
 * operation order of the reference kept, so that on the same inputs it
 * reproduces the reference bit for bit when both are compiled with
 * -O2 -ffp-contract=off (no fused multiply-add, IEEE sqrt and division).
 * Each function names the reference lines it follows.  It is pinned by
 * tests/test_oracle.py against golden vectors written by the reference
 * itself (tests/golden/) and against closed-form checks.
 *
 * Extension for tests (not in the reference): particle identities.  When
 * orc_track_ids(spec) is called the oracle keeps an int64 id per particle,
 * moved along with every reorder (cell sort, absorbing-boundary removal),
 * so particles can be matched one-to-one against the GPU library, which
 * stores its particles in a different (tile) order.
 */
#include <assert.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "simulation.h"
#include "random.h"
#include "timer.h"

/* ------------------------------------------------------------------------
 * timers (reference em2d/timer.c:24-84)
 * ---------------------------------------------------------------------- */
uint64_t timer_ticks(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (uint64_t)tv.tv_sec * 1000000u + (uint64_t)tv.tv_usec;
}

double timer_interval_seconds(uint64_t start, uint64_t end)
{
    return (double)(end - start) * 1.0e-6;
}

double timer_cpu_seconds(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + (double)tv.tv_usec * 1.0e-6;
}

double timer_resolution(void)
{
    struct timeval a, b;
    gettimeofday(&a, NULL);
    do { gettimeofday(&b, NULL); } while (a.tv_usec == b.tv_usec);
    if (a.tv_sec < b.tv_sec) b.tv_usec += 1000000;
    return (double)(b.tv_usec - a.tv_usec) * 1.0e-6;
}

/* ------------------------------------------------------------------------
 * random numbers (reference em2d/random.c:16-101)
 * Marsaglia's two 16-bit multiply-with-carry halves, and the polar
 * Box-Muller method that hands out its second deviate on the next call.
 * ---------------------------------------------------------------------- */
static uint32_t rng_w = 12345u, rng_z = 67890u;
static int rng_have_spare = 0;
static double rng_spare = 0.0;

void set_rand_seed(uint32_t w, uint32_t z)
{
    /* the cached spare deviate is deliberately kept (em2d/random.c:25-29) */
    rng_w = w;
    rng_z = z;
}

/* test extension: drop the cached deviate so independent runs in one
 * process start from the same generator state */
void orc_rng_reset(void) { rng_have_spare = 0; }

static uint32_t mwc32(void)
{
    rng_z = 36969u * (rng_z & 65535u) + (rng_z >> 16);
    rng_w = 18000u * (rng_w & 65535u) + (rng_w >> 16);
    return (rng_z << 16) + rng_w;
}

double rand_norm(void)
{
    if (rng_have_spare) {
        rng_have_spare = 0;
        return rng_spare;
    }
    double a, b, r2;
    for (;;) {
        a = ((double)mwc32() + 0.5) / 2147483649.0 - 1.0;
        b = ((double)mwc32() + 0.5) / 2147483649.0 - 1.0;
        r2 = a * a + b * b;
        if (r2 != 0.0 && r2 < 1.0) break;
    }
    double f = sqrt(-2.0 * log(r2) / r2);
    rng_spare = a * f;
    rng_have_spare = 1;
    return b * f;
}

/* ------------------------------------------------------------------------
 * particle ids (test extension)
 * ---------------------------------------------------------------------- */
struct zpic_dev_species {
    int64_t *id;
    int cap;
    int64_t next;
};

void orc_track_ids(t_species *s)
{
    if (s->dev) return;
    struct zpic_dev_species *d = calloc(1, sizeof(*d));
    d->cap = s->np_max > 0 ? s->np_max : 1;
    d->id = malloc(sizeof(int64_t) * (size_t)d->cap);
    for (int i = 0; i < s->np; i++) d->id[i] = i;
    d->next = s->np;
    s->dev = d;
}

int64_t *orc_ids(t_species *s) { return s->dev ? s->dev->id : NULL; }

static void ids_reserve(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    if (d && d->cap < s->np_max) {
        d->cap = s->np_max;
        d->id = realloc(d->id, sizeof(int64_t) * (size_t)d->cap);
    }
}

/* ------------------------------------------------------------------------
 * species construction and injection (reference em2d/particles.c:60-609)
 * ---------------------------------------------------------------------- */
static double _spec_time = 0.0;
static uint64_t _spec_npush = 0;

double spec_time(void) { return _spec_time; }
uint64_t spec_npush(void) { return _spec_npush; }
double spec_perf(void) { return _spec_npush ? _spec_time / (double)_spec_npush : -1.0; }

static float density_one(float x, void *data) { (void)x; (void)data; return 1.0f; }

void spec_grow_buffer(t_species *s, const int size)
{
    /* em2d/particles.c:460-466: grow in whole blocks of 1024 */
    if (size > s->np_max) {
        s->np_max = (size / 1024 + 1) * 1024;
        s->part = realloc(s->part, sizeof(t_part) * (size_t)s->np_max);
        ids_reserve(s);
    }
}

/* Thermal momenta, per-cell mean removed, plus the fluid drift
 * (em2d/particles.c:85-149).  The cell index uses nx[1] as its row stride,
 * exactly as the reference does.  The reference sizes its per-cell arrays
 * nx[0] * nx[1] (:104-108), so for nx[1] > nx[0] its indices run past them
 * (undefined behaviour); here the arrays cover every index the stride
 * makes, nx[1] * max(nx[0], nx[1]), which gives the reference's values
 * wherever its indices stay in bounds. */
static void inject_momenta(t_species *s, int first, int last)
{
    for (int i = first; i <= last; i++) {
        s->part[i].ux = s->uth[0] * rand_norm();
        s->part[i].uy = s->uth[1] * rand_norm();
        s->part[i].uz = s->uth[2] * rand_norm();
    }
    const int stride = s->nx[1];
    const int ncell = stride * (s->nx[0] > s->nx[1] ? s->nx[0] : s->nx[1]);
    float *mean = calloc((size_t)ncell * 3, sizeof(float));
    int *count = calloc((size_t)ncell, sizeof(int));
    for (int i = first; i <= last; i++) {
        int c = s->part[i].ix + stride * s->part[i].iy;
        mean[3 * c + 0] += s->part[i].ux;
        mean[3 * c + 1] += s->part[i].uy;
        mean[3 * c + 2] += s->part[i].uz;
        count[c] += 1;
    }
    for (int c = 0; c < ncell; c++) {
        float w = (count[c] > 0) ? 1.0f / count[c] : 0;
        mean[3 * c + 0] *= w;
        mean[3 * c + 1] *= w;
        mean[3 * c + 2] *= w;
    }
    for (int i = first; i <= last; i++) {
        int c = s->part[i].ix + stride * s->part[i].iy;
        s->part[i].ux += s->ufl[0] - mean[3 * c + 0];
        s->part[i].uy += s->ufl[1] - mean[3 * c + 1];
        s->part[i].uz += s->ufl[2] - mean[3 * c + 2];
    }
    free(count);
    free(mean);
}

static void put_particle(t_species *s, int *ip, int i, int j, float x, float y)
{
    t_part *p = &s->part[*ip];
    p->ix = i;
    p->iy = j;
    p->x = x;
    p->y = y;
    (*ip)++;
}

/* Particle positions for a cell range (em2d/particles.c:159-354). */
static void inject_positions(t_species *s, const int range[2][2])
{
    const int npc = s->ppc[0] * s->ppc[1];
    const float dpx = 1.0f / s->ppc[0];
    const float dpy = 1.0f / s->ppc[1];
    float *cell = malloc(sizeof(float) * 2 * (size_t)npc);
    int k = 0;
    for (int b = 0; b < s->ppc[1]; b++)
        for (int a = 0; a < s->ppc[0]; a++) {
            cell[k++] = dpx * (a + 0.5);
            cell[k++] = dpy * (b + 0.5);
        }

    int ip = s->np;
    switch (s->density.type) {
    case STEP: {
        float edge = s->density.start / s->dx[0] - s->n_move;
        for (int j = range[1][0]; j <= range[1][1]; j++)
            for (int i = range[0][0]; i <= range[0][1]; i++)
                for (int m = 0; m < npc; m++)
                    if (i + cell[2 * m] > edge) put_particle(s, &ip, i, j, cell[2 * m], cell[2 * m + 1]);
        break;
    }
    case SLAB: {
        float lo = s->density.start / s->dx[0] - s->n_move;
        float hi = s->density.end / s->dx[0] - s->n_move;
        for (int j = range[1][0]; j <= range[1][1]; j++)
            for (int i = range[0][0]; i <= range[0][1]; i++)
                for (int m = 0; m < npc; m++)
                    if (i + cell[2 * m] > lo && i + cell[2 * m] < hi)
                        put_particle(s, &ip, i, j, cell[2 * m], cell[2 * m + 1]);
        break;
    }
    case CUSTOM: {
        /* Inverse of the piecewise-linear cumulative density, cell by cell,
         * continuing along x from the previous injection (moving window). */
        const double hx = s->dx[0], hy = s->dx[1];
        const double px = 1.0 / s->ppc[0], py = 1.0 / s->ppc[1];
        const double threshold = 4 * px * py;
        const int shift = s->n_move;
        const int i0 = range[0][0], j0 = range[1][0];
        unsigned long kx = s->density.custom_x_total_part;
        double cx1 = s->density.custom_x_total_q, cx0;
        double nx1 = (*s->density.custom_x)((i0 + shift) * hx, s->density.custom_data_x), nx0;
        const double ny_first = (*s->density.custom_y)(j0 * hy, s->density.custom_data_y);
        for (int i = i0; i <= range[0][1]; i++) {
            nx0 = nx1;
            nx1 = (*s->density.custom_x)((i + 1 + shift) * hx, s->density.custom_data_x);
            cx0 = cx1;
            cx1 += 0.5 * (nx0 + nx1);
            double tx;
            while ((tx = (kx + 0.5) * px) < cx1) {
                double xp = 2 * (tx - cx0) / (sqrt(nx0 * nx0 + 2 * (nx1 - nx0) * (tx - cx0)) + nx0);
                double dens_x = (0.5 - xp) * nx0 + (0.5 + xp) * nx1;
                unsigned long ky = 0;
                double ny0, ny1 = ny_first, cy0, cy1 = 0;
                for (int j = j0; j <= range[1][1]; j++) {
                    ny0 = ny1;
                    ny1 = (*s->density.custom_y)((j + 1) * hy, s->density.custom_data_y);
                    cy0 = cy1;
                    cy1 += 0.5 * (ny0 + ny1);
                    double ty;
                    while ((ty = (ky + 0.5) * py) < cy1) {
                        double yp = 2 * (ty - cy0) / (sqrt(ny0 * ny0 + 2 * (ny1 - ny0) * (ty - cy0)) + ny0);
                        double dens_y = (0.5 - yp) * ny0 + (0.5 + yp) * ny1;
                        if (dens_x * dens_y > threshold) put_particle(s, &ip, i, j, xp, yp);
                        ky++;
                    }
                }
                kx++;
            }
        }
        s->density.custom_x_total_q = cx1;
        s->density.custom_x_total_part = kx;
        break;
    }
    case EMPTY:
        break;
    default:
        for (int j = range[1][0]; j <= range[1][1]; j++)
            for (int i = range[0][0]; i <= range[0][1]; i++)
                for (int m = 0; m < npc; m++) put_particle(s, &ip, i, j, cell[2 * m], cell[2 * m + 1]);
    }
    s->np = ip;
    free(cell);
}

/* Upper bound of the particles a range receives (em2d/particles.c:367-449),
 * including the reference's use of dx[1] for the slab's upper edge. */
static int inject_count_bound(t_species *s, const int range[2][2])
{
    int n;
    const int rows = range[1][1] - range[1][0] + 1;
    switch (s->density.type) {
    case STEP: {
        int i0 = s->density.start / s->dx[0] - s->n_move;
        if (i0 > range[0][1]) n = 0;
        else {
            if (i0 < range[0][0]) i0 = range[0][0];
            n = (range[0][1] - i0 + 1) * s->ppc[0];
        }
        n *= rows * s->ppc[1];
        break;
    }
    case SLAB: {
        int i0 = s->density.start / s->dx[0] - s->n_move;
        int i1 = s->density.end / s->dx[1] - s->n_move;
        if (i0 > range[0][1] || i1 < range[0][0]) n = 0;
        else {
            if (i0 < range[0][0]) i0 = range[0][0];
            if (i1 > range[0][1]) i1 = range[0][1];
            n = (i1 - i0 + 1) * s->ppc[0];
        }
        n *= rows * s->ppc[1];
        break;
    }
    case CUSTOM: {
        double x = (range[0][0] + s->n_move) * s->dx[0];
        double qx = (*s->density.custom_x)(x, s->density.custom_data_x);
        x = (range[0][1] + 1 + s->n_move) * s->dx[0];
        qx += (*s->density.custom_x)(x, s->density.custom_data_x);
        qx *= 0.5;
        for (int i = range[0][0] + 1; i <= range[0][1]; i++) {
            x = (i + s->n_move) * s->dx[0];
            qx += (*s->density.custom_x)(x, s->density.custom_data_x);
        }
        double y = range[1][0] * s->dx[1];
        double qy = (*s->density.custom_y)(y, s->density.custom_data_y);
        y = (range[1][1] + 1) * s->dx[1];
        qy += (*s->density.custom_y)(y, s->density.custom_data_y);
        qy *= 0.5;
        for (int j = range[1][0] + 1; j <= range[1][1]; j++) {
            y = j * s->dx[1];
            qy += (*s->density.custom_y)(y, s->density.custom_data_y);
        }
        n = ceil(qx * s->ppc[0]) * ceil(qy * s->ppc[1]);
        break;
    }
    case EMPTY:
        n = 0;
        break;
    default:
        n = (range[0][1] - range[0][0] + 1) * s->ppc[0] * rows * s->ppc[1];
    }
    return n;
}

static void inject(t_species *s, const int range[2][2])
{
    const int first = s->np;
    spec_grow_buffer(s, s->np + inject_count_bound(s, range));
    inject_positions(s, range);
    inject_momenta(s, first, s->np - 1);
    if (s->dev) {
        for (int i = first; i < s->np; i++) s->dev->id[i] = s->dev->next++;
    }
}

void spec_new(t_species *s, char name[], const float m_q, const int ppc[],
              const float ufl[], const float uth[],
              const int nx[], float box[], const float dt, t_density *density)
{
    /* em2d/particles.c:535-609 */
    strncpy(s->name, name, MAX_SPNAME_LEN);
    s->name[MAX_SPNAME_LEN] = 0;
    int npc = 1;
    for (int d = 0; d < 2; d++) {
        s->nx[d] = nx[d];
        s->ppc[d] = ppc[d];
        npc *= ppc[d];
        s->box[d] = box[d];
        s->dx[d] = box[d] / nx[d];
    }
    s->m_q = m_q;
    s->q = copysign(1.0f, m_q) / npc;
    s->dt = dt;
    s->np_max = 0;
    s->part = NULL;
    s->dev = NULL;
    if (density) {
        s->density = *density;
    } else {
        memset(&s->density, 0, sizeof(s->density));
        s->density.type = UNIFORM;
        s->density.n = 1.0;
    }
    if (s->density.n == 0.) s->density.n = 1.0;
    if (s->density.type == CUSTOM) {
        if (!s->density.custom_x) s->density.custom_x = density_one;
        if (!s->density.custom_y) s->density.custom_y = density_one;
        s->density.custom_x_total_part = 0;
        s->density.custom_x_total_q = 0;
    }
    s->q *= fabsf(s->density.n);
    for (int d = 0; d < 3; d++) {
        s->ufl[d] = ufl ? ufl[d] : 0;
        s->uth[d] = uth ? uth[d] : 0;
    }
    s->iter = 0;
    s->moving_window = 0;
    s->n_move = 0;
    s->np = 0;
    const int range[2][2] = {{0, nx[0] - 1}, {0, nx[1] - 1}};
    inject(s, range);
    s->n_sort = 16;
}

void spec_delete(t_species *s)
{
    free(s->part);
    s->part = NULL;
    if (s->dev) {
        free(s->dev->id);
        free(s->dev);
        s->dev = NULL;
    }
    s->np = -1;
}

void spec_move_window(t_species *s)
{
    /* em2d/particles.c:619-640 */
    if ((s->iter * s->dt) > (s->dx[0] * (s->n_move + 1))) {
        for (int i = 0; i < s->np; i++) s->part[i].ix--;
        s->n_move++;
        const int range[2][2] = {{s->nx[0] - 1, s->nx[0] - 1}, {0, s->nx[1] - 1}};
        inject(s, range);
    }
}

/* ------------------------------------------------------------------------
 * charge-conserving current deposit (reference em2d/particles.c:773-924)
 * The motion is cut at cell edges into at most three straight segments;
 * each segment adds Jx at two nodes, Jy at two nodes and Jz at four.
 * ---------------------------------------------------------------------- */
typedef struct {
    float x0, x1, y0, y1, dx, dy, qvz;
    int ix, iy;
} segment;

static void deposit_segments(int ix, int iy, int di, int dj, float x0, float y0,
                             float dx, float dy, float qnx, float qny, float qvz,
                             float3 *J, int nrow)
{
    segment sg[3];
    int n = 1;
    sg[0].x0 = x0;
    sg[0].y0 = y0;
    sg[0].dx = dx;
    sg[0].dy = dy;
    sg[0].x1 = x0 + dx;
    sg[0].y1 = y0 + dy;
    sg[0].qvz = qvz / 2.0;
    sg[0].ix = ix;
    sg[0].iy = iy;

    if (di != 0) {
        /* cut at the x edge the particle crossed */
        int edge = (di == 1);
        float f = (x0 + dx - edge) / dx;
        sg[1].x0 = 1 - edge;
        sg[1].x1 = (x0 + dx) - di;
        sg[1].dx = dx * f;
        sg[1].ix = ix + di;
        float yc = y0 + dy * (1.0f - f);
        sg[1].y0 = yc;
        sg[1].y1 = sg[0].y1;
        sg[1].dy = dy * f;
        sg[1].iy = iy;
        sg[1].qvz = sg[0].qvz * f;
        sg[0].x1 = edge;
        sg[0].dx *= (1.0f - f);
        sg[0].dy *= (1.0f - f);
        sg[0].y1 = yc;
        sg[0].qvz *= (1.0f - f);
        n++;
    }
    if (dj != 0) {
        /* the segment that crosses the y edge: the first one if it already
         * leaves the cell in y, the second one otherwise */
        int c = 1 - (sg[0].y1 < 0.0f || sg[0].y1 >= 1.0f);
        int edge = (dj == 1);
        float f = (sg[c].y1 - edge) / sg[c].dy;
        sg[n].y0 = 1 - edge;
        sg[n].y1 = sg[c].y1 - dj;
        sg[n].dy = sg[c].dy * f;
        sg[n].iy = sg[c].iy + dj;
        float xc = sg[c].x0 + sg[c].dx * (1.0f - f);
        sg[n].x0 = xc;
        sg[n].x1 = sg[c].x1;
        sg[n].dx = sg[c].dx * f;
        sg[n].ix = sg[c].ix;
        sg[n].qvz = sg[c].qvz * f;
        sg[c].y1 = edge;
        sg[c].dy *= (1.0f - f);
        sg[c].dx *= (1.0f - f);
        sg[c].x1 = xc;
        sg[c].qvz *= (1.0f - f);
        if (c < n - 1) {
            /* the x-cut segment lies after the y crossing */
            sg[1].y0 -= dj;
            sg[1].y1 -= dj;
            sg[1].iy += dj;
        }
        n++;
    }

    for (int k = 0; k < n; k++) {
        const segment *g = &sg[k];
        float ax0 = 1.0f - g->x0, bx0 = g->x0;
        float ax1 = 1.0f - g->x1, bx1 = g->x1;
        float ay0 = 1.0f - g->y0, by0 = g->y0;
        float ay1 = 1.0f - g->y1, by1 = g->y1;
        float jx = qnx * g->dx;
        float jy = qny * g->dy;
        float wy0 = 0.5f * (ay0 + ay1), wy1 = 0.5f * (by0 + by1);
        float wx0 = 0.5f * (ax0 + ax1), wx1 = 0.5f * (bx0 + bx1);
        float3 *c00 = &J[g->ix + nrow * g->iy];
        float3 *c10 = c00 + 1;
        float3 *c01 = c00 + nrow;
        float3 *c11 = c01 + 1;
        c00->x += jx * wy0;
        c01->x += jx * wy1;
        c00->y += jy * wx0;
        c10->y += jy * wx1;
        c00->z += g->qvz * (ax0 * ay0 + ax1 * ay1 + (ax0 * ay1 - ax1 * ay0) / 2.0f);
        c10->z += g->qvz * (bx0 * ay0 + bx1 * ay1 + (bx0 * ay1 - bx1 * ay0) / 2.0f);
        c01->z += g->qvz * (ax0 * by0 + ax1 * by1 + (ax0 * by1 - ax1 * by0) / 2.0f);
        c11->z += g->qvz * (bx0 * by0 + bx1 * by1 + (bx0 * by1 - bx1 * by0) / 2.0f);
    }
}

/* ------------------------------------------------------------------------
 * cell sort (reference em2d/particles.c:942-1007): stable counting sort on
 * the key ix + iy*nx[0].  Any stable sort yields the same order; this one
 * scatters through a scratch buffer.
 * ---------------------------------------------------------------------- */
static void spec_sort(t_species *s)
{
    const int ncell = s->nx[0] * s->nx[1];
    int *start = calloc((size_t)ncell + 1, sizeof(int));
    for (int i = 0; i < s->np; i++) start[s->part[i].ix + s->part[i].iy * s->nx[0] + 1]++;
    for (int c = 0; c < ncell; c++) start[c + 1] += start[c];
    t_part *out = malloc(sizeof(t_part) * (size_t)(s->np > 0 ? s->np : 1));
    int64_t *oid = s->dev ? malloc(sizeof(int64_t) * (size_t)(s->np > 0 ? s->np : 1)) : NULL;
    for (int i = 0; i < s->np; i++) {
        int dst = start[s->part[i].ix + s->part[i].iy * s->nx[0]]++;
        out[dst] = s->part[i];
        if (oid) oid[dst] = s->dev->id[i];
    }
    memcpy(s->part, out, sizeof(t_part) * (size_t)s->np);
    if (oid) {
        memcpy(s->dev->id, oid, sizeof(int64_t) * (size_t)s->np);
        free(oid);
    }
    free(out);
    free(start);
}

/* ------------------------------------------------------------------------
 * field gather on the staggered grid (reference em2d/particles.c:1029-1071)
 * Ex, By live at (i+1/2, j); Ey, Bx at (i, j+1/2); Ez at (i, j);
 * Bz at (i+1/2, j+1/2).
 * ---------------------------------------------------------------------- */
static void gather(const float3 *E, const float3 *B, int nrow, const t_part *p,
                   float3 *e, float3 *b)
{
    const int i = p->ix, j = p->iy;
    const float wx = p->x, wy = p->y;
    const int ih = i + ((wx < 0.5f) ? -1 : 0);
    const int jh = j + ((wy < 0.5f) ? -1 : 0);
    const float wxh = wx + ((wx < 0.5f) ? 0.5f : -0.5f);
    const float wyh = wy + ((wy < 0.5f) ? 0.5f : -0.5f);

#define ROW(F, c, a, bb, w) (F[(a) + (bb) * nrow].c * (1.0f - (w)) + F[(a) + 1 + (bb) * nrow].c * (w))
    e->x = ROW(E, x, ih, j, wxh) * (1.0f - wy) + ROW(E, x, ih, j + 1, wxh) * wy;
    e->y = ROW(E, y, i, jh, wx) * (1.0f - wyh) + ROW(E, y, i, jh + 1, wx) * wyh;
    e->z = ROW(E, z, i, j, wx) * (1.0f - wy) + ROW(E, z, i, j + 1, wx) * wy;
    b->x = ROW(B, x, i, jh, wx) * (1.0f - wyh) + ROW(B, x, i, jh + 1, wx) * wyh;
    b->y = ROW(B, y, ih, j, wxh) * (1.0f - wy) + ROW(B, y, ih, j + 1, wxh) * wy;
    b->z = ROW(B, z, ih, jh, wxh) * (1.0f - wyh) + ROW(B, z, ih, jh + 1, wxh) * wyh;
#undef ROW
}

static inline int cells_moved(float x) { return (x >= 1.0f) - (x < 0.0f); }

/* ------------------------------------------------------------------------
 * one species step (reference em2d/particles.c:1104-1269)
 * ---------------------------------------------------------------------- */
void spec_advance(t_species *s, t_emf *emf, t_current *current)
{
    uint64_t t0 = timer_ticks();
    const float tem = 0.5 * s->dt / s->m_q;
    const float dt_dx = s->dt / s->dx[0];
    const float dt_dy = s->dt / s->dx[1];
    const float qnx = s->q * s->dx[0] / s->dt;
    const float qny = s->q * s->dx[1] / s->dt;
    const int nx0 = s->nx[0], nx1 = s->nx[1];
    double kin = 0;

    for (int i = 0; i < s->np; i++) {
        t_part *p = &s->part[i];
        float3 e, b;
        gather(emf->E_part, emf->B_part, emf->nrow, p, &e, &b);

        /* relativistic Boris rotation between two half electric kicks */
        e.x *= tem;
        e.y *= tem;
        e.z *= tem;
        float ax = p->ux + e.x, ay = p->uy + e.y, az = p->uz + e.z;
        float usq = ax * ax + ay * ay + az * az;
        float g = sqrtf(1.0f + usq);
        kin += usq / (g + 1);
        float tg = tem / g;
        b.x *= tg;
        b.y *= tg;
        b.z *= tg;
        float s2 = 2.0f / (1.0f + b.x * b.x + b.y * b.y + b.z * b.z);
        float vx = ax + ay * b.z - az * b.y;
        float vy = ay + az * b.x - ax * b.z;
        float vz = az + ax * b.y - ay * b.x;
        b.x *= s2;
        b.y *= s2;
        b.z *= s2;
        ax += vy * b.z - vz * b.y;
        ay += vz * b.x - vx * b.z;
        az += vx * b.y - vy * b.x;
        float ux = ax + e.x, uy = ay + e.y, uz = az + e.z;
        p->ux = ux;
        p->uy = uy;
        p->uz = uz;

        /* move, and count the cells crossed (at most one per axis) */
        float rg = 1.0f / sqrtf(1.0f + ux * ux + uy * uy + uz * uz);
        float mx = dt_dx * rg * ux;
        float my = dt_dy * rg * uy;
        float x1 = p->x + mx, y1 = p->y + my;
        int di = cells_moved(x1), dj = cells_moved(y1);
        x1 -= di;
        y1 -= dj;
        float qvz = s->q * uz * rg;
        deposit_segments(p->ix, p->iy, di, dj, p->x, p->y, mx, my, qnx, qny, qvz,
                         current->J, current->nrow);
        p->x = x1;
        p->y = y1;
        p->ix += di;
        p->iy += dj;
    }
    s->energy = s->q * s->m_q * kin * s->dx[0] * s->dx[1];
    s->iter += 1;

    if (s->moving_window) {
        spec_move_window(s);
        /* absorbing in x: the last particle fills the hole; periodic in y */
        int i = 0;
        while (i < s->np) {
            if (s->part[i].ix < 0 || s->part[i].ix >= nx0) {
                --s->np;
                s->part[i] = s->part[s->np];
                if (s->dev) s->dev->id[i] = s->dev->id[s->np];
                continue;
            }
            s->part[i].iy += ((s->part[i].iy < 0) ? nx1 : 0) - ((s->part[i].iy >= nx1) ? nx1 : 0);
            i++;
        }
    } else {
        for (int i = 0; i < s->np; i++) {
            s->part[i].ix += ((s->part[i].ix < 0) ? nx0 : 0) - ((s->part[i].ix >= nx0) ? nx0 : 0);
            s->part[i].iy += ((s->part[i].iy < 0) ? nx1 : 0) - ((s->part[i].iy >= nx1) ? nx1 : 0);
        }
    }
    if (s->n_sort > 0 && !(s->iter % s->n_sort)) spec_sort(s);

    _spec_npush += (uint64_t)s->np;
    _spec_time += timer_interval_seconds(t0, timer_ticks());
}

/* ------------------------------------------------------------------------
 * diagnostics from particles (reference em2d/particles.c:1289-1324, 1569-1632)
 * ---------------------------------------------------------------------- */
void spec_deposit_charge(const t_species *s, float *rho)
{
    const int nrow = s->nx[0] + 1;
    const float q = s->q;
    for (int i = 0; i < s->np; i++) {
        int c = s->part[i].ix + nrow * s->part[i].iy;
        float wx = s->part[i].x, wy = s->part[i].y;
        rho[c] += (1.0f - wx) * (1.0f - wy) * q;
        rho[c + 1] += (wx) * (1.0f - wy) * q;
        rho[c + nrow] += (1.0f - wx) * (wy) * q;
        rho[c + 1 + nrow] += (wx) * (wy) * q;
    }
    if (!s->moving_window)
        for (int j = 0; j < s->nx[1] + 1; j++) rho[j * nrow] += rho[s->nx[0] + j * nrow];
    for (int i = 0; i < s->nx[0] + 1; i++) rho[i] += rho[i + s->nx[1] * nrow];
}

static float pha_value(const t_species *s, int i, int quant)
{
    switch (quant) {
    case X1: return (s->part[i].x + s->part[i].ix) * s->dx[0];
    case X2: return (s->part[i].y + s->part[i].iy) * s->dx[1];
    case U1: return s->part[i].ux;
    case U2: return s->part[i].uy;
    case U3: return s->part[i].uz;
    }
    return 0;
}

void spec_deposit_pha(const t_species *s, const int rep_type, const int pha_nx[],
                      const float pha_range[][2], float *buf)
{
    const int nrow = pha_nx[0];
    const int qa = rep_type & 0x000F, qb = (rep_type & 0x00F0) >> 4;
    const float amin = pha_range[0][0], bmin = pha_range[1][0];
    const float ra = pha_nx[0] / (pha_range[0][1] - pha_range[0][0]);
    const float rb = pha_nx[1] / (pha_range[1][1] - pha_range[1][0]);
    for (int i = 0; i < s->np; i++) {
        float na = (pha_value(s, i, qa) - amin) * ra;
        float nb = (pha_value(s, i, qb) - bmin) * rb;
        int ia = (int)(na + 0.5f), ib = (int)(nb + 0.5f);
        float wa = na - ia + 0.5f, wb = nb - ib + 0.5f;
        int c = ia + nrow * ib;
        if (ib >= 0 && ib < pha_nx[1]) {
            if (ia >= 0 && ia < pha_nx[0]) buf[c] += (1.0f - wa) * (1.0f - wb) * s->q;
            if (ia + 1 >= 0 && ia + 1 < pha_nx[0]) buf[c + 1] += wa * (1.0f - wb) * s->q;
        }
        c += nrow;
        if (ib + 1 >= 0 && ib + 1 < pha_nx[1]) {
            if (ia >= 0 && ia < pha_nx[0]) buf[c] += (1.0f - wa) * wb * s->q;
            if (ia + 1 >= 0 && ia + 1 < pha_nx[0]) buf[c + 1] += wa * wb * s->q;
        }
    }
}

/* ZDF file output is outside the time step; the oracle writes nothing. */
void spec_report(const t_species *s, const int rep_type, const int pha_nx[], const float pha_range[][2])
{
    (void)s; (void)rep_type; (void)pha_nx; (void)pha_range;
}

/* ------------------------------------------------------------------------
 * current density (reference em2d/current.c:30-459)
 * ---------------------------------------------------------------------- */
static size_t grid_cells(const int nx[2], const int gc[2][2])
{
    return (size_t)(gc[0][0] + nx[0] + gc[0][1]) * (size_t)(gc[1][0] + nx[1] + gc[1][1]);
}

void current_new(t_current *c, int nx[], float box[], float dt)
{
    const int gc[2][2] = {{1, 2}, {1, 2}};
    for (int d = 0; d < 2; d++) {
        c->nx[d] = nx[d];
        c->gc[d][0] = gc[d][0];
        c->gc[d][1] = gc[d][1];
        c->box[d] = box[d];
        c->dx[d] = box[d] / nx[d];
    }
    c->nrow = gc[0][0] + nx[0] + gc[0][1];
    c->J_buf = malloc(grid_cells(c->nx, c->gc) * sizeof(float3));
    assert(c->J_buf);
    c->J = c->J_buf + gc[0][0] + gc[1][0] * c->nrow;
    c->smooth.xtype = NONE;
    c->smooth.ytype = NONE;
    c->smooth.xlevel = 0;
    c->smooth.ylevel = 0;
    c->iter = 0;
    c->dt = dt;
    c->moving_window = 0;
    c->dev = NULL;
    current_zero(c);
}

void current_delete(t_current *c)
{
    free(c->J_buf);
    c->J_buf = NULL;
}

void current_zero(t_current *c)
{
    memset(c->J_buf, 0, grid_cells(c->nx, c->gc) * sizeof(float3));
}

static void f3_add(float3 *a, const float3 *b)
{
    a->x += b->x;
    a->y += b->y;
    a->z += b->z;
}

/* Periodic fold of the guard cells (em2d/current.c:118-159): the lower
 * cells receive the upper images, the upper images are then copies. */
static void current_fold(t_current *c)
{
    const int nrow = c->nrow;
    float3 *J = c->J;
    if (!c->moving_window) {
        for (int j = -c->gc[1][0]; j < c->nx[1] + c->gc[1][1]; j++) {
            for (int i = -c->gc[0][0]; i < c->gc[0][1]; i++) f3_add(&J[i + j * nrow], &J[c->nx[0] + i + j * nrow]);
            for (int i = -c->gc[0][0]; i < c->gc[0][1]; i++) J[c->nx[0] + i + j * nrow] = J[i + j * nrow];
        }
    }
    for (int i = -c->gc[0][0]; i < c->nx[0] + c->gc[0][1]; i++) {
        for (int j = -c->gc[1][0]; j < c->gc[1][1]; j++) f3_add(&J[i + j * nrow], &J[i + (c->nx[1] + j) * nrow]);
        for (int j = -c->gc[1][0]; j < c->gc[1][1]; j++) J[i + (c->nx[1] + j) * nrow] = J[i + j * nrow];
    }
}

/* Compensator weights for n binomial passes (em2d/current.c:297-304). */
static void compensator(int n, float *sa, float *sb)
{
    float a = -1;
    float b = (4.0 + 2.0 * n) / n;
    float t = 2 * a + b;
    *sa = a / t;
    *sb = b / t;
}

static inline float3 stencil3(float sa, float sb, float3 l, float3 m, float3 u)
{
    float3 r;
    r.x = sa * l.x + sb * m.x + sa * u.x;
    r.y = sa * l.y + sb * m.y + sa * u.y;
    r.z = sa * l.z + sb * m.z + sa * u.z;
    return r;
}

/* [sa,sb,sa] along x on the interior rows, then the x guards of those rows
 * (em2d/current.c:316-354). */
static void smooth_x(t_current *c, float sa, float sb)
{
    float3 *J = c->J;
    const int nrow = c->nrow;
    for (int j = 0; j < c->nx[1]; j++) {
        float3 *row = J + j * nrow;
        float3 left = row[-1], mid = row[0];
        for (int i = 0; i < c->nx[0]; i++) {
            float3 right = row[i + 1];
            row[i] = stencil3(sa, sb, left, mid, right);
            left = mid;
            mid = right;
        }
        if (!c->moving_window) {
            for (int i = -c->gc[0][0]; i < 0; i++) row[i] = row[c->nx[0] + i];
            for (int i = 0; i < c->gc[0][1]; i++) row[c->nx[0] + i] = row[i];
        }
    }
}

/* [sa,sb,sa] along y on the interior, then the y guards of every column
 * (em2d/current.c:366-413). */
static void smooth_y(t_current *c, float sa, float sb)
{
    float3 *J = c->J;
    const int nrow = c->nrow;
    float3 *below = malloc(sizeof(float3) * (size_t)c->nx[0]);
    for (int i = 0; i < c->nx[0]; i++) below[i] = J[i - nrow];
    for (int j = 0; j < c->nx[1]; j++) {
        for (int i = 0; i < c->nx[0]; i++) {
            float3 mid = J[j * nrow + i];
            float3 up = J[j * nrow + i + nrow];
            J[j * nrow + i] = stencil3(sa, sb, below[i], mid, up);
            below[i] = mid;
        }
    }
    free(below);
    for (int i = -c->gc[0][0]; i < c->nx[0] + c->gc[0][1]; i++) {
        for (int j = -c->gc[1][0]; j < 0; j++) J[i + j * nrow] = J[i + (c->nx[1] + j) * nrow];
        for (int j = 0; j < c->gc[1][1]; j++) J[i + (c->nx[1] + j) * nrow] = J[i + j * nrow];
    }
}

static void current_smooth(t_current *c)
{
    /* em2d/current.c:427-459; the y binomial pass count is xlevel, as in
     * the reference (current.c:449) */
    float sa, sb;
    if (c->smooth.xtype != NONE) {
        for (int k = 0; k < c->smooth.xlevel; k++) smooth_x(c, 0.25, 0.5);
        if (c->smooth.xtype == COMPENSATED) {
            compensator(c->smooth.xlevel, &sa, &sb);
            smooth_x(c, sa, sb);
        }
    }
    if (c->smooth.ytype != NONE) {
        for (int k = 0; k < c->smooth.xlevel; k++) smooth_y(c, 0.25, 0.5);
        if (c->smooth.ytype == COMPENSATED) {
            compensator(c->smooth.ylevel, &sa, &sb);
            smooth_y(c, sa, sb);
        }
    }
}

void current_update(t_current *c)
{
    current_fold(c);
    current_smooth(c);
    c->iter++;
}

void current_report(const t_current *c, const int jc)
{
    if (jc < 0 || jc > 2) {
        fprintf(stderr, "(*error*) Invalid current component (jc) selected, returning\n");
        return;
    }
    (void)c;
}

/* ------------------------------------------------------------------------
 * EM fields (reference em2d/emf.c:56-980)
 * ---------------------------------------------------------------------- */
static double _emf_time = 0.0;
double emf_time(void) { return _emf_time; }

static void emf_update_gc(t_emf *emf);
static void emf_update_part_fld(t_emf *emf);

void emf_new(t_emf *emf, int nx[], float box[], const float dt)
{
    const int gc[2][2] = {{1, 2}, {1, 2}};
    for (int d = 0; d < 2; d++) {
        emf->nx[d] = nx[d];
        emf->gc[d][0] = gc[d][0];
        emf->gc[d][1] = gc[d][1];
        emf->box[d] = box[d];
        emf->dx[d] = box[d] / nx[d];
    }
    emf->nrow = gc[0][0] + nx[0] + gc[0][1];
    size_t bytes = grid_cells(emf->nx, emf->gc) * sizeof(float3);
    emf->E_buf = calloc(1, bytes);
    emf->B_buf = calloc(1, bytes);
    assert(emf->E_buf && emf->B_buf);
    emf->E = emf->E_buf + gc[0][0] + gc[1][0] * emf->nrow;
    emf->B = emf->B_buf + gc[0][0] + gc[1][0] * emf->nrow;
    emf->dt = dt;
    emf->iter = 0;
    emf->moving_window = 0;
    emf->n_move = 0;
    memset(&emf->ext_fld, 0, sizeof(emf->ext_fld));
    emf->ext_fld.E_type = EMF_FLD_TYPE_NONE;
    emf->ext_fld.B_type = EMF_FLD_TYPE_NONE;
    emf->E_part = emf->E;
    emf->B_part = emf->B;
    emf->dev = NULL;
}

void emf_delete(t_emf *emf)
{
    free(emf->E_buf);
    free(emf->B_buf);
    emf->E_buf = emf->B_buf = NULL;
    if (emf->ext_fld.E_type > EMF_FLD_TYPE_NONE) free(emf->ext_fld.E_part_buf);
    if (emf->ext_fld.B_type > EMF_FLD_TYPE_NONE) free(emf->ext_fld.B_part_buf);
    emf->E_part = emf->B_part = NULL;
}

/* Gaussian beam phase and transverse profile (em2d/emf.c:158-170). */
static float beam_phase(const t_emf_laser *L, const float zpos, const float r)
{
    float z = zpos - L->focus;
    float zr = L->omega0 * (L->W0 * L->W0) / 2;
    float r2 = r * r;
    float curv = r2 * z / (zr * zr + z * z);
    float w2 = (zr * zr) / (zr * zr + z * z);
    float gouy = atan2(z, zr);
    return sqrt(sqrt(w2)) * exp(-r2 * w2 / (L->W0 * L->W0)) * cos(L->omega0 * (z + curv) - gouy);
}

/* sin^2 rise, flat top, sin^2 fall (em2d/emf.c:179-202). */
static float pulse_envelope(const t_emf_laser *L, const float z)
{
    if (z > L->start) return 0.0;
    if (z > L->start - L->rise) {
        float u = z - L->start;
        float e = sin(M_PI_2 * u / L->rise);
        return e * e;
    }
    if (z > L->start - (L->rise + L->flat)) return 1.0;
    if (z > L->start - (L->rise + L->flat + L->fall)) {
        float u = z - (L->start - L->rise - L->flat - L->fall);
        float e = sin(M_PI_2 * u / L->fall);
        return e * e;
    }
    return 0.0;
}

/* Longitudinal components from div E = div B = 0, integrated from the right
 * edge (em2d/emf.c:213-231). */
static void longitudinal_fields(t_emf *emf)
{
    float3 *E = emf->E, *B = emf->B;
    const int nrow = emf->nrow;
    const double ratio = emf->dx[0] / emf->dx[1];
    for (int j = 0; j < emf->nx[1]; j++) {
        double ex = 0.0, bx = 0.0;
        for (int i = emf->nx[0] - 1; i >= 0; i--) {
            ex += ratio * (E[i + 1 + j * nrow].y - E[i + 1 + (j - 1) * nrow].y);
            E[i + j * nrow].x = ex;
            bx += ratio * (B[i + (j + 1) * nrow].y - B[i + j * nrow].y);
            B[i + j * nrow].x = bx;
        }
    }
}

void emf_add_laser(t_emf *const emf, t_emf_laser *L)
{
    /* em2d/emf.c:242-350 */
    if (L->fwhm != 0) {
        if (L->fwhm <= 0) {
            fprintf(stderr, "Invalid laser FWHM, must be > 0, aborting.\n");
            exit(-1);
        }
        L->rise = L->fwhm;
        L->fall = L->fwhm;
        L->flat = 0.;
    }
    if (L->rise <= 0) {
        fprintf(stderr, "Invalid laser RISE, must be > 0, aborting.\n");
        exit(-1);
    }
    if (L->flat < 0) {
        fprintf(stderr, "Invalid laser FLAT, must be >= 0, aborting.\n");
        exit(-1);
    }
    if (L->fall <= 0) {
        fprintf(stderr, "Invalid laser FALL, must be > 0, aborting.\n");
        exit(-1);
    }
    float3 *E = emf->E, *B = emf->B;
    const int nrow = emf->nrow;
    const float hx = emf->dx[0], hy = emf->dx[1];
    const float axis = L->axis;
    const float amp = L->omega0 * L->a0;
    const float cp = cos(L->polarization);
    const float sp = sin(L->polarization);
    if (L->type == PLANE) {
        const float k = L->omega0;
        for (int i = 0; i < emf->nx[0]; i++) {
            float z = i * hx;
            float zh = z + hx / 2;
            float env = amp * pulse_envelope(L, z);
            float envh = amp * pulse_envelope(L, zh);
            for (int j = 0; j < emf->nx[1]; j++) {
                E[i + j * nrow].y += +env * cos(k * z) * cp;
                E[i + j * nrow].z += +env * cos(k * z) * sp;
                B[i + j * nrow].y += -envh * cos(k * zh) * sp;
                B[i + j * nrow].z += +envh * cos(k * zh) * cp;
            }
        }
    } else if (L->type == GAUSSIAN) {
        for (int i = 0; i < emf->nx[0]; i++) {
            float z = i * hx;
            float zh = z + hx / 2;
            float env = amp * pulse_envelope(L, z);
            float envh = amp * pulse_envelope(L, zh);
            for (int j = 0; j < emf->nx[1]; j++) {
                float r = j * hy - axis;
                float rh = r + hy / 2;
                E[i + j * nrow].y += +env * beam_phase(L, z, rh) * cp;
                E[i + j * nrow].z += +env * beam_phase(L, z, r) * sp;
                B[i + j * nrow].y += -envh * beam_phase(L, zh, r) * sp;
                B[i + j * nrow].z += +envh * beam_phase(L, zh, rh) * cp;
            }
        }
        longitudinal_fields(emf);
    }
    emf_update_gc(emf);
}

void emf_report(const t_emf *emf, const char field, const int fc)
{
    if (fc < 0 || fc > 2) {
        fprintf(stderr, "(*error*) Invalid field component (fc) selected, returning\n");
        return;
    }
    (void)emf; (void)field;
}

/* B += dt * (-curl E), over i in [-1,nx], j in [-1,ny] (em2d/emf.c:500-522) */
static void yee_b(t_emf *emf, const float dt)
{
    float3 *B = emf->B;
    const float3 *E = emf->E;
    const float dt_dx = dt / emf->dx[0], dt_dy = dt / emf->dx[1];
    const int nrow = emf->nrow;
    for (int j = -1; j <= emf->nx[1]; j++)
        for (int i = -1; i <= emf->nx[0]; i++) {
            const float3 *e = &E[i + j * nrow];
            const float3 *ex = e + 1, *ey = e + nrow;
            float3 *b = &B[i + j * nrow];
            b->x += (-dt_dy * (ey->z - e->z));
            b->y += (dt_dx * (ex->z - e->z));
            b->z += (-dt_dx * (ex->y - e->y) + dt_dy * (ey->x - e->x));
        }
}

/* E += dt * (curl B - J), over i in [0,nx+1], j in [0,ny+1] (em2d/emf.c:531-562) */
static void yee_e(t_emf *emf, const t_current *cur, const float dt)
{
    float3 *E = emf->E;
    const float3 *B = emf->B, *J = cur->J;
    const float dt_dx = dt / emf->dx[0], dt_dy = dt / emf->dx[1];
    const int nrow = emf->nrow, nrj = cur->nrow;
    for (int j = 0; j <= emf->nx[1] + 1; j++)
        for (int i = 0; i <= emf->nx[0] + 1; i++) {
            const float3 *b = &B[i + j * nrow];
            const float3 *bx = b - 1, *by = b - nrow;
            const float3 *jc = &J[i + j * nrj];
            float3 *e = &E[i + j * nrow];
            e->x += (+dt_dy * (b->z - by->z)) - dt * jc->x;
            e->y += (-dt_dx * (b->z - bx->z)) - dt * jc->y;
            e->z += (+dt_dx * (b->y - bx->y) - dt_dy * (b->x - by->x)) - dt * jc->z;
        }
}

/* Periodic guard copies: x (unless the window moves), then y over every
 * column (em2d/emf.c:573-637). */
static void emf_update_gc(t_emf *emf)
{
    const int nrow = emf->nrow;
    float3 *E = emf->E, *B = emf->B;
    if (!emf->moving_window) {
        for (int j = -emf->gc[1][0]; j < emf->nx[1] + emf->gc[1][1]; j++) {
            for (int i = -emf->gc[0][0]; i < 0; i++) {
                E[i + j * nrow] = E[emf->nx[0] + i + j * nrow];
                B[i + j * nrow] = B[emf->nx[0] + i + j * nrow];
            }
            for (int i = 0; i < emf->gc[0][1]; i++) {
                E[emf->nx[0] + i + j * nrow] = E[i + j * nrow];
                B[emf->nx[0] + i + j * nrow] = B[i + j * nrow];
            }
        }
    }
    for (int i = -emf->gc[0][0]; i < emf->nx[0] + emf->gc[0][1]; i++) {
        for (int j = -emf->gc[1][0]; j < 0; j++) {
            E[i + j * nrow] = E[i + (emf->nx[1] + j) * nrow];
            B[i + j * nrow] = B[i + (emf->nx[1] + j) * nrow];
        }
        for (int j = 0; j < emf->gc[1][1]; j++) {
            E[i + (emf->nx[1] + j) * nrow] = E[i + j * nrow];
            B[i + (emf->nx[1] + j) * nrow] = B[i + j * nrow];
        }
    }
}

/* Shift every row one cell left and clear the last columns
 * (em2d/emf.c:648-675). */
static void emf_move_window(t_emf *emf)
{
    if ((emf->iter * emf->dt) > emf->dx[0] * (emf->n_move + 1)) {
        const int nrow = emf->nrow;
        float3 *E = emf->E, *B = emf->B;
        const float3 zero = {0., 0., 0.};
        for (int j = -emf->gc[1][0]; j < emf->nx[1] + emf->gc[1][1]; j++) {
            for (int i = -emf->gc[0][0]; i < emf->nx[0] + emf->gc[0][1] - 1; i++) {
                E[i + j * nrow] = E[i + j * nrow + 1];
                B[i + j * nrow] = B[i + j * nrow + 1];
            }
            for (int i = emf->nx[0] - 1; i < emf->nx[0] + emf->gc[0][1]; i++) {
                E[i + j * nrow] = zero;
                B[i + j * nrow] = zero;
            }
        }
        emf->n_move++;
    }
}

void emf_advance(t_emf *emf, const t_current *cur)
{
    uint64_t t0 = timer_ticks();
    const float dt = emf->dt;
    yee_b(emf, dt / 2.0f);
    yee_e(emf, cur, dt);
    yee_b(emf, dt / 2.0f);
    emf_update_gc(emf);
    emf_update_part_fld(emf);
    emf->iter += 1;
    if (emf->moving_window) emf_move_window(emf);
    _emf_time += timer_interval_seconds(t0, timer_ticks());
}

void emf_get_energy(const t_emf *emf, double energy[])
{
    /* em2d/emf.c:729-750 */
    const float3 *E = emf->E, *B = emf->B;
    const int nrow = emf->nrow;
    for (int k = 0; k < 6; k++) energy[k] = 0;
    for (int j = 0; j < emf->nx[1]; j++)
        for (int i = 0; i < emf->nx[0]; i++) {
            const float3 e = E[i + j * nrow], b = B[i + j * nrow];
            energy[0] += e.x * e.x;
            energy[1] += e.y * e.y;
            energy[2] += e.z * e.z;
            energy[3] += b.x * b.x;
            energy[4] += b.y * b.y;
            energy[5] += b.z * b.z;
        }
    for (int k = 0; k < 6; k++) energy[k] *= 0.5 * emf->dx[0] * emf->dx[1];
}

static float3 *part_grid(t_emf *emf)
{
    float3 *buf = malloc(grid_cells(emf->nx, emf->gc) * sizeof(float3));
    assert(buf);
    return buf;
}

void emf_set_ext_fld(t_emf *const emf, t_emf_ext_fld *ext)
{
    /* em2d/emf.c:764-832 */
    emf->ext_fld.E_type = ext->E_type;
    if (emf->ext_fld.E_type == EMF_FLD_TYPE_NONE) {
        emf->E_part = emf->E;
        emf->ext_fld.E_part_buf = NULL;
    } else {
        if (ext->E_type == EMF_FLD_TYPE_UNIFORM) emf->ext_fld.E_0 = ext->E_0;
        else if (ext->E_type == EMF_FLD_TYPE_CUSTOM) {
            emf->ext_fld.E_custom = ext->E_custom;
            emf->ext_fld.E_custom_data = ext->E_custom_data;
        } else {
            fprintf(stderr, "Invalid external field type, aborting.\n");
            exit(-1);
        }
        emf->ext_fld.E_part_buf = part_grid(emf);
        emf->E_part = emf->ext_fld.E_part_buf + emf->gc[0][0] + emf->gc[1][0] * emf->nrow;
    }
    emf->ext_fld.B_type = ext->B_type;
    if (emf->ext_fld.B_type == EMF_FLD_TYPE_NONE) {
        emf->B_part = emf->B;
        emf->ext_fld.B_part_buf = NULL;
    } else {
        if (ext->B_type == EMF_FLD_TYPE_UNIFORM) emf->ext_fld.B_0 = ext->B_0;
        else if (ext->B_type == EMF_FLD_TYPE_CUSTOM) {
            emf->ext_fld.B_custom = ext->B_custom;
            emf->ext_fld.B_custom_data = ext->B_custom_data;
        } else {
            fprintf(stderr, "Invalid external field type, aborting.\n");
            exit(-1);
        }
        emf->ext_fld.B_part_buf = part_grid(emf);
        emf->B_part = emf->ext_fld.B_part_buf + emf->gc[0][0] + emf->gc[1][0] * emf->nrow;
    }
    emf_update_part_fld(emf);
}

/* E_part = E + external, B_part = B + external over the whole buffer
 * (em2d/emf.c:838-914). */
static void emf_update_part_fld(t_emf *emf)
{
    const int nrow = emf->nrow;
    const t_emf_ext_fld *x = &emf->ext_fld;
    for (int f = 0; f < 2; f++) {
        enum emf_fld_type type = f ? x->B_type : x->E_type;
        if (type == EMF_FLD_TYPE_NONE) continue;
        const float3 *src = f ? emf->B : emf->E;
        float3 *dst = f ? emf->B_part : emf->E_part;
        for (int j = -emf->gc[1][0]; j < emf->nx[1] + emf->gc[1][1]; j++)
            for (int i = -emf->gc[0][0]; i < emf->nx[0] + emf->gc[0][1]; i++) {
                float3 add;
                if (type == EMF_FLD_TYPE_UNIFORM) add = f ? x->B_0 : x->E_0;
                else if (f) add = (*x->B_custom)(i, emf->dx[0], j, emf->dx[1], x->B_custom_data);
                else add = (*x->E_custom)(i, emf->dx[0], j, emf->dx[1], x->E_custom_data);
                float3 v = src[i + j * nrow];
                v.x += add.x;
                v.y += add.y;
                v.z += add.z;
                dst[i + j * nrow] = v;
            }
    }
}

void emf_init_fld(t_emf *const emf, t_emf_init_fld *init)
{
    /* em2d/emf.c:922-980 */
    if (emf->iter != 0) {
        fprintf(stderr, "emf_init_fld should only be called at initialization, aborting...\n");
        exit(-1);
    }
    const int nrow = emf->nrow;
    for (int f = 0; f < 2; f++) {
        enum emf_fld_type type = f ? init->B_type : init->E_type;
        if (type == EMF_FLD_TYPE_NONE) continue;
        float3 *dst = f ? emf->B : emf->E;
        for (int j = -emf->gc[1][0]; j < emf->nx[1] + emf->gc[1][1]; j++)
            for (int i = -emf->gc[0][0]; i < emf->nx[0] + emf->gc[0][1]; i++) {
                if (type == EMF_FLD_TYPE_UNIFORM) dst[j * nrow + i] = f ? init->B_0 : init->E_0;
                else if (f) dst[j * nrow + i] = (init->B_custom)(i, emf->dx[0], j, emf->dx[1], init->B_custom_data);
                else dst[j * nrow + i] = (init->E_custom)(i, emf->dx[0], j, emf->dx[1], init->E_custom_data);
            }
    }
}

/* ------------------------------------------------------------------------
 * simulation driver (reference em2d/simulation.c:25-223)
 * ---------------------------------------------------------------------- */
int report(int n, int ndump)
{
    return (ndump > 0) ? !(n % ndump) : 0;
}

void sim_iter(t_simulation *sim)
{
    current_zero(&sim->current);
    for (int i = 0; i < sim->n_species; i++) spec_advance(&sim->species[i], &sim->emf, &sim->current);
    current_update(&sim->current);
    emf_advance(&sim->emf, &sim->current);
}

void sim_timings(t_simulation *sim, uint64_t t0, uint64_t t1)
{
    (void)sim;
    fprintf(stderr, "Time for spec. advance = %f s\n", spec_time());
    fprintf(stderr, "Time for emf   advance = %f s\n", emf_time());
    fprintf(stderr, "Total simulation time  = %f s\n", timer_interval_seconds(t0, t1));
    fprintf(stderr, "\n");
    double perf = spec_perf();
    if (perf > 0) {
        fprintf(stderr, "Particle advance [nsec/part] = %f \n", 1.e9 * perf);
        fprintf(stderr, "Particle advance [Mpart/sec] = %f \n", 1.e-6 / perf);
    }
}

void sim_new(t_simulation *sim, int nx[], float box[], float dt, float tmax, int ndump,
             t_species *species, int n_species)
{
    sim->dt = dt;
    sim->tmax = tmax;
    sim->ndump = ndump;
    emf_new(&sim->emf, nx, box, dt);
    current_new(&sim->current, nx, box, dt);
    sim->n_species = n_species;
    sim->species = species;
    sim->moving_window = 0;
    float cour = sqrtf(1.0f / (1.0f / (sim->emf.dx[0] * sim->emf.dx[0]) + 1.0f / (sim->emf.dx[1] * sim->emf.dx[1])));
    if (dt >= cour) {
        fprintf(stderr, "Invalid timestep, courant condition violation, dtmax = %f \n", cour);
        exit(-1);
    }
}

void sim_add_laser(t_simulation *sim, t_emf_laser *laser) { emf_add_laser(&sim->emf, laser); }

void sim_set_smooth(t_simulation *sim, t_smooth *smooth)
{
    if ((smooth->xtype != NONE) && (smooth->xlevel <= 0)) {
        fprintf(stderr, "Invalid smooth level along x direction\n");
        exit(-1);
    }
    if ((smooth->ytype != NONE) && (smooth->ylevel <= 0)) {
        fprintf(stderr, "Invalid smooth level along y direction\n");
        exit(-1);
    }
    sim->current.smooth = *smooth;
}

void sim_set_moving_window(t_simulation *sim)
{
    sim->emf.moving_window = 1;
    sim->current.moving_window = 1;
    for (int i = 0; i < sim->n_species; i++) sim->species[i].moving_window = 1;
}

void sim_set_ext_fld(t_simulation *sim, t_emf_ext_fld *ext) { emf_set_ext_fld(&sim->emf, ext); }

void sim_report_energy(t_simulation *sim)
{
    /* the first field component is counted twice, as in
     * em2d/simulation.c:191-194 */
    double e[6];
    emf_get_energy(&sim->emf, e);
    double fld = e[0];
    for (int k = 0; k < 6; k++) fld += e[k];
    double kin = 0;
    for (int i = 0; i < sim->n_species; i++) kin += sim->species[i].energy;
    printf("Energy (fields | particles | total) = %e %e %e\n", fld, kin, fld + kin);
}

void sim_delete(t_simulation *sim)
{
    for (int i = 0; i < sim->n_species; i++) spec_delete(&sim->species[i]);
    free(sim->species);
    current_delete(&sim->current);
    emf_delete(&sim->emf);
}
