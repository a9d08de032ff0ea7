/*
 * particles.c - particle species of the em2d time step on the GPU.
 *
 * Host side of the reference em2d/particles.c: species construction and
 * particle injection run on the host with the reference's arithmetic and
 * random-number order (so initial states and moving-window injections are
 * bit-identical to the reference); the push, deposit, boundary conditions
 * and sort of spec_advance run as HIP kernels on particles kept in device
 * memory, grouped by tile.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "random.h"
#include "timer.h"
#include "zpic_internal.h"
#include "zpic_rng.h"
#include "zpic_zdf.h"

/* all live species, for the process-wide push counter; spec_new may run
 * on several host threads at once (independent simulations initialised
 * in parallel), so the list is guarded */
static struct zpic_dev_species *g_live = NULL;
static unsigned long long g_npush_deleted = 0;
static pthread_mutex_t g_live_lock = PTHREAD_MUTEX_INITIALIZER;

/* ------------------------------------------------------------------------
 * injection on the host (em2d/particles.c:85-492)
 * ---------------------------------------------------------------------- */
static float unit_density(float x, void *data)
{
    (void)x;
    (void)data;
    return 1.0f;
}

/* Upper bound of the particles a cell range receives (em2d/particles.c:367-449);
 * the slab's upper edge divides by dx[1], as the reference does. */
static int inject_bound(const t_species *s, const int r[2][2])
{
    const int rows = r[1][1] - r[1][0] + 1;
    switch (s->density.type) {
    case STEP: {
        int i0 = s->density.start / s->dx[0] - s->n_move;
        if (i0 > r[0][1]) return 0;
        if (i0 < r[0][0]) i0 = r[0][0];
        return (r[0][1] - i0 + 1) * s->ppc[0] * rows * s->ppc[1];
    }
    case SLAB: {
        int i0 = s->density.start / s->dx[0] - s->n_move;
        int i1 = s->density.end / s->dx[1] - s->n_move;
        if (i0 > r[0][1] || i1 < r[0][0]) return 0;
        if (i0 < r[0][0]) i0 = r[0][0];
        if (i1 > r[0][1]) i1 = r[0][1];
        return (i1 - i0 + 1) * s->ppc[0] * rows * s->ppc[1];
    }
    case CUSTOM: {
        const t_density *d = &s->density;
        /* trapezoid sums: end points halved, then the inner points */
        double qx = (*d->custom_x)((r[0][0] + s->n_move) * s->dx[0], d->custom_data_x);
        qx += (*d->custom_x)((r[0][1] + 1 + s->n_move) * s->dx[0], d->custom_data_x);
        qx *= 0.5;
        for (int i = r[0][0] + 1; i <= r[0][1]; i++) qx += (*d->custom_x)((i + s->n_move) * s->dx[0], d->custom_data_x);
        double qy = (*d->custom_y)(r[1][0] * s->dx[1], d->custom_data_y);
        qy += (*d->custom_y)((r[1][1] + 1) * s->dx[1], d->custom_data_y);
        qy *= 0.5;
        for (int j = r[1][0] + 1; j <= r[1][1]; j++) qy += (*d->custom_y)(j * s->dx[1], d->custom_data_y);
        return ceil(qx * s->ppc[0]) * ceil(qy * s->ppc[1]);
    }
    case EMPTY:
        return 0;
    default:
        return (r[0][1] - r[0][0] + 1) * s->ppc[0] * rows * s->ppc[1];
    }
}

/* Lattice positions of a cell range for the density profile; particles are
 * written at out[0..), the count is returned (em2d/particles.c:159-354). */
static int inject_positions(t_species *s, const int r[2][2], t_part *out)
{
    const int npc = s->ppc[0] * s->ppc[1];
    const float hx = 1.0f / s->ppc[0], hy = 1.0f / s->ppc[1];
    float *lat = malloc(sizeof(float) * 2 * (size_t)(npc > 0 ? npc : 1));
    for (int b = 0, k = 0; b < s->ppc[1]; b++)
        for (int a = 0; a < s->ppc[0]; a++, k++) {
            lat[2 * k] = hx * (a + 0.5);
            lat[2 * k + 1] = hy * (b + 0.5);
        }
    int n = 0;
#define EMIT(I, J, X, Y)                                                                                               \
    do {                                                                                                               \
        out[n].ix = (I);                                                                                               \
        out[n].iy = (J);                                                                                               \
        out[n].x = (X);                                                                                                \
        out[n].y = (Y);                                                                                                \
        n++;                                                                                                           \
    } while (0)
    if (s->density.type == STEP || s->density.type == SLAB) {
        const float lo = s->density.start / s->dx[0] - s->n_move;
        const float hi = s->density.end / s->dx[0] - s->n_move;
        const int slab = (s->density.type == SLAB);
        for (int j = r[1][0]; j <= r[1][1]; j++)
            for (int i = r[0][0]; i <= r[0][1]; i++)
                for (int k = 0; k < npc; k++)
                    if (i + lat[2 * k] > lo && (!slab || i + lat[2 * k] < hi)) EMIT(i, j, lat[2 * k], lat[2 * k + 1]);
    } else if (s->density.type == CUSTOM) {
        /* invert the trapezoidal cumulative density along x, continuing
         * from earlier injections, and along y from the bottom */
        t_density *d = &s->density;
        const double hxd = s->dx[0], hyd = s->dx[1];
        const double px = 1.0 / s->ppc[0], py = 1.0 / s->ppc[1];
        const double thr = 4 * px * py;
        unsigned long kx = d->custom_x_total_part;
        double cx1 = d->custom_x_total_q;
        double fx1 = (*d->custom_x)((r[0][0] + s->n_move) * hxd, d->custom_data_x);
        const double fy_first = (*d->custom_y)(r[1][0] * hyd, d->custom_data_y);
        for (int i = r[0][0]; i <= r[0][1]; i++) {
            const double fx0 = fx1;
            fx1 = (*d->custom_x)((i + 1 + s->n_move) * hxd, d->custom_data_x);
            const double cx0 = cx1;
            cx1 += 0.5 * (fx0 + fx1);
            double tx;
            while ((tx = (kx + 0.5) * px) < cx1) {
                const double xp = 2 * (tx - cx0) / (sqrt(fx0 * fx0 + 2 * (fx1 - fx0) * (tx - cx0)) + fx0);
                const double wxd = (0.5 - xp) * fx0 + (0.5 + xp) * fx1;
                unsigned long ky = 0;
                double fy1 = fy_first, cy1 = 0;
                for (int j = r[1][0]; j <= r[1][1]; j++) {
                    const double fy0 = fy1;
                    fy1 = (*d->custom_y)((j + 1) * hyd, d->custom_data_y);
                    const double cy0 = cy1;
                    cy1 += 0.5 * (fy0 + fy1);
                    double ty;
                    while ((ty = (ky + 0.5) * py) < cy1) {
                        const double yp = 2 * (ty - cy0) / (sqrt(fy0 * fy0 + 2 * (fy1 - fy0) * (ty - cy0)) + fy0);
                        const double wyd = (0.5 - yp) * fy0 + (0.5 + yp) * fy1;
                        if (wxd * wyd > thr) EMIT(i, j, (float)xp, (float)yp);
                        ky++;
                    }
                }
                kx++;
            }
        }
        d->custom_x_total_q = cx1;
        d->custom_x_total_part = kx;
    } else if (s->density.type != EMPTY) {
        for (int j = r[1][0]; j <= r[1][1]; j++)
            for (int i = r[0][0]; i <= r[0][1]; i++)
                for (int k = 0; k < npc; k++) EMIT(i, j, lat[2 * k], lat[2 * k + 1]);
    }
#undef EMIT
    free(lat);
    return n;
}

/* Thermal momenta with the per-cell mean removed, plus the fluid drift
 * (em2d/particles.c:85-149).  The cell key uses nx[1] as the row stride,
 * as the reference does. */
/* Momenta of n injected particles (em2d/particles.c:85-149); w: scratch of
 * the cell means kept between calls (NULL: a temporary one). */
static void inject_momenta(const t_species *s, t_part *p, int n, zpic_inj_scratch *w)
{
    for (int i = 0; i < n; i++) {
        p[i].ux = s->uth[0] * rand_norm();
        p[i].uy = s->uth[1] * rand_norm();
        p[i].uz = s->uth[2] * rand_norm();
    }
    /* the row stride is nx[1], as in the reference; the arrays cover every
     * index it makes (the reference's nx[0] * nx[1] arrays are overrun
     * when nx[1] > nx[0]).  Only the cells that hold new particles are
     * used, so only those are cleared and scaled (the reference clears
     * and scales every cell, which gives the same values for these). */
    const int cells = s->nx[1] * (s->nx[0] > s->nx[1] ? s->nx[0] : s->nx[1]);
    zpic_inj_scratch local = {NULL, NULL, 0};
    if (!w) w = &local;
    if (w->cells < cells) {
        free(w->mean);
        free(w->count);
        w->mean = malloc(sizeof(float3) * (size_t)cells);
        w->count = malloc(sizeof(int) * (size_t)cells);
        if (!w->mean || !w->count) zpic_fatal("inject_momenta: out of host memory");
        w->cells = cells;
    }
    float3 *mean = w->mean;
    int *count = w->count;
    for (int i = 0; i < n; i++) {
        const int c = p[i].ix + s->nx[1] * p[i].iy;
        mean[c].x = mean[c].y = mean[c].z = 0.0f;
        count[c] = 0;
    }
    for (int i = 0; i < n; i++) {
        const int c = p[i].ix + s->nx[1] * p[i].iy;
        mean[c].x += p[i].ux;
        mean[c].y += p[i].uy;
        mean[c].z += p[i].uz;
        count[c] += 1;
    }
    for (int i = 0; i < n; i++) {
        const int c = p[i].ix + s->nx[1] * p[i].iy;
        if (count[c] > 0) { /* scale each cell once; the sign marks it done */
            const float wgt = 1.0f / count[c];
            mean[c].x *= wgt;
            mean[c].y *= wgt;
            mean[c].z *= wgt;
            count[c] = -count[c];
        }
    }
    for (int i = 0; i < n; i++) {
        const int c = p[i].ix + s->nx[1] * p[i].iy;
        p[i].ux += s->ufl[0] - mean[c].x;
        p[i].uy += s->ufl[1] - mean[c].y;
        p[i].uz += s->ufl[2] - mean[c].z;
    }
    if (w == &local) {
        free(local.mean);
        free(local.count);
    }
}

static void grow_sets(struct zpic_dev_species *d, int need);

/* Logs of the polar method that the device initialisation could not
 * certify as correctly rounded (expected to stay 0). */
static long long g_init_ambiguous = 0;
long long zpic_device_init_ambiguous(void) { return g_init_ambiguous + zpic_rng_device_ambiguous(); }

#define ZPIC_MZ 2422800383u /* 36969 * 2^16 - 1 */
#define ZPIC_MW 1179647999u /* 18000 * 2^16 - 1 */

static uint32_t host_powmod(uint32_t a, uint64_t e, uint32_t m)
{
    uint64_t r = 1 % m, b = a % m;
    while (e) {
        if (e & 1u) r = (r * b) % m;
        b = (b * b) % m;
        e >>= 1;
    }
    return (uint32_t)r;
}

/* The particles of spec_new made on the device, for the uniform, step and
 * slab profiles: the same positions, thermal momenta (the generator's
 * stream continued from its present state, which is advanced past the
 * deviates used) and mean removal as inject_positions and inject_momenta,
 * left unsorted in the device set until the species' tiles are chosen
 * (zpic_species_ensure).  The host mirror is filled on spec_sync_host. */
static void device_init(t_species *s, const int r[2][2])
{
    struct zpic_dev_species *d = s->dev;
    const int npc = s->ppc[0] * s->ppc[1];
    const float hx = 1.0f / s->ppc[0], hy = 1.0f / s->ppc[1];
    const int c0 = r[0][0], ncols = r[0][1] - r[0][0] + 1, rows = r[1][1] - r[1][0] + 1;
    float *lat = malloc(sizeof(float) * 2 * (size_t)(npc > 0 ? npc : 1));
    int *tab_i = malloc(sizeof(int) * (size_t)ncols * (size_t)(npc > 0 ? npc : 1));
    int *tab_k = malloc(sizeof(int) * (size_t)ncols * (size_t)(npc > 0 ? npc : 1));
    int *cstart = malloc(sizeof(int) * (size_t)ncols), *ccount = malloc(sizeof(int) * (size_t)ncols);
    if (!lat || !tab_i || !tab_k || !cstart || !ccount) zpic_fatal("spec_new: out of host memory");
    for (int b = 0, k = 0; b < s->ppc[1]; b++)
        for (int a = 0; a < s->ppc[0]; a++, k++) {
            lat[2 * k] = hx * (a + 0.5);
            lat[2 * k + 1] = hy * (b + 0.5);
        }
    /* one row of cells: the (column, lattice point) of each particle, in
     * the order inject_positions emits them */
    const float lo = s->density.start / s->dx[0] - s->n_move;
    const float hi = s->density.end / s->dx[0] - s->n_move;
    const int cut = s->density.type != UNIFORM, slab = s->density.type == SLAB;
    int perrow = 0;
    for (int i = r[0][0]; i <= r[0][1]; i++) {
        cstart[i - c0] = perrow;
        for (int k = 0; k < npc; k++)
            if (!cut || (i + lat[2 * k] > lo && (!slab || i + lat[2 * k] < hi))) {
                tab_i[perrow] = i;
                tab_k[perrow] = k;
                perrow++;
            }
        ccount[i - c0] = perrow - cstart[i - c0];
    }
    const long long np = (long long)perrow * rows;
    if (np > ZK_MAX_PARTS) zpic_fatal("species '%s' needs %lld particles; at most %d are supported", s->name, np,
                                      ZK_MAX_PARTS);
    s->np = (int)np;
    spec_grow_buffer(s, s->np);
    d->cap = 0;
    d->cur = 0;
    grow_sets(d, s->np > 0 ? s->np : 1);

    if (np > 0) {
        zk_init_args a;
        memset(&a, 0, sizeof(a));
        a.np = np;
        a.perrow = perrow;
        a.row0 = r[1][0];
        a.rows = rows;
        a.col0 = c0;
        a.ncols = ncols;
        int *dtab = zpic_dmalloc(sizeof(int) * 2 * (size_t)perrow + sizeof(int) * 2 * (size_t)ncols);
        float *dlat = zpic_dmalloc(sizeof(float) * 2 * (size_t)npc);
        zpic_h2d(dtab, tab_i, sizeof(int) * (size_t)perrow);
        zpic_h2d(dtab + perrow, tab_k, sizeof(int) * (size_t)perrow);
        zpic_h2d(dtab + 2 * perrow, cstart, sizeof(int) * (size_t)ncols);
        zpic_h2d(dtab + 2 * perrow + ncols, ccount, sizeof(int) * (size_t)ncols);
        zpic_h2d(dlat, lat, sizeof(float) * 2 * (size_t)npc);
        a.tab_i = dtab;
        a.tab_k = dtab + perrow;
        a.col_start = dtab + 2 * perrow;
        a.col_count = dtab + 2 * perrow + ncols;
        a.lat = dlat;
        const zk_parts *P = &d->set[d->cur];
        a.cell = P->cell;
        a.cbits = P->cbits;
        a.x = P->x;
        a.y = P->y;
        a.ux = P->ux;
        a.uy = P->uy;
        a.uz = P->uz;
        a.stride = s->nx[1];
        for (int c = 0; c < 3; c++) {
            a.uth[c] = s->uth[c];
            a.ufl[c] = s->ufl[c];
        }
        ZCHECK(zk_init_positions(&a, zpic_stream()));

        /* thermal momenta: 3 np deviates, the first one kept from before */
        zpic_rng_state *st = zpic_rng();
        const long long D = 3 * np;
        a.z0 = st->z;
        a.w0 = st->w;
        a.z1 = 36969u * (st->z & 0xffffu) + (st->z >> 16);
        a.w1 = 18000u * (st->w & 0xffffu) + (st->w >> 16);
        a.kept0 = st->have_kept ? 1 : 0;
        a.kept_in = st->kept;
        a.pairs = (D - a.kept0 + 1) / 2;
        const int per = 64;
        long long nat = (long long)(a.pairs * 1.2732395447351628 * 1.002) + 65536; /* 4/pi attempts per pair */
        for (;;) {
            const long long nthr = (nat + per - 1) / per;
            const int nblk = (int)((nthr + 255) / 256);
            int *thr = zpic_dmalloc(sizeof(int) * 256 * (size_t)nblk);
            int *bcnt = zpic_dmalloc(sizeof(int) * (size_t)nblk);
            ZCHECK(zk_init_rng_count(&a, nat, per, thr, bcnt, nblk, zpic_stream()));
            int *hb = malloc(sizeof(int) * (size_t)nblk);
            long long *hoff = malloc(sizeof(long long) * (size_t)nblk);
            zpic_d2h(hb, bcnt, sizeof(int) * (size_t)nblk);
            long long tot = 0;
            for (int b = 0; b < nblk; b++) {
                hoff[b] = tot;
                tot += hb[b];
            }
            if (tot >= a.pairs) {
                long long *doff = zpic_dmalloc(sizeof(long long) * (size_t)nblk);
                zk_init_out *dout = zpic_dmalloc(sizeof(zk_init_out));
                const zk_init_out zero = {-1, 0.0, 0, 0};
                zpic_h2d(dout, &zero, sizeof(zero));
                zpic_h2d(doff, hoff, sizeof(long long) * (size_t)nblk);
                ZCHECK(zk_init_rng_emit(&a, nat, per, thr, doff, nblk, dout, zpic_stream()));
                zk_init_out out;
                zpic_d2h(&out, dout, sizeof(out));
                zpic_dfree(doff);
                zpic_dfree(dout);
                if (out.last_attempt < 0) zpic_fatal("spec_new: device initialisation lost its last deviate pair");
                /* the generator after the calls of the last attempt used */
                const uint64_t n = 2 * (uint64_t)out.last_attempt + 2;
                st->z = (uint32_t)(((uint64_t)host_powmod(36969u, n - 1, ZPIC_MZ) * a.z1) % ZPIC_MZ);
                st->w = (uint32_t)(((uint64_t)host_powmod(18000u, n - 1, ZPIC_MW) * a.w1) % ZPIC_MW);
                st->have_kept = out.have_kept;
                st->kept = out.have_kept ? out.kept : 0.0;
                g_init_ambiguous += out.ambiguous;
            }
            free(hb);
            free(hoff);
            zpic_dfree(thr);
            zpic_dfree(bcnt);
            if (tot >= a.pairs) break;
            nat = nat + nat / 16 + 65536;
        }
        /* mean removal: keys ix + nx[1] iy */
        const long long nkeys = (long long)(r[0][1] + 1) + (long long)s->nx[1] * r[1][1];
        float *mean = zpic_dmalloc(sizeof(float) * 3 * (size_t)nkeys);
        ZCHECK(zk_init_mean(&a, mean, (int)nkeys, zpic_stream()));
        zpic_sync();
        zpic_dfree(mean);
        zpic_dfree(dtab);
        zpic_dfree(dlat);
    }
    free(lat);
    free(tab_i);
    free(tab_k);
    free(cstart);
    free(ccount);
    d->np_upper = s->np;
    d->dev_unsorted = 1;
    d->host_dirty = 0;
    d->host_valid = 0;
}

void spec_grow_buffer(t_species *s, const int size)
{
    if (size > s->np_max) {
        s->np_max = (size / 1024 + 1) * 1024;
        s->part = realloc(s->part, sizeof(t_part) * (size_t)s->np_max);
        if (!s->part) zpic_fatal("out of host memory for %d particles", s->np_max);
    }
}

/* ------------------------------------------------------------------------
 * construction
 * ---------------------------------------------------------------------- */
void spec_new(t_species *s, char name[], const float m_q, const int ppc[], const float ufl[], const float uth[],
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
    if (density) {
        s->density = *density;
    } else {
        memset(&s->density, 0, sizeof(s->density));
        s->density.type = UNIFORM;
        s->density.n = 1.0;
    }
    if (s->density.n == 0.) s->density.n = 1.0;
    if (s->density.type == CUSTOM) {
        if (!s->density.custom_x) s->density.custom_x = unit_density;
        if (!s->density.custom_y) s->density.custom_y = unit_density;
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
    s->energy = 0;
    s->np = 0;

    const int r[2][2] = {{0, nx[0] - 1}, {0, nx[1] - 1}};
    struct zpic_dev_species *d = calloc(1, sizeof(*d));
    s->dev = d;
    d->cbits = zpic_cell_bits(nx[0], nx[1]);
    const int dens = s->density.type;
    if (zpic_device_init() && (dens == UNIFORM || dens == STEP || dens == SLAB)) {
        device_init(s, r);
    } else {
        spec_grow_buffer(s, inject_bound(s, r));
        s->np = inject_positions(s, r, s->part);
        inject_momenta(s, s->part, s->np, NULL);
        d->host_dirty = 1;
        d->host_valid = 1;
    }
    s->n_sort = 16;

    d->auto_sort = 1;
    d->owner = s;
    pthread_mutex_lock(&g_live_lock);
    d->next_live = g_live;
    g_live = d;
    pthread_mutex_unlock(&g_live_lock);
}

/* The species' device counters.  A push that met a tile with more
 * particles than the fixed-point current windows sum exactly
 * (ZK_FX_TILE_MAX, all species of the launch) has counted it: the current
 * of that step is not exact, so the run stops here. */
static void ctr_check(const zk_spec_ctr *c)
{
    if (c->n_fx_range > 0)
        zpic_fatal("a tile holds more than %d particles, the range of the fixed-point current windows; "
                   "use smaller tiles (zpic_set_tiling)", ZK_FX_TILE_MAX);
}

static void ctr_read(struct zpic_dev_species *d, zk_spec_ctr *c)
{
    zpic_d2h(c, d->ctr, sizeof(*c));
    ctr_check(c);
}

static void species_free_device(struct zpic_dev_species *d)
{
    if (d->ready) {
        zk_spec_ctr c;
        zpic_d2h(&c, d->ctr, sizeof(c));
        d->npush_retired += c.npush;
    }
    for (int k = 0; k < 2; k++) {
        zk_parts *p = &d->set[k];
        zpic_dfree(p->cell);
        zpic_dfree(p->x);
        zpic_dfree(p->y);
        zpic_dfree(p->ux);
        zpic_dfree(p->uy);
        zpic_dfree(p->uz);
        zpic_dfree(p->id);
        memset(p, 0, sizeof(*p));
    }
    zpic_dfree(d->ovf_list);
    d->ovf_list = NULL;
    zpic_dfree(d->tile_n);
    zpic_dfree(d->tile_n2);
    zpic_dfree(d->sstay);
    zpic_dfree(d->sleave);
    d->tile_n = d->tile_n2 = d->sstay = d->sleave = NULL;
    d->counted = 0;
    zpic_dfree(d->tile_off);
    zpic_dfree(d->tile_off2);
    zpic_dfree(d->stay);
    zpic_dfree(d->nbr);
    zpic_dfree(d->tailn);
    d->tile_off2 = d->stay = d->nbr = d->tailn = NULL;
    zpic_dfree(d->cnt);
    zpic_dfree(d->cursor);
    zpic_dfree(d->epart);
    zpic_dfree(d->stage);
    zpic_dfree(d->ctr);
    if (d->probe_ev) zpic_event_wait(d->probe_ev);
    zpic_event_free(d->probe_ev);
    zpic_pinned_free(d->probe_host);
    d->probe_ev = d->probe_host = NULL;
    d->probe_ready = 0;
    for (int k = 0; k < ZPIC_PIN_SLOTS; k++) {
        if (d->pin_ev[k]) zpic_event_wait(d->pin_ev[k]);
        zpic_event_free(d->pin_ev[k]);
        zpic_pinned_free(d->pin[k]);
        d->pin_ev[k] = d->pin[k] = d->pin_dev[k] = NULL;
        d->pin_cap[k] = 0;
    }
    d->pin_next = 0;
    d->tile_off = d->cnt = d->cursor = NULL;
    d->epart = NULL;
    d->stage = NULL;
    d->ctr = NULL;
    d->ready = 0;
}

static unsigned long long device_npush(struct zpic_dev_species *d)
{
    if (!d->ready) return d->npush_retired;
    zk_spec_ctr c;
    ctr_read(d, &c);
    return d->npush_retired + c.npush;
}

void spec_delete(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    if (d) {
        zpic_sync();
        species_free_device(d);
        pthread_mutex_lock(&g_live_lock);
        g_npush_deleted += d->npush_retired;
        struct zpic_dev_species **pp = &g_live;
        while (*pp && *pp != d) pp = &(*pp)->next_live;
        if (*pp) *pp = d->next_live;
        pthread_mutex_unlock(&g_live_lock);
        free(d->host_ids);
        free(d->inj.mean);
        free(d->inj.count);
        free(d);
        s->dev = NULL;
    }
    free(s->part);
    s->part = NULL;
    s->np = -1;
}

uint64_t spec_npush(void)
{
    pthread_mutex_lock(&g_live_lock);
    unsigned long long n = g_npush_deleted;
    for (struct zpic_dev_species *d = g_live; d; d = d->next_live) n += device_npush(d);
    pthread_mutex_unlock(&g_live_lock);
    return n;
}

double spec_time(void) { return zpic_timer_total(ZPIC_T_SPEC); }

double spec_perf(void)
{
    const uint64_t n = spec_npush();
    return n > 0 ? spec_time() / (double)n : -1.0;
}

/* ------------------------------------------------------------------------
 * device residency
 * ---------------------------------------------------------------------- */
static void alloc_set(zk_parts *p, int cap, int ids, int cbits)
{
    p->cell = zpic_dmalloc(sizeof(int) * (size_t)cap);
    p->cbits = cbits;
    p->x = zpic_dmalloc(sizeof(float) * (size_t)cap);
    p->y = zpic_dmalloc(sizeof(float) * (size_t)cap);
    p->ux = zpic_dmalloc(sizeof(float) * (size_t)cap);
    p->uy = zpic_dmalloc(sizeof(float) * (size_t)cap);
    p->uz = zpic_dmalloc(sizeof(float) * (size_t)cap);
    p->id = ids ? zpic_dmalloc(sizeof(int64_t) * (size_t)cap) : NULL;
}

static void ensure_stage(struct zpic_dev_species *d, size_t bytes)
{
    if (bytes <= d->stage_cap) return;
    zpic_sync();
    zpic_dfree(d->stage);
    d->stage = zpic_dmalloc(bytes);
    d->stage_cap = bytes;
}

/* Grow both buffer sets to hold at least `need` particles. */
static void grow_sets(struct zpic_dev_species *d, int need)
{
    if (need <= d->cap) return;
    if (need > ZK_MAX_PARTS)
        zpic_fatal("species needs %d particle slots; at most %d per species are supported", need, ZK_MAX_PARTS);
    long long want = (long long)need + need / 8 + 4096;
    const int cap = want > ZK_MAX_PARTS ? ZK_MAX_PARTS : (int)want;
    zpic_sync();
    for (int k = 0; k < 2; k++) {
        zk_parts old = d->set[k], nw;
        alloc_set(&nw, cap, d->track_ids, d->cbits);
        if (k == d->cur && old.cell && d->np_upper > 0) {
            const size_t n = (size_t)d->np_upper;
            zpic_d2d_async(nw.cell, old.cell, 4 * n);
            zpic_d2d_async(nw.x, old.x, 4 * n);
            zpic_d2d_async(nw.y, old.y, 4 * n);
            zpic_d2d_async(nw.ux, old.ux, 4 * n);
            zpic_d2d_async(nw.uy, old.uy, 4 * n);
            zpic_d2d_async(nw.uz, old.uz, 4 * n);
            if (old.id && nw.id)
                zpic_d2d_async(nw.id, old.id, 8 * n);
        }
        zpic_sync();
        zpic_dfree(old.cell);
        zpic_dfree(old.x);
        zpic_dfree(old.y);
        zpic_dfree(old.ux);
        zpic_dfree(old.uy);
        zpic_dfree(old.uz);
        zpic_dfree(old.id);
        d->set[k] = nw;
    }
    zpic_dfree(d->ovf_list);
    d->ovf_list = zpic_dmalloc(sizeof(int) * (size_t)cap);
    d->cap = cap;
}

/* Slots for n particles in tile ranges with slack (tile_cap in sort_core.h:
 * about half as many more, the mean count per tile, room for growth, plus
 * 32 per tile) and the staging area of the staged re-sort behind them
 * (half the ranges). */
static int layout_slots(long long n, int ntiles)
{
    const long long r = 3 * n + 32LL * ntiles;
    long long want = r + (r >> 1) + 4096;
    return want > ZK_MAX_PARTS ? ZK_MAX_PARTS : (int)want;
}

static long long g_staged[4]; /* staged re-sorts, full re-sorts after them, unplaced, overfull */
long long zpic_staged_sorts(void) { return g_staged[0]; }
void zpic_staged_stats(long long out[4])
{
    for (int k = 0; k < 4; k++) out[k] = g_staged[k];
}

/* Counters after a sort: slots in use, the tail, free slots, particles;
 * returns the particles a staged re-sort could not put in their tile. */
static int sort_readback(struct zpic_dev_species *d)
{
    zk_spec_ctr c;
    ctr_read(d, &c);
    d->np_upper = c.np_total;
    d->tail_upper = c.np_total - c.n_sorted;
    d->n_slack = c.n_slack;
    d->np_live = c.np_total - c.n_slack - c.n_dead;
    d->counted = 1;
    g_staged[2] += c.n_unplaced;
    g_staged[3] += c.n_overfull;
    return c.n_unplaced + c.n_overfull;
}

/* Full sort into the other buffer set: `spread` for an unsorted upload
 * (counting sort with cell interleaving), else the re-sort of tile ranges
 * plus a tail; with `slack` the new ranges keep free slots for the staged
 * re-sorts that follow. */
static void sort_now_from(t_species *s, int spread, int slack, const int *before);
static void sort_now(t_species *s, int spread, int slack)
{
    sort_now_from(s, spread, slack, s->dev->counted ? s->dev->tile_n : NULL);
}

/* before: the tile counts of the previous sort (NULL: unknown) */
static void sort_now_from(t_species *s, int spread, int slack, const int *before)
{
    struct zpic_dev_species *d = s->dev;
    /* free slots only serve the staged re-sort (mode 2: one per range) */
    if (slack) slack = zpic_staged_sort() == 2 ? 2 : zpic_staged_sort() ? 1 : 0;
    if (slack) {
        const int want = layout_slots(d->np_live, d->t.ntiles);
        if (want > d->cap) grow_sets(d, want);
    }
    const int other = 1 - d->cur;
    zpic_ktimer_begin(ZPIC_T_SORT);
    if (spread) {
        ZCHECK(zk_sort(&d->set[d->cur], &d->set[other], d->ctr, d->cnt, d->tile_off, d->cursor, d->tile_n, &d->t,
                       d->np_upper, spread, slack, d->cap, zpic_stream()));
    } else {
        /* the particles are in tile ranges (plus a tail): keep the stayers */
        ZCHECK(zk_sort_tiles(&d->set[d->cur], &d->set[other], d->ctr, d->cnt, d->stay, d->nbr, d->tailn, d->tile_off,
                             d->counted ? d->tile_n : NULL, before, d->tile_off2, d->tile_n2, d->cursor, &d->t,
                             d->tail_upper, slack, d->cap, zpic_stream()));
        int *tmp = d->tile_off;
        d->tile_off = d->tile_off2;
        d->tile_off2 = tmp;
        tmp = d->tile_n;
        d->tile_n = d->tile_n2;
        d->tile_n2 = tmp;
    }
    zpic_ktimer_end(ZPIC_T_SORT);
    d->cur = other;
    d->moved_since_sort = 0;
    d->steps_since_sort = 0;
    d->probe_ready = 0;
    if (slack || zpic_staged_sort()) {
        /* the staged re-sorts need the counters; reading them waits for the
         * queue, so the plain re-sort (packed ranges) does without */
        sort_readback(d);
    } else {
        d->tail_upper = 0; /* np_upper stays an upper bound: a sort only drops slots */
        d->counted = 1;
        d->n_slack = 0;
    }
}

/* Whether the push of this step may sort (staged re-sort): the fixed-point
 * push, counted ranges with free slots, and no unsorted tail. */
static int staged_ready(const struct zpic_dev_species *d)
{
    return zpic_staged_sort() && zpic_fixed_deposit() && d->counted && d->n_slack > 0 && d->tail_upper == 0;
}

/* Second half of a staged re-sort (the push wrote set[1 - cur]). */
static void resort_staged(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    g_staged[0]++;
    zpic_ktimer_begin(ZPIC_T_SORT);
    ZCHECK(zk_resort_staged(&d->set[1 - d->cur], d->ctr, d->tile_off, d->tile_n, d->sstay, d->sleave, d->cnt,
                            d->tailn, d->stay, d->cursor, d->nbr, &d->t, zpic_stream()));
    zpic_ktimer_end(ZPIC_T_SORT);
    d->cur = 1 - d->cur;
    d->probe_ready = 0;
    /* particles left outside their tile: a full re-sort puts every one in
     * its own tile's range */
    if (sort_readback(d) > 0) {
        g_staged[1]++;
        /* sized by the growth since the sort before the staged one (the
         * counts the staged re-sort left in sstay) */
        sort_now_from(s, 0, 1, d->sstay);
    }
}

/* a sort into packed ranges (no free slots), as batches and checkpoints
 * expect */
void zpic_spec_sort(t_species *s) { sort_now(s, 0, 0); }

/* Upload the host mirror into the device buffers and sort it into tiles. */
static void upload(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    const int np = s->np > 0 ? s->np : 0;
    grow_sets(d, np);
    ensure_stage(d, sizeof(t_part) * (size_t)(np > 0 ? np : 1));
    zpic_h2d(d->stage, s->part, sizeof(t_part) * (size_t)np);
    ZCHECK(zk_unpack(d->stage, np, &d->set[d->cur], 0, d->ctr, 0, zpic_stream()));
    if (d->track_ids && np > 0) zpic_h2d(d->set[d->cur].id, d->host_ids, sizeof(int64_t) * (size_t)np);
    d->np_upper = np;
    d->tail_upper = np;
    d->np_live = np;
    /* packed ranges: the first scheduled sort makes room for staged ones
     * (batches and checkpoints keep packed ranges and never need it) */
    sort_now(s, 1, 0);
    d->host_dirty = 0;
}

/* The kernels index grids by cell: refuse particles outside the grid
 * before anything reaches the device. */
static void validate_host_particles(const t_species *s)
{
    for (int i = 0; i < s->np; i++) {
        const t_part *p = &s->part[i];
        if (p->ix < 0 || p->ix >= s->nx[0] || p->iy < 0 || p->iy >= s->nx[1])
            zpic_fatal("species '%s': particle %d has cell (%d, %d) outside the %d x %d grid", s->name, i, p->ix,
                       p->iy, s->nx[0], s->nx[1]);
    }
}

void zpic_species_ensure(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    if (!d) zpic_fatal("species '%s' was not created with spec_new", s->name);
    if (d->host_dirty) validate_host_particles(s);
    if (!d->ready) {
        zpic_tiles_pref(s->nx[0], s->nx[1], d->pref[0], d->pref[1], d->pref[2], &d->t);
        int want = s->np > 0 ? s->np : 1;
        if (s->moving_window) {
            /* a full box of particles plus a few columns of injections */
            long long full = (long long)s->nx[0] * s->nx[1] * s->ppc[0] * s->ppc[1];
            if (full > want) want = (int)full;
        }
        if (!d->dev_unsorted) {
            d->cap = 0;
            d->cur = 0;
        }
        grow_sets(d, want);
        d->ctr = zpic_dmalloc(sizeof(zk_spec_ctr));
        zk_spec_ctr zero = {0};
        if (d->dev_unsorted) zero.np_total = d->np_upper; /* made by spec_new, not yet sorted */
        zpic_h2d(d->ctr, &zero, sizeof(zero));
        d->tile_off = zpic_dmalloc(sizeof(int) * (size_t)(d->t.ntiles + 2));
        d->tile_off2 = zpic_dmalloc(sizeof(int) * (size_t)(d->t.ntiles + 2));
        d->stay = zpic_dmalloc(sizeof(int) * (size_t)d->t.ntiles);
        d->nbr = zpic_dmalloc(sizeof(int) * 13 * (size_t)d->t.ntiles);
        d->tailn = zpic_dmalloc(sizeof(int) * (size_t)d->t.ntiles);
        ZCHECK(zk_fill_f32((float *)d->tailn, (size_t)d->t.ntiles, 0.0f, zpic_stream())); /* int zeros */
        d->cnt = zpic_dmalloc(sizeof(int) * (size_t)d->t.ntiles);
        d->cursor = zpic_dmalloc(sizeof(int) * (size_t)d->t.ntiles);
        ZCHECK(zk_fill_f32((float *)d->cnt, (size_t)d->t.ntiles, 0.0f, zpic_stream())); /* int zeros */
        d->epart_cap = d->t.ntiles + ZPIC_GLOBAL_BLOCKS;
        d->epart = zpic_dmalloc(sizeof(double) * (size_t)d->epart_cap);
        d->ready = 1;
        d->tile_n = zpic_dmalloc(sizeof(int) * (size_t)d->t.ntiles);
        d->tile_n2 = zpic_dmalloc(sizeof(int) * (size_t)d->t.ntiles);
        d->sstay = zpic_dmalloc(sizeof(int) * (size_t)d->t.ntiles);
        d->sleave = zpic_dmalloc(sizeof(int) * (size_t)d->t.ntiles);
        d->counted = 0;
        d->n_slack = 0;
        if (d->dev_unsorted) {
            d->tail_upper = d->np_upper;
            d->np_live = d->np_upper;
            sort_now(s, 1, 0);
            d->dev_unsorted = 0;
            d->host_dirty = 0;
        } else {
            d->host_dirty = 1;
        }
    }
    if (d->host_dirty) upload(s);
}

void spec_mark_host_dirty(t_species *s)
{
    if (s->dev) {
        s->dev->host_dirty = 1;
        s->dev->host_valid = 1;
    }
}

void spec_sync_host(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    if (d && d->dev_unsorted) zpic_species_ensure(s);
    if (!d || !d->ready || d->host_valid || d->host_dirty) return;
    zk_spec_ctr c;
    ctr_read(d, &c);
    if (c.n_dead > 0 || c.n_sorted != c.np_total || c.n_slack > 0) {
        /* packed ranges: every slot holds a particle */
        sort_now(s, 0, 0);
        ctr_read(d, &c);
    }
    const int np = c.np_total;
    d->np_upper = np;
    spec_grow_buffer(s, np);
    ensure_stage(d, sizeof(t_part) * (size_t)(np > 0 ? np : 1));
    ZCHECK(zk_pack(&d->set[d->cur], np, d->stage, zpic_stream()));
    zpic_d2h(s->part, d->stage, sizeof(t_part) * (size_t)np);
    if (d->track_ids) {
        if (d->host_ids_cap < s->np_max) {
            d->host_ids = realloc(d->host_ids, sizeof(int64_t) * (size_t)s->np_max);
            d->host_ids_cap = s->np_max;
        }
        zpic_d2h(d->host_ids, d->set[d->cur].id, sizeof(int64_t) * (size_t)np);
    }
    s->np = np;
    s->energy = c.energy;
    d->host_valid = 1;
}

void spec_track_ids(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    if (!d || d->track_ids) return;
    spec_sync_host(s);
    d->host_ids_cap = s->np_max > 0 ? s->np_max : 1;
    d->host_ids = realloc(d->host_ids, sizeof(int64_t) * (size_t)d->host_ids_cap);
    for (int i = 0; i < s->np; i++) d->host_ids[i] = i;
    d->next_id = s->np;
    d->track_ids = 1;
    if (d->ready) {
        /* re-create the buffers with id arrays */
        species_free_device(d);
    }
    d->host_dirty = 1;
}

void spec_get_ids(t_species *s, int64_t *ids)
{
    struct zpic_dev_species *d = s->dev;
    if (!d || !d->track_ids) zpic_fatal("spec_get_ids: ids are not tracked for species '%s'", s->name);
    spec_sync_host(s);
    memcpy(ids, d->host_ids, sizeof(int64_t) * (size_t)s->np);
}

void zpic_spec_sync_energy(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    if (!d || !d->ready || d->host_valid) return;
    zk_spec_ctr c;
    ctr_read(d, &c);
    s->energy = c.energy;
}

int spec_overflow_count(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    if (!d || !d->ready) return 0;
    zk_spec_ctr c;
    ctr_read(d, &c);
    return c.n_overflow;
}

/* ------------------------------------------------------------------------
 * time step (em2d/particles.c:1104-1269)
 * ---------------------------------------------------------------------- */
/* Inject the rightmost column after a window move (em2d/particles.c:619-640),
 * appending to the unsorted tail on the device.  The column is built in a
 * page-locked slot that the append kernel reads directly (a copy on the
 * stream would hold the GPU for the copy engine's start-up, about 25 us per
 * move on the 4000 x 256 LWFA deck); the host waits only when it reuses a
 * slot whose kernel has not run yet.  Columns of more than ZK_APPEND_MAX
 * particles are copied and unpacked by the grid-wide kernels. */
static void inject_column_device(t_species *s, int k, zk_rng_state *rng);

static void inject_column(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    const int r[2][2] = {{s->nx[0] - 1, s->nx[0] - 1}, {0, s->nx[1] - 1}};
    const int bound = inject_bound(s, r);
    const int k = d->pin_next;
    d->pin_next = (k + 1) % ZPIC_PIN_SLOTS;
    if (!d->pin_ev[k]) d->pin_ev[k] = zpic_event_new();
    zpic_event_wait(d->pin_ev[k]);
    const int npc = s->ppc[0] * s->ppc[1];
    const int dens = s->density.type;
    if (zpic_device_inject() && (dens == UNIFORM || dens == STEP || dens == SLAB) && bound <= ZK_APPEND_MAX &&
        sizeof(int) * (size_t)npc + sizeof(float) * 2 * (size_t)npc <= ZK_INJECT_TABLE_MAX) {
        zk_rng_state *rng = zpic_rng_device();
        if (rng) {
            inject_column_device(s, k, rng);
            return;
        }
    }
    const size_t need = sizeof(t_part) * (size_t)(bound > 0 ? bound : 1);
    if (need > d->pin_cap[k]) {
        zpic_pinned_free(d->pin[k]);
        d->pin[k] = zpic_pinned_alloc(need);
        d->pin_dev[k] = zpic_pinned_device(d->pin[k]);
        d->pin_cap[k] = need;
    }
    t_part *tmp = (t_part *)d->pin[k];
    const int n = inject_positions(s, r, tmp);
    inject_momenta(s, tmp, n, &d->inj);
    if (n > 0) {
        grow_sets(d, d->np_upper + n);
        if (n <= ZK_APPEND_MAX) {
            ZCHECK(zk_unpack_append(d->pin_dev[k], n, &d->set[d->cur], d->next_id, d->ctr, zpic_stream()));
            zpic_event_record(d->pin_ev[k]);
        } else {
            ensure_stage(d, sizeof(t_part) * (size_t)n);
            zpic_h2d_async(d->stage, tmp, sizeof(t_part) * (size_t)n);
            zpic_event_record(d->pin_ev[k]);
            ZCHECK(zk_unpack(d->stage, n, &d->set[d->cur], d->next_id, d->ctr, 1, zpic_stream()));
        }
        d->next_id += n;
        d->np_upper += n;
        d->tail_upper += n;
        d->np_live += n;
    }
}

/* The column made on the device (zk_inject_column): the host lists the
 * lattice points of a row that the density profile keeps (as
 * inject_positions does for each cell of the column) in page-locked slot
 * k, and the kernel draws the momenta from the device's copy of the
 * generator state. */
static void inject_column_device(t_species *s, int k, zk_rng_state *rng)
{
    struct zpic_dev_species *d = s->dev;
    const int npc = s->ppc[0] * s->ppc[1];
    const size_t need = sizeof(int) * (size_t)(npc > 0 ? npc : 1) + sizeof(float) * 2 * (size_t)(npc > 0 ? npc : 1);
    if (need > d->pin_cap[k]) {
        zpic_pinned_free(d->pin[k]);
        d->pin[k] = zpic_pinned_alloc(need);
        d->pin_dev[k] = zpic_pinned_device(d->pin[k]);
        d->pin_cap[k] = need;
    }
    float *lat = (float *)d->pin[k];
    int *tab = (int *)(lat + 2 * npc);
    const float hx = 1.0f / s->ppc[0], hy = 1.0f / s->ppc[1];
    for (int b = 0, m = 0; b < s->ppc[1]; b++)
        for (int a = 0; a < s->ppc[0]; a++, m++) {
            lat[2 * m] = hx * (a + 0.5);
            lat[2 * m + 1] = hy * (b + 0.5);
        }
    const int i = s->nx[0] - 1;
    const float lo = s->density.start / s->dx[0] - s->n_move;
    const float hi = s->density.end / s->dx[0] - s->n_move;
    const int cut = s->density.type != UNIFORM, slab = s->density.type == SLAB;
    int perrow = 0;
    for (int m = 0; m < npc; m++)
        if (!cut || (i + lat[2 * m] > lo && (!slab || i + lat[2 * m] < hi))) tab[perrow++] = m;
    const int n = perrow * s->nx[1];
    if (n > 0) {
        grow_sets(d, d->np_upper + n);
        zk_inject_args a;
        memset(&a, 0, sizeof(a));
        a.perrow = perrow;
        a.rows = s->nx[1];
        a.row0 = 0;
        a.col = i;
        a.tab_k = (const int *)((const float *)d->pin_dev[k] + 2 * npc);
        a.lat = (const float *)d->pin_dev[k];
        a.nlat = npc;
        a.dst = d->set[d->cur];
        a.ctr = d->ctr;
        a.id0 = d->next_id;
        for (int c = 0; c < 3; c++) {
            a.uth[c] = s->uth[c];
            a.ufl[c] = s->ufl[c];
        }
        a.rng = rng;
        ZCHECK(zk_inject_column(&a, zpic_stream()));
        zpic_event_record(d->pin_ev[k]);
        d->next_id += n;
        d->np_upper += n;
        d->tail_upper += n;
        d->np_live += n;
    }
}

/* Arguments of the push of species s (em2d/particles.c:1104-1117); sets
 * *move when the window moves this step.  The species, fields and current
 * must be on the device. */
static void push_args(t_species *s, t_emf *emf, t_current *cur, zk_push_args *pa, int *move)
{
    struct zpic_dev_species *d = s->dev;
    struct zpic_dev_current *dc = cur->dev;
    zk_push_args a;
    memset(&a, 0, sizeof(a));
    a.p = d->set[d->cur];
    a.tile_off = d->tile_off;
    a.tile_n = d->counted ? d->tile_n : NULL;
    a.ctr = d->ctr;
    a.E = zpic_emf_particle_E(emf);
    a.B = zpic_emf_particle_B(emf);
    a.nrow = emf->nrow;
    a.region = dc->region;
    a.fresh = dc->fresh;
    zpic_current_jov_register(cur, s->q);
    a.Jov = dc->Jov;
    a.jov_scale = zpic_current_jov_scale(cur);
    a.epart = d->epart;
    a.t = d->t;
    a.moving_window = s->moving_window;
    a.open_y = d->open_y;
    /* constants with the reference's precision (em2d/particles.c:1111-1117) */
    a.tem = 0.5 * s->dt / s->m_q;
    a.dt_dx = s->dt / s->dx[0];
    a.dt_dy = s->dt / s->dx[1];
    a.qnx = s->q * s->dx[0] / s->dt;
    a.qny = s->q * s->dx[1] / s->dt;
    a.q = s->q;
    const int next_iter = s->iter + 1;
    *move = s->moving_window && ((next_iter * s->dt) > (s->dx[0] * (s->n_move + 1)));
    a.shift = *move;
    /* the global pass pushes the unsorted tail (injected particles), if any */
    a.ovf_list = d->ovf_list;
    a.global_blocks = d->tail_upper > 0 ? ZPIC_GLOBAL_BLOCKS : 0;
    a.waves = zpic_push_waves();
    a.fixed = zpic_fixed_deposit();
    zpic_fixed_point_scale(s->q, &a.fx_scale, &a.fx_unscale);
    *pa = a;
}

/* The rest of spec_advance after its push: iteration count, window
 * injection and energy (push_done; with fin set the energy is left to
 * the caller's zk_spec_finalize_set, which gets this species' entry). */
static void push_done(t_species *s, int move, int tb, zk_finalize_set *fin)
{
    struct zpic_dev_species *d = s->dev;
    s->iter++;
    if (move) {
        s->n_move++;
        inject_column(s);
        d->moved_since_sort++;
    }
    const float qm = s->q * s->m_q;
    if (fin) {
        const int k = fin->n++;
        fin->ctr[k] = d->ctr;
        fin->epart[k] = d->epart;
        fin->nparts[k] = d->t.ntiles + tb;
        fin->qm[k] = (double)qm;
        fin->dx0[k] = (double)s->dx[0];
        fin->dx1[k] = (double)s->dx[1];
    } else {
        ZCHECK(zk_spec_finalize(d->ctr, d->epart, d->t.ntiles + tb, (double)qm, (double)s->dx[0],
                                (double)s->dx[1], zpic_stream()));
    }
    d->host_valid = 0;
}

/* Whether the tiles are sorted after the push of this step (move: the
 * window moves in it), decided before the push so that the push itself
 * can sort. */
static int sort_due(t_species *s, int move)
{
    struct zpic_dev_species *d = s->dev;
    /* the species' schedule (em2d/particles.c:1262-1264); particles that
     * left their tile windows take the slow global path, so the tiles are
     * also re-sorted early when that path carries too many pushes (the
     * order of the particles is not part of the results) */
    /* a window move shifts every particle one cell towards its tile's lower
     * x margin: after H moves the tiles are sorted again */
    const int iter = s->iter + 1, moved = d->moved_since_sort + (move ? 1 : 0);
    int sort = (s->n_sort > 0 && !(iter % s->n_sort)) || moved >= (d->t.H > 1 ? d->t.H : 1);
    d->steps_since_sort++;
    if (!sort && d->auto_sort && d->steps_since_sort % ZPIC_SORT_PROBE == 0) {
        zk_spec_ctr c;
        if (d->probe_ready) {
            /* the copy queued behind the last push (probe_prefetch) */
            zpic_event_wait(d->probe_ev);
            memcpy(&c, d->probe_host, sizeof(c));
            ctr_check(&c);
        } else {
            ctr_read(d, &c);
        }
        d->probe_ready = 0;
        /* global-path pushes since the later of the last probe and the last sort */
        long long base = (long long)(unsigned)c.ovf_sorted;
        if (base < d->ovf_seen) base = d->ovf_seen;
        const long long novf = (long long)(unsigned)c.n_overflow - base;
        d->ovf_seen = (long long)(unsigned)c.n_overflow;
        if (novf * 1000 > (long long)ZPIC_SORT_OVF_PERMILLE * ZPIC_SORT_PROBE * (long long)c.np_total) sort = 1;
    }
    return sort;
}

/* At the end of a step whose successor probes the counters (sort_due),
 * queue their copy into page-locked memory behind this step's push: the
 * probe then reads it while the GPU runs the rest of the step (current
 * fold, filters, field update) instead of draining the queue and leaving
 * the GPU idle until the next push is launched.  A sort clears the copy
 * (probe_ready), since it changes the counters. */
static void probe_prefetch(t_species *s)
{
    struct zpic_dev_species *d = s->dev;
    if (!d->auto_sort || (d->steps_since_sort + 1) % ZPIC_SORT_PROBE != 0) return;
    if (!d->probe_host) d->probe_host = zpic_pinned_alloc(sizeof(zk_spec_ctr));
    if (!d->probe_ev) d->probe_ev = zpic_event_new_read();
    zpic_d2h_async(d->probe_host, d->ctr, sizeof(zk_spec_ctr));
    zpic_event_record(d->probe_ev);
    d->probe_ready = 1;
}

/* Lets the push of this step sort the species (staged re-sort). */
static void push_sorting(t_species *s, zk_push_args *a)
{
    struct zpic_dev_species *d = s->dev;
    a->sorting = 1;
    a->sdst = d->set[1 - d->cur];
    a->sstay = d->sstay;
    a->sleave = d->sleave;
}

void spec_advance(t_species *s, t_emf *emf, t_current *cur)
{
    zpic_species_ensure(s);
    zpic_emf_ensure(emf);
    zpic_current_ensure(cur, &s->dev->t);

    zpic_timer_begin(ZPIC_T_SPEC);
    zk_push_args a;
    int move;
    push_args(s, emf, cur, &a, &move);
    const int sort = sort_due(s, move);
    const int staged = sort && staged_ready(s->dev) && a.fixed;
    if (staged) push_sorting(s, &a);
    zpic_ktimer_begin(ZPIC_T_PUSH);
    ZCHECK(zk_push(&a, zpic_stream()));
    zpic_ktimer_end(ZPIC_T_PUSH);
    cur->dev->fresh = 0;
    /* a staged re-sort completes before the window's injections are
     * appended to the (new) set */
    if (staged) resort_staged(s);
    push_done(s, move, a.global_blocks, NULL);
    if (staged) {
        s->dev->moved_since_sort = 0;
        s->dev->steps_since_sort = 0;
    } else if (sort) {
        sort_now(s, 0, 1);
    }
    probe_prefetch(s);
    zpic_timer_end(ZPIC_T_SPEC);
}

static int same_tiles(const zk_tiles *a, const zk_tiles *b)
{
    return a->nx == b->nx && a->ny == b->ny && a->TX == b->TX && a->TY == b->TY && a->H == b->H &&
           a->WX == b->WX && a->WY == b->WY && a->ntx == b->ntx && a->nty == b->nty && a->ntiles == b->ntiles;
}

void zpic_spec_advance_all(t_species *spec, int n, t_emf *emf, t_current *cur)
{
    int fuse = n >= 2 && n <= ZK_FUSE_MAX && zpic_species_fusion() && zpic_fixed_deposit();
    if (fuse) {
        zpic_emf_ensure(emf);
        for (int k = 0; k < n; k++) {
            zpic_species_ensure(&spec[k]);
            zpic_current_ensure(cur, &spec[k].dev->t);
        }
        zk_push_multi m;
        int move[ZK_FUSE_MAX];
        memset(&m, 0, sizeof(m));
        m.n = n;
        /* every species' charge is covered by the overflow grid's scale
         * before any arguments are formed */
        for (int k = 0; k < n; k++) zpic_current_jov_register(cur, spec[k].q);
        for (int k = 0; k < n; k++) push_args(&spec[k], emf, cur, &m.s[k], &move[k]);
        for (int k = 1; k < n && fuse; k++)
            fuse = same_tiles(&m.s[k].t, &m.s[0].t) && m.s[k].fx_scale == m.s[0].fx_scale;
        if (fuse) {
            zpic_timer_begin(ZPIC_T_SPEC);
            /* the sort decisions (whose counter reads wait for the queue)
             * are taken before the push, which sorts where it can */
            int due[ZK_FUSE_MAX], staged[ZK_FUSE_MAX];
            for (int k = 0; k < n; k++) {
                due[k] = sort_due(&spec[k], move[k]);
                staged[k] = due[k] && staged_ready(spec[k].dev);
                if (staged[k]) push_sorting(&spec[k], &m.s[k]);
            }
            zpic_ktimer_begin(ZPIC_T_PUSH);
            ZCHECK(zk_push_species(&m, zpic_stream()));
            zpic_ktimer_end(ZPIC_T_PUSH);
            cur->dev->fresh = 0;
            for (int k = 0; k < n; k++)
                if (staged[k]) resort_staged(&spec[k]);
            zk_finalize_set fin;
            memset(&fin, 0, sizeof(fin));
            for (int k = 0; k < n; k++) push_done(&spec[k], move[k], m.s[k].global_blocks, &fin);
            ZCHECK(zk_spec_finalize_set(&fin, zpic_stream()));
            for (int k = 0; k < n; k++) {
                if (staged[k]) {
                    spec[k].dev->moved_since_sort = 0;
                    spec[k].dev->steps_since_sort = 0;
                } else if (due[k]) {
                    sort_now(&spec[k], 0, 1);
                }
            }
            for (int k = 0; k < n; k++) probe_prefetch(&spec[k]);
            zpic_timer_end(ZPIC_T_SPEC);
            return;
        }
    }
    for (int k = 0; k < n; k++) spec_advance(&spec[k], emf, cur);
}

void spec_move_window(t_species *s)
{
    /* standalone form of the window move (em2d/particles.c:619-640), on the
     * host mirror; spec_advance applies the shift inside the push kernel */
    if ((s->iter * s->dt) > (s->dx[0] * (s->n_move + 1))) {
        spec_sync_host(s);
        for (int i = 0; i < s->np; i++) s->part[i].ix--;
        s->n_move++;
        const int r[2][2] = {{s->nx[0] - 1, s->nx[0] - 1}, {0, s->nx[1] - 1}};
        spec_grow_buffer(s, s->np + inject_bound(s, r));
        const int n = inject_positions(s, r, s->part + s->np);
        inject_momenta(s, s->part + s->np, n, NULL);
        if (s->dev && s->dev->track_ids) {
            struct zpic_dev_species *d = s->dev;
            if (d->host_ids_cap < s->np_max) {
                d->host_ids = realloc(d->host_ids, sizeof(int64_t) * (size_t)s->np_max);
                d->host_ids_cap = s->np_max;
            }
            for (int i = 0; i < n; i++) d->host_ids[s->np + i] = d->next_id++;
        }
        s->np += n;
        spec_mark_host_dirty(s);
    }
}

/* ------------------------------------------------------------------------
 * diagnostics
 * ---------------------------------------------------------------------- */

/* spec_deposit_charge without any allocation: the deposit and the caller's
 * density are staged in the library's file-scope scratch array (two grids
 * of at most ZK_NOALLOC_CHARGE_MAX nodes); the species must already be on
 * the device with no pending host edits (an upload may allocate).  The
 * scratch is shared by every call on a device: calls that overlap on one
 * GPU race. */
int spec_deposit_charge_noalloc(const t_species *cs, float *charge)
{
    const t_species *s = cs;
    const struct zpic_dev_species *d = s->dev;
    const int nrow = s->nx[0] + 1;
    const size_t n = (size_t)nrow * (size_t)(s->nx[1] + 1);
    if (n > (size_t)ZK_NOALLOC_CHARGE_MAX) {
        fprintf(stderr, "spec_deposit_charge_noalloc: %zu charge nodes exceed the maximum of %d\n", n,
                ZK_NOALLOC_CHARGE_MAX);
        return -1;
    }
    if (!d || !d->ready || d->host_dirty) {
        fprintf(stderr, "spec_deposit_charge_noalloc: species '%s' has no up-to-date device state "
                        "(call it after a step, or use spec_deposit_charge)\n", s->name);
        return -1;
    }
    float *scr = NULL;
    const int e = zk_noalloc_scratch(&scr);
    if (e) {
        fprintf(stderr, "spec_deposit_charge_noalloc: no scratch on this device (HIP error %d)\n", e);
        return -1;
    }
    float *rho = scr, *acc = scr + ZK_NOALLOC_CHARGE_MAX;
    zpic_h2d_async(acc, charge, sizeof(float) * n);
    /* the scratch keeps the previous call's contents: clear the deposit */
    ZCHECK(zk_fill_f32(rho, n, 0.0f, zpic_stream()));
    ZCHECK(zk_charge(&d->set[d->cur], d->np_upper, d->ctr, s->q, s->nx[0], rho, zpic_stream()));
    ZCHECK(zk_add_f32(acc, rho, n, zpic_stream()));
    zpic_d2h(charge, acc, sizeof(float) * n);
    if (!s->moving_window)
        for (int j = 0; j < s->nx[1] + 1; j++) charge[j * nrow] += charge[s->nx[0] + j * nrow];
    if (!d->open_y)
        for (int i = 0; i < s->nx[0] + 1; i++) charge[i] += charge[i + s->nx[1] * nrow];
    return 0;
}
void spec_deposit_charge(const t_species *cs, float *charge)
{
    /* CIC deposit on the device, boundary fold on the host
     * (em2d/particles.c:1289-1324) */
    t_species *s = (t_species *)cs;
    zpic_species_ensure(s);
    struct zpic_dev_species *d = s->dev;
    const int nrow = s->nx[0] + 1;
    const size_t n = (size_t)nrow * (size_t)(s->nx[1] + 1);
    float *rho = zpic_dmalloc(sizeof(float) * n);
    ZCHECK(zk_fill_f32(rho, n, 0.0f, zpic_stream()));
    ZCHECK(zk_charge(&d->set[d->cur], d->np_upper, d->ctr, s->q, s->nx[0], rho, zpic_stream()));
    float *h = malloc(sizeof(float) * n);
    zpic_d2h(h, rho, sizeof(float) * n);
    zpic_dfree(rho);
    for (size_t k = 0; k < n; k++) charge[k] += h[k];
    free(h);
    if (!s->moving_window)
        for (int j = 0; j < s->nx[1] + 1; j++) charge[j * nrow] += charge[s->nx[0] + j * nrow];
    if (!d->open_y)
        for (int i = 0; i < s->nx[0] + 1; i++) charge[i] += charge[i + s->nx[1] * nrow];
}

static float pha_quantity(const t_species *s, int i, int q)
{
    switch (q) {
    case X1: return (s->part[i].x + s->part[i].ix) * s->dx[0];
    case X2: return (s->part[i].y + s->part[i].iy) * s->dx[1];
    case U1: return s->part[i].ux;
    case U2: return s->part[i].uy;
    case U3: return s->part[i].uz;
    }
    return 0;
}

void spec_deposit_pha(const t_species *cs, const int rep_type, const int pha_nx[], const float pha_range[][2],
                      float *buf)
{
    /* phase-space density (em2d/particles.c:1569-1632): the reference's
     * weights and bounds per particle, added into buf.  On the device once
     * the particles live there; from the host buffer while it is the
     * authoritative copy (before the first step, or after host edits not
     * yet uploaded) */
    t_species *s = (t_species *)cs;
    struct zpic_dev_species *d = s->dev;
    const int na = pha_nx[0], nb = pha_nx[1];
    if (na <= 0 || nb <= 0) return;
    const int qa = rep_type & 0x000F, qb = (rep_type & 0x00F0) >> 4;
    const float amin = pha_range[0][0], bmin = pha_range[1][0];
    const float ra = pha_nx[0] / (pha_range[0][1] - pha_range[0][0]);
    const float rb = pha_nx[1] / (pha_range[1][1] - pha_range[1][0]);
    if (!d || !d->ready || d->host_dirty) {
        for (int i = 0; i < s->np; i++) {
            const float va = (pha_quantity(s, i, qa) - amin) * ra;
            const float vb = (pha_quantity(s, i, qb) - bmin) * rb;
            const int ia = (int)(va + 0.5f), ib = (int)(vb + 0.5f);
            const float wa = va - ia + 0.5f, wb = vb - ib + 0.5f;
            int c = ia + na * ib;
            if (ib >= 0 && ib < nb) {
                if (ia >= 0 && ia < na) buf[c] += (1.0f - wa) * (1.0f - wb) * s->q;
                if (ia + 1 >= 0 && ia + 1 < na) buf[c + 1] += wa * (1.0f - wb) * s->q;
            }
            c += na;
            if (ib + 1 >= 0 && ib + 1 < nb) {
                if (ia >= 0 && ia < na) buf[c] += (1.0f - wa) * wb * s->q;
                if (ia + 1 >= 0 && ia + 1 < na) buf[c + 1] += wa * wb * s->q;
            }
        }
        return;
    }
    zpic_species_ensure(s);
    const size_t n = (size_t)na * (size_t)nb;
    float *dbuf = zpic_dmalloc(sizeof(float) * n);
    ZCHECK(zk_fill_f32(dbuf, n, 0.0f, zpic_stream()));
    ZCHECK(zk_pha(&d->set[d->cur], d->np_upper, d->ctr, qa, qb, s->dx[0], s->dx[1], na, nb, amin, bmin, ra, rb, s->q,
                  dbuf, zpic_stream()));
    float *h = malloc(sizeof(float) * n);
    zpic_d2h(h, dbuf, sizeof(float) * n);
    zpic_dfree(dbuf);
    for (size_t k = 0; k < n; k++) buf[k] += h[k];
    free(h);
}

/* em2d/particles.c:1339-1416: positions (from the box corner) and
 * momenta of every particle, to PARTICLES/<name>/particles-<name>-<iter>.zdf;
 * the particles are in this library's tile order */
static void rep_particles(t_species *s)
{
    spec_sync_host(s);
    const int np = s->np > 0 ? s->np : 0;
    float *d[5];
    for (int k = 0; k < 5; k++) d[k] = malloc(sizeof(float) * (size_t)(np > 0 ? np : 1));
    for (int i = 0; i < np; i++) {
        const t_part *p = &s->part[i];
        d[0][i] = (s->n_move + p->ix + p->x) * s->dx[0];
        d[1][i] = (p->iy + p->y) * s->dx[1];
        d[2][i] = p->ux;
        d[3][i] = p->uy;
        d[4][i] = p->uz;
    }
    static const char *const quants[] = {"x", "y", "ux", "uy", "uz"};
    static const char *const qlabels[] = {"x", "y", "u_x", "u_y", "u_z"};
    static const char *const qunits[] = {"c/\\omega_p", "c/\\omega_p", "c", "c", "c"};
    const zdf_iter it = {"ITERATION", s->iter, s->iter * s->dt, "1/\\omega_p"};
    char path[1024];
    snprintf(path, 1024, "PARTICLES/%s", s->name);
    zpic_zdf_save_particles(s->name, s->name, (uint64_t)np, 5, quants, qlabels, qunits, (const float *const *)d, &it,
                            path);
    for (int k = 0; k < 5; k++) free(d[k]);
}

/* em2d/particles.c:1423-1497: charge density, interior nodes, to
 * CHARGE/<name>/<name>-charge-<iter>.zdf */
static void rep_charge(t_species *s)
{
    const size_t n1 = (size_t)(s->nx[0] + 1) * (size_t)(s->nx[1] + 1);
    float *charge = calloc(n1, sizeof(float));
    spec_deposit_charge(s, charge);
    float *buf = malloc(sizeof(float) * (size_t)s->nx[0] * (size_t)s->nx[1]);
    for (int j = 0; j < s->nx[1]; j++)
        for (int i = 0; i < s->nx[0]; i++) buf[i + j * s->nx[0]] = charge[i + j * (s->nx[0] + 1)];
    free(charge);
    const zdf_axis axis[2] = {{"x", "x", "c/\\omega_p", s->n_move * s->dx[0], s->box[0] + s->n_move * s->dx[0]},
                              {"y", "y", "c/\\omega_p", 0.0, s->box[1]}};
    char name[128], label[128];
    snprintf(name, 128, "%s-charge", s->name);
    snprintf(label, 128, "%s \\rho", s->name);
    const zdf_grid g = {name, label, "n_e", 2, {(uint64_t)s->nx[0], (uint64_t)s->nx[1], 0}, axis};
    const zdf_iter it = {"ITERATION", s->iter, s->iter * s->dt, "1/\\omega_p"};
    char path[1024];
    snprintf(path, 1024, "CHARGE/%s", s->name);
    zpic_zdf_save_grid(buf, &g, &it, path);
    free(buf);
}

/* em2d/particles.c:1640-1711: phase-space density to
 * PHASESPACE/<name>/<name>-<q1><q2>-<iter>.zdf */
static void rep_pha(t_species *s, int rep_type, const int pha_nx[], const float pha_range[][2])
{
    float *buf = calloc((size_t)pha_nx[0] * (size_t)pha_nx[1], sizeof(float));
    spec_deposit_pha(s, rep_type, pha_nx, pha_range, buf);
    const int q1 = rep_type & 0x000F, q2 = (rep_type & 0x00F0) >> 4;
    static const char *const ax_name[] = {"x1", "x2", "x3", "u1", "u2", "u3"};
    static const char *const ax_label[] = {"x", "y", "z", "u_x", "u_y", "u_z"};
    const char *u1 = (q1 == X1 || q1 == X2) ? "c/\\omega_p" : ((q1 >= U1 && q1 <= U3) ? "m_e c" : "");
    const char *u2 = (q2 == X1 || q2 == X2) ? "c/\\omega_p" : ((q2 >= U1 && q2 <= U3) ? "m_e c" : "");
    const zdf_axis axis[2] = {{ax_name[q1 - 1], ax_label[q1 - 1], u1, pha_range[0][0], pha_range[0][1]},
                              {ax_name[q2 - 1], ax_label[q2 - 1], u2, pha_range[1][0], pha_range[1][1]}};
    char name[64], label[64];
    snprintf(name, 64, "%s-%s%s", s->name, ax_name[q1 - 1], ax_name[q2 - 1]);
    snprintf(label, 64, "%s %s-%s", s->name, ax_label[q1 - 1], ax_label[q2 - 1]);
    const zdf_grid g = {name, label, "a.u.", 2, {(uint64_t)pha_nx[0], (uint64_t)pha_nx[1], 0}, axis};
    const zdf_iter it = {"ITERATION", s->iter, s->iter * s->dt, "1/\\omega_p"};
    char path[1024];
    snprintf(path, 1024, "PHASESPACE/%s", s->name);
    zpic_zdf_save_grid(buf, &g, &it, path);
    free(buf);
}

void spec_report(const t_species *cs, const int rep_type, const int pha_nx[], const float pha_range[][2])
{
    /* em2d/particles.c:1724-1740 */
    t_species *s = (t_species *)cs;
    switch (rep_type & 0xF000) {
    case CHARGE:
        rep_charge(s);
        break;
    case PHA:
        rep_pha(s, rep_type, pha_nx, pha_range);
        break;
    case PARTICLES:
        rep_particles(s);
        break;
    }
}
