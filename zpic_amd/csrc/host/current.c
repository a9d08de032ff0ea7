/*
 * current.c - electric current density on the GPU.
 *
 * Host side of the reference em2d/current.c.  current_zero opens a step
 * (the next deposit overwrites the per-tile windows); current_update sums
 * the windows into J with the periodic fold of the reference's guard cells
 * and applies the digital filter passes (em2d/current.c:171-183, 427-459).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zpic_internal.h"

#include "zpic_zdf.h"

/* 0: one launch per filter pass (for comparisons) */
#ifndef ZPIC_SMOOTH_X_FUSE
#define ZPIC_SMOOTH_X_FUSE 1
#endif

void current_new(t_current *c, int nx[], float box[], float dt)
{
    /* em2d/current.c:30-79 */
    for (int d = 0; d < 2; d++) {
        c->nx[d] = nx[d];
        c->gc[d][0] = 1;
        c->gc[d][1] = 2;
        c->box[d] = box[d];
        c->dx[d] = box[d] / nx[d];
    }
    c->nrow = nx[0] + 3;
    c->J_buf = calloc(zpic_grid_nodes(c->nx), sizeof(float3));
    if (!c->J_buf) zpic_fatal("out of host memory for the current grid");
    c->J = c->J_buf + 1 + c->nrow;
    c->smooth.xtype = NONE;
    c->smooth.ytype = NONE;
    c->smooth.xlevel = 0;
    c->smooth.ylevel = 0;
    c->iter = 0;
    c->dt = dt;
    c->moving_window = 0;
    c->dev = calloc(1, sizeof(struct zpic_dev_current));
    c->dev->host_valid = 1;
    c->dev->fresh = 1;
}

void current_delete(t_current *c)
{
    struct zpic_dev_current *d = c->dev;
    if (d) {
        zpic_sync();
        zpic_dfree(d->J[0]);
        zpic_dfree(d->J[1]);
        zpic_dfree(d->region);
        zpic_dfree(d->Jov);
        free(d);
        c->dev = NULL;
    }
    free(c->J_buf);
    c->J_buf = NULL;
}

void zpic_current_ensure(t_current *c, const zk_tiles *t)
{
    struct zpic_dev_current *d = c->dev;
    if (!d) zpic_fatal("current density was not created with current_new");
    if (!d->ready) {
        zk_tiles tt;
        if (t) tt = *t;
        else zpic_tiles_pref(c->nx[0], c->nx[1], d->pref[0], d->pref[1], d->pref[2], &tt);
        d->t = tt;
        const size_t nodes = zpic_grid_nodes(c->nx);
        d->J[0] = zpic_dmalloc(sizeof(float3) * nodes);
        d->J[1] = zpic_dmalloc(sizeof(float3) * nodes);
        d->Jov = zpic_dmalloc(sizeof(long long) * 3 * nodes);
        d->region = zpic_dmalloc(sizeof(float) * 3 * (size_t)tt.WX * (size_t)tt.WY * (size_t)tt.ntiles);
        ZCHECK(zk_fill_f32(d->J[0], 3 * nodes, 0.0f, zpic_stream()));
        ZCHECK(zk_fill_f32((float *)d->Jov, 6 * nodes, 0.0f, zpic_stream()));
        d->jcur = 0;
        d->ready = 1;
    } else if (t && (t->TX != d->t.TX || t->TY != d->t.TY || t->H != d->t.H || t->nx != d->t.nx ||
                     t->ny != d->t.ny)) {
        zpic_fatal("species and current density use different grids or tilings");
    }
}

void current_zero(t_current *c)
{
    /* em2d/current.c:98-107: every deposit of the step starts from zero */
    zpic_current_ensure(c, NULL);
    struct zpic_dev_current *d = c->dev;
    /* zero bits are the fixed-point zero */
    ZCHECK(zk_fill_f32((float *)d->Jov, 6 * zpic_grid_nodes(c->nx), 0.0f, zpic_stream()));
    d->fresh = 1;
    d->host_valid = 0;
}

/* The overflow grid sums out-of-window deposits in 64-bit fixed point with
 * one scale for every species: Q_J = 2^jov_k, the largest Q among the
 * species that deposit there (each contribution is O(|q|) <= Q_J).  A
 * species of larger charge raises Q_J; what the grid already holds is then
 * divided by the power-of-two ratio with an arithmetic shift, which is the
 * same in every run. */
void zpic_current_jov_register(t_current *c, float q)
{
    struct zpic_dev_current *d = c->dev;
    const int k = zpic_fixed_point_exp(q);
    if (d->jov_set && k <= d->jov_k) return;
    if (d->jov_set && d->ready)
        ZCHECK(zk_rescale_fx(d->Jov, 3 * zpic_grid_nodes(c->nx), k - d->jov_k, zpic_stream()));
    d->jov_k = k;
    d->jov_set = 1;
}

float zpic_current_jov_scale(const t_current *c) { return ldexpf(1.0f, 40 - (c->dev->jov_set ? c->dev->jov_k : 0)); }

float zpic_current_jov_unscale(const t_current *c) { return ldexpf(1.0f, (c->dev->jov_set ? c->dev->jov_k : 0) - 40); }

/* Weights of the compensating pass after n binomial passes
 * (em2d/current.c:297-304). */
static void compensator(int n, float *sa, float *sb)
{
    float a = -1;
    float b = (4.0 + 2.0 * n) / n;
    float total = 2 * a + b;
    *sa = a / total;
    *sb = b / total;
}

/* The filter passes of current_update in order: binomial passes, then the
 * compensator, along x and then y; the y pass count is xlevel, as in the
 * reference (em2d/current.c:427-459, 449).  Returns the number of passes. */
int zpic_current_passes(const t_current *c, int axis[], float sa[], float sb[], int max)
{
    int n = 0;
#define PASS(AX, A, B)                                                                                                 \
    do {                                                                                                               \
        if (n >= max) zpic_fatal("more than %d current filter passes", max);                                          \
        axis[n] = (AX);                                                                                                \
        sa[n] = (A);                                                                                                   \
        sb[n] = (B);                                                                                                   \
        n++;                                                                                                           \
    } while (0)
    float a, b;
    if (c->smooth.xtype != NONE) {
        for (int k = 0; k < c->smooth.xlevel; k++) PASS(0, 0.25f, 0.5f);
        if (c->smooth.xtype == COMPENSATED) {
            compensator(c->smooth.xlevel, &a, &b);
            PASS(0, a, b);
        }
    }
    if (c->smooth.ytype != NONE) {
        for (int k = 0; k < c->smooth.xlevel; k++) PASS(1, 0.25f, 0.5f);
        if (c->smooth.ytype == COMPENSATED) {
            compensator(c->smooth.ylevel, &a, &b);
            PASS(1, a, b);
        }
    }
#undef PASS
    return n;
}

void current_update(t_current *c)
{
    zpic_current_ensure(c, NULL);
    struct zpic_dev_current *d = c->dev;
    zk_reduce_args a;
    a.region = d->fresh ? NULL : d->region;
    a.Jov = d->Jov;
    a.jov_unscale = zpic_current_jov_unscale(c);
    a.J = d->J[d->jcur];
    a.nrow = c->nrow;
    a.t = d->t;
    a.periodic_x = !c->moving_window;
    a.open_y = d->open_y;
    zpic_ktimer_begin(ZPIC_T_CURRENT);
    ZCHECK(zk_current_reduce(&a, zpic_stream()));
    const int maxp = 2 * (c->smooth.xlevel > 0 ? c->smooth.xlevel : 0) + 2;
    int axis[maxp];
    float sa[maxp], sb[maxp];
    const int np = zpic_current_passes(c, axis, sa, sb, maxp);
    if (ZPIC_SMOOTH_X_FUSE && np == 2 && axis[0] == 0 && axis[1] == 1 && !d->open_y) {
        /* one x and one y pass: one launch */
        ZCHECK(zk_smooth_xy(d->J[d->jcur], d->J[1 - d->jcur], c->nx[0], c->nx[1], c->nrow, sa[0], sb[0], sa[1], sb[1],
                            !c->moving_window, zpic_stream()));
        d->jcur = 1 - d->jcur;
        zpic_ktimer_end(ZPIC_T_CURRENT);
        c->iter++;
        d->host_valid = 0;
        return;
    }
    for (int k = 0; k < np;) {
        /* consecutive x passes run in one launch */
        int run = 1;
        while (ZPIC_SMOOTH_X_FUSE && k + run < np && run < ZK_MAX_PASSES && axis[k] == 0 && axis[k + run] == 0)
            run++;
        if (axis[k] == 0 && run > 1) {
            ZCHECK(zk_smooth_x_passes(d->J[d->jcur], d->J[1 - d->jcur], c->nx[0], c->nx[1], c->nrow, run, sa + k,
                                      sb + k, !c->moving_window, zpic_stream()));
            k += run;
        } else {
            ZCHECK(zk_smooth(d->J[d->jcur], d->J[1 - d->jcur], c->nx[0], c->nx[1], c->nrow, axis[k], sa[k], sb[k],
                             !c->moving_window, d->open_y, zpic_stream()));
            k++;
        }
        d->jcur = 1 - d->jcur;
    }
    zpic_ktimer_end(ZPIC_T_CURRENT);
    c->iter++;
    d->host_valid = 0;
}

void current_sync_host(t_current *c)
{
    struct zpic_dev_current *d = c->dev;
    if (!d || !d->ready || d->host_valid) return;
    zpic_d2h(c->J_buf, d->J[d->jcur], sizeof(float3) * zpic_grid_nodes(c->nx));
    d->host_valid = 1;
}

const float *zpic_current_J(t_current *c) { return c->dev->J[c->dev->jcur]; }

void current_report(const t_current *cc, const int jc)
{
    /* em2d/current.c:195-281: one component of J, interior nodes, to
     * CURRENT/J<jc>-<iter>.zdf */
    t_current *c = (t_current *)cc;
    if (jc < 0 || jc > 2) {
        fprintf(stderr, "(*error*) Invalid current component (jc) selected, returning\n");
        return;
    }
    current_sync_host(c);
    float *buf = malloc(sizeof(float) * (size_t)c->nx[0] * (size_t)c->nx[1]);
    for (int j = 0; j < c->nx[1]; j++)
        for (int i = 0; i < c->nx[0]; i++) {
            const float3 v = c->J[i + j * c->nrow];
            buf[i + j * c->nx[0]] = jc == 0 ? v.x : (jc == 1 ? v.y : v.z);
        }
    const char comp[] = {'x', 'y', 'z'};
    char name[16], label[16];
    snprintf(name, 3, "J%1u", (unsigned)jc);
    snprintf(label, 4, "J_%c", comp[jc]);
    const zdf_axis axis[2] = {{"x", "x", "c/\\omega_p", 0.0, c->box[0]}, {"y", "y", "c/\\omega_p", 0.0, c->box[1]}};
    const zdf_grid g = {name, label, "e \\omega_p^2 / c", 2, {(uint64_t)c->nx[0], (uint64_t)c->nx[1], 0}, axis};
    const zdf_iter it = {"ITERATION", c->iter, c->iter * c->dt, "1/\\omega_p"};
    zpic_zdf_save_grid(buf, &g, &it, "CURRENT");
    free(buf);
}
