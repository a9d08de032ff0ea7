/*
 * adjoint.c - differentiable runs: the loss of a run and its gradient with
 * respect to the initial state, by the adjoint method with checkpoints.
 *
 * No counterpart in the reference.  The forward run is the ordinary time
 * step (sim_iter, em2d/simulation.c:45-56) on the GPU; every `ckpt` steps
 * the particles are tile-sorted and the state (particles with their ids,
 * tile ranges, E, B) is copied aside.  The backward pass walks the
 * segments from the last: it restores the segment's checkpoint, re-runs
 * its steps while recording the state at the start of each, then applies
 * the reverse step (adjoint.hip) from the last step of the segment to the
 * first.  Particles keep their slots inside a segment; between segments
 * the particle adjoints are carried across the checkpoint's sort through
 * the ids.
 *
 * With recording on (zpic_diff_record_all) the forward run itself keeps
 * the state at the start of every step, and the backward pass goes
 * straight to the reverse steps: no segment is run twice, at the cost of
 * the memory of the whole trajectory (28 bytes per particle and step plus
 * two grids per step).
 *
 * The loss is  L = sum over steps n in [loss_first, loss_last] of
 * dt * (magnetic terms of emf_get_energy at step n), FP64; its gradient
 * with respect to the grid B is dt * dx * dy * B on interior nodes.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../kernels/zk_adjoint.h"
#include "zpic_internal.h"

typedef struct {
    zk_parts p;     /* particles (with ids) */
    int *tile_off;  /* tile ranges */
    zk_spec_ctr ctr;
} spec_ckpt;

struct zpic_diff {
    t_simulation *sim;
    int nsteps, ckpt, nck, loss_first, loss_last;
    int ns;
    int record;        /* 1: the forward run records every step */
    int nrec;          /* particle-state slots per species in rp */
    size_t nodes;
    int *np;           /* particles per species */
    spec_ckpt *cs;     /* [(nck + 1) * ns]: checkpoints, then the final state */
    float **cE, **cB;  /* [nck + 1] */
    zk_parts *rp;      /* [nrec * ns]: particle state at the start of each step of a segment
                          (of the run when recording) */
    float **rE, **rB;  /* [nrec + 1]: fields at the start of each step, then at the end */
    zk_adj_vec *lam, *lam2; /* per species */
    int **pos;
    float *lE, *lB, *lJ, *aE, *aB, *Lam, *region;
    int splits;        /* workgroups per tile in the reverse push */
    unsigned *amax;    /* [ns] running maxima of the gathered-field adjoints (float bits) */
    float *out;        /* device staging of one species' adjoints by id */
    int forward_done;
    int iter0;
};

/* workgroups the reverse push aims for (tiles times splits) */
#define ZPIC_ADJ_BLOCKS 1024

#define CS(D, c, k) (&(D)->cs[(size_t)(c) * (D)->ns + (k)])

static void parts_alloc(zk_parts *p, int n, int ids, int cbits)
{
    const size_t m = (size_t)(n > 0 ? n : 1);
    p->cell = zpic_dmalloc(4 * m);
    p->cbits = cbits;
    p->x = zpic_dmalloc(4 * m);
    p->y = zpic_dmalloc(4 * m);
    p->ux = zpic_dmalloc(4 * m);
    p->uy = zpic_dmalloc(4 * m);
    p->uz = zpic_dmalloc(4 * m);
    p->id = ids ? zpic_dmalloc(8 * m) : NULL;
}

static void parts_free(zk_parts *p)
{
    zpic_dfree(p->cell);
    zpic_dfree(p->x);
    zpic_dfree(p->y);
    zpic_dfree(p->ux);
    zpic_dfree(p->uy);
    zpic_dfree(p->uz);
    zpic_dfree(p->id);
    memset(p, 0, sizeof(*p));
}

static void parts_copy(const zk_parts *dst, const zk_parts *src, int n)
{
    ZCHECK(zk_copy_parts(dst, src, n, zpic_stream()));
}

static void vec_alloc(zk_adj_vec *v, int n)
{
    const size_t m = 4 * (size_t)(n > 0 ? n : 1);
    v->lx = zpic_dmalloc(m);
    v->ly = zpic_dmalloc(m);
    v->lux = zpic_dmalloc(m);
    v->luy = zpic_dmalloc(m);
    v->luz = zpic_dmalloc(m);
}

static void vec_free(zk_adj_vec *v)
{
    zpic_dfree(v->lx);
    zpic_dfree(v->ly);
    zpic_dfree(v->lux);
    zpic_dfree(v->luy);
    zpic_dfree(v->luz);
}

static void vec_zero(const zk_adj_vec *v, int n)
{
    const zk_stream_t s = zpic_stream();
    ZCHECK(zk_fill_f32(v->lx, (size_t)n, 0.0f, s));
    ZCHECK(zk_fill_f32(v->ly, (size_t)n, 0.0f, s));
    ZCHECK(zk_fill_f32(v->lux, (size_t)n, 0.0f, s));
    ZCHECK(zk_fill_f32(v->luy, (size_t)n, 0.0f, s));
    ZCHECK(zk_fill_f32(v->luz, (size_t)n, 0.0f, s));
}

/* record slots for nrec steps */
static void records_alloc(zpic_diff *D, int nrec)
{
    const size_t gb = sizeof(float) * 3 * D->nodes;
    D->nrec = nrec;
    D->rp = calloc((size_t)nrec * (D->ns > 0 ? D->ns : 1), sizeof(zk_parts));
    for (int n = 0; n < nrec; n++)
        for (int k = 0; k < D->ns; k++)
            parts_alloc(&D->rp[(size_t)n * D->ns + k], D->np[k], 0, D->sim->species[k].dev->cbits);
    D->rE = calloc((size_t)nrec + 1, sizeof(float *));
    D->rB = calloc((size_t)nrec + 1, sizeof(float *));
    for (int n = 0; n <= nrec; n++) {
        D->rE[n] = zpic_dmalloc(gb);
        D->rB[n] = zpic_dmalloc(gb);
    }
}

static void records_free(zpic_diff *D)
{
    for (int n = 0; n < D->nrec; n++)
        for (int k = 0; k < D->ns; k++) parts_free(&D->rp[(size_t)n * D->ns + k]);
    for (int n = 0; n <= D->nrec; n++) {
        zpic_dfree(D->rE[n]);
        zpic_dfree(D->rB[n]);
    }
    free(D->rp);
    free(D->rE);
    free(D->rB);
    D->rp = NULL;
    D->rE = D->rB = NULL;
    D->nrec = 0;
}

/* record slot of step n of the segment starting at s0 */
static int rslot(const zpic_diff *D, int n, int s0) { return D->record ? n : n - s0; }

zpic_diff *zpic_diff_new(t_simulation *sim, int nsteps, int ckpt, int loss_first, int loss_last)
{
    if (!sim || !sim->emf.dev) zpic_fatal("zpic_diff_new: simulation was not created with sim_new");
    if (nsteps <= 0 || ckpt <= 0) zpic_fatal("zpic_diff_new: steps and checkpoint interval must be positive");
    if (sim->moving_window || sim->emf.moving_window)
        zpic_fatal("zpic_diff_new: differentiable runs do not support a moving window");
    if (sim->emf.ext_fld.E_type != EMF_FLD_TYPE_NONE || sim->emf.ext_fld.B_type != EMF_FLD_TYPE_NONE)
        zpic_fatal("zpic_diff_new: differentiable runs do not support external fields");
    if (sim->emf.dev && sim->emf.dev->open_y)
        zpic_fatal("zpic_diff_new: differentiable runs do not support open y boundaries");
    if (sim->current.smooth.xtype != NONE || sim->current.smooth.ytype != NONE)
        zpic_fatal("zpic_diff_new: differentiable runs do not support current filtering");
    zpic_diff *D = calloc(1, sizeof(*D));
    D->sim = sim;
    D->nsteps = nsteps;
    D->ckpt = ckpt;
    D->nck = (nsteps + ckpt - 1) / ckpt;
    D->loss_first = loss_first;
    D->loss_last = loss_last;
    D->ns = sim->n_species;
    D->nodes = zpic_grid_nodes(sim->emf.nx);
    D->np = calloc((size_t)(D->ns > 0 ? D->ns : 1), sizeof(int));
    for (int k = 0; k < D->ns; k++) {
        t_species *s = &sim->species[k];
        if (s->moving_window) zpic_fatal("zpic_diff_new: differentiable runs do not support a moving window");
        spec_track_ids(s);
        zpic_species_ensure(s);
        zpic_current_ensure(&sim->current, &s->dev->t);
        zpic_spec_sort(s);
        zk_spec_ctr c;
        zpic_d2h(&c, s->dev->ctr, sizeof(c));
        if (c.n_dead || c.n_sorted != c.np_total) zpic_fatal("zpic_diff_new: species '%s' is not sorted", s->name);
        D->np[k] = c.np_total;
    }
    zpic_emf_ensure(&sim->emf);
    zpic_current_ensure(&sim->current, NULL);
    const size_t gb = sizeof(float) * 3 * D->nodes;
    const zk_tiles *t = &sim->current.dev->t;
    /* the reverse push keeps E, B and the 64-bit adjoint windows of a tile
     * in LDS: 84 bytes per window node */
    if ((size_t)84 * (size_t)t->WX * (size_t)t->WY > (size_t)160 * 1024)
        zpic_fatal("zpic_diff_new: tiles of %d x %d cells with margin %d need more LDS than a workgroup has; "
                   "choose smaller tiles (zpic_sim_set_tiling)", t->TX, t->TY, t->H);

    D->cs = calloc((size_t)(D->nck + 1) * (D->ns > 0 ? D->ns : 1), sizeof(spec_ckpt));
    D->cE = calloc((size_t)D->nck + 1, sizeof(float *));
    D->cB = calloc((size_t)D->nck + 1, sizeof(float *));
    for (int c = 0; c <= D->nck; c++) {
        D->cE[c] = zpic_dmalloc(gb);
        D->cB[c] = zpic_dmalloc(gb);
        for (int k = 0; k < D->ns; k++) {
            parts_alloc(&CS(D, c, k)->p, D->np[k], 1, D->sim->species[k].dev->cbits);
            CS(D, c, k)->tile_off = zpic_dmalloc(sizeof(int) * (size_t)(t->ntiles + 2));
        }
    }
    records_alloc(D, D->ckpt < D->nsteps ? D->ckpt : D->nsteps);
    D->lam = calloc((size_t)(D->ns > 0 ? D->ns : 1), sizeof(zk_adj_vec));
    D->lam2 = calloc((size_t)(D->ns > 0 ? D->ns : 1), sizeof(zk_adj_vec));
    D->pos = calloc((size_t)(D->ns > 0 ? D->ns : 1), sizeof(int *));
    int npmax = 1;
    for (int k = 0; k < D->ns; k++) {
        vec_alloc(&D->lam[k], D->np[k]);
        vec_alloc(&D->lam2[k], D->np[k]);
        D->pos[k] = zpic_dmalloc(sizeof(int) * (size_t)(D->np[k] > 0 ? D->np[k] : 1));
        if (D->np[k] > npmax) npmax = D->np[k];
    }
    D->lE = zpic_dmalloc(gb);
    D->lB = zpic_dmalloc(gb);
    D->lJ = zpic_dmalloc(gb);
    D->aE = zpic_dmalloc(gb);
    D->aB = zpic_dmalloc(gb);
    D->Lam = zpic_dmalloc(gb);
    /* enough workgroups to fill the GPU on small grids (a few per CU) */
    D->splits = t->ntiles > 0 ? (ZPIC_ADJ_BLOCKS + t->ntiles - 1) / t->ntiles : 1;
    if (D->splits > 8) D->splits = 8;
    D->region = zpic_dmalloc(sizeof(float) * 6 * (size_t)t->WX * t->WY * t->ntiles * D->splits);
    D->amax = zpic_dmalloc(sizeof(unsigned) * (size_t)(D->ns > 0 ? D->ns : 1));
    D->out = zpic_dmalloc(sizeof(float) * 5 * (size_t)npmax);
    return D;
}

/* copy the simulation state into slot c (sorting the particles first) */
static void save_state(zpic_diff *D, int c)
{
    t_simulation *sim = D->sim;
    for (int k = 0; k < D->ns; k++) {
        t_species *s = &sim->species[k];
        struct zpic_dev_species *d = s->dev;
        zpic_spec_sort(s);
        spec_ckpt *C = CS(D, c, k);
        parts_copy(&C->p, &d->set[d->cur], D->np[k]);
        zpic_d2d_async(C->tile_off, d->tile_off, sizeof(int) * (size_t)(d->t.ntiles + 2));
        zpic_d2h(&C->ctr, d->ctr, sizeof(zk_spec_ctr));
        if (C->ctr.np_total != D->np[k] || C->ctr.n_sorted != C->ctr.np_total || C->ctr.n_dead)
            zpic_fatal("zpic_diff: species '%s' changed its particle count", s->name);
    }
    struct zpic_dev_emf *de = sim->emf.dev;
    ZCHECK(zk_copy_pair(D->cE[c], de->E[de->cur], D->cB[c], de->B[de->cur], 3 * D->nodes, zpic_stream()));
}

/* put slot c back into the simulation, at iteration `iter` */
static void restore_state(zpic_diff *D, int c, int iter)
{
    t_simulation *sim = D->sim;
    for (int k = 0; k < D->ns; k++) {
        t_species *s = &sim->species[k];
        struct zpic_dev_species *d = s->dev;
        spec_ckpt *C = CS(D, c, k);
        parts_copy(&d->set[d->cur], &C->p, D->np[k]);
        zpic_d2d_async(d->tile_off, C->tile_off, sizeof(int) * (size_t)(d->t.ntiles + 2));
        zpic_h2d(d->ctr, &C->ctr, sizeof(zk_spec_ctr));
        d->probe_ready = 0;
        /* packed ranges (saved after zpic_spec_sort) */
        d->counted = 0;
        d->n_slack = 0;
        d->tail_upper = 0;
        d->np_upper = C->ctr.np_total;
        d->host_valid = 0;
        s->iter = iter;
    }
    struct zpic_dev_emf *de = sim->emf.dev;
    ZCHECK(zk_copy_pair(de->E[de->cur], D->cE[c], de->B[de->cur], D->cB[c], 3 * D->nodes, zpic_stream()));
    de->host_valid = 0;
    sim->emf.iter = iter;
    sim->current.iter = iter;
    sim->current.dev->host_valid = 0;
}

void zpic_diff_record_all(zpic_diff *D, int on)
{
    if (!D) zpic_fatal("zpic_diff_record_all: no run");
    if (D->forward_done) zpic_fatal("zpic_diff_record_all: call before zpic_diff_forward");
    on = on ? 1 : 0;
    if (on == D->record) return;
    zpic_sync();
    records_free(D);
    D->record = on;
    records_alloc(D, on ? D->nsteps : (D->ckpt < D->nsteps ? D->ckpt : D->nsteps));
}

/* the state at the start of step n (segment start s0) into its record slot */
static void record_state(zpic_diff *D, int n, int s0)
{
    struct zpic_dev_emf *de = D->sim->emf.dev;
    const int r = rslot(D, n, s0);
    for (int k = 0; k < D->ns; k++) {
        struct zpic_dev_species *d = D->sim->species[k].dev;
        parts_copy(&D->rp[(size_t)r * D->ns + k], &d->set[d->cur], D->np[k]);
    }
    ZCHECK(zk_copy_pair(D->rE[r], de->E[de->cur], D->rB[r], de->B[de->cur], 3 * D->nodes, zpic_stream()));
}

/* the fields at the end of the segment [s0, s1) into their record slot */
static void record_fields(zpic_diff *D, int s1, int s0)
{
    struct zpic_dev_emf *de = D->sim->emf.dev;
    const int r = rslot(D, s1, s0);
    ZCHECK(zk_copy_pair(D->rE[r], de->E[de->cur], D->rB[r], de->B[de->cur], 3 * D->nodes, zpic_stream()));
}

static int in_loss(const zpic_diff *D, int n) { return n >= D->loss_first && n <= D->loss_last; }

static double b_energy(t_simulation *sim)
{
    double e[6];
    emf_get_energy(&sim->emf, e);
    return e[3] + e[4] + e[5];
}

/* sim_iter with the species' sorts switched off (the checkpoints sort),
 * so particles keep their slots through a segment */
static void step_unsorted(zpic_diff *D)
{
    t_simulation *sim = D->sim;
    int saved[64], saved_auto[64];
    const int ns = D->ns < 64 ? D->ns : 64;
    for (int k = 0; k < ns; k++) {
        saved[k] = sim->species[k].n_sort;
        saved_auto[k] = sim->species[k].dev->auto_sort;
        sim->species[k].n_sort = 0;
        sim->species[k].dev->auto_sort = 0;
    }
    sim_iter(sim);
    for (int k = 0; k < ns; k++) {
        sim->species[k].n_sort = saved[k];
        sim->species[k].dev->auto_sort = saved_auto[k];
    }
}

double zpic_diff_forward(zpic_diff *D)
{
    if (!D) zpic_fatal("zpic_diff_forward: no run");
    if (D->ns > 64) zpic_fatal("zpic_diff_forward: at most 64 species");
    t_simulation *sim = D->sim;
    const double dt = sim->dt;
    D->iter0 = sim->emf.iter;
    double loss = 0.0;
    if (in_loss(D, 0)) loss += dt * b_energy(sim);
    for (int n = 0; n < D->nsteps; n++) {
        if (n % D->ckpt == 0) save_state(D, n / D->ckpt);
        if (D->record) record_state(D, n, 0);
        step_unsorted(D);
        if (in_loss(D, n + 1)) loss += dt * b_energy(sim);
    }
    if (D->record) record_fields(D, D->nsteps, 0);
    save_state(D, D->nck);
    D->forward_done = 1;
    return loss;
}

void zpic_diff_backward(zpic_diff *D, float *dE, float *dB, float *const *dpart)
{
    if (!D || !D->forward_done) zpic_fatal("zpic_diff_backward: run zpic_diff_forward first");
    t_simulation *sim = D->sim;
    t_emf *emf = &sim->emf;
    const zk_stream_t st = zpic_stream();
    const size_t gb = sizeof(float) * 3 * D->nodes;
    const int nx = emf->nx[0], ny = emf->nx[1], nrow = emf->nrow;
    const float scale = (float)((double)sim->dt * emf->dx[0] * emf->dx[1]);
    ZCHECK(zk_fill_f32(D->lE, 3 * D->nodes, 0.0f, st));
    ZCHECK(zk_fill_f32(D->lB, 3 * D->nodes, 0.0f, st));
    for (int k = 0; k < D->ns; k++) vec_zero(&D->lam[k], D->np[k]);
    ZCHECK(zk_fill_f32((float *)D->amax, (size_t)(D->ns > 0 ? D->ns : 1), 0.0f, st));

    for (int c = D->nck - 1; c >= 0; c--) {
        const int s0 = c * D->ckpt, s1 = (s0 + D->ckpt < D->nsteps) ? s0 + D->ckpt : D->nsteps;
        if (!D->record) {
            /* re-run the segment, recording the state at the start of each step */
            restore_state(D, c, D->iter0 + s0);
            for (int n = s0; n < s1; n++) {
                record_state(D, n, s0);
                step_unsorted(D);
            }
            record_fields(D, s1, s0);
        }
        /* particle adjoints: from the order of the next checkpoint (sorted
         * after this segment) to this segment's order */
        if (c < D->nck - 1) {
            for (int k = 0; k < D->ns; k++) {
                ZCHECK(zk_adj_permute(&D->lam[k], &D->lam2[k], CS(D, c + 1, k)->p.id, CS(D, c, k)->p.id, D->pos[k],
                                      D->np[k], st));
                const zk_adj_vec tmp = D->lam[k];
                D->lam[k] = D->lam2[k];
                D->lam2[k] = tmp;
            }
        }
        for (int n = s1 - 1; n >= s0; n--) {
            const int r = rslot(D, n, s0);
            if (in_loss(D, n + 1)) ZCHECK(zk_adj_loss(D->lB, D->rB[r + 1], nx, ny, nrow, scale, st));
            ZCHECK(zk_adj_fields(D->lE, D->lB, D->lJ, D->aE, D->aB, nx, ny, nrow, sim->dt, emf->dx[0], emf->dx[1],
                                 st));
            ZCHECK(zk_adj_images(D->lJ, D->Lam, nx, ny, nrow, st));
            for (int k = 0; k < D->ns; k++) {
                t_species *s = &sim->species[k];
                struct zpic_dev_species *d = s->dev;
                zk_adj_push_args a;
                memset(&a, 0, sizeof(a));
                a.p = D->rp[(size_t)r * D->ns + k];
                a.tile_off = CS(D, c, k)->tile_off;
                a.E = D->rE[r];
                a.B = D->rB[r];
                a.lam = D->Lam;
                a.nrow = nrow;
                a.lin = D->lam[k];
                a.lout = D->lam2[k];
                a.amax = D->amax + k;
                a.region = D->region;
                a.lE = D->lE;
                a.lB = D->lB;
                a.t = d->t;
                a.ovf_list = d->ovf_list;
                a.splits = D->splits;
                /* the push constants of spec_advance */
                a.tem = 0.5 * s->dt / s->m_q;
                a.dt_dx = s->dt / s->dx[0];
                a.dt_dy = s->dt / s->dx[1];
                a.qnx = s->q * s->dx[0] / s->dt;
                a.qny = s->q * s->dx[1] / s->dt;
                a.q = s->q;
                ZCHECK(zk_adj_push(&a, st));
                const zk_adj_vec tmp = D->lam[k];
                D->lam[k] = D->lam2[k];
                D->lam2[k] = tmp;
            }
        }
    }
    if (in_loss(D, 0)) ZCHECK(zk_adj_loss(D->lB, D->cB[0], nx, ny, nrow, scale, st));

    if (dE) zpic_d2h(dE, D->lE, gb);
    if (dB) zpic_d2h(dB, D->lB, gb);
    for (int k = 0; k < D->ns; k++) {
        if (!dpart || !dpart[k]) continue;
        ZCHECK(zk_adj_by_id(&D->lam[k], CS(D, 0, k)->p.id, D->np[k], D->out, st));
        zpic_d2h(dpart[k], D->out, sizeof(float) * 5 * (size_t)D->np[k]);
    }
    /* leave the simulation in its state after the forward run */
    restore_state(D, D->nck, D->iter0 + D->nsteps);
    zpic_sync();
}

void zpic_diff_delete(zpic_diff *D)
{
    if (!D) return;
    zpic_sync();
    for (int c = 0; c <= D->nck; c++) {
        zpic_dfree(D->cE[c]);
        zpic_dfree(D->cB[c]);
        for (int k = 0; k < D->ns; k++) {
            parts_free(&CS(D, c, k)->p);
            zpic_dfree(CS(D, c, k)->tile_off);
        }
    }
    records_free(D);
    for (int k = 0; k < D->ns; k++) {
        vec_free(&D->lam[k]);
        vec_free(&D->lam2[k]);
        zpic_dfree(D->pos[k]);
    }
    zpic_dfree(D->lE);
    zpic_dfree(D->lB);
    zpic_dfree(D->lJ);
    zpic_dfree(D->aE);
    zpic_dfree(D->aB);
    zpic_dfree(D->Lam);
    zpic_dfree(D->region);
    zpic_dfree(D->amax);
    zpic_dfree(D->out);
    free(D->cs);
    free(D->cE);
    free(D->cB);
    free(D->lam);
    free(D->lam2);
    free(D->pos);
    free(D->np);
    free(D);
}
