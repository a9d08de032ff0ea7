/*
 * batch.c - many independent simulations advanced together.
 *
 * The reference has no batching; a parameter scan runs its simulations
 * one after another (mods/em1d_param_scan/main.c:50-98).  Here every phase
 * of the step of sim_iter (em2d/simulation.c:45-56) is one kernel launch
 * over all simulations of the batch: the work of a phase is a table of
 * (simulation or species, block) items, and the kernels run the same
 * per-block device code as the one-simulation launches, so each
 * simulation evolves as if stepped on its own.  The buffer parities
 * (fields, current, particle set) and the per-species sort schedule live
 * in device descriptors that the last launch of a step advances, so the
 * step is captured once into a HIP graph and replayed: a step costs one
 * graph launch whatever the number of simulations.
 *
 * Two graphs are captured: the plain step and the step with the tile sort
 * (the sort launches skip species whose schedule does not ask for it).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zpic_internal.h"

/* blocks per species of the out-of-window pass: one per this many
 * particles, within [4, ZPIC_GLOBAL_BLOCKS] */
/* particles per block of the sort's count and scatter passes */

typedef struct {
    int off, n; /* items [off, off + n) of the table, in pairs */
} item_range;

struct zpic_batch {
    int nsims, nspecs, max_spec, max_passes;
    t_simulation **sims;
    t_species **spec;   /* all species, simulation by simulation */
    int *spec_sim;      /* simulation of each species */
    int *spec_k;        /* index of each species in its simulation */
    zk_bsim *hsims;     /* host copies of the descriptors */
    zk_bspec *hspecs;
    zk_bsim *dsims;     /* device descriptors */
    zk_bspec *dspecs;
    int *ditems;        /* device item table (int pairs) */
    item_range fill, node, yee, guard, sort;
    item_range *tile, *glob; /* per species index */
    int *gblocks;            /* per species: blocks of the out-of-window pass */
    int *ntiles;             /* per simulation, at creation */
    size_t lds;
    int waves, fixed;
    void *graph[2];     /* 0: plain step, 1: step with the tile sort */
};

/* ---- descriptors ------------------------------------------------------- */

static void fill_sim(zpic_batch *b, int i)
{
    t_simulation *sim = b->sims[i];
    t_emf *emf = &sim->emf;
    t_current *cur = &sim->current;
    struct zpic_dev_emf *de = emf->dev;
    struct zpic_dev_current *dc = cur->dev;
    zk_bsim *M = &b->hsims[i];
    memset(M, 0, sizeof(*M));
    for (int k = 0; k < 2; k++) {
        M->E[k] = de->E[k];
        M->B[k] = de->B[k];
        M->J[k] = dc->J[k];
    }
    M->region = dc->region;
    for (int k = 0; k < sim->n_species; k++) zpic_current_jov_register(cur, sim->species[k].q);
    M->Jov = dc->Jov;
    M->jov_unscale = zpic_current_jov_unscale(cur);
    M->ecur = de->cur;
    M->jcur = dc->jcur;
    M->npass = zpic_current_passes(cur, M->pass_axis, M->pass_sa, M->pass_sb, ZK_MAX_PASSES);
    M->deposited = sim->n_species > 0;
    M->nx = emf->nx[0];
    M->ny = emf->nx[1];
    M->nrow = emf->nrow;
    M->t = dc->t;
    /* constants of the one-simulation field solver (zk_yee) */
    const float dt2 = emf->dt / 2.0f;
    M->yk.dtx2 = dt2 / emf->dx[0];
    M->yk.dty2 = dt2 / emf->dx[1];
    M->yk.dtx = emf->dt / emf->dx[0];
    M->yk.dty = emf->dt / emf->dx[1];
    M->yk.dt = emf->dt;
}

static void fill_spec(zpic_batch *b, int k)
{
    t_species *s = b->spec[k];
    struct zpic_dev_species *d = s->dev;
    const int si = b->spec_sim[k];
    t_current *cur = &b->sims[si]->current;
    struct zpic_dev_current *dc = cur->dev;
    zk_bspec *S = &b->hspecs[k];
    memset(S, 0, sizeof(*S));
    zk_push_args *a = &S->a;
    /* the same arguments as spec_advance, minus the fields and the
     * particle set, which the kernels select through the parities */
    a->tile_off = d->tile_off;
    a->ctr = d->ctr;
    a->nrow = cur->nrow;
    a->region = dc->region;
    a->fresh = (b->spec_k[k] == 0);
    a->Jov = dc->Jov;
    a->jov_scale = zpic_current_jov_scale(cur);
    a->epart = d->epart;
    a->t = d->t;
    a->moving_window = 0;
    a->shift = 0;
    a->ovf_list = d->ovf_list;
    a->global_blocks = b->gblocks[k];
    a->waves = b->waves;
    a->fixed = b->fixed;
    zpic_fixed_point_scale(s->q, &a->fx_scale, &a->fx_unscale);
    a->tem = 0.5 * s->dt / s->m_q;
    a->dt_dx = s->dt / s->dx[0];
    a->dt_dy = s->dt / s->dx[1];
    a->qnx = s->q * s->dx[0] / s->dt;
    a->qny = s->q * s->dx[1] / s->dt;
    a->q = s->q;
    S->set[0] = d->set[0];
    S->set[1] = d->set[1];
    S->pcur = d->cur;
    S->sim = si;
    S->nparts = d->t.ntiles + a->global_blocks;
    S->iter = s->iter;
    S->n_sort = s->n_sort;
    S->offs[0] = d->tile_off;
    S->offs[1] = d->tile_off2;
    S->ocur = 0;
    S->cnt = d->cnt;
    S->cursor = d->cursor;
    S->stay = d->stay;
    S->nbr = d->nbr;
    S->tailn = d->tailn;
    const float qm = s->q * s->m_q;
    S->qm = qm;
    S->dx0 = s->dx[0];
    S->dx1 = s->dx[1];
}

/* Bring every simulation to the device and refresh the descriptors from
 * the host-side state (buffer parities, iteration counts). */
static void refresh(zpic_batch *b)
{
    for (int i = 0; i < b->nsims; i++) {
        t_simulation *sim = b->sims[i];
        zpic_emf_ensure(&sim->emf);
        zpic_current_ensure(&sim->current, NULL);
    }
    for (int k = 0; k < b->nspecs; k++) {
        t_species *s = b->spec[k];
        zpic_species_ensure(s);
        struct zpic_dev_species *d = s->dev;
        if (d->moved_since_sort) zpic_fatal("batch: species '%s' has an unsorted window shift", s->name);
        /* the batched re-sort keeps tile ranges only: no unsorted tail */
        if (d->tail_upper > 0 || d->n_slack > 0) zpic_spec_sort(s);
        if (d->epart_cap < d->t.ntiles + ZPIC_GLOBAL_BLOCKS) zpic_fatal("batch: energy partials too small");
        zpic_current_ensure(&b->sims[b->spec_sim[k]]->current, &d->t);
    }
    for (int i = 0; i < b->nsims; i++) fill_sim(b, i);
    for (int k = 0; k < b->nspecs; k++) fill_spec(b, k);
    zpic_h2d(b->dsims, b->hsims, sizeof(zk_bsim) * (size_t)b->nsims);
    if (b->nspecs) zpic_h2d(b->dspecs, b->hspecs, sizeof(zk_bspec) * (size_t)b->nspecs);
}

/* ---- item tables --------------------------------------------------------- */

typedef struct {
    int *v;
    int n, cap; /* in pairs */
} table;

static void put(table *t, int desc, int block)
{
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 4096;
        t->v = realloc(t->v, sizeof(int) * 2 * (size_t)t->cap);
        if (!t->v) zpic_fatal("batch: out of host memory");
    }
    t->v[2 * t->n] = desc;
    t->v[2 * t->n + 1] = block;
    t->n++;
}

static int blocks_of(size_t n, int per) { return (int)((n + (size_t)per - 1) / (size_t)per); }

/* order of work: heavier descriptors first, so the long blocks start early */
static const double *g_weight;
static int by_weight(const void *x, const void *y)
{
    const int a = *(const int *)x, c = *(const int *)y;
    if (g_weight[a] > g_weight[c]) return -1;
    if (g_weight[a] < g_weight[c]) return 1;
    return a - c;
}

static void build_items(zpic_batch *b, table *T)
{
    int *order = malloc(sizeof(int) * (size_t)(b->nsims > b->nspecs ? b->nsims : b->nspecs) + sizeof(int));
    double *w = malloc(sizeof(double) * (size_t)(b->nsims > b->nspecs ? b->nsims : b->nspecs) + sizeof(double));

    /* grid phases: simulations by node count */
    for (int i = 0; i < b->nsims; i++) {
        order[i] = i;
        w[i] = (double)b->hsims[i].t.nx * b->hsims[i].t.ny;
    }
    g_weight = w;
    qsort(order, (size_t)b->nsims, sizeof(int), by_weight);
#define GRID_PHASE(range, nblocks_expr)                                                                                \
    do {                                                                                                               \
        (range).off = T->n;                                                                                            \
        for (int q = 0; q < b->nsims; q++) {                                                                           \
            const int i = order[q];                                                                                    \
            const zk_bsim *M = &b->hsims[i];                                                                           \
            const int nb = (nblocks_expr);                                                                             \
            for (int k = 0; k < nb; k++) put(T, i, k);                                                                 \
        }                                                                                                              \
        (range).n = T->n - (range).off;                                                                                \
    } while (0)
    GRID_PHASE(b->fill, blocks_of(3 * (size_t)(M->nx + 3) * (M->ny + 3), 256));
    GRID_PHASE(b->node, blocks_of((size_t)(M->nx + 3) * (M->ny + 3), 256));
    GRID_PHASE(b->yee, ((M->nx + 3 + ZK_YEE_BX - 1) / ZK_YEE_BX) * ((M->ny + 3 + ZK_YEE_BY - 1) / ZK_YEE_BY));
    GRID_PHASE(b->guard, blocks_of(3 * (size_t)(M->nx + 3) + 3 * (size_t)M->ny, 256));
#undef GRID_PHASE

    /* particle phases: species by particles per tile (the tile pass) or
     * particles (the other passes) */
    for (int k = 0; k < b->nspecs; k++) {
        order[k] = k;
        w[k] = (double)b->spec[k]->np / b->hspecs[k].a.t.ntiles;
    }
    qsort(order, (size_t)b->nspecs, sizeof(int), by_weight);
    for (int p = 0; p < b->max_spec; p++) {
        b->tile[p].off = T->n;
        for (int q = 0; q < b->nspecs; q++) {
            const int k = order[q];
            if (b->spec_k[k] != p) continue;
            for (int t = 0; t < b->hspecs[k].a.t.ntiles; t++) put(T, k, t);
        }
        b->tile[p].n = T->n - b->tile[p].off;
        b->glob[p].off = T->n;
        for (int q = 0; q < b->nspecs; q++) {
            const int k = order[q];
            if (b->spec_k[k] != p) continue;
            for (int g = 0; g < b->hspecs[k].a.global_blocks; g++) put(T, k, g);
        }
        b->glob[p].n = T->n - b->glob[p].off;
    }
    for (int k = 0; k < b->nspecs; k++) {
        order[k] = k;
        w[k] = (double)b->spec[k]->np;
    }
    qsort(order, (size_t)b->nspecs, sizeof(int), by_weight);
    /* the re-sort: one workgroup per tile of every species */
    b->sort.off = T->n;
    for (int q = 0; q < b->nspecs; q++) {
        const int k = order[q];
        for (int t = 0; t < b->hspecs[k].a.t.ntiles; t++) put(T, k, t);
    }
    b->sort.n = T->n - b->sort.off;
    free(order);
    free(w);
}

/* ---- the step ------------------------------------------------------------ */

#define ITEMS(r) (b->ditems + 2 * (size_t)(r).off), (r).n

static void enqueue_step(zpic_batch *b, int with_sort)
{
    const zk_stream_t s = zpic_stream();
    /* current_zero (em2d/current.c:98-107) */
    ZCHECK(zk_batch_fill_jov(b->dsims, ITEMS(b->fill), s));
    /* spec_advance of the p-th species of every simulation; the species
     * of one simulation deposit into one set of windows, in order */
    for (int p = 0; p < b->max_spec; p++)
        ZCHECK(zk_batch_push(b->dspecs, b->dsims, ITEMS(b->tile[p]), ITEMS(b->glob[p]), b->lds, b->waves, b->fixed,
                             s));
    ZCHECK(zk_batch_finalize(b->dspecs, b->nspecs, s));
    if (with_sort) ZCHECK(zk_batch_sort(b->dspecs, b->nspecs, ITEMS(b->sort), s));
    /* current_update, emf_advance */
    ZCHECK(zk_batch_current(b->dsims, ITEMS(b->node), b->max_passes, s));
    ZCHECK(zk_batch_fields(b->dsims, ITEMS(b->yee), ITEMS(b->guard), s));
    if (with_sort) ZCHECK(zk_batch_end_step_sorted(b->dsims, b->nsims, b->dspecs, b->nspecs, s));
    else ZCHECK(zk_batch_end_step(b->dsims, b->nsims, b->dspecs, b->nspecs, s));
}

zpic_batch *zpic_batch_new(t_simulation *const *sims, int n)
{
    if (n <= 0 || !sims) zpic_fatal("zpic_batch_new: no simulations");
    zpic_batch *b = calloc(1, sizeof(*b));
    b->nsims = n;
    b->sims = malloc(sizeof(t_simulation *) * (size_t)n);
    for (int i = 0; i < n; i++) {
        t_simulation *sim = sims[i];
        if (!sim || !sim->emf.dev || !sim->current.dev) zpic_fatal("zpic_batch_new: simulation %d was not created with sim_new", i);
        if (sim->moving_window || sim->emf.moving_window || sim->current.moving_window)
            zpic_fatal("zpic_batch_new: simulation %d uses a moving window, which batches do not support", i);
        if (sim->emf.ext_fld.E_type != EMF_FLD_TYPE_NONE || sim->emf.ext_fld.B_type != EMF_FLD_TYPE_NONE)
            zpic_fatal("zpic_batch_new: simulation %d has external fields, which batches do not support", i);
        if (sim->emf.dev && sim->emf.dev->open_y)
            zpic_fatal("zpic_batch_new: simulation %d has open y boundaries, which batches do not support", i);
        for (int j = 0; j < sim->n_species; j++)
            if (sim->species[j].moving_window)
                zpic_fatal("zpic_batch_new: simulation %d uses a moving window, which batches do not support", i);
        for (int j = 0; j < i; j++)
            if (sims[j] == sim) zpic_fatal("zpic_batch_new: simulation %d appears twice", i);
        b->sims[i] = sim;
        b->nspecs += sim->n_species;
        if (sim->n_species > b->max_spec) b->max_spec = sim->n_species;
    }
    const int ns = b->nspecs > 0 ? b->nspecs : 1;
    b->spec = malloc(sizeof(t_species *) * (size_t)ns);
    b->spec_sim = malloc(sizeof(int) * (size_t)ns);
    b->spec_k = malloc(sizeof(int) * (size_t)ns);
    for (int i = 0, k = 0; i < n; i++)
        for (int j = 0; j < b->sims[i]->n_species; j++, k++) {
            b->spec[k] = &b->sims[i]->species[j];
            b->spec_sim[k] = i;
            b->spec_k[k] = j;
        }
    b->tile = calloc((size_t)(b->max_spec > 0 ? b->max_spec : 1), sizeof(item_range));
    b->glob = calloc((size_t)(b->max_spec > 0 ? b->max_spec : 1), sizeof(item_range));
    b->hsims = calloc((size_t)n, sizeof(zk_bsim));
    b->hspecs = calloc((size_t)ns, sizeof(zk_bspec));
    b->dsims = zpic_dmalloc(sizeof(zk_bsim) * (size_t)n);
    b->dspecs = zpic_dmalloc(sizeof(zk_bspec) * (size_t)ns);
    b->waves = zpic_push_waves();
    b->fixed = zpic_fixed_deposit();
    /* block counts of the particle passes are fixed with the item tables */
    b->gblocks = malloc(sizeof(int) * (size_t)ns);
    /* the tail pass pushes particles appended since the last sort; batch
     * members have none (no moving window, and refresh sorts any tail) */
    for (int k = 0; k < b->nspecs; k++) b->gblocks[k] = 0;

    refresh(b);
    b->ntiles = malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) b->ntiles[i] = b->hsims[i].t.ntiles;
    b->lds = 0;
    for (int k = 0; k < b->nspecs; k++) {
        const zk_tiles *t = &b->hspecs[k].a.t;
        /* E, B window and the current window (64-bit with the fixed-point deposit) */
        const size_t l = (size_t)t->WX * t->WY * (6 * sizeof(float) + 3 * (b->fixed ? 8 : 4));
        if (l > b->lds) b->lds = l;
    }
    for (int i = 0; i < n; i++)
        if (b->hsims[i].npass > b->max_passes) b->max_passes = b->hsims[i].npass;
    ZCHECK(zk_batch_prepare_push(b->lds));

    table T = {0};
    build_items(b, &T);
    b->ditems = zpic_dmalloc(sizeof(int) * 2 * (size_t)(T.n > 0 ? T.n : 1));
    zpic_h2d(b->ditems, T.v, sizeof(int) * 2 * (size_t)T.n);
    free(T.v);

    for (int g = 0; g < 2; g++) {
        zpic_sync();
        zpic_capture_begin();
        enqueue_step(b, g);
        b->graph[g] = zpic_capture_end();
    }
    return b;
}

/* does any species of the batch sort after the step that starts at its
 * current iteration count? */
static int step_sorts(const zpic_batch *b)
{
    for (int k = 0; k < b->nspecs; k++) {
        const t_species *s = b->spec[k];
        if (s->n_sort > 0 && (s->iter + 1) % s->n_sort == 0) return 1;
    }
    return 0;
}

void zpic_batch_run(zpic_batch *b, int steps)
{
    if (!b) zpic_fatal("zpic_batch_run: no batch");
    if (steps <= 0) return;
    refresh(b);
    for (int i = 0; i < b->nsims; i++)
        if (b->hsims[i].npass > b->max_passes || b->hsims[i].t.ntiles != b->ntiles[i])
            zpic_fatal("zpic_batch_run: the filter or grid of simulation %d changed after zpic_batch_new", i);
    for (int step = 0; step < steps; step++) {
        zpic_graph_launch(b->graph[step_sorts(b)]);
        /* the host-side state follows the device: iteration counts, buffer
         * parities, and which particle set holds each species */
        for (int k = 0; k < b->nspecs; k++) {
            t_species *s = b->spec[k];
            s->iter++;
            s->dev->probe_ready = 0; /* the batch's pushes change the counters */
            if (s->n_sort > 0 && s->iter % s->n_sort == 0) {
                /* the re-sort wrote the other particle set and tile ranges */
                struct zpic_dev_species *d = s->dev;
                int *tmp = d->tile_off;
                d->tile_off = d->tile_off2;
                d->tile_off2 = tmp;
                d->counted = 0; /* packed ranges; tile_n is not kept */
                d->cur ^= 1;
                d->steps_since_sort = 0;
            }
        }
        for (int i = 0; i < b->nsims; i++) {
            t_simulation *sim = b->sims[i];
            struct zpic_dev_current *dc = sim->current.dev;
            struct zpic_dev_emf *de = sim->emf.dev;
            dc->jcur = (dc->jcur + b->hsims[i].npass) & 1;
            sim->current.iter++;
            de->cur ^= 1;
            sim->emf.iter++;
        }
    }
    for (int i = 0; i < b->nsims; i++) {
        t_simulation *sim = b->sims[i];
        sim->current.dev->fresh = sim->n_species == 0;
        sim->current.dev->host_valid = 0;
        sim->emf.dev->host_valid = 0;
    }
    for (int k = 0; k < b->nspecs; k++) b->spec[k]->dev->host_valid = 0;
}

void zpic_batch_delete(zpic_batch *b)
{
    if (!b) return;
    zpic_sync();
    zpic_graph_free(b->graph[0]);
    zpic_graph_free(b->graph[1]);
    zpic_dfree(b->dsims);
    zpic_dfree(b->dspecs);
    zpic_dfree(b->ditems);
    free(b->sims);
    free(b->spec);
    free(b->spec_sim);
    free(b->spec_k);
    free(b->tile);
    free(b->glob);
    free(b->gblocks);
    free(b->ntiles);
    free(b->hsims);
    free(b->hspecs);
    free(b);
}
