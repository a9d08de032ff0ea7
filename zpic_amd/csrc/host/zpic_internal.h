/*
 * zpic_internal.h - device-side state behind the em2d structs.
 *
 * Each t_species / t_emf / t_current carries a `dev` pointer to one of the
 * structs below.  The device copy is authoritative once a step has run;
 * the host buffers of the reference structs (part, E_buf, B_buf, J_buf)
 * are mirrors, refreshed by the *_sync_host calls and uploaded again after
 * the *_mark_host_dirty calls (or automatically before the first step).
 */
#ifndef ZPIC_INTERNAL_H
#define ZPIC_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include "simulation.h"
#include "zpic_gpu.h"
#include "zpic_rt.h"

/* blocks of the global (out-of-window and tail) push pass */
#define ZPIC_GLOBAL_BLOCKS 512
/* early tile sort: every ZPIC_SORT_PROBE pushes the overflow counter is
 * read; if more than ZPIC_SORT_OVF_PERMILLE per mille of the pushes since
 * the previous probe took the global path, the tiles are re-sorted */
#define ZPIC_SORT_PROBE 4
#define ZPIC_SORT_OVF_PERMILLE 2
/* page-locked staging slots of the injected columns: the host may run this
 * many window moves ahead of the GPU */
#define ZPIC_PIN_SLOTS 8

/* host scratch of the per-cell mean momenta removed from injected particles */
typedef struct {
    float3 *mean;
    int *count;
    int cells;
} zpic_inj_scratch;

struct zpic_dev_species {
    int ready;            /* device buffers allocated and filled */
    int host_dirty;       /* host mirror edited: upload before the next push */
    int host_valid;       /* host mirror equals the device contents */
    int cap;              /* particle slots per buffer set */
    int cur;              /* buffer set holding the particles */
    zk_parts set[2];
    int track_ids;
    int64_t next_id;
    zk_spec_ctr *ctr;     /* device counters */
    zk_tiles t;
    int *tile_off, *cnt, *cursor;
    int *tile_off2, *stay; /* second range array and stayer counts of the re-sort */
    int *nbr, *tailn;      /* re-sort: mover runs per neighbour slot, farther movers per tile */
    int *ovf_list;        /* slots listed by the tile pass for the global pass */
    double *epart;
    int epart_cap;
    void *stage;          /* device AoS staging for uploads/downloads */
    size_t stage_cap;
    void *pin[ZPIC_PIN_SLOTS];   /* page-locked staging of injected columns */
    size_t pin_cap[ZPIC_PIN_SLOTS];
    void *pin_dev[ZPIC_PIN_SLOTS]; /* the slots' device addresses */
    void *pin_ev[ZPIC_PIN_SLOTS]; /* recorded after the last read of each slot */
    int pin_next;
    zpic_inj_scratch inj; /* host scratch of the injected columns' cell means */
    int np_upper;         /* upper bound of slots in use */
    int tail_upper;       /* upper bound of unsorted slots */
    int moved_since_sort; /* window moves since the last sort */
    int steps_since_sort; /* pushes since the last tile sort */
    int auto_sort;        /* 1: also sort early when many particles left their tiles */
    long long ovf_seen;   /* overflow count at the last probe */
    void *probe_host;     /* page-locked copy of the counters for the next probe */
    void *probe_ev;       /* recorded after that copy */
    int probe_ready;      /* 1: the copy was queued after the last push */
    unsigned long long npush_retired; /* pushes counted before buffers were freed */
    int64_t *host_ids;    /* ids in host-mirror order (when tracked) */
    int host_ids_cap;
    struct zpic_dev_species *next_live;
    t_species *owner;
    int pref[3];          /* preferred tile (tx, ty, halo), 0: process-wide setting */
    int open_y;           /* particles leaving through y are removed */
    int cbits;            /* bits of ix in the packed cell of the device sets */
    int dev_unsorted;     /* particles made on the device by spec_new, in set[cur],
                             not yet sorted into tiles (no tiles chosen yet) */
    /* staged re-sort (the push of a sort step sorts; only the particles
     * that changed tile are moved afterwards) */
    int *tile_n, *tile_n2; /* particles at the front of each tile range */
    int counted;          /* tile_n describes the current ranges (ranges may
                             end in free slots); 0 after code that only knows
                             packed ranges (batches, adjoint restores) */
    int n_slack;          /* free slots of the ranges at the last sort */
    int np_live;          /* upper bound of the particles (not slots) */
    int *sstay, *sleave;  /* per tile: stayers written, leavers staged */
};

struct zpic_dev_current {
    int ready;
    int host_valid;
    zk_tiles t;
    float *J[2];
    int jcur;
    float *region;        /* per-tile current windows */
    long long *Jov;       /* raw current of out-of-window pushes (fixed point) */
    int jov_k;            /* Jov holds multiples of 2^(jov_k - 40) */
    int jov_set;          /* jov_k was set by a species */
    int fresh;            /* no deposit since current_zero */
    int pref[3];          /* preferred tile (tx, ty, halo), 0: process-wide setting */
    int open_y;           /* no fold along y (open y boundaries) */
};

struct zpic_dev_emf {
    int ready;
    int host_dirty;
    int host_valid;
    float *E[2], *B[2];
    int cur;
    float *Ep, *Bp;       /* fields seen by particles (external fields on) */
    float *Eext, *Bext;   /* device grids of custom external fields */
    float3 *Eext_host, *Bext_host; /* the same, evaluated on the host */
    double *epartial, *energy6;
    int nblk;
    int open_y;           /* Mur absorbing boundaries at the y edges */
};

/* helpers shared by the host modules */
void zpic_species_ensure(t_species *spec);
/* spec_advance of n species of one simulation, in one fused tile launch
 * when they share tiling and fixed-point scale (zpic_set_species_fusion) */
void zpic_spec_advance_all(t_species *spec, int n, t_emf *emf, t_current *cur);
void zpic_emf_ensure(t_emf *emf);
void zpic_current_ensure(t_current *cur, const zk_tiles *t);
const float *zpic_emf_particle_E(t_emf *emf);
const float *zpic_emf_particle_B(t_emf *emf);
const float *zpic_current_J(t_current *cur);
/* the overflow grid's fixed-point scale covers a species of charge q
 * (raised, with the grid's sums rescaled exactly, when q needs more) */
void zpic_current_jov_register(t_current *cur, float q);
float zpic_current_jov_scale(const t_current *cur);   /* 2^(40 - jov_k) */
float zpic_current_jov_unscale(const t_current *cur); /* 2^(jov_k - 40) */

void zpic_spec_sync_energy(t_species *s);
void zpic_spec_sort(t_species *s); /* tile sort of the device particles now */
int zpic_current_passes(const t_current *c, int axis[], float sa[], float sb[], int max);

#endif
