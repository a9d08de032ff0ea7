/*
 * zk.h - launchers of the em2d HIP kernels (gfx950), callable from C.
 *
 * Every launcher takes plain device pointers, sizes and scalars in a small
 * argument struct and a HIP stream; it returns the int of the
 * launch (0 on success).  The host layer (zpic_amd/csrc/host) owns all allocations.
 *
 * Grids are float3 arrays of (nx+3) x (ny+3) nodes including guard cells
 * (one lower, two upper guards per axis), row stride nrow = nx + 3, and are
 * addressed from the buffer origin: node (i, j), i,j >= -1, sits at
 * (i + 1) + (j + 1) * nrow.  Particles are stored as six separate arrays
 * (packed cell index, x, y, ux, uy, uz), grouped by tile.
 */
#ifndef ZK_H
#define ZK_H

#include <stddef.h>
#include <stdint.h>

/* HIP stream handle (same type as hipStream_t) and launch status (a
 * int value, 0 on success); kept opaque so that C code including
 * the em2d headers need not include the HIP headers, whose vector types
 * clash with the em2d float3. */
typedef struct ihipStream_t *zk_stream_t;

#ifdef __cplusplus
extern "C" {
#endif

/* Cell index ix that marks a removed particle (absorbing boundary) in the
 * kernels' registers; in memory a removed particle's packed cell is -1. */
#define ZK_DEAD (-0x40000000)

/* Most particle slots of one species: the push addresses the particle
 * arrays with 32-bit byte offsets (4 * index). */
#define ZK_MAX_PARTS (1 << 30)

/* Fixed-point current windows (scale 2^32 / Q, sums read back from 51
 * bits): every particle adds at most Q to a window node, so the sums are
 * exact for tiles of at most this many particles (all species of a launch,
 * free slots included). */
#define ZK_FX_TILE_MAX ((1 << 18) - 16)

/* Per-species counters kept in device memory. */
typedef struct zk_spec_ctr {
    int np_total;               /* slots in use: tile ranges + unsorted tail */
    int n_sorted;               /* end of the tile ranges                     */
    int n_dead;                 /* removed particles still occupying slots    */
    int n_overflow;             /* pushes on the global path, accumulated     */
    int n_slack;                /* free slots at the ends of the tile ranges
                                   (capacity left for the staged re-sort)     */
    int n_unplaced;             /* leavers the last sorting push found no
                                   staging room for; 0 after a full sort      */
    int n_overfull;             /* arrivals the last staged re-sort found no
                                   free slot for; 0 after a full sort         */
    int ovf_sorted;             /* n_overflow at the last sort               */
    int n_fx_range;             /* tile pushes over more than ZK_FX_TILE_MAX
                                   particles (the fixed-point windows' exact
                                   range); must stay 0                        */
    unsigned long long npush;   /* accumulated particle pushes                */
    double energy;              /* kinetic energy of the last push            */
} zk_spec_ctr;

/* Particle arrays of one buffer set: the cell index pair is packed into
 * one word, cell = ix | iy << cbits (nx <= 2^cbits, cbits + bits of ny <=
 * 31), -1 for a removed particle; 24 bytes per particle. */
typedef struct zk_parts {
    int *cell;
    float *x, *y, *ux, *uy, *uz;
    int64_t *id;                /* optional identities (NULL when untracked) */
    int cbits;
} zk_parts;

/* Tile geometry shared by the push, the sort and the current reduction.
 * Tiles are TX x TY cells (the last tile of a row or column may be
 * narrower: ntx = ceil(nx / TX)).  A tile's window is
 * its cells plus H cells of drift margin plus the interpolation/deposit
 * stencil (1 below, 2 above): WX = TX + 2H + 3 nodes, WY likewise. */
typedef struct zk_tiles {
    int nx, ny, TX, TY, H, WX, WY, ntx, nty, ntiles;
} zk_tiles;

typedef struct zk_push_args {
    zk_parts p;
    const int *tile_off;        /* [ntiles + 1] */
    const int *tile_n;          /* [ntiles] slots in use at the front of each
                                   range (NULL: range t ends at tile_off[t + 1]) */
    zk_spec_ctr *ctr;
    const float *E, *B;         /* fields seen by particles, buffer origin */
    int nrow;
    float *region;              /* [ntiles][3][WY][WX] current windows */
    int fresh;                  /* 1: overwrite region, 0: accumulate    */
    long long *Jov;             /* raw current of out-of-window pushes, 64-bit
                                   fixed point at jov_scale (order-free sums) */
    double *epart;              /* energy partials, one per block        */
    zk_tiles t;
    int moving_window;          /* x absorbing + window shift           */
    int open_y;                 /* y absorbing (particles leaving are removed) */
    int shift;                  /* 1 when the window moves this step    */
    int *ovf_list;              /* slots of out-of-window particles: tile t
                                   lists its own at tile_off[t] onwards */
    int global_blocks;          /* blocks of the global (tail) pass      */
    int waves;                  /* tile-pass variant: 0, 2 or 8           */
    int fixed;                  /* 1: fixed-point (int64) current window  */
    float fx_scale;             /* fixed point: 2^32 / Q, Q = 2^ceil(log2|q|) */
    float fx_unscale;           /* fixed point: Q / 2^32                  */
    float jov_scale;            /* fixed point of Jov: 2^40 / Q_J, Q_J the
                                   largest Q of the species depositing there */
    float tem, dt_dx, dt_dy, qnx, qny, q;
    /* staged re-sort: with `sorting` set the push writes every particle
     * that survives the step to the other buffer set `sdst`: packed at the
     * front of its tile's range while it stays in the tile, else to its
     * source tile's staging region behind the ranges (slots S + [off[t] / 2,
     * off[t + 1] / 2), S = tile_off[ntiles]); sstay and sleave receive the
     * counts per tile.  zk_resort_staged then moves the staged particles
     * to their tiles. */
    int sorting;
    zk_parts sdst;
    int *sstay, *sleave;
} zk_push_args;

/* Species of one simulation pushed by one tile launch: same tiling,
 * fields, current windows and fixed-point scale; s[0] supplies those. */
#define ZK_FUSE_MAX 4
typedef struct zk_push_multi {
    zk_push_args s[ZK_FUSE_MAX];
    int n;
} zk_push_multi;

int zk_fill_f32(float *p, size_t n, float v, zk_stream_t s);
/* divide the n fixed-point values of p by 2^shift (arithmetic shift) */
int zk_rescale_fx(long long *p, size_t n, int shift, zk_stream_t s);
int zk_push(const zk_push_args *a, zk_stream_t s);
/* fused tile pass over m->n species (fixed-point windows), then the
 * global pass of each species that has an unsorted tail */
int zk_push_species(const zk_push_multi *m, zk_stream_t s);
/* energy and push count of up to ZK_FUSE_MAX species in one launch */
typedef struct zk_finalize_set {
    int n;
    zk_spec_ctr *ctr[ZK_FUSE_MAX];
    const double *epart[ZK_FUSE_MAX];
    int nparts[ZK_FUSE_MAX];
    double qm[ZK_FUSE_MAX], dx0[ZK_FUSE_MAX], dx1[ZK_FUSE_MAX];
} zk_finalize_set;
int zk_spec_finalize_set(const zk_finalize_set *f, zk_stream_t s);
int zk_spec_finalize(zk_spec_ctr *ctr, const double *epart, int nparts,
                            double qm, double dx0, double dx1, zk_stream_t s);

/* Tile ranges.  Range t of a sorted set starts at off[t]; its first
 * tile_n[t] slots hold the tile's particles and the rest, up to
 * off[t + 1], are free slots marked removed (cell -1).  A sort with
 * `slack` gives every tile room for count / 2 + 32 more particles when
 * the set (of cap_slots slots) can also hold a staging area of half the
 * ranges behind them; without slack, or without the room, off[t + 1] =
 * off[t] + tile_n[t]. */
/* tile counting sort: src -> dst; off has ntiles + 2 entries, cnt, cursor
 * and tile_n ntiles */
int zk_sort(const zk_parts *src, const zk_parts *dst, zk_spec_ctr *ctr,
                   int *cnt, int *off, int *cursor, int *tile_n, const zk_tiles *t,
                   int np_upper, int spread, int slack, int cap_slots, zk_stream_t s);
/* re-sort of a tile-sorted input: old_off (ntiles + 2, the last entry
 * scratch) and old_n -> new_off (ntiles + 1) and new_n; `before` (or NULL)
 * the counts of the previous sort, whose growth sizes the free slots; move
 * and tailn (ntiles, zero on entry and exit), stay and cursor (ntiles),
 * nbr (13 ntiles) */
int zk_sort_tiles(const zk_parts *src, const zk_parts *dst, zk_spec_ctr *ctr, int *move, int *stay, int *nbr,
                  int *tailn, int *old_off, const int *old_n, const int *before, int *new_off, int *new_n,
                  int *cursor, const zk_tiles *t, int tail_upper, int slack, int cap_slots, zk_stream_t s);
/* Second half of a staged re-sort, after a push with `sorting` set: the
 * staged leavers of every tile (sleave[t] of them in the staging region of
 * dst) go to the free slots of their new tile's range behind its sstay[t]
 * stayers; tile_n receives the new counts and the free slots are marked
 * removed.  Arrivals for a full range stay in the staging region, which
 * then becomes the unsorted tail (n_overfull counts them).  arr and far
 * (ntiles, zero on entry and exit), arr0, farcur (ntiles) and nbr (13
 * ntiles) are scratch. */
int zk_resort_staged(const zk_parts *dst, zk_spec_ctr *ctr, const int *tile_off, int *tile_n, int *sstay,
                     const int *sleave, int *arr, int *far, int *arr0, int *farcur, int *nbr, const zk_tiles *t,
                     zk_stream_t s);
/* AoS staging <-> SoA */
int zk_unpack(const void *aos, int n, const zk_parts *dst, int64_t id0,
                     zk_spec_ctr *ctr, int append, zk_stream_t s);
int zk_pack(const zk_parts *src, int n, void *aos, zk_stream_t s);
/* append n <= ZK_APPEND_MAX particles (t_part AoS, device-accessible, e.g.
 * page-locked host memory) to the unsorted tail in one launch */
#define ZK_APPEND_MAX 16384
int zk_unpack_append(const void *aos, int n, const zk_parts *dst, int64_t id0, zk_spec_ctr *ctr, zk_stream_t s);

/* current: window reduction + raw overflow -> folded J; smoothing passes */
typedef struct zk_reduce_args {
    const float *region;
    const long long *Jov;
    float jov_unscale;          /* Q_J / 2^40 */
    float *J;
    int nrow;
    zk_tiles t;
    int periodic_x;
    int open_y;                 /* no fold along y: the raw guard rows are kept */
} zk_reduce_args;
int zk_current_reduce(const zk_reduce_args *a, zk_stream_t s);
int zk_smooth(const float *in, float *out, int nx, int ny, int nrow, int axis,
                     float sa, float sb, int periodic_x, int open_y, zk_stream_t s);
/* npass (<= ZK_MAX_PASSES) consecutive x passes in one launch, bit-identical
 * to npass zk_smooth launches along x */
/* one x pass, then one y pass with periodic y, in one launch (the same
 * values as two zk_smooth launches) */
int zk_smooth_xy(const float *in, float *out, int nx, int ny, int nrow, float sax, float sbx, float say, float sby,
                 int periodic_x, zk_stream_t s);
int zk_smooth_x_passes(const float *in, float *out, int nx, int ny, int nrow, int npass, const float *sa,
                       const float *sb, int periodic_x, zk_stream_t s);

/* fields */
/* open_y: first-order Mur absorbing boundaries at the y edges; shift: the
 * output is the grid after the step and a window move (one column left,
 * the rightmost columns zeroed), as zk_move_window would make it */
int zk_yee(const float *E, const float *B, const float *J, float *Eo, float *Bo,
                  int nx, int ny, int nrow, float dt, float dx, float dy, int open_y, int shift,
                  zk_stream_t s);
int zk_field_guards(float *E, float *B, int nx, int ny, int nrow, int periodic_x,
                           int open_y, zk_stream_t s);
int zk_move_window(const float *E, const float *B, float *Eo, float *Bo, int nx,
                          int ny, int nrow, zk_stream_t s);
/* laser pulse added to E, B (buffer origins; em2d/emf.c:242-350): type 0
 * plane, 1 gaussian (then the longitudinal components from div E = div B
 * = 0, ratio = dx / dy); guards are left to zk_field_guards */
typedef struct zk_laser_k {
    float start, rise, flat, fall, focus, omega0, W0, axis, amp, cp, sp, hx, hy;
    int type;
} zk_laser_k;
int zk_laser(float *E, float *B, int nx, int ny, int nrow, const zk_laser_k *L, double ratio, zk_stream_t s);
/* out = F + ext[node] when ext is given, else F + (ux,uy,uz), over every node */
int zk_add_ext(const float *F, const float *ext, float ux, float uy, float uz,
                      float *out, size_t nodes, zk_stream_t s);
int zk_emf_energy(const float *E, const float *B, int nx, int ny, int nrow,
                         double *partial, int nblk, double *out6, double scale, zk_stream_t s);
int zk_charge(const zk_parts *p, int np_upper, const zk_spec_ctr *ctr, float q,
                     int nx, float *rho, zk_stream_t s);
/* Scratch of the allocation-free charge deposit: a file-scope device array
 * of 2 * ZK_NOALLOC_CHARGE_MAX floats (64 MiB) on every device the library
 * runs on.  zk_noalloc_scratch returns its address on the current device
 * (looked up once per device).  zk_add_f32: acc[k] += v[k]. */
#define ZK_NOALLOC_CHARGE_MAX (8 << 20)
int zk_noalloc_scratch(float **ptr);
int zk_add_f32(float *acc, const float *v, size_t n, zk_stream_t s);
/* out (na x nb, row length na) += phase-space density of quantities qa, qb
 * (em2d/particles.c:1569-1632 weights): amin/bmin the lower range bounds,
 * ra/rb the bins per unit */
int zk_pha(const zk_parts *p, int np_upper, const zk_spec_ctr *ctr, int qa, int qb, float dx0, float dx1, int na,
           int nb, float amin, float bmin, float ra, float rb, float q, float *out, zk_stream_t s);

/* ---- particle initialisation on the device (init.hip) ---- */
typedef struct zk_init_args {
    long long np;                       /* particles of the injected range */
    int perrow;                         /* particles per row of cells */
    int row0, rows, col0, ncols;        /* injected cell range */
    const int *tab_i, *tab_k;           /* [perrow] column and lattice point of a row's particles */
    const int *col_start, *col_count;   /* [ncols] first entry and entries of a column in a row */
    const float *lat;                   /* [2 ppc] lattice points in the cell */
    int *cell;                          /* outputs, particle order (cell packed, cbits) */
    int cbits;
    float *x, *y, *ux, *uy, *uz;
    int stride;                         /* cell key ix + stride iy of the mean removal (nx[1]) */
    float uth[3], ufl[3];
    uint32_t z0, w0;                    /* generator state before the species */
    uint32_t z1, w1;                    /* ... and after its first step */
    int kept0;                          /* 1: a deviate kept from before is deviate 0 */
    double kept_in;
    long long pairs;                    /* pairs of deviates drawn */
} zk_init_args;
/* The injection generator's state on the device (the host keeps its own
 * copy in random.c and hands the state over in either direction). */
typedef struct zk_rng_state {
    uint32_t w, z;                      /* multiply-with-carry halves */
    double kept;                        /* deviate kept by the polar method */
    int have_kept;
    int ambiguous;                      /* logs not certified as correctly rounded, accumulated */
} zk_rng_state;

/* One injected column of a moving window (em2d/particles.c:476-492 with
 * spec_set_x / spec_set_u, :85-354), made on the device and appended to the
 * unsorted tail: rows x perrow particles, the row's lattice points tab_k,
 * thermal momenta from the generator state *rng (continued there), the
 * per-cell mean removed, the drift added. */
typedef struct zk_inject_args {
    int perrow, rows, row0, col;
    const int *tab_k;                   /* [perrow] lattice point of each particle of a row */
    const float *lat;                   /* [2 nlat] lattice points */
    int nlat;
    zk_parts dst;
    zk_spec_ctr *ctr;
    int64_t id0;
    float uth[3], ufl[3];
    zk_rng_state *rng;
} zk_inject_args;
/* at most ZK_APPEND_MAX particles and ZK_INJECT_TABLE_MAX bytes of
 * tab_k and lattice */
#define ZK_INJECT_TABLE_MAX (48 * 1024)
int zk_inject_column(const zk_inject_args *a, zk_stream_t s);

typedef struct zk_init_out {
    long long last_attempt;             /* attempt that gave the last pair */
    double kept;                        /* deviate left over (have_kept) */
    int have_kept;
    int ambiguous;                      /* log values too close to a rounding boundary */
} zk_init_out;
int zk_init_positions(const zk_init_args *a, zk_stream_t s);
/* attempts [0, nat) of the polar method, `per` per thread, nblocks blocks
 * of 256 threads: per-thread prefix in the block, per-block count */
int zk_init_rng_count(const zk_init_args *a, long long nat, int per, int *thr_off, int *blk_cnt, int nblocks,
                      zk_stream_t s);
/* blk_off: exclusive prefix of the block counts */
int zk_init_rng_emit(const zk_init_args *a, long long nat, int per, const int *thr_off, const long long *blk_off,
                     int nblocks, zk_init_out *out, zk_stream_t s);
/* mean (3 nkeys floats) per cell key, then u += ufl - mean */
int zk_init_mean(const zk_init_args *a, float *mean, int nkeys, zk_stream_t s);

/* ---- batches of independent simulations (batch.hip) ---- */
#define ZK_MAX_PASSES 16

/* Yee constants: (dt/2)/dx, (dt/2)/dy, dt/dx, dt/dy, dt, and the Mur
 * coefficient (dt - dy)/(dt + dy) of open y boundaries */
typedef struct zk_yee_k {
    float dtx2, dty2, dtx, dty, dt, mur_y;
} zk_yee_k;

/* One simulation of a batch.  ecur/jcur select the current E,B and J
 * buffers and are advanced on the device at the end of every step. */
typedef struct zk_bsim {
    float *E[2], *B[2], *J[2];
    float *region;
    long long *Jov;
    float jov_unscale;
    int ecur, jcur, npass, deposited;
    int nx, ny, nrow;
    zk_tiles t;
    zk_yee_k yk;
    int pass_axis[ZK_MAX_PASSES];
    float pass_sa[ZK_MAX_PASSES], pass_sb[ZK_MAX_PASSES];
} zk_bsim;

/* One species of a batch: the static push arguments (a.global_blocks is
 * its block count in the out-of-window pass), both particle buffer sets
 * (pcur selects the current one), its simulation, its sort arrays and
 * schedule: it is sorted after the push that makes iter + 1 a multiple
 * of n_sort; iter counts the steps done, advanced on the device. */
typedef struct zk_bspec {
    zk_push_args a;
    zk_parts set[2];
    int pcur, sim, nparts, iter, n_sort;
    int *offs[2];            /* tile ranges: offs[ocur] is the push's, the re-sort writes the other */
    int ocur;
    int *cnt, *cursor, *stay, *nbr, *tailn; /* re-sort scratch (zk_sort_tiles) */
    double qm, dx0, dx1;
} zk_bspec;

/* Item tables hold (descriptor index, block index) pairs of ints; the
 * launches are one workgroup per item. */
int zk_batch_fill_jov(const zk_bsim *sims, const int *items, int n, zk_stream_t s);
int zk_batch_prepare_push(size_t lds);
int zk_batch_push(const zk_bspec *specs, const zk_bsim *sims, const int *tile_items, int ntile,
                  const int *global_items, int nglobal, size_t lds, int waves, int fixed, zk_stream_t s);
int zk_batch_finalize(const zk_bspec *specs, int nspecs, zk_stream_t s);
int zk_batch_current(const zk_bsim *sims, const int *node_items, int n, int max_passes, zk_stream_t s);
int zk_batch_fields(const zk_bsim *sims, const int *yee_items, int nyee, const int *guard_items, int nguard,
                    zk_stream_t s);
int zk_batch_sort(const zk_bspec *specs, int nspecs, const int *items, int n, zk_stream_t s);
int zk_batch_end_step(zk_bsim *sims, int nsims, zk_bspec *specs, int nspecs, zk_stream_t s);
int zk_batch_end_step_sorted(zk_bsim *sims, int nsims, zk_bspec *specs, int nspecs, zk_stream_t s);
/* output-node tiles of the Yee kernel */
#define ZK_YEE_BX 32
#define ZK_YEE_BY 16

#ifdef __cplusplus
}
#endif

#endif
