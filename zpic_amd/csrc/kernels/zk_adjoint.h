/*
 * zk_adjoint.h - launchers of the reverse-mode kernels (adjoint.hip).
 *
 * Layouts are those of zk.h: guard-inclusive float3 grids of
 * (nx + 3) x (ny + 3) nodes with row stride nrow, particles as seven
 * arrays grouped by tile.  Particle adjoints are five arrays: the
 * adjoints of the cell offsets x, y (equal to those of the continuous
 * positions) and of the momenta ux, uy, uz.
 */
#ifndef ZK_ADJOINT_H
#define ZK_ADJOINT_H

#include "zk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zk_adj_vec {
    float *lx, *ly, *lux, *luy, *luz;
} zk_adj_vec;

typedef struct zk_adj_push_args {
    zk_parts p;                 /* particle state at the start of the step (read) */
    const int *tile_off;        /* [ntiles + 1] tile ranges of p */
    const float *E, *B;         /* fields at the start of the step */
    const float *lam;           /* adjoint of the raw current, per buffer node */
    int nrow;
    zk_adj_vec lin;             /* particle adjoints after the step (read) */
    zk_adj_vec lout;            /* particle adjoints before the step (written) */
    float *region;              /* [ntiles * splits][6][WY][WX] field-adjoint windows */
    float *lE, *lB;             /* field adjoints (accumulated) */
    zk_tiles t;
    int *ovf_list;              /* slots outside their tile window, listed at the start
                                   of the listing workgroup's share of the tile range */
    int splits;                 /* workgroups per tile (>= 1) */
    unsigned *amax;             /* float bits: running maximum of the gathered-field
                                   adjoints of this species (0 at the start) */
    float tem, dt_dx, dt_dy, qnx, qny, q;
} zk_adj_push_args;

/* dst = src for the seven particle arrays (and ids when both have them) */
int zk_copy_parts(const zk_parts *dst, const zk_parts *src, int n, zk_stream_t s);
/* a = sa and b = sb, n floats each */
int zk_copy_pair(float *a, const float *sa, float *b, const float *sb, size_t n, zk_stream_t s);
/* lB(interior) += scale * B(interior) */
int zk_adj_loss(float *lB, const float *B, int nx, int ny, int nrow, float scale, zk_stream_t s);
/* Reverse of zk_yee + zk_field_guards (periodic): on entry lE, lB hold the
 * adjoints of the step's output fields, on exit those of its input fields;
 * lJ receives the adjoint of the current.  aE, aB are scratch grids. */
int zk_adj_fields(float *lE, float *lB, float *lJ, float *aE, float *aB, int nx, int ny, int nrow, float dt,
                  float dx, float dy, zk_stream_t s);
/* Reverse of the periodic current fold: lam(node) = sum of lJ over the
 * buffer nodes with the same periodic image */
int zk_adj_images(const float *lJ, float *lam, int nx, int ny, int nrow, zk_stream_t s);
/* Reverse of the particle push of one species (tile pass, global pass,
 * field-adjoint windows added into lE, lB) */
int zk_adj_push(const zk_adj_push_args *a, zk_stream_t s);
/* dst[i] = src[pos[ids_dst[i]]] with pos[ids_src[k]] = k: particle
 * adjoints carried from one particle order to another */
int zk_adj_permute(const zk_adj_vec *src, const zk_adj_vec *dst, const int64_t *ids_src, const int64_t *ids_dst,
                   int *pos, int n, zk_stream_t s);
/* out[5 id + c] = component c of slot i (x, y, ux, uy, uz) */
int zk_adj_by_id(const zk_adj_vec *v, const int64_t *ids, int n, float *out, zk_stream_t s);

#ifdef __cplusplus
}
#endif

#endif
