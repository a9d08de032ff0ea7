/*
 * current.h - electric current density grid of the em2d time step.
 *
 * Drop-in for the reference header em2d/current.h.  Type names, field
 * names, field order and enum values match the reference so that decks and
 * the Python binding compile unchanged; one trailing field (`dev`) is added
 * to hold the device-resident state of this library.
 */
#ifndef CURRENT_H_DROPIN
#define CURRENT_H_DROPIN

#include "zpic.h"

/* Digital filter kinds (reference em2d/current.h:19-23). */
enum smooth_type {
    NONE,
    BINOMIAL,
    COMPENSATED
};

/* Filter parameters (reference em2d/current.h:30-35). */
typedef struct Smooth {
    enum smooth_type xtype;
    enum smooth_type ytype;
    int xlevel;
    int ylevel;
} t_smooth;

struct zpic_dev_current;

/*
 * Current density (reference em2d/current.h:41-70).
 *
 * J_buf is the host mirror of the guard-inclusive grid, (gc[0][0]+nx[0]+gc[0][1])
 * columns by (gc[1][0]+nx[1]+gc[1][1]) rows of float3, row stride `nrow`;
 * J points at cell [0][0].  With this library the authoritative copy lives
 * in device memory: J_buf is filled by current_sync_host().
 */
typedef struct Current {
    float3 *J;
    float3 *J_buf;
    int nx[2];
    int nrow;
    int gc[2][2];
    float box[2];
    float dx[2];
    t_smooth smooth;
    float dt;
    int iter;
    int moving_window;
    /* device state owned by libzpic_em2d (not in the reference struct) */
    struct zpic_dev_current *dev;
} t_current;

/* reference em2d/current.h:80 */
void current_new(t_current *current, int nx[], float box[], float dt);
/* reference em2d/current.h:87 */
void current_delete(t_current *current);
/* reference em2d/current.h:94 - sets J (guards included) to zero */
void current_zero(t_current *current);
/* reference em2d/current.h:101 - guard fold, smoothing, iter++ */
void current_update(t_current *current);
/* reference em2d/current.h:109 - ZDF diagnostic writer */
void current_report(const t_current *current, const int jc);

#endif
