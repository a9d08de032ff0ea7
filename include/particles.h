/*
 * particles.h - particle species of the em2d time step.
 *
 * Drop-in for the reference header em2d/particles.h.  Type names, field
 * names, field order, enum values and diagnostic codes match the reference;
 * one trailing field (`dev`) holds the device-resident state of this library.
 */
#ifndef PARTICLES_H_DROPIN
#define PARTICLES_H_DROPIN

#include <stdint.h>

#include "zpic.h"
#include "emf.h"
#include "current.h"

/* reference em2d/particles.h:23 */
#define MAX_SPNAME_LEN 32

/*
 * One particle (reference em2d/particles.h:29-37): cell index (ix,iy),
 * position inside the cell in [0,1), generalized velocity u = gamma*v.
 * This is the layout of the host mirror `part`; on the device the same
 * seven quantities are stored as separate arrays (structure of arrays).
 */
typedef struct Particle {
    int ix;
    int iy;
    float x;
    float y;
    float ux;
    float uy;
    float uz;
} t_part;

/* Density profile kinds (reference em2d/particles.h:43-49). */
enum density_type {
    UNIFORM,
    EMPTY,
    STEP,
    SLAB,
    CUSTOM
};

/* Density profile (reference em2d/particles.h:55-78). */
typedef struct Density {
    float n;
    enum density_type type;
    float start;
    float end;
    float (*custom_x)(float, void *);
    void *custom_data_x;
    float (*custom_y)(float, void *);
    void *custom_data_y;
    unsigned long custom_x_total_part;
    double custom_x_total_q;
} t_density;

struct zpic_dev_species;

/*
 * Particle species (reference em2d/particles.h:85-132).  `part[0..np)` is
 * the host mirror of the particle buffer; with this library it is filled by
 * spec_sync_host() and uploaded again after spec_mark_host_dirty().  The
 * order of particles in the buffer is not part of the contract (the
 * reference reorders it by its cell sort and by absorbing-boundary removal).
 */
typedef struct Species {
    char name[MAX_SPNAME_LEN + 1];
    t_part *part;
    int np;
    int np_max;
    float m_q;
    double energy;
    float q;
    int ppc[2];
    t_density density;
    float ufl[3];
    float uth[3];
    int nx[2];
    float dx[2];
    float box[2];
    float dt;
    int iter;
    int moving_window;
    int n_move;
    int n_sort;
    /* device state owned by libzpic_em2d (not in the reference struct) */
    struct zpic_dev_species *dev;
} t_species;

/* reference em2d/particles.h:151 */
void spec_new(t_species *spec, char name[], const float m_q, const int ppc[],
              const float ufl[], const float uth[],
              const int nx[], float box[], const float dt, t_density *density);
/* reference em2d/particles.h:160 */
void spec_delete(t_species *spec);
/* reference em2d/particles.h:171 - capacity rounded up to multiples of 1024 */
void spec_grow_buffer(t_species *spec, const int size);
/* reference em2d/particles.h:191 - push, deposit, BCs, window, sort */
void spec_advance(t_species *spec, t_emf *emf, t_current *current);
/* reference em2d/particles.h:201 */
void spec_move_window(t_species *spec);
/* reference em2d/particles.h:208-222 - push counters and timers */
uint64_t spec_npush(void);
double spec_time(void);
double spec_perf(void);

/* Diagnostic codes (reference em2d/particles.h:228-245). */
#define CHARGE    0x1000
#define PHA       0x2000
#define PARTICLES 0x3000
#define X1        0x0001
#define X2        0x0002
#define U1        0x0004
#define U2        0x0005
#define U3        0x0006
#define PHASESPACE(a, b) ((a) + (b) * 16 + PHA)

/* reference em2d/particles.h:257 */
void spec_deposit_pha(const t_species *spec, const int rep_type,
                      const int pha_nx[], const float pha_range[][2], float *buf);
/* reference em2d/particles.h:270 - ZDF diagnostic writer */
void spec_report(const t_species *spec, const int rep_type,
                 const int pha_nx[], const float pha_range[][2]);
/* reference em2d/particles.h:282 - CIC charge on an (nx+1)*(ny+1) grid */
void spec_deposit_charge(const t_species *spec, float *charge);

#endif
