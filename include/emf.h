/*
 * emf.h - electromagnetic fields of the em2d time step.
 *
 * Drop-in for the reference header em2d/emf.h.  Type names, field names,
 * field order and enum values match the reference; one trailing field
 * (`dev`) holds the device-resident state of this library.
 */
#ifndef EMF_H_DROPIN
#define EMF_H_DROPIN

#include "zpic.h"
#include "current.h"

/* External / initial field kinds (reference em2d/emf.h:20-24). */
enum emf_fld_type {
    EMF_FLD_TYPE_NONE,
    EMF_FLD_TYPE_UNIFORM,
    EMF_FLD_TYPE_CUSTOM
};

/* External field parameters (reference em2d/emf.h:30-45). */
typedef struct EMF_ExternalField {
    enum emf_fld_type E_type;
    enum emf_fld_type B_type;
    float3 E_0;
    float3 B_0;
    float3 (*E_custom)(int, float, int, float, void *);
    float3 (*B_custom)(int, float, int, float, void *);
    void *E_custom_data;
    void *B_custom_data;
    float3 *E_part_buf;
    float3 *B_part_buf;
} t_emf_ext_fld;

/* Initial field parameters (reference em2d/emf.h:51-65). */
typedef struct EMF_InitialField {
    enum emf_fld_type E_type;
    enum emf_fld_type B_type;
    float3 E_0;
    float3 B_0;
    float3 (*E_custom)(int, float, int, float, void *);
    float3 (*B_custom)(int, float, int, float, void *);
    void *E_custom_data;
    void *B_custom_data;
} t_emf_init_fld;

/* Diagnostic selectors (reference em2d/emf.h:71-76). */
enum emf_diag {
    EFLD,
    BFLD,
    EPART,
    BPART
};

struct zpic_dev_emf;

/*
 * EM fields (reference em2d/emf.h:83-120).  E_buf/B_buf are host mirrors of
 * the guard-inclusive grids (gc = {{1,2},{1,2}}, row stride nrow = nx[0]+3);
 * E/B point at cell [0][0].  E_part/B_part alias E/B unless external fields
 * are set.  The authoritative copies live in device memory; emf_sync_host()
 * fills the mirrors.
 */
typedef struct EMF {
    float3 *E;
    float3 *B;
    float3 *E_buf;
    float3 *B_buf;
    float3 *E_part;
    float3 *B_part;
    int nx[2];
    int nrow;
    int gc[2][2];
    float box[2];
    float dx[2];
    float dt;
    int iter;
    int moving_window;
    int n_move;
    t_emf_ext_fld ext_fld;
    /* device state owned by libzpic_em2d (not in the reference struct) */
    struct zpic_dev_emf *dev;
} t_emf;

/* Laser pulse kinds and parameters (reference em2d/emf.h:126-156). */
enum emf_laser_type {
    PLANE,
    GAUSSIAN
};

typedef struct EMF_Laser {
    enum emf_laser_type type;
    float start;
    float fwhm;
    float rise;
    float flat;
    float fall;
    float a0;
    float omega0;
    float polarization;
    float W0;
    float focus;
    float axis;
} t_emf_laser;

/* reference em2d/emf.h:164 - FP64 sums of E_c^2, B_c^2 times 0.5*dx*dy */
void emf_get_energy(const t_emf *emf, double energy[]);
/* reference em2d/emf.h:177 */
void emf_new(t_emf *emf, int nx[], float box[], const float dt);
/* reference em2d/emf.h:187 */
void emf_delete(t_emf *emf);
/* reference em2d/emf.h:199 - ZDF diagnostic writer */
void emf_report(const t_emf *emf, const char field, const int fc);
/* reference em2d/emf.h:210 */
void emf_add_laser(t_emf *const emf, t_emf_laser *laser);
/* reference em2d/emf.h:218 */
void emf_init_fld(t_emf *const emf, t_emf_init_fld *init_fld);
/* reference em2d/emf.h:226 */
void emf_set_ext_fld(t_emf *const emf, t_emf_ext_fld *ext_fld);
/* reference em2d/emf.h:239 - Yee B/2, E, B/2; guards; ext fields; window */
void emf_advance(t_emf *emf, const t_current *current);
/* reference em2d/emf.h:246 */
double emf_time(void);

#endif
