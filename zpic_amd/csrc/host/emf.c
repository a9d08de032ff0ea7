/*
 * emf.c - electromagnetic fields on the GPU.
 *
 * Host side of the reference em2d/emf.c.  Field initialisation (laser
 * pulses, initial and external fields) runs on the host mirrors E_buf/B_buf
 * with the reference's arithmetic; the time step (fused Yee update, guard
 * cells, particle fields, moving window) and the energy sums run on the
 * device.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zpic_internal.h"
#include "zpic_zdf.h"

void emf_new(t_emf *emf, int nx[], float box[], const float dt)
{
    /* em2d/emf.c:56-113 */
    for (int d = 0; d < 2; d++) {
        emf->nx[d] = nx[d];
        emf->gc[d][0] = 1;
        emf->gc[d][1] = 2;
        emf->box[d] = box[d];
        emf->dx[d] = box[d] / nx[d];
    }
    emf->nrow = nx[0] + 3;
    const size_t nodes = zpic_grid_nodes(emf->nx);
    emf->E_buf = calloc(nodes, sizeof(float3));
    emf->B_buf = calloc(nodes, sizeof(float3));
    if (!emf->E_buf || !emf->B_buf) zpic_fatal("out of host memory for the field grids");
    emf->E = emf->E_buf + 1 + emf->nrow;
    emf->B = emf->B_buf + 1 + emf->nrow;
    emf->dt = dt;
    emf->iter = 0;
    emf->moving_window = 0;
    emf->n_move = 0;
    memset(&emf->ext_fld, 0, sizeof(emf->ext_fld));
    emf->E_part = emf->E;
    emf->B_part = emf->B;
    emf->dev = calloc(1, sizeof(struct zpic_dev_emf));
    emf->dev->host_dirty = 1;
    emf->dev->host_valid = 1;
}

void emf_delete(t_emf *emf)
{
    struct zpic_dev_emf *d = emf->dev;
    if (d) {
        zpic_sync();
        for (int k = 0; k < 2; k++) {
            zpic_dfree(d->E[k]);
            zpic_dfree(d->B[k]);
        }
        zpic_dfree(d->Ep);
        zpic_dfree(d->Bp);
        zpic_dfree(d->Eext);
        zpic_dfree(d->Bext);
        free(d->Eext_host);
        free(d->Bext_host);
        zpic_dfree(d->epartial);
        zpic_dfree(d->energy6);
        free(d);
        emf->dev = NULL;
    }
    free(emf->E_buf);
    free(emf->B_buf);
    emf->E_buf = emf->B_buf = NULL;
    if (emf->ext_fld.E_type > EMF_FLD_TYPE_NONE) free(emf->ext_fld.E_part_buf);
    if (emf->ext_fld.B_type > EMF_FLD_TYPE_NONE) free(emf->ext_fld.B_part_buf);
    emf->E_part = emf->B_part = NULL;
}

/* ---- host-side field set-up (runs on the mirrors) ---- */

/* Periodic guard copies on the host mirrors (em2d/emf.c:573-637). */
static void host_update_gc(t_emf *emf)
{
    const int nrow = emf->nrow, nx = emf->nx[0], ny = emf->nx[1];
    float3 *E = emf->E, *B = emf->B;
    if (!emf->moving_window) {
        for (int j = -1; j < ny + 2; j++) {
            E[-1 + j * nrow] = E[nx - 1 + j * nrow];
            B[-1 + j * nrow] = B[nx - 1 + j * nrow];
            for (int i = 0; i < 2; i++) {
                E[nx + i + j * nrow] = E[i + j * nrow];
                B[nx + i + j * nrow] = B[i + j * nrow];
            }
        }
    }
    for (int i = -1; i < nx + 2; i++) {
        E[i - nrow] = E[i + (ny - 1) * nrow];
        B[i - nrow] = B[i + (ny - 1) * nrow];
        for (int j = 0; j < 2; j++) {
            E[i + (ny + j) * nrow] = E[i + j * nrow];
            B[i + (ny + j) * nrow] = B[i + j * nrow];
        }
    }
}

/* Fields seen by particles on the host mirrors (em2d/emf.c:838-914). */
static void host_part_fld(t_emf *emf)
{
    const int nrow = emf->nrow;
    const t_emf_ext_fld *x = &emf->ext_fld;
    for (int f = 0; f < 2; f++) {
        const enum emf_fld_type type = f ? x->B_type : x->E_type;
        if (type == EMF_FLD_TYPE_NONE) continue;
        const float3 *src = f ? emf->B : emf->E;
        float3 *dst = f ? emf->B_part : emf->E_part;
        for (int j = -1; j < emf->nx[1] + 2; j++)
            for (int i = -1; i < emf->nx[0] + 2; i++) {
                float3 add;
                if (type == EMF_FLD_TYPE_UNIFORM) add = f ? x->B_0 : x->E_0;
                else if (f) add = (*x->B_custom)(i, emf->dx[0], j, emf->dx[1], x->B_custom_data);
                else add = (*x->E_custom)(i, emf->dx[0], j, emf->dx[1], x->E_custom_data);
                float3 v = src[i + j * nrow];
                v.x += add.x;
                v.y += add.y;
                v.z += add.z;
                dst[i + j * nrow] = v;
            }
    }
}

/* Make the host mirrors current before a host-side edit. */
static void host_edit_begin(t_emf *emf)
{
    emf_sync_host(emf);
    emf->dev->host_dirty = 1;
}

/* Gaussian beam phase and transverse profile (em2d/emf.c:158-170). */
static float beam_phase(const t_emf_laser *L, const float zpos, const float r)
{
    const float z = zpos - L->focus;
    const float zr = L->omega0 * (L->W0 * L->W0) / 2;
    const float r2 = r * r;
    const float curv = r2 * z / (zr * zr + z * z);
    const float w2 = (zr * zr) / (zr * zr + z * z);
    const float gouy = atan2(z, zr);
    return sqrt(sqrt(w2)) * exp(-r2 * w2 / (L->W0 * L->W0)) * cos(L->omega0 * (z + curv) - gouy);
}

/* sin^2 rise, flat top, sin^2 fall (em2d/emf.c:179-202). */
static float pulse_envelope(const t_emf_laser *L, const float z)
{
    if (z > L->start) return 0.0;
    if (z > L->start - L->rise) {
        const float u = z - L->start;
        const float e = sin(M_PI_2 * u / L->rise);
        return e * e;
    }
    if (z > L->start - (L->rise + L->flat)) return 1.0;
    if (z > L->start - (L->rise + L->flat + L->fall)) {
        const float u = z - (L->start - L->rise - L->flat - L->fall);
        const float e = sin(M_PI_2 * u / L->fall);
        return e * e;
    }
    return 0.0;
}

void emf_add_laser(t_emf *const emf, t_emf_laser *L)
{
    /* em2d/emf.c:242-350 */
    if (L->fwhm != 0) {
        if (L->fwhm <= 0) {
            fprintf(stderr, "Invalid laser FWHM, must be > 0, aborting.\n");
            exit(-1);
        }
        L->rise = L->fwhm;
        L->fall = L->fwhm;
        L->flat = 0.;
    }
    if (L->rise <= 0) {
        fprintf(stderr, "Invalid laser RISE, must be > 0, aborting.\n");
        exit(-1);
    }
    if (L->flat < 0) {
        fprintf(stderr, "Invalid laser FLAT, must be >= 0, aborting.\n");
        exit(-1);
    }
    if (L->fall <= 0) {
        fprintf(stderr, "Invalid laser FALL, must be > 0, aborting.\n");
        exit(-1);
    }
    const float amp = L->omega0 * L->a0;
    const float cp = cos(L->polarization);
    const float sp = sin(L->polarization);
    if (zpic_device_init() && (L->type == PLANE || L->type == GAUSSIAN)) {
        /* on the device (zpic_set_device_init): the same expressions, the
         * guards refreshed as host_update_gc does */
        zpic_emf_ensure(emf);
        struct zpic_dev_emf *d = emf->dev;
        zk_laser_k k;
        k.start = L->start;
        k.rise = L->rise;
        k.flat = L->flat;
        k.fall = L->fall;
        k.focus = L->focus;
        k.omega0 = L->omega0;
        k.W0 = L->W0;
        k.axis = L->axis;
        k.amp = amp;
        k.cp = cp;
        k.sp = sp;
        k.hx = emf->dx[0];
        k.hy = emf->dx[1];
        k.type = L->type == PLANE ? 0 : 1;
        ZCHECK(zk_laser(d->E[d->cur], d->B[d->cur], emf->nx[0], emf->nx[1], emf->nrow, &k,
                        (double)(emf->dx[0] / emf->dx[1]), zpic_stream()));
        ZCHECK(zk_field_guards(d->E[d->cur], d->B[d->cur], emf->nx[0], emf->nx[1], emf->nrow, !emf->moving_window,
                               0, zpic_stream()));
        d->host_valid = 0;
        return;
    }
    host_edit_begin(emf);
    float3 *E = emf->E, *B = emf->B;
    const int nrow = emf->nrow;
    const float hx = emf->dx[0], hy = emf->dx[1];
    const float axis = L->axis;
    if (L->type == PLANE) {
        const float k = L->omega0;
        for (int i = 0; i < emf->nx[0]; i++) {
            const float z = i * hx;
            const float zh = z + hx / 2;
            const float env = amp * pulse_envelope(L, z);
            const float envh = amp * pulse_envelope(L, zh);
            for (int j = 0; j < emf->nx[1]; j++) {
                E[i + j * nrow].y += +env * cos(k * z) * cp;
                E[i + j * nrow].z += +env * cos(k * z) * sp;
                B[i + j * nrow].y += -envh * cos(k * zh) * sp;
                B[i + j * nrow].z += +envh * cos(k * zh) * cp;
            }
        }
    } else if (L->type == GAUSSIAN) {
        for (int i = 0; i < emf->nx[0]; i++) {
            const float z = i * hx;
            const float zh = z + hx / 2;
            const float env = amp * pulse_envelope(L, z);
            const float envh = amp * pulse_envelope(L, zh);
            for (int j = 0; j < emf->nx[1]; j++) {
                const float r = j * hy - axis;
                const float rh = r + hy / 2;
                E[i + j * nrow].y += +env * beam_phase(L, z, rh) * cp;
                E[i + j * nrow].z += +env * beam_phase(L, z, r) * sp;
                B[i + j * nrow].y += -envh * beam_phase(L, zh, r) * sp;
                B[i + j * nrow].z += +envh * beam_phase(L, zh, rh) * cp;
            }
        }
        /* longitudinal components from div E = div B = 0, integrated from
         * the right edge (em2d/emf.c:213-231) */
        const double ratio = emf->dx[0] / emf->dx[1];
        for (int j = 0; j < emf->nx[1]; j++) {
            double ex = 0.0, bx = 0.0;
            for (int i = emf->nx[0] - 1; i >= 0; i--) {
                ex += ratio * (E[i + 1 + j * nrow].y - E[i + 1 + (j - 1) * nrow].y);
                E[i + j * nrow].x = ex;
                bx += ratio * (B[i + (j + 1) * nrow].y - B[i + j * nrow].y);
                B[i + j * nrow].x = bx;
            }
        }
    }
    host_update_gc(emf);
}

void emf_init_fld(t_emf *const emf, t_emf_init_fld *init)
{
    /* em2d/emf.c:922-980 */
    if (emf->iter != 0) {
        fprintf(stderr, "emf_init_fld should only be called at initialization, aborting...\n");
        exit(-1);
    }
    host_edit_begin(emf);
    const int nrow = emf->nrow;
    for (int f = 0; f < 2; f++) {
        const enum emf_fld_type type = f ? init->B_type : init->E_type;
        if (type == EMF_FLD_TYPE_NONE) continue;
        float3 *dst = f ? emf->B : emf->E;
        for (int j = -1; j < emf->nx[1] + 2; j++)
            for (int i = -1; i < emf->nx[0] + 2; i++) {
                if (type == EMF_FLD_TYPE_UNIFORM) dst[j * nrow + i] = f ? init->B_0 : init->E_0;
                else if (f) dst[j * nrow + i] = (init->B_custom)(i, emf->dx[0], j, emf->dx[1], init->B_custom_data);
                else dst[j * nrow + i] = (init->E_custom)(i, emf->dx[0], j, emf->dx[1], init->E_custom_data);
            }
    }
}

/* Host grid of a custom external field, evaluated once (the callback takes
 * only the node position, em2d/emf.h:37-38); uploaded with the fields. */
static float3 *custom_grid(t_emf *emf, float3 (*fn)(int, float, int, float, void *), void *data)
{
    const size_t nodes = zpic_grid_nodes(emf->nx);
    float3 *h = malloc(sizeof(float3) * nodes);
    if (!h) zpic_fatal("out of host memory for an external field grid");
    for (int j = -1; j < emf->nx[1] + 2; j++)
        for (int i = -1; i < emf->nx[0] + 2; i++) h[(i + 1) + (j + 1) * emf->nrow] = fn(i, emf->dx[0], j, emf->dx[1], data);
    return h;
}

void emf_set_ext_fld(t_emf *const emf, t_emf_ext_fld *ext)
{
    /* em2d/emf.c:764-832 */
    host_edit_begin(emf);
    struct zpic_dev_emf *d = emf->dev;
    const size_t nodes = zpic_grid_nodes(emf->nx);
    for (int f = 0; f < 2; f++) {
        const enum emf_fld_type type = f ? ext->B_type : ext->E_type;
        if (f) emf->ext_fld.B_type = type;
        else emf->ext_fld.E_type = type;
        if (type == EMF_FLD_TYPE_NONE) {
            if (f) {
                emf->B_part = emf->B;
                emf->ext_fld.B_part_buf = NULL;
            } else {
                emf->E_part = emf->E;
                emf->ext_fld.E_part_buf = NULL;
            }
            continue;
        }
        if (type == EMF_FLD_TYPE_UNIFORM) {
            if (f) emf->ext_fld.B_0 = ext->B_0;
            else emf->ext_fld.E_0 = ext->E_0;
        } else if (type == EMF_FLD_TYPE_CUSTOM) {
            if (f) {
                emf->ext_fld.B_custom = ext->B_custom;
                emf->ext_fld.B_custom_data = ext->B_custom_data;
            } else {
                emf->ext_fld.E_custom = ext->E_custom;
                emf->ext_fld.E_custom_data = ext->E_custom_data;
            }
        } else {
            fprintf(stderr, "Invalid external field type, aborting.\n");
            exit(-1);
        }
        float3 *buf = malloc(sizeof(float3) * nodes);
        if (!buf) zpic_fatal("out of host memory for the particle field grid");
        if (f) {
            emf->ext_fld.B_part_buf = buf;
            emf->B_part = buf + 1 + emf->nrow;
        } else {
            emf->ext_fld.E_part_buf = buf;
            emf->E_part = buf + 1 + emf->nrow;
        }
        if (type == EMF_FLD_TYPE_CUSTOM) {
            float3 *g = f ? custom_grid(emf, ext->B_custom, ext->B_custom_data)
                          : custom_grid(emf, ext->E_custom, ext->E_custom_data);
            if (f) {
                free(d->Bext_host);
                d->Bext_host = g;
            } else {
                free(d->Eext_host);
                d->Eext_host = g;
            }
        }
    }
    host_part_fld(emf);
}

/* ---- device residency ---- */

static void part_fields(t_emf *emf)
{
    struct zpic_dev_emf *d = emf->dev;
    const size_t nodes = zpic_grid_nodes(emf->nx);
    const t_emf_ext_fld *x = &emf->ext_fld;
    if (x->E_type != EMF_FLD_TYPE_NONE) {
        if (!d->Ep) d->Ep = zpic_dmalloc(sizeof(float3) * nodes);
        ZCHECK(zk_add_ext(d->E[d->cur], x->E_type == EMF_FLD_TYPE_CUSTOM ? d->Eext : NULL, x->E_0.x, x->E_0.y,
                          x->E_0.z, d->Ep, nodes, zpic_stream()));
    }
    if (x->B_type != EMF_FLD_TYPE_NONE) {
        if (!d->Bp) d->Bp = zpic_dmalloc(sizeof(float3) * nodes);
        ZCHECK(zk_add_ext(d->B[d->cur], x->B_type == EMF_FLD_TYPE_CUSTOM ? d->Bext : NULL, x->B_0.x, x->B_0.y,
                          x->B_0.z, d->Bp, nodes, zpic_stream()));
    }
}

void zpic_emf_ensure(t_emf *emf)
{
    struct zpic_dev_emf *d = emf->dev;
    if (!d) zpic_fatal("fields were not created with emf_new");
    const size_t nodes = zpic_grid_nodes(emf->nx);
    if (!d->ready) {
        for (int k = 0; k < 2; k++) {
            d->E[k] = zpic_dmalloc(sizeof(float3) * nodes);
            d->B[k] = zpic_dmalloc(sizeof(float3) * nodes);
        }
        d->nblk = 256;
        d->epartial = zpic_dmalloc(sizeof(double) * 6 * (size_t)d->nblk);
        d->energy6 = zpic_dmalloc(sizeof(double) * 6);
        d->cur = 0;
        d->ready = 1;
        d->host_dirty = 1;
    }
    if (d->host_dirty) {
        for (int f = 0; f < 2; f++) {
            float3 **host = f ? &d->Bext_host : &d->Eext_host;
            float **dev = f ? &d->Bext : &d->Eext;
            if (*host) {
                if (!*dev) *dev = zpic_dmalloc(sizeof(float3) * nodes);
                zpic_h2d(*dev, *host, sizeof(float3) * nodes);
            }
        }
        zpic_h2d(d->E[d->cur], emf->E_buf, sizeof(float3) * nodes);
        zpic_h2d(d->B[d->cur], emf->B_buf, sizeof(float3) * nodes);
        /* the particle fields are re-formed from E and B; the mirror
         * E_part_buf/B_part_buf was computed the same way */
        part_fields(emf);
        d->host_dirty = 0;
    }
}

void emf_mark_host_dirty(t_emf *emf)
{
    if (emf->dev) {
        emf->dev->host_dirty = 1;
        emf->dev->host_valid = 1;
    }
}

void emf_sync_host(t_emf *emf)
{
    struct zpic_dev_emf *d = emf->dev;
    if (!d || !d->ready || d->host_valid || d->host_dirty) return;
    const size_t bytes = sizeof(float3) * zpic_grid_nodes(emf->nx);
    zpic_d2h(emf->E_buf, d->E[d->cur], bytes);
    zpic_d2h(emf->B_buf, d->B[d->cur], bytes);
    if (emf->ext_fld.E_type != EMF_FLD_TYPE_NONE && d->Ep) zpic_d2h(emf->ext_fld.E_part_buf, d->Ep, bytes);
    if (emf->ext_fld.B_type != EMF_FLD_TYPE_NONE && d->Bp) zpic_d2h(emf->ext_fld.B_part_buf, d->Bp, bytes);
    d->host_valid = 1;
}

const float *zpic_emf_particle_E(t_emf *emf)
{
    struct zpic_dev_emf *d = emf->dev;
    return emf->ext_fld.E_type != EMF_FLD_TYPE_NONE ? d->Ep : d->E[d->cur];
}

const float *zpic_emf_particle_B(t_emf *emf)
{
    struct zpic_dev_emf *d = emf->dev;
    return emf->ext_fld.B_type != EMF_FLD_TYPE_NONE ? d->Bp : d->B[d->cur];
}

/* ---- time step (em2d/emf.c:688-716) ---- */

void emf_advance(t_emf *emf, const t_current *ccur)
{
    t_current *cur = (t_current *)ccur;
    zpic_emf_ensure(emf);
    zpic_current_ensure(cur, NULL);
    struct zpic_dev_emf *d = emf->dev;
    zpic_timer_begin(ZPIC_T_EMF);
    const int o = 1 - d->cur;
    /* a window move after this step (em2d/emf.c:648-650, with the
     * iteration count the step ends with) is written by the field update
     * itself, unless external fields are formed from the unshifted grid in
     * between (em2d/emf.c:703) */
    const int move = emf->moving_window && ((emf->iter + 1) * emf->dt) > emf->dx[0] * (emf->n_move + 1);
    const int fused = move && zpic_window_shift_fusion() && emf->ext_fld.E_type == EMF_FLD_TYPE_NONE &&
                      emf->ext_fld.B_type == EMF_FLD_TYPE_NONE;
    zpic_ktimer_begin(ZPIC_T_FIELD);
    ZCHECK(zk_yee(d->E[d->cur], d->B[d->cur], zpic_current_J(cur), d->E[o], d->B[o], emf->nx[0], emf->nx[1],
                  emf->nrow, emf->dt, emf->dx[0], emf->dx[1], d->open_y, fused, zpic_stream()));
    ZCHECK(zk_field_guards(d->E[o], d->B[o], emf->nx[0], emf->nx[1], emf->nrow, !emf->moving_window, d->open_y,
                           zpic_stream()));
    zpic_ktimer_end(ZPIC_T_FIELD);
    d->cur = o;
    part_fields(emf);
    emf->iter += 1;
    if (fused) {
        emf->n_move++;
    } else if (move) {
        const int m = 1 - d->cur;
        ZCHECK(zk_move_window(d->E[d->cur], d->B[d->cur], d->E[m], d->B[m], emf->nx[0], emf->nx[1], emf->nrow,
                              zpic_stream()));
        d->cur = m;
        emf->n_move++;
    }
    zpic_timer_end(ZPIC_T_EMF);
    d->host_valid = 0;
}

double emf_time(void) { return zpic_timer_total(ZPIC_T_EMF); }

void emf_get_energy(const t_emf *cemf, double energy[])
{
    t_emf *emf = (t_emf *)cemf;
    zpic_emf_ensure(emf);
    struct zpic_dev_emf *d = emf->dev;
    const double scale = 0.5 * emf->dx[0] * emf->dx[1];
    ZCHECK(zk_emf_energy(d->E[d->cur], d->B[d->cur], emf->nx[0], emf->nx[1], emf->nrow, d->epartial, d->nblk,
                         d->energy6, scale, zpic_stream()));
    zpic_d2h(energy, d->energy6, sizeof(double) * 6);
}

void emf_report(const t_emf *cemf, const char field, const int fc)
{
    /* em2d/emf.c:368-480: one component of E, B or the particle fields,
     * interior nodes, to EMF/<name>-<iter>.zdf */
    t_emf *emf = (t_emf *)cemf;
    if (fc < 0 || fc > 2) {
        fprintf(stderr, "(*error*) Invalid field component (fc) selected, returning\n");
        return;
    }
    const char comp[] = {'x', 'y', 'z'};
    char name[16], label[16];
    const float3 *f;
    emf_sync_host(emf);
    switch (field) {
    case EFLD:
        f = emf->E;
        snprintf(name, 16, "E%1d", fc);
        snprintf(label, 16, "E_%c", comp[fc]);
        break;
    case BFLD:
        f = emf->B;
        snprintf(name, 16, "B%1d", fc);
        snprintf(label, 16, "B_%c", comp[fc]);
        break;
    case EPART:
        f = emf->E_part;
        snprintf(name, 16, "E%1d-part", fc);
        snprintf(label, 16, "E_{%cp}", comp[fc]);
        break;
    case BPART:
        f = emf->B_part;
        snprintf(name, 16, "B%1d-part", fc);
        snprintf(label, 16, "B_{%cp}", comp[fc]);
        break;
    default:
        fprintf(stderr, "Invalid field type selected, returning\n");
        return;
    }
    float *buf = malloc(sizeof(float) * (size_t)emf->nx[0] * (size_t)emf->nx[1]);
    for (int j = 0; j < emf->nx[1]; j++)
        for (int i = 0; i < emf->nx[0]; i++) {
            const float3 v = f[i + j * emf->nrow];
            buf[i + j * emf->nx[0]] = fc == 0 ? v.x : (fc == 1 ? v.y : v.z);
        }
    const zdf_axis axis[2] = {{"x", "x", "c/\\omega_p", 0.0, emf->box[0]}, {"y", "y", "c/\\omega_p", 0.0, emf->box[1]}};
    const zdf_grid g = {name, label, "m_e c \\omega_p e^{-1}", 2, {(uint64_t)emf->nx[0], (uint64_t)emf->nx[1], 0}, axis};
    const zdf_iter it = {"ITERATION", emf->iter, emf->iter * emf->dt, "1/\\omega_p"};
    zpic_zdf_save_grid(buf, &g, &it, "EMF");
    free(buf);
}
