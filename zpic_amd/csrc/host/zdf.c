/*
 * zdf.c - ZDF diagnostic files (grids and particle sets).
 *
 * Writes the ZDF format of the reference's diagnostics (em2d/zdf.c,
 * em2d/zdf.h) so the *_report functions produce the same files: a 4-byte
 * magic "ZDF1", then records of (uint32 id, string name, uint64 length,
 * payload), little endian, every item padded to 4 bytes; strings are a
 * uint32 length and the bytes.  Grid files hold a TYPE string "grid",
 * the grid info, the iteration and one dataset; particle files hold TYPE
 * "particles", the particle-set info, the iteration and one 1D float32
 * dataset per quantity.  File names: <path>/<name>-<%06u iter>.zdf and
 * <path>/particles-<name>-<%06u iter>.zdf (zdf_open_grid_file and
 * zdf_open_part_file, em2d/zdf.c:1500-1600).
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include "zpic_zdf.h"

/* record ids (id << 16 | version), em2d/zdf.c:76-86 */
enum {
    REC_STRING = 0x00030000,
    REC_DATASET = 0x00100002,
    REC_ITERATION = 0x00200001,
    REC_GRID_INFO = 0x00210001,
    REC_PART_INFO = 0x00220002
};
enum { ZDF_FLOAT32 = 9 }; /* position of float32 in the reference's type list, em2d/zdf.h:65-77 */

static size_t pad4(size_t n) { return (n + 3) & ~(size_t)3; }

typedef struct {
    FILE *fp;
    int ok;
    uint32_t ndatasets;
} zfile;

static void put(zfile *z, const void *p, size_t n)
{
    if (z->ok && n && fwrite(p, 1, n, z->fp) != n) z->ok = 0;
}
static void put_u32(zfile *z, uint32_t v) { put(z, &v, 4); }
static void put_i32(zfile *z, int32_t v) { put(z, &v, 4); }
static void put_u64(zfile *z, uint64_t v) { put(z, &v, 8); }
static void put_f64(zfile *z, double v) { put(z, &v, 8); }

static size_t str_size(const char *s)
{
    const size_t n = s ? strlen(s) : 0;
    return 4 + pad4(n);
}

static void put_str(zfile *z, const char *s)
{
    const size_t n = s ? strlen(s) : 0;
    static const char zero[4] = {0, 0, 0, 0};
    put_u32(z, (uint32_t)n);
    put(z, s, n);
    put(z, zero, pad4(n) - n);
}

static void put_record(zfile *z, uint32_t id, const char *name, uint64_t length)
{
    put_u32(z, id);
    put_str(z, name);
    put_u64(z, length);
}

/* mkdir -p */
static void make_path(const char *path)
{
    char buf[1024];
    snprintf(buf, sizeof buf, "%s", path);
    for (char *p = buf + 1; *p; p++) {
        if (*p != '/') continue;
        *p = 0;
        mkdir(buf, 0755);
        *p = '/';
    }
    mkdir(buf, 0755);
}

static int open_file(zfile *z, const char *filename)
{
    z->fp = fopen(filename, "w+b");
    z->ok = z->fp != NULL;
    z->ndatasets = 0;
    if (!z->ok) {
        fprintf(stderr, "(*error*) Unable to open ZDF file %s: %s\n", filename, strerror(errno));
        return 0;
    }
    put(z, "ZDF1", 4);
    return z->ok;
}

static int close_file(zfile *z)
{
    const int ok = z->ok;
    if (z->fp && fclose(z->fp)) return 0;
    z->fp = NULL;
    return ok;
}

static void put_string_rec(zfile *z, const char *name, const char *value)
{
    put_record(z, REC_STRING, name, str_size(value));
    put_str(z, value);
}

static void put_iteration(zfile *z, const zdf_iter *it)
{
    put_record(z, REC_ITERATION, it->name, 4 + 8 + str_size(it->time_units));
    put_i32(z, it->n);
    put_f64(z, it->t);
    put_str(z, it->time_units);
}

static void put_dataset(zfile *z, const char *name, const float *data, int ndims, const uint64_t *count)
{
    uint64_t n = 1;
    for (int i = 0; i < ndims; i++) n *= count[i];
    const uint64_t header = 4 + 4 + 4 + 8 * (uint64_t)ndims;
    put_record(z, REC_DATASET, name, header + 4 * n);
    put_u32(z, ++z->ndatasets);
    put_i32(z, ZDF_FLOAT32);
    put_u32(z, (uint32_t)ndims);
    for (int i = 0; i < ndims; i++) put_u64(z, count[i]);
    put(z, data, 4 * (size_t)n);
}

int zpic_zdf_save_grid(const float *data, const zdf_grid *g, const zdf_iter *it, const char *path)
{
    char filename[1200];
    make_path(path);
    snprintf(filename, sizeof filename, "%s/%s-%06u.zdf", path, g->name, (unsigned)it->n);
    zfile z;
    if (!open_file(&z, filename)) return 0;
    put_string_rec(&z, "TYPE", "grid");
    /* grid info, em2d/zdf.c:994-1058 */
    size_t len = 4 + 8 * (size_t)g->ndims + str_size(g->label) + str_size(g->units) + 4;
    for (int i = 0; i < g->ndims; i++)
        len += str_size(g->axis[i].name) + 4 + 16 + str_size(g->axis[i].label) + str_size(g->axis[i].units);
    put_record(&z, REC_GRID_INFO, g->name, len);
    put_u32(&z, (uint32_t)g->ndims);
    for (int i = 0; i < g->ndims; i++) put_u64(&z, g->count[i]);
    put_str(&z, g->label);
    put_str(&z, g->units);
    put_i32(&z, 1); /* axis information follows */
    for (int i = 0; i < g->ndims; i++) {
        put_str(&z, g->axis[i].name);
        put_i32(&z, 0); /* linear axis */
        put_f64(&z, g->axis[i].min);
        put_f64(&z, g->axis[i].max);
        put_str(&z, g->axis[i].label);
        put_str(&z, g->axis[i].units);
    }
    put_iteration(&z, it);
    put_dataset(&z, g->name, data, g->ndims, g->count);
    return close_file(&z);
}

int zpic_zdf_save_particles(const char *name, const char *label, uint64_t np, int nquants,
                            const char *const *quants, const char *const *qlabels, const char *const *qunits,
                            const float *const *data, const zdf_iter *it, const char *path)
{
    char filename[1200];
    make_path(path);
    snprintf(filename, sizeof filename, "%s/particles-%s-%06u.zdf", path, name, (unsigned)it->n);
    zfile z;
    if (!open_file(&z, filename)) return 0;
    put_string_rec(&z, "TYPE", "particles");
    /* particle-set info, em2d/zdf.c:1063-1111 */
    size_t len = str_size(label) + 8 + 4;
    for (int i = 0; i < nquants; i++) len += str_size(quants[i]) + str_size(qlabels[i]) + str_size(qunits[i]);
    put_record(&z, REC_PART_INFO, name, len);
    put_str(&z, label);
    put_u64(&z, np);
    put_u32(&z, (uint32_t)nquants);
    for (int i = 0; i < nquants; i++) put_str(&z, quants[i]);
    for (int i = 0; i < nquants; i++) put_str(&z, qlabels[i]);
    for (int i = 0; i < nquants; i++) put_str(&z, qunits[i]);
    put_iteration(&z, it);
    for (int i = 0; i < nquants; i++) {
        const uint64_t c = np;
        put_dataset(&z, quants[i], data[i], 1, &c);
    }
    return close_file(&z);
}
