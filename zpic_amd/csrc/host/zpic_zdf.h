/*
 * zpic_zdf.h - writer of ZDF diagnostic files (see zdf.c).
 */
#ifndef ZPIC_ZDF_H
#define ZPIC_ZDF_H

#include <stdint.h>

typedef struct {
    const char *name, *label, *units;
    double min, max;
} zdf_axis;

typedef struct {
    const char *name, *label, *units;
    int ndims;
    uint64_t count[3];
    const zdf_axis *axis;
} zdf_grid;

typedef struct {
    const char *name;
    int n;
    double t;
    const char *time_units;
} zdf_iter;

int zpic_zdf_save_grid(const float *data, const zdf_grid *grid, const zdf_iter *iter, const char *path);
int zpic_zdf_save_particles(const char *name, const char *label, uint64_t np, int nquants,
                            const char *const *quants, const char *const *qlabels, const char *const *qunits,
                            const float *const *data, const zdf_iter *iter, const char *path);

#endif
