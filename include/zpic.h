/*
 * zpic.h - basic vector type shared by every em2d grid.
 *
 * Drop-in for the reference header em2d/zpic.h:18-22 (float3) and its
 * math constants (em2d/zpic.h:24-35).  The layout of float3 is part of the
 * ABI: every field/current grid is an array of float3 {x,y,z}, 12 bytes,
 * with no padding.
 */
#ifndef ZPIC_H_DROPIN
#define ZPIC_H_DROPIN

typedef struct Float3 {
    float x;
    float y;
    float z;
} float3;

#ifndef M_PI
#define M_PI   3.14159265358979323846264338327950288
#endif
#ifndef M_PI_2
#define M_PI_2 1.57079632679489661923132169163975144
#endif
#ifndef M_PI_4
#define M_PI_4 0.785398163397448309615660845819875721
#endif

#endif
