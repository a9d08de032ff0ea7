/*
 * zpic_rng.h - state of the injection generator (random.c), shared with
 * the device-side initialisation, which continues the same sequence.
 */
#ifndef ZPIC_RNG_H
#define ZPIC_RNG_H

#include <stdint.h>

typedef struct {
    uint32_t w, z;   /* multiply-with-carry halves (em2d/random.c:16-17) */
    double kept;     /* deviate kept by the polar method */
    int have_kept;
} zpic_rng_state;

/* the state rand_norm() draws from on the calling thread */
zpic_rng_state *zpic_rng(void);

#endif
