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

/* the state rand_norm() draws from on the calling thread (taken back from
 * the device first when injections there drew last) */
zpic_rng_state *zpic_rng(void);

/* The process-global state on the device for the column injections of
 * zk_inject_column, handed over from the host when the host holds the
 * newer one; NULL when the calling thread draws from a private state. */
struct zk_rng_state;
struct zk_rng_state *zpic_rng_device(void);
/* uncertified logs of the device draws so far (takes the state back) */
int zpic_rng_device_ambiguous(void);

#endif
