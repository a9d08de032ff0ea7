/*
 * random.c - normal deviates for particle injection.
 *
 * Same generator and sequence as the reference em2d/random.c:16-101:
 * George Marsaglia's multiply-with-carry pair (x <- 36969 x + carry,
 * y <- 18000 y + carry, 16-bit halves, result (x << 16) + y) feeding the
 * polar Box-Muller method, which returns one deviate and keeps the other
 * for the next call.  The kept deviate survives set_rand_seed, as in the
 * reference, so a run reproduces the reference's injected momenta exactly.
 *
 * The generator state is process-global, as in the reference.  A thread
 * that initialises an independent simulation while other threads do the
 * same first calls zpic_rand_thread_private(1): from then on that thread
 * draws from a state of its own (default seeds, no kept deviate) and
 * leaves the process-global one alone.
 */
#include <math.h>
#include <stdint.h>

#include "random.h"
#include "zpic_gpu.h"
#include "zpic_rng.h"

/* reference default seeds, em2d/random.c:16-17 */
static zpic_rng_state g_state = {12345u, 67890u, 0.0, 0};
static _Thread_local zpic_rng_state t_state;
static _Thread_local int t_private = 0;

zpic_rng_state *zpic_rng(void) { return t_private ? &t_state : &g_state; }

void zpic_rand_thread_private(int on)
{
    if (on && !t_private) {
        t_state.w = 12345u;
        t_state.z = 67890u;
        t_state.kept = 0.0;
        t_state.have_kept = 0;
    }
    t_private = on ? 1 : 0;
}

void set_rand_seed(uint32_t w, uint32_t z)
{
    zpic_rng_state *s = zpic_rng();
    s->w = w;
    s->z = z;
}

void zpic_rand_clear(void) { zpic_rng()->have_kept = 0; }

static uint32_t next_u32(zpic_rng_state *s)
{
    s->z = 36969u * (s->z & 0xffffu) + (s->z >> 16);
    s->w = 18000u * (s->w & 0xffffu) + (s->w >> 16);
    return (s->z << 16) + s->w;
}

double rand_norm(void)
{
    zpic_rng_state *st = zpic_rng();
    if (st->have_kept) {
        st->have_kept = 0;
        return st->kept;
    }
    double u, v, s;
    do {
        u = (next_u32(st) + 0.5) / 2147483649.0 - 1.0;
        v = (next_u32(st) + 0.5) / 2147483649.0 - 1.0;
        s = u * u + v * v;
    } while (s == 0.0 || s >= 1.0);
    const double f = sqrt(-2.0 * log(s) / s);
    st->kept = u * f;
    st->have_kept = 1;
    return v * f;
}
