/*
 * random.c - process-global normal deviates for particle injection.
 *
 * Same generator and sequence as the reference em2d/random.c:16-101:
 * George Marsaglia's multiply-with-carry pair (x <- 36969 x + carry,
 * y <- 18000 y + carry, 16-bit halves, result (x << 16) + y) feeding the
 * polar Box-Muller method, which returns one deviate and keeps the other
 * for the next call.  The kept deviate survives set_rand_seed, as in the
 * reference, so a run reproduces the reference's injected momenta exactly.
 *
 * The generator state is per thread (the reference's is process-global):
 * a single-threaded program sees the reference's behaviour, and several
 * host threads can initialise independent simulations at once.
 */
#include <math.h>
#include <stdint.h>

#include "random.h"
#include "zpic_gpu.h"

static _Thread_local uint32_t g_w = 12345u; /* reference default seeds, em2d/random.c:16-17 */
static _Thread_local uint32_t g_z = 67890u;
static _Thread_local double g_kept;
static _Thread_local int g_have_kept = 0;

void set_rand_seed(uint32_t w, uint32_t z)
{
    g_w = w;
    g_z = z;
}

void zpic_rand_clear(void) { g_have_kept = 0; }

static uint32_t next_u32(void)
{
    g_z = 36969u * (g_z & 0xffffu) + (g_z >> 16);
    g_w = 18000u * (g_w & 0xffffu) + (g_w >> 16);
    return (g_z << 16) + g_w;
}

double rand_norm(void)
{
    if (g_have_kept) {
        g_have_kept = 0;
        return g_kept;
    }
    double u, v, s;
    do {
        u = (next_u32() + 0.5) / 2147483649.0 - 1.0;
        v = (next_u32() + 0.5) / 2147483649.0 - 1.0;
        s = u * u + v * v;
    } while (s == 0.0 || s >= 1.0);
    const double f = sqrt(-2.0 * log(s) / s);
    g_kept = u * f;
    g_have_kept = 1;
    return v * f;
}
