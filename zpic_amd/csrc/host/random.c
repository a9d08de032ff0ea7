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
 * The generator state is process-global, as in the reference (with
 * zpic_set_device_inject it is handed to the device, where the moving
 * window's columns are injected, and taken back before the host draws).  A thread
 * that initialises an independent simulation while other threads do the
 * same first calls zpic_rand_thread_private(1): from then on that thread
 * draws from a state of its own (default seeds, no kept deviate) and
 * leaves the process-global one alone.
 */
#include <math.h>
#include <stdint.h>

#include "random.h"
#include "../kernels/zk.h"
#include "zpic_gpu.h"
#include "zpic_rng.h"
#include "zpic_rt.h"

/* reference default seeds, em2d/random.c:16-17 */
static zpic_rng_state g_state = {12345u, 67890u, 0.0, 0};
static _Thread_local zpic_rng_state t_state;
static _Thread_local int t_private = 0;

/* The process-global state can also live on the device, where injected
 * columns draw from it (zk_inject_column): g_on_device says that the
 * device copy is the newer one; the host takes it back before drawing. */
static zk_rng_state *g_dev, *g_stage;
static int g_on_device;
static int g_dev_ambiguous; /* uncertified logs of the device draws */

zpic_rng_state *zpic_rng(void)
{
    if (t_private) return &t_state;
    if (g_on_device) {
        zk_rng_state d;
        zpic_d2h(&d, g_dev, sizeof(d));
        g_state.w = d.w;
        g_state.z = d.z;
        g_state.kept = d.kept;
        g_state.have_kept = d.have_kept;
        g_dev_ambiguous = d.ambiguous;
        g_on_device = 0;
    }
    return &g_state;
}

zk_rng_state *zpic_rng_device(void)
{
    if (t_private) return NULL;
    if (!g_dev) {
        g_dev = zpic_dmalloc(sizeof(zk_rng_state));
        g_stage = zpic_pinned_alloc(sizeof(zk_rng_state));
    }
    if (!g_on_device) {
        /* the staging buffer is rewritten only after the state came back,
         * which waited for this copy */
        g_stage->w = g_state.w;
        g_stage->z = g_state.z;
        g_stage->kept = g_state.kept;
        g_stage->have_kept = g_state.have_kept;
        g_stage->ambiguous = g_dev_ambiguous;
        zpic_h2d_async(g_dev, g_stage, sizeof(*g_stage));
        g_on_device = 1;
    }
    return g_dev;
}

int zpic_rng_device_ambiguous(void)
{
    (void)zpic_rng(); /* takes the state back if the device holds it */
    return g_dev_ambiguous;
}

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
