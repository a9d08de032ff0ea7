/*
 * bench_driver.c - CPU timing driver for the em2d time step.
 *
 * TEST INFRASTRUCTURE ONLY (CPU baseline).  Written against the em2d C API
 * only, so the same file is compiled against the reference sources
 * (oracle/_ref/ref_bench) and against the oracle restatement
 * (oracle/_build/oracle_bench).  It builds a two-species counter-streaming
 * Weibel setup (reference deck em2d/input/weibel.c:10-42 with the sizes
 * taken from the command line), runs `steps` calls of sim_iter and prints
 * one JSON line with the push count and the timers.
 *
 * usage: bench nx ny ppcx ppcy steps n_sort smooth_level [dx dt uz uth m_q2]
 * (m_q2 is the mass-to-charge ratio of the second, counter-drifting species;
 * +1 as in the deck, -1 for two electron populations)
 */
#include <stdio.h>
#include <stdlib.h>

#include "simulation.h"
#include "timer.h"

int main(int argc, char **argv)
{
    if (argc < 8) {
        fprintf(stderr, "usage: %s nx ny ppcx ppcy steps n_sort smooth_level [dx dt uz uth m_q2]\n", argv[0]);
        return 2;
    }
    int nx[2] = {atoi(argv[1]), atoi(argv[2])};
    int ppc[2] = {atoi(argv[3]), atoi(argv[4])};
    int steps = atoi(argv[5]);
    int n_sort = atoi(argv[6]);
    int level = atoi(argv[7]);
    float h = argc > 8 ? (float)atof(argv[8]) : 0.1f;
    float dt = argc > 9 ? (float)atof(argv[9]) : 0.07f;
    float uz = argc > 10 ? (float)atof(argv[10]) : 0.6f;
    float uth = argc > 11 ? (float)atof(argv[11]) : 0.1f;
    float mq2 = argc > 12 ? (float)atof(argv[12]) : 1.0f;
    float box[2] = {nx[0] * h, nx[1] * h};

    uint64_t ti0 = timer_ticks();
    t_species *sp = (t_species *)malloc(2 * sizeof(t_species));
    float ufl[3] = {0.0f, 0.0f, uz};
    float th[3] = {uth, uth, uth};
    spec_new(&sp[0], "electrons", -1.0f, ppc, ufl, th, nx, box, dt, NULL);
    ufl[2] = -uz;
    spec_new(&sp[1], mq2 > 0 ? "positrons" : "electrons2", mq2, ppc, ufl, th, nx, box, dt, NULL);
    sp[0].n_sort = n_sort;
    sp[1].n_sort = n_sort;
    t_simulation sim;
    sim_new(&sim, nx, box, dt, steps * dt, 0, sp, 2);
    if (level > 0) {
        t_smooth s = {BINOMIAL, BINOMIAL, level, level};
        sim_set_smooth(&sim, &s);
    }
    uint64_t ti1 = timer_ticks();

    uint64_t t0 = timer_ticks();
    for (int n = 0; n < steps; n++) sim_iter(&sim);
    uint64_t t1 = timer_ticks();

    double loop = timer_interval_seconds(t0, t1);
    printf("{\"np\": %d, \"steps\": %d, \"npush\": %llu, \"spec_time\": %.6f, \"emf_time\": %.6f, "
           "\"loop_time\": %.6f, \"init_time\": %.6f, \"particle_steps_per_s\": %.6e}\n",
           sim.species[0].np + sim.species[1].np, steps, (unsigned long long)spec_npush(), spec_time(),
           emf_time(), loop, timer_interval_seconds(ti0, ti1), (double)spec_npush() / loop);
    sim_delete(&sim);
    return 0;
}
