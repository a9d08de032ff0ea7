/*
 * simulation.c - em2d simulation object and time-step driver.
 *
 * Same sequence and checks as the reference em2d/simulation.c:25-223; the
 * step functions it calls run on the GPU.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "timer.h"
#include "zpic_internal.h"

int report(int n, int ndump) { return (ndump > 0) ? !(n % ndump) : 0; }

void sim_iter(t_simulation *sim)
{
    /* em2d/simulation.c:45-56 */
    current_zero(&sim->current);
    /* spec_advance of every species; species that share their tiling are
     * pushed by one launch (same results up to the FP32 rounding of the
     * current windows) */
    zpic_spec_advance_all(sim->species, sim->n_species, &sim->emf, &sim->current);
    current_update(&sim->current);
    emf_advance(&sim->emf, &sim->current);
}

void sim_timings(t_simulation *sim, uint64_t t0, uint64_t t1)
{
    (void)sim;
    zpic_sync();
    fprintf(stderr, "Time for spec. advance = %f s\n", spec_time());
    fprintf(stderr, "Time for emf   advance = %f s\n", emf_time());
    fprintf(stderr, "Total simulation time  = %f s\n", timer_interval_seconds(t0, t1));
    fprintf(stderr, "\n");
    const double perf = spec_perf();
    if (perf > 0) {
        fprintf(stderr, "Particle advance [nsec/part] = %f \n", 1.e9 * perf);
        fprintf(stderr, "Particle advance [Mpart/sec] = %f \n", 1.e-6 / perf);
    }
}

void sim_new(t_simulation *sim, int nx[], float box[], float dt, float tmax, int ndump, t_species *species,
             int n_species)
{
    /* em2d/simulation.c:93-112 */
    sim->dt = dt;
    sim->tmax = tmax;
    sim->ndump = ndump;
    emf_new(&sim->emf, nx, box, dt);
    current_new(&sim->current, nx, box, dt);
    sim->n_species = n_species;
    sim->species = species;
    sim->moving_window = 0;
    const float cour = sqrtf(1.0f / (1.0f / (sim->emf.dx[0] * sim->emf.dx[0]) +
                                     1.0f / (sim->emf.dx[1] * sim->emf.dx[1])));
    if (dt >= cour) {
        fprintf(stderr, "Invalid timestep, courant condition violation, dtmax = %f \n", cour);
        exit(-1);
    }
}

void sim_add_laser(t_simulation *sim, t_emf_laser *laser) { emf_add_laser(&sim->emf, laser); }

void sim_set_smooth(t_simulation *sim, t_smooth *smooth)
{
    /* em2d/simulation.c:134-147 */
    if ((smooth->xtype != NONE) && (smooth->xlevel <= 0)) {
        fprintf(stderr, "Invalid smooth level along x direction\n");
        exit(-1);
    }
    if ((smooth->ytype != NONE) && (smooth->ylevel <= 0)) {
        fprintf(stderr, "Invalid smooth level along y direction\n");
        exit(-1);
    }
    sim->current.smooth = *smooth;
}

void sim_set_moving_window(t_simulation *sim)
{
    sim->emf.moving_window = 1;
    sim->current.moving_window = 1;
    for (int i = 0; i < sim->n_species; i++) sim->species[i].moving_window = 1;
}

void zpic_sim_set_open_y(t_simulation *sim)
{
    /* no counterpart in em2d, whose y boundary is always periodic; the
     * model is the open boundary of the reference's 1D code (Mur fields,
     * em1d/emf.c:379-406; particle removal, em1d/particles.h:77-80) */
    if (sim->emf.iter > 0) zpic_fatal("zpic_sim_set_open_y must be called before the first step");
    sim->emf.dev->open_y = 1;
    sim->current.dev->open_y = 1;
    for (int i = 0; i < sim->n_species; i++) sim->species[i].dev->open_y = 1;
}

void sim_set_ext_fld(t_simulation *sim, t_emf_ext_fld *ext_fld) { emf_set_ext_fld(&sim->emf, ext_fld); }

void sim_sync_host(t_simulation *sim)
{
    for (int i = 0; i < sim->n_species; i++) spec_sync_host(&sim->species[i]);
    emf_sync_host(&sim->emf);
    current_sync_host(&sim->current);
}

void sim_report_energy(t_simulation *sim)
{
    /* the first field component is counted twice, as in the reference
     * (em2d/simulation.c:183-205) */
    double e[6];
    emf_get_energy(&sim->emf, e);
    double fld = e[0];
    for (int k = 0; k < 6; k++) fld += e[k];
    double kin = 0;
    for (int i = 0; i < sim->n_species; i++) {
        zpic_spec_sync_energy(&sim->species[i]);
        kin += sim->species[i].energy;
    }
    printf("Energy (fields | particles | total) = %e %e %e\n", fld, kin, fld + kin);
}

void sim_delete(t_simulation *sim)
{
    for (int i = 0; i < sim->n_species; i++) spec_delete(&sim->species[i]);
    free(sim->species);
    current_delete(&sim->current);
    emf_delete(&sim->emf);
}

void zpic_sim_set_tiling(t_simulation *sim, int tx, int ty, int halo)
{
    /* takes effect for device state not created yet (before the first step) */
    for (int k = 0; k < sim->n_species; k++) {
        struct zpic_dev_species *d = sim->species[k].dev;
        if (!d) continue;
        if (d->ready) zpic_fatal("zpic_sim_set_tiling: species '%s' is already on the device", sim->species[k].name);
        d->pref[0] = tx;
        d->pref[1] = ty;
        d->pref[2] = halo;
    }
    struct zpic_dev_current *c = sim->current.dev;
    if (c) {
        if (c->ready) zpic_fatal("zpic_sim_set_tiling: the current density is already on the device");
        c->pref[0] = tx;
        c->pref[1] = ty;
        c->pref[2] = halo;
    }
}
