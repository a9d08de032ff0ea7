/*
 * simulation.h - em2d simulation object and the time-step driver.
 *
 * Drop-in for the reference header em2d/simulation.h.  The struct layout
 * and every function signature match the reference (em2d/simulation.h:13-65).
 */
#ifndef SIMULATION_H_DROPIN
#define SIMULATION_H_DROPIN

#include <stdint.h>

#include "particles.h"
#include "emf.h"
#include "current.h"

/* reference em2d/simulation.h:13-29 */
typedef struct Simulation {
    float dt;
    float tmax;
    int ndump;
    int n_species;
    t_species *species;
    t_emf emf;
    t_current current;
    int moving_window;
} t_simulation;

/* Deck callbacks, supplied by the user program (reference em2d/simulation.h:40,52). */
void sim_init(t_simulation *sim);
void sim_report(t_simulation *sim);

/* reference em2d/simulation.h:54 - current_zero, spec_advance per species,
 * current_update, emf_advance */
void sim_iter(t_simulation *sim);
/* reference em2d/simulation.h:55 */
void sim_report_energy(t_simulation *sim);
/* reference em2d/simulation.h:57 - takes ownership of `species` */
void sim_new(t_simulation *sim, int nx[], float box[], float dt, float tmax,
             int ndump, t_species *species, int n_species);
/* reference em2d/simulation.h:58 */
int report(int n, int ndump);
/* reference em2d/simulation.h:59 */
void sim_timings(t_simulation *sim, uint64_t t0, uint64_t t1);
/* reference em2d/simulation.h:60 */
void sim_add_laser(t_simulation *sim, t_emf_laser *laser);
/* reference em2d/simulation.h:61 */
void sim_delete(t_simulation *sim);
/* reference em2d/simulation.h:62 */
void sim_set_moving_window(t_simulation *sim);
/* reference em2d/simulation.h:63 */
void sim_set_smooth(t_simulation *sim, t_smooth *smooth);
/* reference em2d/simulation.h:64 */
void sim_set_ext_fld(t_simulation *sim, t_emf_ext_fld *ext_fld);

#endif
