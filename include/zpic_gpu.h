/*
 * zpic_gpu.h - additions of libzpic_em2d to the em2d C API.
 *
 * None of these functions exists in the reference.  The em2d entry points
 * (sim_iter, spec_advance, current_update, emf_advance, ...) keep their
 * reference signatures and run on the GPU; these calls manage the host
 * mirrors of the device-resident state and expose tuning and test hooks.
 *
 * Host mirrors: after a step the particle buffer `spec->part`/`spec->np`,
 * `spec->energy` and the grids `E_buf`, `B_buf`, `J_buf` are stale until
 * the matching *_sync_host call.  A caller that edits a mirror calls the
 * matching *_mark_host_dirty so the edit is uploaded before the next step.
 * The reference's Python module handed out views of these buffers
 * (python/source/em2d.pyx:305-312, 1044-1150, 1567-1619); its replacement
 * calls these functions around every access.
 */
#ifndef ZPIC_GPU_H
#define ZPIC_GPU_H

#include <stdint.h>

#include "simulation.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Select the HIP device before the first GPU call (default: current device). */
void zpic_set_device(int device);

/* Tile shape for species/currents created afterwards: tx x ty cells with a
 * drift margin of `halo` cells; 0 keeps the default (16, 16, 3).  Any grid
 * size is accepted: a side that the preferred width does not divide uses a
 * divisor within a quarter of that width, or else tiles of that width with
 * a narrower last tile.  A tile whose window (tile + 2 halo + 3 nodes per
 * side) would not fit the LDS of one workgroup is shrunk until it does. */
void zpic_set_tiling(int tx, int ty, int halo);
/* The same for one simulation only (its species and current), before its
 * first step; 0 entries keep the process-wide setting. */
void zpic_sim_set_tiling(t_simulation *sim, int tx, int ty, int halo);

/* Drop the deviate the normal generator keeps for its next call.
 * set_rand_seed keeps it, as the reference does (em2d/random.c:25-29,
 * 69-70); clearing it lets independent runs in one process start from
 * identical generator states. */
void zpic_rand_clear(void);

/* Generator state of the calling thread: 0 (default) draws from the
 * process-global state, as the reference's single generator
 * (em2d/random.c:16-17); 1 gives this thread a state of its own, starting
 * from the default seeds with no kept deviate, so that several threads can
 * initialise independent simulations at once. */
void zpic_rand_thread_private(int on);

/* Variant of the float-atomic push kernel (zpic_set_fixed_deposit(0)):
 * 0 (default) keeps every value in registers, 8 requests 8 waves per SIMD
 * with a few register spills, 2 pushes two particles per thread and
 * iteration.  The fixed-point push (the default) has one variant. */
void zpic_set_push_waves(int waves);

/* Current deposit of the push: 1 (default; environment variable
 * ZPIC_FIXED_DEPOSIT=0 turns it off) sums each tile window exactly in
 * 64-bit fixed point (units of 2^-40 Q, Q the power of two at or above the
 * particle charge) with integer LDS atomics, independent of the order of
 * the atomics, and rounds each window node once; 0 sums the windows with
 * float LDS atomics. */
void zpic_set_fixed_deposit(int on);

/* sim_iter pushes all species of a simulation in one tile launch when they
 * share the tiling and the fixed-point scale (same |q| rounded up to a
 * power of two) and the fixed-point deposit is on: each tile's fields are
 * staged and its current window is stored once per step instead of once
 * per species.  1 (default; environment variable ZPIC_FUSE_SPECIES=0
 * turns it off) or 0 (one launch per species, as spec_advance).  The
 * fused window is rounded to FP32 once for all species together, so the
 * current differs from the per-species launches at FP32 rounding. */
void zpic_set_species_fusion(int on);

/* Window moves written by the field update: on a step that ends in a move
 * of the simulation window (and with no external fields) the fused Yee
 * kernel writes the shifted grid (one column left, rightmost columns
 * zeroed) instead of a separate pass copying the whole grid after it.
 * 1 (default; environment variable ZPIC_FUSE_WINDOW_SHIFT=0 turns it off)
 * or 0.  Both give the same fields bit for bit. */
void zpic_set_window_shift_fusion(int on);

/* Staged re-sorts (no counterpart in em2d, whose spec_sort orders the
 * particles by cell on the host): the push of a step that ends in a tile
 * sort writes the particles that stayed in their tile packed to the front
 * of the tile's range of the other buffer set and the others to a staging
 * area, and only those are then moved to their new tiles (tile ranges keep
 * free slots for them; arrivals that do not fit wait in an unsorted tail
 * until the next full sort).  0 (default; every sort reads and writes all
 * particles, ranges without free slots), 1 (environment variable
 * ZPIC_STAGED_SORT=1) or 2 (one free slot per range, so that ranges fill
 * up: for tests).  Results do not depend on it: the order of the particles
 * is not part of them and the current is summed exactly.  On the 2048^2
 * Weibel benchmark the sorting pushes (+0.4 ms each) and the spread-out
 * ranges (+3 % per push) cost what the cheaper re-sorts save, so it is off
 * by default. */
void zpic_set_staged_sort(int mode);
/* Staged re-sorts run so far in this process (a diagnostic), and with
 * zpic_staged_stats: those, the full re-sorts that followed them (when a
 * particle was left outside its tile), and the particles so left: leavers
 * that found no staging room, arrivals that found no free slot. */
long long zpic_staged_sorts(void);
void zpic_staged_stats(long long out[4]);

/* Open (absorbing) y boundaries for a simulation, before its first step.
 * No counterpart in em2d, whose y boundary is always periodic; the model
 * is the open boundary of the reference's 1D code (em1d/emf.c:379-406,
 * em1d/particles.h:77-80):
 *  - fields: first-order Mur conditions on the tangential components Ex,
 *    Ez of the guard rows j = -1 and j = ny, applied after the E update of
 *    each step: E_g(n+1) = E_in(n) + S (E_in(n+1) - E_g(n)) with E_in the
 *    adjacent interior row and S = (dt - dy) / (dt + dy); the guard rows
 *    are no longer periodic copies;
 *  - particles: a particle whose cell leaves [0, ny) is removed;
 *  - current: no fold along y (the raw guard rows are kept, and y filter
 *    passes leave the guard rows unchanged); spec_deposit_charge does not
 *    fold along y.
 * The x boundary stays periodic, or absorbing with the moving window.
 * Batches and differentiable runs refuse such simulations. */
void zpic_sim_set_open_y(t_simulation *sim);

/* spec_deposit_charge (em2d/particles.c:1289-1324) without allocating: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the deposit is staged in a file-scope device array of
 * the library (64 MiB per device, held until the process ends) instead of a
 * grid allocated per call.  Same arguments and result as
 * spec_deposit_charge (the density is added to `charge`, (nx+1) x (ny+1)
 * floats, then folded at the periodic boundaries).  Limits: at most
 * 8,388,608 charge nodes ((nx+1)(ny+1), e.g. 2896 x 2896 cells), and the
 * species must already be on the device without pending host edits (after
 * its first step); otherwise it prints why and returns -1 without touching
 * `charge`.  Returns 0 on success.  Every call on a GPU uses that GPU's one
 * array: calls that overlap on one GPU (from several threads) race. */
int spec_deposit_charge_noalloc(const t_species *spec, float *charge);

/* Particle initialisation of spec_new on the device, for the uniform, step
 * and slab densities (custom densities are always initialised on the
 * host): 1 on, 0 (default; environment variable ZPIC_DEVICE_INIT=1 turns
 * it on) off.  The device makes the same lattice positions, draws the
 * thermal momenta from the reference generator's own stream (continued
 * from its present state, which is then advanced past the deviates used,
 * em2d/random.c:48-101) and removes the same per-cell means
 * (em2d/particles.c:85-354).  The polar method's log is correctly rounded
 * on the device, which the C library's is not everywhere (it differs in
 * about 0.06 % of the logs), so a momentum can in rare cases differ from
 * the host initialisation's in its last bit.  The particles stay on the
 * device: spec->part is filled by spec_sync_host (before the first step
 * as after it).  emf_add_laser then also makes its pulse on the device
 * (em2d/emf.c:158-350, the same expressions). */
void zpic_set_device_init(int on);
/* the setting in force (ZPIC_DEVICE_INIT=1 in the environment turns it on
 * when zpic_set_device_init has not been called) */
int zpic_device_init(void);

/* Moving-window injection on the device: with 1 (environment variable
 * ZPIC_DEVICE_INJECT=1) each column a window move injects (uniform, step
 * and slab profiles, em2d/particles.c:476-492) is made by one kernel,
 * which draws the thermal momenta from a device copy of the generator
 * state; the host takes the state back before it draws itself (rand_norm,
 * spec_new, injection on the host), so the sequence is the reference's.
 * 0 (default): the column is made on the host and read by the device from
 * page-locked memory, which overlaps with the GPU's work. */
void zpic_set_device_inject(int on);
/* logs of the device initialisation whose correct rounding the
 * double-double evaluation could not certify (expected 0) */
long long zpic_device_init_ambiguous(void);

/* Block until all queued GPU work is done. */
void zpic_synchronize(void);

/* Refresh host mirrors from the device. */
void spec_sync_host(t_species *spec);
void emf_sync_host(t_emf *emf);
void current_sync_host(t_current *current);
void sim_sync_host(t_simulation *sim);

/* Declare a host mirror edited: it is uploaded before the next step. */
void spec_mark_host_dirty(t_species *spec);
void emf_mark_host_dirty(t_emf *emf);

/* Particle identities for one-to-one comparisons: after spec_track_ids the
 * particles present get ids 0..np-1 in host-buffer order and injected ones
 * continue the count.  spec_get_ids copies the ids in the order of the
 * host mirror filled by the last spec_sync_host. */
void spec_track_ids(t_species *spec);
void spec_get_ids(t_species *spec, int64_t *ids);

/* Pushes, accumulated since the species was created, that found their
 * particle outside the tile window (and took the global-memory path); a
 * diagnostic of the tile margin. */
int spec_overflow_count(t_species *spec);

/* Per-kernel GPU time, measured with HIP events on the library's stream
 * around each launch while enabled.  Classes: 0 push and deposit, 1 tile
 * sort, 2 current reduction and smoothing, 3 field solver. */
enum zpic_kernel_class { ZPIC_K_PUSH = 0, ZPIC_K_SORT = 1, ZPIC_K_CURRENT = 2, ZPIC_K_FIELD = 3 };
void zpic_kernel_timing(int enable);
double zpic_kernel_time(int kclass, long long *launches); /* seconds */
void zpic_kernel_timing_reset(void);

/* Adapter for field callbacks that cannot return a float3 by value (e.g.
 * from Python): set E_custom/B_custom to zpic_field_callback and the data
 * pointer to a zpic_field_cb whose fn writes the three components. */
typedef struct zpic_field_cb {
    void (*fn)(int ix, float dx, int iy, float dy, float *out3, void *user);
    void *user;
} zpic_field_cb;
float3 zpic_field_callback(int ix, float dx, int iy, float dy, void *data);

/* Batches of independent simulations (parameter scans).  No counterpart
 * in the reference, whose nearest relative runs the simulations of a scan
 * one after another (mods/em1d_param_scan/main.c:50-98).
 *
 * zpic_batch_new takes n simulations made with sim_new (any grid size,
 * species count, particle count, filter and sort period; periodic, no
 * moving window, no external fields) and zpic_batch_run advances every
 * one of them by `steps` iterations: each iteration is exactly sim_iter
 * of each simulation, but all simulations advance together, one kernel
 * launch per phase of the step, replayed from a captured HIP graph.
 * Afterwards every simulation is in the state sim_iter would leave it in
 * (host mirrors refreshed by the usual *_sync_host calls).  A simulation
 * may be stepped with sim_iter, or edited, between batch runs; its grid,
 * species and filter must not change while the batch exists.  Free the
 * batch before the simulations. */
typedef struct zpic_batch zpic_batch;
zpic_batch *zpic_batch_new(t_simulation *const *sims, int n);
void zpic_batch_run(zpic_batch *batch, int steps);
void zpic_batch_delete(zpic_batch *batch);

/* Differentiable runs.  No counterpart in the reference.
 *
 * zpic_diff_new prepares `steps` iterations of `sim` (periodic, no moving
 * window, no external fields, no current filter) from its present state,
 * with a checkpoint every `ckpt` steps; the particles are tile-sorted at
 * each checkpoint instead of on the species' own schedule.  The loss is
 *     L = sum over n in [loss_first, loss_last] of dt * (E_B3 + E_B4 + E_B5)(n)
 * with E_B the six terms of emf_get_energy after n of the steps (FP64).
 * zpic_diff_forward runs the steps and returns L; zpic_diff_backward then
 * returns dL/d(initial state): dE and dB as guard-inclusive grids laid out
 * like E_buf/B_buf ((nx+3)*(ny+3) float3), and for species k, dpart[k]
 * holds 5 floats per particle (x, y, ux, uy, uz; position adjoints are
 * per cell width) in the order of spec->part when zpic_diff_new was
 * called.  Any pointer may be NULL.  Afterwards the simulation is in its
 * state after the forward run. */
typedef struct zpic_diff zpic_diff;
zpic_diff *zpic_diff_new(t_simulation *sim, int steps, int ckpt, int loss_first, int loss_last);
double zpic_diff_forward(zpic_diff *run);
void zpic_diff_backward(zpic_diff *run, float *dE, float *dB, float *const *dpart);
void zpic_diff_delete(zpic_diff *run);
/* Before zpic_diff_forward: on = 1 makes the forward run keep the state at
 * the start of every step (28 bytes per particle and step, plus E and B
 * per step) so the backward pass runs no segment twice; on = 0 (default)
 * keeps checkpoints only and re-runs each segment during the backward
 * pass.  The gradients are the same either way. */
void zpic_diff_record_all(zpic_diff *run, int on);

#ifdef __cplusplus
}
#endif

#endif
