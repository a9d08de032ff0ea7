/*
 * zpic_rt.h - HIP runtime services used by the host layer, in plain C.
 *
 * Only runtime.c includes the HIP headers; everything else goes through
 * these functions, so the em2d headers (which define their own float3) and
 * the HIP vector types never meet in one translation unit.  Every failure
 * prints a message and exits with status -1, the reference's convention
 * for fatal errors.
 */
#ifndef ZPIC_RT_H
#define ZPIC_RT_H

#include <stddef.h>

#include "../kernels/zk.h"

#define ZCHECK(call) zpic_check((call), #call, __FILE__, __LINE__)
void zpic_check(int status, const char *what, const char *file, int line);
void zpic_fatal(const char *fmt, ...);

zk_stream_t zpic_stream(void);
void *zpic_dmalloc(size_t bytes);
void zpic_dfree(void *p);
void zpic_h2d(void *dst, const void *src, size_t bytes);      /* synchronous */
void zpic_d2h(void *dst, const void *src, size_t bytes);      /* synchronous */
void zpic_d2d_async(void *dst, const void *src, size_t bytes); /* on the stream */
void zpic_sync(void);
void zpic_h2d_async(void *dst, const void *src, size_t bytes); /* host buffer must outlive the copy */
void zpic_d2h_async(void *dst, const void *src, size_t bytes); /* into page-locked memory; wait on an
                                                                  event from zpic_event_new_read */

/* page-locked host memory (copies from it run asynchronously) and events
 * on the library's stream that the host can wait for */
void *zpic_pinned_alloc(size_t bytes);
void zpic_pinned_free(void *p);
void *zpic_pinned_device(void *p); /* device address of page-locked memory */
void *zpic_event_new(void);
void *zpic_event_new_read(void); /* for data the host reads after waiting */
void zpic_event_record(void *ev);
void zpic_event_wait(void *ev); /* returns at once for an event never recorded */
void zpic_event_free(void *ev);

/* stream capture into a HIP graph; zpic_capture_end returns the
 * instantiated graph, replayed on the stream by zpic_graph_launch */
void zpic_capture_begin(void);
void *zpic_capture_end(void);
void zpic_graph_launch(void *exec);
void zpic_graph_free(void *exec);
void zpic_tiles_for(int nx, int ny, zk_tiles *t);
/* the same with a preferred tile (0 entries: the process-wide setting) */
void zpic_tiles_pref(int nx, int ny, int want_x, int want_y, int want_h, zk_tiles *t);
int zpic_push_waves(void);
int zpic_fixed_deposit(void);
int zpic_device_init(void);
int zpic_species_fusion(void);
int zpic_window_shift_fusion(void);
int zpic_device_inject(void);
int zpic_staged_sort(void);
/* fixed-point deposit constants for particle charge q (see push_core.h) */
void zpic_fixed_point_scale(float q, float *scale, float *unscale);
int zpic_fixed_point_exp(float q); /* k with Q = 2^k the power of two at or above |q| */
int zpic_cell_bits(int nx, int ny); /* bits of ix in the packed particle cell */
size_t zpic_grid_nodes(const int nx[2]); /* (nx+3)*(ny+3) guard-inclusive nodes */

/* GPU timers: spec_time() / emf_time(), and per-kernel-class timers that
 * are recorded only while zpic_kernel_timing(1) is on */
enum {
    ZPIC_T_SPEC = 0,
    ZPIC_T_EMF = 1,
    ZPIC_T_PUSH = 2,
    ZPIC_T_SORT = 3,
    ZPIC_T_CURRENT = 4,
    ZPIC_T_FIELD = 5,
    ZPIC_NTIMERS = 6
};
void zpic_ktimer_begin(int which); /* no-op unless kernel timing is on */
void zpic_ktimer_end(int which);
void zpic_timer_begin(int which);
void zpic_timer_end(int which);
double zpic_timer_total(int which);

#endif
