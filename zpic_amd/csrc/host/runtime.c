/*
 * runtime.c - HIP device, stream, memory, timers and tile choice.
 *
 * The em2d API is single-threaded and not reentrant, like the reference
 * (process-global timers and RNG, em2d/particles.c:26-27, em2d/emf.c:27,
 * em2d/random.c:16-17), so one process-wide stream carries all work.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include <hip/hip_runtime_api.h>

#include "zpic_rt.h"

static int g_device = -1;
static int g_ready = 0;
static hipStream_t g_stream;

void zpic_set_device(int device);
void zpic_set_tiling(int tx, int ty, int halo);
void zpic_synchronize(void);

static int g_tx = 16, g_ty = 16, g_halo = 3;
static int g_push_waves = 0;

void zpic_set_push_waves(int waves) { g_push_waves = (waves == 8 || waves == 2) ? waves : 0; }
int zpic_push_waves(void) { return g_push_waves; }

/* -1: not chosen yet (environment variable ZPIC_DEVICE_INIT, default off) */
static int g_device_init = -1;
void zpic_set_device_init(int on) { g_device_init = on ? 1 : 0; }
int zpic_device_init(void)
{
    if (g_device_init < 0) {
        const char *e = getenv("ZPIC_DEVICE_INIT");
        g_device_init = (e && e[0] == '1') ? 1 : 0;
    }
    return g_device_init;
}

/* -1: not chosen yet (environment variable ZPIC_FIXED_DEPOSIT, default on) */
static int g_fixed_deposit = -1;
void zpic_set_fixed_deposit(int on) { g_fixed_deposit = on ? 1 : 0; }
int zpic_fixed_deposit(void)
{
    if (g_fixed_deposit < 0) {
        const char *e = getenv("ZPIC_FIXED_DEPOSIT");
        g_fixed_deposit = (e && e[0] == '0') ? 0 : 1;
    }
    return g_fixed_deposit;
}

/* -1: not chosen yet (environment variable ZPIC_STAGED_SORT, default off) */
static int g_staged_sort = -1;
void zpic_set_staged_sort(int mode) { g_staged_sort = mode == 2 ? 2 : mode ? 1 : 0; }
int zpic_staged_sort(void)
{
    if (g_staged_sort < 0) {
        const char *e = getenv("ZPIC_STAGED_SORT");
        g_staged_sort = (e && e[0] == '2') ? 2 : (e && e[0] == '1') ? 1 : 0;
    }
    return g_staged_sort;
}

/* -1: not chosen yet (environment variable ZPIC_FUSE_SPECIES, default on) */
static int g_fuse_species = -1;
void zpic_set_species_fusion(int on) { g_fuse_species = on ? 1 : 0; }
int zpic_species_fusion(void)
{
    if (g_fuse_species < 0) {
        const char *e = getenv("ZPIC_FUSE_SPECIES");
        g_fuse_species = (e && e[0] == '0') ? 0 : 1;
    }
    return g_fuse_species;
}

/* -1: not chosen yet (environment variable ZPIC_DEVICE_INJECT, default off) */
static int g_device_inject = -1;
void zpic_set_device_inject(int on) { g_device_inject = on ? 1 : 0; }
int zpic_device_inject(void)
{
    if (g_device_inject < 0) {
        const char *e = getenv("ZPIC_DEVICE_INJECT");
        g_device_inject = (e && e[0] == '1') ? 1 : 0;
    }
    return g_device_inject;
}

/* -1: not chosen yet (environment variable ZPIC_FUSE_WINDOW_SHIFT, default on) */
static int g_fuse_shift = -1;
void zpic_set_window_shift_fusion(int on) { g_fuse_shift = on ? 1 : 0; }
int zpic_window_shift_fusion(void)
{
    if (g_fuse_shift < 0) {
        const char *e = getenv("ZPIC_FUSE_WINDOW_SHIFT");
        g_fuse_shift = (e && e[0] == '0') ? 0 : 1;
    }
    return g_fuse_shift;
}

/* Bits of the x cell index in the packed particle cell (ix | iy << bits):
 * the least b with nx <= 2^b; the pair must fit 31 bits. */
int zpic_cell_bits(int nx, int ny)
{
    int bx = 0, by = 0;
    while ((1 << bx) < nx) bx++;
    while ((1 << by) < ny) by++;
    if (bx + by > 31) zpic_fatal("grid %d x %d has too many cells for 31-bit cell indices", nx, ny);
    return bx;
}

int zpic_fixed_point_exp(float q)
{
    /* Q = 2^k, the power of two at or above |q| */
    int k = 0;
    const float aq = fabsf(q);
    if (aq > 0.0f) {
        frexpf(aq, &k); /* aq = f 2^k, f in [0.5, 1) */
        if (ldexpf(1.0f, k - 1) == aq) k -= 1;
    }
    return k;
}

/* Scale of the push's fixed-point current windows: 2^32 / Q.  A window
 * node's sum must stay below 2^50 (zk_fx_read), i.e. 2^18 Q, which tiles
 * of at most ZK_FX_TILE_MAX particles guarantee (the push counts the tiles
 * over it); contributions are rounded to 2^-33 Q. */
void zpic_fixed_point_scale(float q, float *scale, float *unscale)
{
    const int k = zpic_fixed_point_exp(q);
    *scale = ldexpf(1.0f, 32 - k);
    *unscale = ldexpf(1.0f, k - 32);
}

void zpic_check(int status, const char *what, const char *file, int line)
{
    if (status != 0) {
        fprintf(stderr, "zpic_em2d: HIP error '%s' in %s (%s:%d), aborting.\n",
                hipGetErrorString((hipError_t)status), what, file, line);
        exit(-1);
    }
}

void zpic_fatal(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "zpic_em2d: ");
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    exit(-1);
}

void zpic_set_device(int device)
{
    if (g_ready && device != g_device) zpic_fatal("zpic_set_device must be called before the first GPU use");
    g_device = device;
}

void zpic_set_tiling(int tx, int ty, int halo)
{
    g_tx = tx > 0 ? tx : 16;
    g_ty = ty > 0 ? ty : 16;
    g_halo = halo > 0 ? halo : 3;
}

static void runtime_init(void)
{
    if (g_ready) return;
    int n = 0;
    ZCHECK(hipGetDeviceCount(&n));
    if (n <= 0) zpic_fatal("no HIP device available");
    if (g_device < 0) ZCHECK(hipGetDevice(&g_device));
    ZCHECK(hipSetDevice(g_device));
    ZCHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_ready = 1;
}

zk_stream_t zpic_stream(void)
{
    runtime_init();
    return g_stream;
}

void *zpic_dmalloc(size_t bytes)
{
    runtime_init();
    void *p = NULL;
    if (bytes == 0) bytes = 256;
    ZCHECK(hipMalloc(&p, bytes));
    return p;
}

void zpic_dfree(void *p)
{
    if (p) ZCHECK(hipFree(p));
}

void zpic_h2d(void *dst, const void *src, size_t bytes)
{
    if (bytes) ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, zpic_stream()));
    ZCHECK(hipStreamSynchronize(zpic_stream()));
}

void zpic_d2h(void *dst, const void *src, size_t bytes)
{
    if (bytes) ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, zpic_stream()));
    ZCHECK(hipStreamSynchronize(zpic_stream()));
}

void zpic_d2h_async(void *dst, const void *src, size_t bytes)
{
    if (bytes) ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, zpic_stream()));
}

void zpic_d2d_async(void *dst, const void *src, size_t bytes)
{
    if (bytes) ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, zpic_stream()));
}

void zpic_sync(void)
{
    if (g_ready) ZCHECK(hipStreamSynchronize(g_stream));
}

void zpic_synchronize(void) { zpic_sync(); }

size_t zpic_grid_nodes(const int nx[2]) { return (size_t)(nx[0] + 3) * (size_t)(nx[1] + 3); }

/* Tile width for one axis: the preferred width when it divides n, else the
 * divisor of n closest to it within a quarter of it, else the preferred
 * width itself, with a narrower last tile (the tile count is rounded up);
 * never wider than n. */
static int pick_tile(int n, int want)
{
    if (want >= n) return n;
    if (n % want == 0) return want;
    int best = -1;
    for (int d = want - want / 4; d <= want + want / 4; d++) {
        if (d < 1 || n % d) continue;
        if (best < 0 || abs(d - want) < abs(best - want)) best = d;
    }
    return best > 0 ? best : want;
}

void zpic_tiles_for(int nx, int ny, zk_tiles *t) { zpic_tiles_pref(nx, ny, 0, 0, 0, t); }

/* Bytes of LDS the fixed-point push needs per window node: E and B (24)
 * and the 64-bit current window (24). */
#define ZPIC_PUSH_LDS_PER_NODE 48
#define ZPIC_LDS_MAX (160 * 1024)

void zpic_tiles_pref(int nx, int ny, int want_x, int want_y, int want_h, zk_tiles *t)
{
    if (nx < 1 || ny < 1) zpic_fatal("grid %d x %d has no cells", nx, ny);
    int wx = want_x > 0 ? want_x : g_tx, wy = want_y > 0 ? want_y : g_ty;
    const int h = want_h > 0 ? want_h : g_halo;
    /* the push's window must fit the LDS of one workgroup: shrink the
     * larger side of the requested tile until it does */
    while ((size_t)ZPIC_PUSH_LDS_PER_NODE * (size_t)(wx + 2 * h + 3) * (size_t)(wy + 2 * h + 3) > ZPIC_LDS_MAX) {
        if (wx <= 1 && wy <= 1) zpic_fatal("tile margin %d leaves no room for a tile in the LDS", h);
        if (wx >= wy) wx = wx > 1 ? wx - 1 : 1;
        else wy = wy > 1 ? wy - 1 : 1;
    }
    const int tx = pick_tile(nx, wx), ty = pick_tile(ny, wy);
    t->nx = nx;
    t->ny = ny;
    t->TX = tx;
    t->TY = ty;
    t->H = h;
    t->WX = tx + 2 * h + 3;
    t->WY = ty + 2 * h + 3;
    t->ntx = (nx + tx - 1) / tx;
    t->nty = (ny + ty - 1) / ty;
    t->ntiles = t->ntx * t->nty;
}

/* ---- GPU timers: event pairs accumulated lazily ---- */
#define ZPIC_EV_PAIRS 64
static struct {
    hipEvent_t ev[ZPIC_EV_PAIRS][2];
    int used;
    int open;
    double total;
    long long count;
    int created;
} g_timer[ZPIC_NTIMERS];
static int g_ktiming = 0;

/* accumulate the oldest n recorded intervals and drop them */
static void timer_drain_n(int w, int n)
{
    if (n <= 0) return;
    ZCHECK(hipEventSynchronize(g_timer[w].ev[n - 1][1]));
    for (int k = 0; k < n; k++) {
        float ms = 0.0f;
        ZCHECK(hipEventElapsedTime(&ms, g_timer[w].ev[k][0], g_timer[w].ev[k][1]));
        g_timer[w].total += 1e-3 * ms;
        g_timer[w].count++;
    }
    /* the drained event pairs move behind the ones still in flight */
    hipEvent_t keep[ZPIC_EV_PAIRS][2];
    const int rest = g_timer[w].used - n;
    for (int k = 0; k < rest; k++) {
        keep[k][0] = g_timer[w].ev[n + k][0];
        keep[k][1] = g_timer[w].ev[n + k][1];
    }
    for (int k = 0; k < n; k++) {
        keep[rest + k][0] = g_timer[w].ev[k][0];
        keep[rest + k][1] = g_timer[w].ev[k][1];
    }
    for (int k = 0; k < g_timer[w].used; k++) {
        g_timer[w].ev[k][0] = keep[k][0];
        g_timer[w].ev[k][1] = keep[k][1];
    }
    g_timer[w].used = rest;
}

static void timer_drain(int w) { timer_drain_n(w, g_timer[w].used); }

void zpic_timer_begin(int w)
{
    hipStream_t s = zpic_stream();
    if (!g_timer[w].created) {
        for (int k = 0; k < ZPIC_EV_PAIRS; k++) {
            /* timing only: no system-scope fence when recorded (with one,
             * each record wrote back and invalidated the caches and held
             * the next kernel for about 6 us) */
            ZCHECK(hipEventCreateWithFlags(&g_timer[w].ev[k][0], hipEventDisableSystemFence));
            ZCHECK(hipEventCreateWithFlags(&g_timer[w].ev[k][1], hipEventDisableSystemFence));
        }
        g_timer[w].created = 1;
    }
    /* full: drain the older half, recorded long enough ago that the host
     * does not wait for the GPU here */
    if (g_timer[w].used == ZPIC_EV_PAIRS) timer_drain_n(w, ZPIC_EV_PAIRS / 2);
    ZCHECK(hipEventRecord(g_timer[w].ev[g_timer[w].used][0], s));
    g_timer[w].open = 1;
}

void zpic_timer_end(int w)
{
    if (!g_timer[w].open) return;
    ZCHECK(hipEventRecord(g_timer[w].ev[g_timer[w].used][1], zpic_stream()));
    g_timer[w].used++;
    g_timer[w].open = 0;
}

double zpic_timer_total(int w)
{
    if (g_ready) timer_drain(w);
    return g_timer[w].total;
}

void zpic_ktimer_begin(int w)
{
    if (g_ktiming) zpic_timer_begin(w);
}

void zpic_ktimer_end(int w)
{
    if (g_ktiming) zpic_timer_end(w);
}

void zpic_kernel_timing(int enable)
{
    if (!enable && g_ktiming) {
        for (int w = ZPIC_T_PUSH; w < ZPIC_NTIMERS; w++) timer_drain(w);
    }
    g_ktiming = enable;
}

double zpic_kernel_time(int kclass, long long *launches)
{
    const int w = ZPIC_T_PUSH + kclass;
    if (w < ZPIC_T_PUSH || w >= ZPIC_NTIMERS) return 0.0;
    if (g_ready) timer_drain(w);
    if (launches) *launches = g_timer[w].count;
    return g_timer[w].total;
}

void zpic_kernel_timing_reset(void)
{
    for (int w = ZPIC_T_PUSH; w < ZPIC_NTIMERS; w++) {
        if (g_ready) timer_drain(w);
        g_timer[w].total = 0.0;
        g_timer[w].count = 0;
    }
}

/* ---- HIP graphs of the batched step ---- */

void zpic_capture_begin(void) { ZCHECK(hipStreamBeginCapture(zpic_stream(), hipStreamCaptureModeThreadLocal)); }

void *zpic_capture_end(void)
{
    hipGraph_t g = NULL;
    hipGraphExec_t x = NULL;
    ZCHECK(hipStreamEndCapture(zpic_stream(), &g));
    ZCHECK(hipGraphInstantiate(&x, g, NULL, NULL, 0));
    ZCHECK(hipGraphDestroy(g));
    return (void *)x;
}

void zpic_graph_launch(void *exec) { ZCHECK(hipGraphLaunch((hipGraphExec_t)exec, zpic_stream())); }

void zpic_graph_free(void *exec)
{
    if (exec) ZCHECK(hipGraphExecDestroy((hipGraphExec_t)exec));
}

void zpic_h2d_async(void *dst, const void *src, size_t bytes)
{
    if (bytes) ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, zpic_stream()));
}

/* Page-locked host memory, mapped and coherent: kernels read it over the
 * bus without caching it, so a kernel sees what the host wrote before its
 * launch (the injected columns are read this way, once each). */
void *zpic_pinned_alloc(size_t bytes)
{
    void *p = NULL;
    ZCHECK(hipHostMalloc(&p, bytes > 0 ? bytes : 1, hipHostMallocMapped | hipHostMallocCoherent));
    return p;
}

void *zpic_pinned_device(void *p)
{
    void *d = NULL;
    ZCHECK(hipHostGetDevicePointer(&d, p, 0));
    return d;
}

void zpic_pinned_free(void *p)
{
    if (p) ZCHECK(hipHostFree(p));
}

/* An event that tells the host when the work before it has run; it makes
 * no device writes visible to the host (no system-scope fence), so it
 * guards only the reuse of buffers that the device reads. */
void *zpic_event_new(void)
{
    hipEvent_t e;
    ZCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
    return (void *)e;
}

/* An event after which the host reads what the work before it wrote (a
 * copy into page-locked memory): recorded with the system-scope fence. */
void *zpic_event_new_read(void)
{
    hipEvent_t e;
    ZCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return (void *)e;
}

void zpic_event_record(void *ev) { ZCHECK(hipEventRecord((hipEvent_t)ev, zpic_stream())); }

void zpic_event_wait(void *ev) { ZCHECK(hipEventSynchronize((hipEvent_t)ev)); }

void zpic_event_free(void *ev)
{
    if (ev) ZCHECK(hipEventDestroy((hipEvent_t)ev));
}
