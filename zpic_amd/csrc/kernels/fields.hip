/*
 * fields.hip - Yee field solver and field bookkeeping (gfx950).
 *
 * zk_yee fuses the three half/whole steps of the reference emf_advance
 * (em2d/emf.c:688-698): B += dt/2 curl E over i,j in [-1,n] (yee_b,
 * :500-522), E += dt (curl B - J) over [0,n+1] (yee_e, :531-562), and
 * B += dt/2 curl E again, into one pass over LDS tiles.  Each node is
 * computed with the reference's expressions and update ranges, so with the
 * same J the result is bit-identical.  Inputs and outputs are distinct
 * buffers (the host swaps them).
 *
 * zk_field_guards copies guard cells from their periodic images
 * (emf_update_gc, :573-637); zk_move_window shifts the grid one cell left
 * and clears the last columns (emf_move_window, :648-675); zk_add_ext
 * forms the fields seen by particles (emf_update_part_fld, :838-914);
 * zk_emf_energy sums E_c^2 and B_c^2 in FP64 (emf_get_energy, :729-750);
 * zk_charge deposits the CIC charge density (spec_deposit_charge,
 * em2d/particles.c:1289-1324, before its boundary fold).
 */
#include <mutex>

#include "fields_core.h"

using namespace zk;

/* Scratch of spec_deposit_charge_noalloc: the deposit (first half) and the
 * caller's density (second half), ZK_NOALLOC_CHARGE_MAX nodes each.  Every
 * device holds its own copy for the life of the process. */
static __device__ __align__(256) float scratch[2 * ZK_NOALLOC_CHARGE_MAX];

namespace {

__global__ void __launch_bounds__(256) yee_kernel(const float *__restrict__ E, const float *__restrict__ B,
                                                  const float *__restrict__ J, float *__restrict__ Eo,
                                                  float *__restrict__ Bo, int nx, int ny, int nrow, YeeK k,
                                                  int open_y, int shift)
{
    yee_tile(E, B, J, Eo, Bo, nx, ny, nrow, k, blockIdx.x, blockIdx.y, open_y, shift);
}

__global__ void guard_kernel(float *E, float *B, int nx, int ny, int nrow, int periodic_x, int open_y)
{
    guard_node(E, B, nx, ny, nrow, periodic_x, blockIdx.x * blockDim.x + threadIdx.x, open_y);
}

__global__ void move_window_kernel(const float *E, const float *B, float *Eo, float *Bo, int nx, int ny, int nrow)
{
    const int bw = nx + 3, bh = ny + 3;
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= bw * bh) return;
    const int i = node % bw - 1;
    const size_t d = 3 * (size_t)node;
    if (i >= nx - 1) {
        for (int c = 0; c < 3; c++) Eo[d + c] = Bo[d + c] = 0.0f;
    } else {
        for (int c = 0; c < 3; c++) {
            Eo[d + c] = E[d + 3 + c];
            Bo[d + c] = B[d + 3 + c];
        }
    }
}

/* ---- laser pulse (em2d/emf.c:158-350), on the device ----
 * The reference's expressions: float operands, the C library's double
 * functions (sin, cos, exp, atan2, sqrt) with the results rounded to float
 * where the reference assigns them; the device's double functions agree
 * with the host's to within an ulp of double, so a float result can differ
 * only where the double value lies that close to a float rounding
 * boundary. */
__device__ inline float laser_envelope(const zk_laser_k &L, const float z)
{
    if (z > L.start) return 0.0f;
    if (z > L.start - L.rise) {
        const float u = z - L.start;
        const float e = (float)sin(1.5707963267948966 * u / L.rise);
        return e * e;
    }
    if (z > L.start - (L.rise + L.flat)) return 1.0f;
    if (z > L.start - (L.rise + L.flat + L.fall)) {
        const float u = z - (L.start - L.rise - L.flat - L.fall);
        const float e = (float)sin(1.5707963267948966 * u / L.fall);
        return e * e;
    }
    return 0.0f;
}

__device__ inline float laser_phase(const zk_laser_k &L, const float zpos, const float r)
{
    const float z = zpos - L.focus;
    const float zr = L.omega0 * (L.W0 * L.W0) / 2;
    const float r2 = r * r;
    const float curv = r2 * z / (zr * zr + z * z);
    const float w2 = (zr * zr) / (zr * zr + z * z);
    const float gouy = (float)atan2((double)z, (double)zr);
    return (float)(sqrt(sqrt((double)w2)) * exp((double)(-r2 * w2 / (L.W0 * L.W0))) *
                   cos((double)(L.omega0 * (z + curv) - gouy)));
}

/* one thread per interior node; E, B from the buffer origin */
__global__ void laser_kernel(float *E, float *B, int nx, int ny, int nrow, zk_laser_k L)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nx * ny) return;
    const int i = node % nx, j = node / nx;
    const size_t g = 3 * ((size_t)(i + 1) + (size_t)(j + 1) * nrow);
    const float z = i * L.hx;
    const float zh = z + L.hx / 2;
    const float env = L.amp * laser_envelope(L, z);
    const float envh = L.amp * laser_envelope(L, zh);
    if (L.type == 0) { /* plane: the products and the sums in double, as in the reference */
        const double c = cos((double)(L.omega0 * z)), ch = cos((double)(L.omega0 * zh));
        E[g + 1] = (float)((double)E[g + 1] + (double)env * c * (double)L.cp);
        E[g + 2] = (float)((double)E[g + 2] + (double)env * c * (double)L.sp);
        B[g + 1] = (float)((double)B[g + 1] + (double)(-envh) * ch * (double)L.sp);
        B[g + 2] = (float)((double)B[g + 2] + (double)envh * ch * (double)L.cp);
    } else {
        const float r = j * L.hy - L.axis;
        const float rh = r + L.hy / 2;
        E[g + 1] += +env * laser_phase(L, z, rh) * L.cp;
        E[g + 2] += +env * laser_phase(L, z, r) * L.sp;
        B[g + 1] += -envh * laser_phase(L, zh, r) * L.sp;
        B[g + 2] += +envh * laser_phase(L, zh, rh) * L.cp;
    }
}

/* E_x, B_x from div E = div B = 0, integrated in double from the right
 * edge of each row (em2d/emf.c:213-231): one thread per row */
__global__ void laser_div_kernel(float *E, float *B, int nx, int ny, int nrow, double ratio)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ny) return;
#define AT(a, b) (3 * ((size_t)((a) + 1) + (size_t)((b) + 1) * nrow))
    double ex = 0.0, bx = 0.0;
    for (int i = nx - 1; i >= 0; i--) {
        ex += ratio * (E[AT(i + 1, j) + 1] - E[AT(i + 1, j - 1) + 1]);
        E[AT(i, j)] = (float)ex;
        bx += ratio * (B[AT(i, j + 1) + 1] - B[AT(i, j) + 1]);
        B[AT(i, j)] = (float)bx;
    }
#undef AT
}

__global__ void add_ext_kernel(const float *F, const float *ext, float ux, float uy, float uz, float *out,
                               size_t nodes)
{
    for (size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x; n < nodes; n += (size_t)gridDim.x * blockDim.x) {
        float a0 = ext ? ext[3 * n + 0] : ux;
        float a1 = ext ? ext[3 * n + 1] : uy;
        float a2 = ext ? ext[3 * n + 2] : uz;
        out[3 * n + 0] = F[3 * n + 0] + a0;
        out[3 * n + 1] = F[3 * n + 1] + a1;
        out[3 * n + 2] = F[3 * n + 2] + a2;
    }
}

__global__ void energy_partial(const float *E, const float *B, int nx, int ny, int nrow, double *partial)
{
    __shared__ double red[16];
    double s[6] = {0, 0, 0, 0, 0, 0};
    const int cells = nx * ny;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) {
        const int i = c % nx, j = c / nx;
        const size_t g = 3 * ((size_t)(i + 1) + (size_t)(j + 1) * nrow);
        for (int q = 0; q < 3; q++) {
            const float e = E[g + q], b = B[g + q];
            s[q] += (double)(e * e);
            s[3 + q] += (double)(b * b);
        }
    }
    for (int q = 0; q < 6; q++) {
        const double v = zk_block_sum(s[q], red);
        if (threadIdx.x == 0) partial[6 * blockIdx.x + q] = v;
    }
}

__global__ void energy_final(const double *partial, int nblk, double *out6, double scale)
{
    if (threadIdx.x < 6) {
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += partial[6 * b + threadIdx.x];
        out6[threadIdx.x] = s * scale;
    }
}

__global__ void charge_kernel(zk_parts p, const zk_spec_ctr *ctr, float q, int nx, float *rho)
{
    const int n = ctr->np_total;
    const int nrow = nx + 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int cell = p.cell[i];
        if (cell < 0) continue;
        const int c = zk_cell_ix(cell, p.cbits) + nrow * zk_cell_iy(cell, p.cbits);
        const float wx = p.x[i], wy = p.y[i];
        unsafeAtomicAdd(&rho[c], (1.0f - wx) * (1.0f - wy) * q);
        unsafeAtomicAdd(&rho[c + 1], (wx) * (1.0f - wy) * q);
        unsafeAtomicAdd(&rho[c + nrow], (1.0f - wx) * (wy) * q);
        unsafeAtomicAdd(&rho[c + 1 + nrow], (wx) * (wy) * q);
    }
}

/* phase-space quantity code (X1 1, X2 2, U1 4, U2 5, U3 6) of particle i,
 * as em2d/particles.c:1569-1632 forms it */
__device__ inline float pha_quantity(const zk_parts &p, int i, int q, float dx0, float dx1)
{
    switch (q) {
    case 1: return (p.x[i] + zk_cell_ix(p.cell[i], p.cbits)) * dx0;
    case 2: return (p.y[i] + zk_cell_iy(p.cell[i], p.cbits)) * dx1;
    case 4: return p.ux[i];
    case 5: return p.uy[i];
    case 6: return p.uz[i];
    }
    return 0.0f;
}

/* Bilinear phase-space density of quantities (qa, qb): the reference's
 * weights and bounds per particle, summed into a block-private LDS copy of
 * the grid when it fits (lds = 1), else straight into the global grid. */
__global__ void pha_kernel(zk_parts p, const zk_spec_ctr *ctr, int qa, int qb, float dx0, float dx1, int na, int nb,
                           float amin, float bmin, float ra, float rb, float q, float *out, int lds)
{
    extern __shared__ float hist[];
    const int nbins = na * nb;
    float *h = lds ? hist : out;
    if (lds) {
        for (int m = threadIdx.x; m < nbins; m += blockDim.x) hist[m] = 0.0f;
        __syncthreads();
    }
    const int n = ctr->np_total;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (p.cell[i] < 0) continue;
        const float va = (pha_quantity(p, i, qa, dx0, dx1) - amin) * ra;
        const float vb = (pha_quantity(p, i, qb, dx0, dx1) - bmin) * rb;
        const int ia = (int)(va + 0.5f), ib = (int)(vb + 0.5f);
        const float wa = va - ia + 0.5f, wb = vb - ib + 0.5f;
        int c = ia + na * ib;
        if (ib >= 0 && ib < nb) {
            if (ia >= 0 && ia < na) atomicAdd(&h[c], (1.0f - wa) * (1.0f - wb) * q);
            if (ia + 1 >= 0 && ia + 1 < na) atomicAdd(&h[c + 1], wa * (1.0f - wb) * q);
        }
        c += na;
        if (ib + 1 >= 0 && ib + 1 < nb) {
            if (ia >= 0 && ia < na) atomicAdd(&h[c], (1.0f - wa) * wb * q);
            if (ia + 1 >= 0 && ia + 1 < na) atomicAdd(&h[c + 1], wa * wb * q);
        }
    }
    if (lds) {
        __syncthreads();
        for (int m = threadIdx.x; m < nbins; m += blockDim.x)
            if (hist[m] != 0.0f) unsafeAtomicAdd(&out[m], hist[m]);
    }
}

__global__ void add_kernel(float *acc, const float *v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc[i] += v[i];
}

#define ZK_MAX_DEVICES 64
float *g_scratch[ZK_MAX_DEVICES];
std::mutex g_scratch_lock;

} // namespace

/* The array's name is not a device address in host code: its address on a
 * device comes from hipGetSymbolAddress, called once per device with that
 * device current, and only that address is passed to kernels and copies. */
extern "C" int zk_noalloc_scratch(float **ptr)
{
    int dev = 0;
    int e = (int)hipGetDevice(&dev);
    if (e) return e;
    if (dev < 0 || dev >= ZK_MAX_DEVICES) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = (int)hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e) return e;
        g_scratch[dev] = (float *)p;
    }
    *ptr = g_scratch[dev];
    return 0;
}

extern "C" int zk_add_f32(float *acc, const float *v, size_t n, zk_stream_t s)
{
    if (n == 0) return 0;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(add_kernel, dim3((unsigned)blocks), dim3(256), 0, s, acc, v, n);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_pha(const zk_parts *p, int np_upper, const zk_spec_ctr *ctr, int qa, int qb, float dx0, float dx1,
                      int na, int nb, float amin, float bmin, float ra, float rb, float q, float *out, zk_stream_t s)
{
    if (na <= 0 || nb <= 0) return 0;
    const size_t lds = (size_t)na * nb * sizeof(float);
    const int use_lds = lds <= 48 * 1024;
    long long b = (np_upper + 255) / 256;
    if (b < 1) b = 1;
    if (b > (use_lds ? 1024 : 4096)) b = use_lds ? 1024 : 4096;
    hipLaunchKernelGGL(pha_kernel, dim3((unsigned)b), dim3(256), use_lds ? lds : 0, s, *p, ctr, qa, qb, dx0, dx1, na,
                       nb, amin, bmin, ra, rb, q, out, use_lds);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_yee(const float *E, const float *B, const float *J, float *Eo, float *Bo, int nx, int ny,
                             int nrow, float dt, float dx, float dy, int open_y, int shift, zk_stream_t s)
{
    YeeK k;
    const float dt2 = dt / 2.0f;
    k.dtx2 = dt2 / dx;
    k.dty2 = dt2 / dy;
    k.dtx = dt / dx;
    k.dty = dt / dy;
    k.dt = dt;
    k.mur_y = (dt - dy) / (dt + dy);
    dim3 grid((nx + 3 + BX - 1) / BX, (ny + 3 + BY - 1) / BY);
    hipLaunchKernelGGL(yee_kernel, grid, dim3(256), 0, s, E, B, J, Eo, Bo, nx, ny, nrow, k, open_y, shift);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_field_guards(float *E, float *B, int nx, int ny, int nrow, int periodic_x, int open_y,
                               zk_stream_t s)
{
    /* a moving window (no x copies) with open y (no y copies): nothing to do */
    if (!periodic_x && open_y) return 0;
    const int n = 3 * (nx + 3) + 3 * ny;
    hipLaunchKernelGGL(guard_kernel, dim3((n + 255) / 256), dim3(256), 0, s, E, B, nx, ny, nrow, periodic_x,
                       open_y);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_move_window(const float *E, const float *B, float *Eo, float *Bo, int nx, int ny, int nrow,
                                     zk_stream_t s)
{
    const int nodes = (nx + 3) * (ny + 3);
    hipLaunchKernelGGL(move_window_kernel, dim3((nodes + 255) / 256), dim3(256), 0, s, E, B, Eo, Bo, nx, ny, nrow);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_add_ext(const float *F, const float *ext, float ux, float uy, float uz, float *out,
                                 size_t nodes, zk_stream_t s)
{
    size_t blocks = (nodes + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(add_ext_kernel, dim3((unsigned)blocks), dim3(256), 0, s, F, ext, ux, uy, uz, out, nodes);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_emf_energy(const float *E, const float *B, int nx, int ny, int nrow, double *partial,
                                    int nblk, double *out6, double scale, zk_stream_t s)
{
    hipLaunchKernelGGL(energy_partial, dim3(nblk), dim3(256), 0, s, E, B, nx, ny, nrow, partial);
    hipLaunchKernelGGL(energy_final, dim3(1), dim3(64), 0, s, partial, nblk, out6, scale);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_charge(const zk_parts *p, int np_upper, const zk_spec_ctr *ctr, float q, int nx, float *rho,
                                zk_stream_t s)
{
    long long b = (np_upper + 255) / 256;
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(charge_kernel, dim3((unsigned)b), dim3(256), 0, s, *p, ctr, q, nx, rho);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_laser(float *E, float *B, int nx, int ny, int nrow, const zk_laser_k *L, double ratio, zk_stream_t s)
{
    const int nodes = nx * ny;
    if (nodes <= 0) return 0;
    hipLaunchKernelGGL(laser_kernel, dim3((nodes + 255) / 256), dim3(256), 0, s, E, B, nx, ny, nrow, *L);
    if (L->type != 0) hipLaunchKernelGGL(laser_div_kernel, dim3((ny + 63) / 64), dim3(64), 0, s, E, B, nx, ny, nrow, ratio);
    ZK_LAUNCH_CHECK();
}
