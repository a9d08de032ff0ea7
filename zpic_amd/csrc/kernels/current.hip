/*
 * current.hip - current density after deposition (gfx950).
 *
 * zk_current_reduce turns the per-tile current windows written by the
 * push kernel (plus the raw current of out-of-window pushes) into the final
 * guard-inclusive J.  It reproduces the reference's deposit-then-fold
 * semantics (em2d/current.c:118-159): deposits are accumulated at raw
 * node positions in [-1, n+1], and every node of a periodic axis receives
 * the sum of all its periodic images, guard copies included.  Along x with
 * a moving window, and along y with open boundaries, there is no fold and
 * the raw guard columns (rows) are kept.
 * One thread per output node; contributions are summed in a fixed order,
 * so the result does not depend on scheduling.
 *
 * zk_smooth applies one [sa, sb, sa] pass along x or y with the guard
 * handling of kernel_x / kernel_y (em2d/current.c:316-413).
 */
#include "current_core.h"

using namespace zk;

namespace {

__global__ void reduce_kernel(zk_reduce_args a) { reduce_node(a, blockIdx.x * blockDim.x + threadIdx.x); }

__global__ void smooth_kernel(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nrow,
                              int axis, float sa, float sb, int periodic_x, int open_y)
{
    smooth_node(in, out, nx, ny, nrow, axis, sa, sb, periodic_x, blockIdx.x * blockDim.x + threadIdx.x, open_y);
}

} // namespace

extern "C" int zk_current_reduce(const zk_reduce_args *a, zk_stream_t s)
{
    const int nodes = (a->t.nx + 3) * (a->t.ny + 3);
    hipLaunchKernelGGL(reduce_kernel, dim3((nodes + 255) / 256), dim3(256), 0, s, *a);
    ZK_LAUNCH_CHECK();
}

extern "C" int zk_smooth(const float *in, float *out, int nx, int ny, int nrow, int axis, float sa, float sb,
                                int periodic_x, int open_y, zk_stream_t s)
{
    const int nodes = (nx + 3) * (ny + 3);
    hipLaunchKernelGGL(smooth_kernel, dim3((nodes + 255) / 256), dim3(256), 0, s, in, out, nx, ny, nrow, axis, sa, sb,
                       periodic_x, open_y);
    ZK_LAUNCH_CHECK();
}
