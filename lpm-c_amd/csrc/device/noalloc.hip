// updateRR with no allocation on any call: lpmx_update_rr's kernels and reductions
// with every intermediate in one library-owned static device array.
#include "common.h"

// Per-particle squares of the residual and reaction norms (2 * LPMX_NOALLOC_MAX_PARTICLES),
// the partials and total of the device sum (LPMX_RED_PARTS + 8), the constrained-DOF
// counter (8 doubles' room): 16.0 MiB for at most 1,048,576 particles.  Each GPU the
// library runs on holds its own copy until the process ends.
#define NA_SQ 0
#define NA_PARTS (2 * (int64_t)LPMX_NOALLOC_MAX_PARTICLES)
#define NA_COUNT (NA_PARTS + LPMX_RED_PARTS + 8)
#define NA_TOTAL (NA_COUNT + 8)
static __device__ __align__(256) double scratch[NA_TOTAL];

// Device address of this device's copy of `scratch`, looked up on the entry point's
// first call on a device, with that device current; the host-side name of the array
// is not a device address and is never passed to a kernel or a copy.
static double *na_ptr[64];

// Clears the constrained-DOF counter (a kernel, not hipMemsetAsync: a caller may
// capture the call in a graph).
__global__ void k_na_zero(int *p) { *p = 0; }

// Every call on a GPU shares that GPU's array: calls that overlap on one GPU race.
extern "C" int lpmx_update_rr_noalloc(double *nres, double *nrea, int *nreaction)
{
    RETCHK(lpmx_require_model());
    const int n = g.n;
    if (n > LPMX_NOALLOC_MAX_PARTICLES)
        return lpmx_fail(LPMX_ERR_ARG, "lpmx_update_rr_noalloc handles at most %d particles (the model has %d)",
                         LPMX_NOALLOC_MAX_PARTICLES, n);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        return lpmx_fail(LPMX_ERR_ARG, "device index %d out of range", dev);
    if (!na_ptr[dev])
        HIPCHK(hipGetSymbolAddress((void **)&na_ptr[dev], HIP_SYMBOL(scratch)));
    double *base = na_ptr[dev];
    double *sq = base + NA_SQ, *parts = base + NA_PARTS;
    int *cnt = (int *)(base + NA_COUNT);
    hipLaunchKernelGGL(k_update_rr, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, sq, sq + n);
    LAUNCHCHK();
    double a = 0, b = 0;
    if (nres)
    {
        RETCHK(lpmx_dev_sum_into(sq, n, 0, parts, &a));
        *nres = sqrt(a);
    }
    if (nrea)
    {
        RETCHK(lpmx_dev_sum_into(sq + n, n, 0, parts, &b));
        *nrea = sqrt(b);
    }
    if (nreaction)
    {
        hipLaunchKernelGGL(k_na_zero, dim3(1), dim3(1), 0, g_stream, cnt);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_count_bc, dim3(nblocks((int64_t)n * g.dim)), dim3(LPMX_BLOCK), 0, g_stream, g.dispBC,
                           n * g.dim, cnt);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(nreaction, cnt, sizeof(int), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    return LPMX_OK;
}
