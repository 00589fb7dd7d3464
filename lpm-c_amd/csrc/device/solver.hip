// Linear solve K disp = residual on the device: block-Jacobi preconditioned CG
// over the sliced-ELL 3x3 (2x2) block matrix, in FP64 or in the mixed form
// (FP32 values/vectors, compensated FP32 dot accumulation, FP64 refinement).
//
// One CG iteration is four launches.  k_cg_spmv forms p = z + beta p_old on the fly
// for every gathered column (p is double-buffered), writes p and q = K p for its
// own rows and one partial p.q per work-group; k_cg_alpha (one work-group) sums
// the partials in a fixed order into alpha.  k_cg_update does x += alpha p,
// r -= alpha q, z = M^-1 r and the partials of r.z and r.r; k_cg_beta forms beta
// and applies the stopping test.  The partials cross work-groups only at kernel
// boundaries: a grid-wide "last work-group" hand-off would need an agent-scope
// release fence in every work-group, which on a multi-XCD device writes back the
// XCD's L2 and costs more than the streaming work itself.  Every kernel returns at
// once when the device flag `done` is set, so the host launches batches of
// iterations and reads the flag once per batch.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "common.h"

int lpmx_bj_inverse(void);

struct CgS
{
    double rz, rr, rr0, thr, alpha, beta, pq;
    int it, maxit, done, pad;
};

// --------------------------------------------------------- compensated sums -----
// Error-free transformations (Knuth TwoSum, FMA-based TwoProduct) for the FP32
// dot products of the mixed form.
struct F2
{
    float hi, lo;
};

__device__ __forceinline__ F2 two_sum(float a, float b)
{
    float s = a + b;
    float bb = s - a;
    float e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__device__ __forceinline__ F2 f2_add(F2 a, F2 b)
{
    F2 s = two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
    return two_sum(s.hi, s.lo);
}

__device__ __forceinline__ F2 f2_dot_acc(F2 acc, float a, float b)
{
    float p = a * b;
    float e = fmaf(a, b, -p);
    F2 s = two_sum(acc.hi, p);
    s.lo += acc.lo + e;
    return s;
}

__device__ __forceinline__ F2 wave_sum_f2(F2 v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
    {
        F2 w = {__shfl_xor(v.hi, o, 64), __shfl_xor(v.lo, o, 64)};
        v = f2_add(v, w);
    }
    return v;
}

// Work-group total as a double: compensated FP32 within wavefronts, the few wave
// totals combined in FP64.
__device__ __forceinline__ double block_sum_f2(F2 v, double *sh)
{
    v = wave_sum_f2(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0)
        sh[w] = (double)v.hi + (double)v.lo;
    __syncthreads();
    double t = 0;
    for (int k = 0; k < nw; k++)
        t += sh[k];
    return t;
}

template <typename T>
struct Acc;
template <>
struct Acc<double>
{
    double v = 0.0;
    __device__ void add(double a, double b) { v += a * b; }
    __device__ double block(double *sh) { return block_sum(v, sh); }
};
template <>
struct Acc<float>
{
    F2 v = {0.0f, 0.0f};
    __device__ void add(float a, float b) { v = f2_dot_acc(v, a, b); }
    __device__ double block(double *sh) { return block_sum_f2(v, sh); }
};

// Sum of parts[stride*k + off], k < np, by one work-group in a fixed order (the
// partials were written by an earlier kernel on the same stream).
__device__ __forceinline__ double group_total(const double *parts, int np, int stride, int off, double *sh)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < np; k += blockDim.x)
        s += parts[(int64_t)stride * k + off];
    return block_sum(s, sh);
}

// ------------------------------------------------------------------ kernels ------
// One SELL slot of a row: the D x D block (nontemporal: the matrix streams through
// once per iteration and must not evict the gathered vector window from L2).
template <int D, typename TV, typename TX>
__device__ __forceinline__ void slot_mac(const TV *__restrict__ kb, const TX *__restrict__ z, int m, TX *acc)
{
    TX zm[D];
#pragma unroll
    for (int c = 0; c < D; c++)
        zm[c] = z[(int64_t)m * D + c];
#pragma unroll
    for (int r = 0; r < D; r++)
#pragma unroll
        for (int c = 0; c < D; c++)
            acc[r] += (TX)__builtin_nontemporal_load(kb + (r * D + c) * 64) * zm[c];
}

// The SpMV of one CG iteration in recurrence form: w = K z over the gathered
// columns (only z is gathered), then for the lane's own row p = z + beta p and
// q = w + beta q in place (q = K p by linearity), and the partial p.q.  U slots are
// loaded ahead per step to keep more matrix loads in flight.
template <int D, typename TV, typename TX, int U>
__global__ __launch_bounds__(256) void k_cg_spmv(int n, const int *__restrict__ sptr, const int *__restrict__ col,
                                                 const TV *__restrict__ K, const TX *__restrict__ z,
                                                 TX *__restrict__ p, TX *__restrict__ q, double *parts,
                                                 const CgS *sc)
{
    __shared__ double sh[4];
    if (sc->done)
        return;
    const int blk = xcd_block();
    const int i = blk * 256 + threadIdx.x;
    const TX beta = (TX)sc->beta;
    Acc<TX> dot;
    if (i < n)
    {
        const int s = i >> 6, l = i & 63;
        const int k0 = sptr[s], k1 = sptr[s + 1];
        TX acc[D];
#pragma unroll
        for (int r = 0; r < D; r++)
            acc[r] = 0;
        int k = k0;
        for (; k + U <= k1; k += U)
        {
            int m[U];
#pragma unroll
            for (int u = 0; u < U; u++)
                m[u] = __builtin_nontemporal_load(col + (int64_t)(k + u) * 64 + l);
#pragma unroll
            for (int u = 0; u < U; u++)
                slot_mac<D, TV, TX>(K + (int64_t)(k + u) * D * D * 64 + l, z, m[u], acc);
        }
        for (; k < k1; k++)
            slot_mac<D, TV, TX>(K + (int64_t)k * D * D * 64 + l, z, __builtin_nontemporal_load(col + (int64_t)k * 64 + l),
                                acc);
#pragma unroll
        for (int c = 0; c < D; c++)
        {
            const int64_t o = (int64_t)i * D + c;
            const TX pi = z[o] + beta * p[o];
            const TX qi = acc[c] + beta * q[o];
            p[o] = pi;
            q[o] = qi;
            dot.add(pi, qi);
        }
    }
    const double part = dot.block(sh);
    if (threadIdx.x == 0)
        parts[blk] = part;
}

// slots loaded ahead per step: FP32 blocks are half the bytes, so more fit in flight
template <typename TV>
constexpr int spmv_unroll() { return sizeof(TV) == 4 ? 4 : 1; }

// alpha = rz / p.q from the SpMV partials (one work-group).
__global__ __launch_bounds__(256) void k_cg_alpha(const double *parts, int np, CgS *sc)
{
    __shared__ double sh[4];
    if (sc->done)
        return;
    const double pq = group_total(parts, np, 1, 0, sh);
    if (threadIdx.x == 0)
    {
        sc->pq = pq;
        if (!(pq > 0.0) || !isfinite(pq))
            sc->done = 3; // breakdown: K not positive definite on p
        else
            sc->alpha = sc->rz / pq;
    }
}

template <int D, typename TX, typename TM>
__global__ __launch_bounds__(256) void k_cg_update(int n, TX *__restrict__ x, const TX *__restrict__ p,
                                                   TX *__restrict__ r, const TX *__restrict__ q,
                                                   TX *__restrict__ z, const TM *__restrict__ Minv, double *parts,
                                                   const CgS *sc)
{
    __shared__ double sh[4];
    if (sc->done)
        return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const TX alpha = (TX)sc->alpha;
    Acc<TX> rz, rr;
    if (i < n)
    {
        TX rv[D];
#pragma unroll
        for (int c = 0; c < D; c++)
        {
            const int64_t o = (int64_t)i * D + c;
            x[o] += alpha * p[o];
            rv[c] = r[o] - alpha * q[o];
            r[o] = rv[c];
        }
#pragma unroll
        for (int a = 0; a < D; a++)
        {
            TX zv = 0;
#pragma unroll
            for (int c = 0; c < D; c++)
                zv += (TX)Minv[(int64_t)(a * D + c) * n + i] * rv[c];
            z[(int64_t)i * D + a] = zv;
            rz.add(rv[a], zv);
            rr.add(rv[a], rv[a]);
        }
    }
    const double prz = rz.block(sh), prr = rr.block(sh);
    if (threadIdx.x == 0)
        parts[2 * blockIdx.x] = prz, parts[2 * blockIdx.x + 1] = prr;
}

// r.z and r.r from the update partials: stopping test, then beta (one work-group).
__global__ __launch_bounds__(256) void k_cg_beta(const double *parts, int np, CgS *sc)
{
    __shared__ double sh[4];
    if (sc->done)
        return;
    const double a = group_total(parts, np, 2, 0, sh);
    const double b = group_total(parts, np, 2, 1, sh);
    if (threadIdx.x == 0)
    {
        sc->it += 1;
        sc->rr = b;
        if (b <= sc->thr)
            sc->done = 1;
        else if (sc->it >= sc->maxit)
            sc->done = 2;
        else
        {
            sc->beta = a / sc->rz;
            sc->rz = a;
        }
    }
}

// r = b (converted), x = 0, z = M^-1 r; partials of rz and rr.  RESTART also sets
// p = q = 0 (a new solve); without it the search direction is kept (residual
// replacement of the mixed form).
template <int D, typename TX, typename TM, bool RESTART>
__global__ __launch_bounds__(256) void k_cg_init(int n, const double *__restrict__ b, TX *__restrict__ r,
                                                 TX *__restrict__ x, TX *__restrict__ p0, TX *__restrict__ q0,
                                                 TX *__restrict__ z, const TM *__restrict__ Minv, double *parts)
{
    __shared__ double sh[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    Acc<TX> rz, rr;
    if (i < n)
    {
        TX rv[D];
#pragma unroll
        for (int c = 0; c < D; c++)
        {
            const int64_t o = (int64_t)i * D + c;
            rv[c] = (TX)b[o];
            r[o] = rv[c];
            x[o] = 0;
            if (RESTART)
            {
                p0[o] = 0; // with beta = 0 the first SpMV gives p = z, q = K z
                q0[o] = 0;
            }
        }
#pragma unroll
        for (int a = 0; a < D; a++)
        {
            TX zv = 0;
#pragma unroll
            for (int c = 0; c < D; c++)
                zv += (TX)Minv[(int64_t)(a * D + c) * n + i] * rv[c];
            z[(int64_t)i * D + a] = zv;
            rz.add(rv[a], zv);
            rr.add(rv[a], rv[a]);
        }
    }
    const double prz = rz.block(sh), prr = rr.block(sh);
    if (threadIdx.x == 0)
        parts[2 * blockIdx.x] = prz, parts[2 * blockIdx.x + 1] = prr;
}

// With an external preconditioner (the AMG V-cycle): the update without z, ...
template <int D, typename TX>
__global__ __launch_bounds__(256) void k_cg_update_np(int n, TX *__restrict__ x, const TX *__restrict__ p,
                                                      TX *__restrict__ r, const TX *__restrict__ q, const CgS *sc)
{
    if (sc->done)
        return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const TX alpha = (TX)sc->alpha;
#pragma unroll
    for (int c = 0; c < D; c++)
    {
        const int64_t o = (int64_t)i * D + c;
        x[o] += alpha * p[o];
        r[o] -= alpha * q[o];
    }
}

// ... the start without z ...
template <int D, typename TX, bool RESTART>
__global__ __launch_bounds__(256) void k_cg_init_np(int n, const double *__restrict__ b, TX *__restrict__ r,
                                                    TX *__restrict__ x, TX *__restrict__ p0, TX *__restrict__ q0)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
#pragma unroll
    for (int c = 0; c < D; c++)
    {
        const int64_t o = (int64_t)i * D + c;
        r[o] = (TX)b[o];
        x[o] = 0;
        if (RESTART)
            p0[o] = 0, q0[o] = 0;
    }
}

// ... and the partials of r.z and r.r once z = B r is there (same layout as the
// block-Jacobi update's).
template <int D, typename TX>
__global__ __launch_bounds__(256) void k_cg_dots(int n, const TX *__restrict__ r, const TX *__restrict__ z,
                                                 double *parts, const CgS *sc)
{
    __shared__ double sh[4];
    if (sc && sc->done)
        return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    Acc<TX> rz, rr;
    if (i < n)
#pragma unroll
        for (int c = 0; c < D; c++)
        {
            const int64_t o = (int64_t)i * D + c;
            rz.add(r[o], z[o]);
            rr.add(r[o], r[o]);
        }
    const double prz = rz.block(sh), prr = rr.block(sh);
    if (threadIdx.x == 0)
        parts[2 * blockIdx.x] = prz, parts[2 * blockIdx.x + 1] = prr;
}

// Solver scalars of a new solve from the init partials (one work-group).
__global__ __launch_bounds__(256) void k_cg_start(const double *parts, int np, CgS *sc, double rtol2, double atol2,
                                                  int maxit)
{
    __shared__ double sh[4];
    const double a = group_total(parts, np, 2, 0, sh);
    const double bb = group_total(parts, np, 2, 1, sh);
    if (threadIdx.x == 0)
    {
        sc->rz = a;
        sc->rr = bb;
        sc->rr0 = bb;
        sc->thr = rtol2 * bb + atol2;
        sc->beta = 0.0;
        sc->alpha = 0.0;
        sc->pq = 0.0;
        sc->it = 0;
        sc->maxit = maxit;
        sc->done = (bb <= sc->thr) ? 1 : (maxit <= 0 ? 2 : 0);
    }
}

// Residual replacement: the CG continues with the replaced residual's r.z and the
// kept search direction, beta = (r.z)_new / (r.z)_old (one work-group).
__global__ __launch_bounds__(256) void k_cg_resume(const double *parts, int np, CgS *sc, double rtol2, double atol2)
{
    __shared__ double sh[4];
    const double a = group_total(parts, np, 2, 0, sh);
    const double bb = group_total(parts, np, 2, 1, sh);
    if (threadIdx.x == 0)
    {
        sc->beta = a / sc->rz;
        sc->rz = a;
        sc->rr = bb;
        sc->thr = rtol2 * bb + atol2;
        sc->done = (bb <= sc->thr) ? 1 : (sc->it >= sc->maxit ? 2 : 0);
    }
}

// r = b - K x in FP64 with |r|^2 (the refinement residual of the mixed form).
template <int D>
__global__ __launch_bounds__(256) void k_residual64(int n, const int *__restrict__ sptr, const int *__restrict__ col,
                                                    const double *__restrict__ K, const double *__restrict__ x,
                                                    const double *__restrict__ b, double *__restrict__ r,
                                                    double *parts)
{
    __shared__ double sh[4];
    const int blk = xcd_block();
    const int i = blk * 256 + threadIdx.x;
    double rr = 0.0;
    if (i < n)
    {
        const int s = i >> 6, l = i & 63;
        double acc[D];
        for (int c = 0; c < D; c++)
            acc[c] = 0;
        for (int k = sptr[s]; k < sptr[s + 1]; k++)
        {
            const int m = col[(int64_t)k * 64 + l];
            const double *kb = K + (int64_t)k * D * D * 64 + l;
            for (int rr_ = 0; rr_ < D; rr_++)
                for (int c = 0; c < D; c++)
                    acc[rr_] += kb[(rr_ * D + c) * 64] * x[(int64_t)m * D + c];
        }
        for (int c = 0; c < D; c++)
        {
            const double v = b[(int64_t)i * D + c] - acc[c];
            r[(int64_t)i * D + c] = v;
            rr += v * v;
        }
    }
    const double part = block_sum(rr, sh);
    if (threadIdx.x == 0)
        parts[blk] = part;
}

// *out = sum of the np partials in a fixed order (one work-group).
__global__ __launch_bounds__(256) void k_parts_total(const double *parts, int np, double *out)
{
    __shared__ double sh[4];
    const double t = group_total(parts, np, 1, 0, sh);
    if (threadIdx.x == 0)
        *out = t;
}

template <int D>
__global__ void k_add_f32(int n, double *x, const float *e)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < (int64_t)n * D)
        x[t] += (double)e[t];
}

// xyz += disp (solver.c:262-268), disp is dim-strided.
__global__ void k_apply_disp(int n, int dim, double *xyz, const double *disp)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    for (int c = 0; c < dim; c++)
        xyz[3 * i + c] += disp[(int64_t)i * dim + c];
}

// Plain y = K x (FP64 or FP32 values), for tests and the standalone bench.
template <int D, typename TV>
__global__ __launch_bounds__(256) void k_spmv_plain(int n, const int *__restrict__ sptr, const int *__restrict__ col,
                                                    const TV *__restrict__ K, const double *__restrict__ x,
                                                    double *__restrict__ y)
{
    const int i = xcd_block() * 256 + threadIdx.x;
    if (i >= n)
        return;
    const int s = i >> 6, l = i & 63;
    double acc[D];
    for (int c = 0; c < D; c++)
        acc[c] = 0;
    for (int k = sptr[s]; k < sptr[s + 1]; k++)
    {
        const int m = col[(int64_t)k * 64 + l];
        const TV *kb = K + (int64_t)k * D * D * 64 + l;
        for (int r = 0; r < D; r++)
            for (int c = 0; c < D; c++)
                acc[r] += (double)kb[(r * D + c) * 64] * x[(int64_t)m * D + c];
    }
    for (int c = 0; c < D; c++)
        y[(int64_t)i * D + c] = acc[c];
}

// ------------------------------------------------------------------ workspace ----
struct Work
{
    int n = 0, dim = 0;
    double *x64 = nullptr, *r64 = nullptr, *z64 = nullptr, *q64 = nullptr, *p64 = nullptr;
    float *x32 = nullptr, *r32 = nullptr, *z32 = nullptr, *q32 = nullptr, *p32 = nullptr;
    double *b = nullptr, *parts = nullptr, *scal = nullptr;
    CgS *sc = nullptr;
    CgS *hsc = nullptr; // pinned host copy
    hipEvent_t ev[2 * 64];
    bool have_ev = false;
} W;

// preconditioner of the solves: 0 block-Jacobi, 1 smoothed-aggregation AMG
static int g_precond = 0;
static int g_amg_epoch = -1;
static int g_amg_last_iters = 8; // iterations of the last AMG-preconditioned CG segment

extern "C" int lpmx_set_preconditioner(int kind)
{
    if (kind < 0 || kind > 2)
        return lpmx_fail(LPMX_ERR_ARG, "preconditioner %d: 0 (block-Jacobi), 1 (AMG) or 2 (one AMG V-cycle)", kind);
    g_precond = kind;
    return LPMX_OK;
}

static int amg_apply(const double *r, double *z, const int *done) { return lpmx_amg_apply_f64(r, z, done); }
static int amg_apply(const float *r, float *z, const int *done) { return lpmx_amg_apply_f32(r, z, done); }

static int g_prof = 0;
static double g_prof_ms = 0.0;
static int64_t g_prof_n = 0;

void lpmx_solver_free(void)
{
    void *ps[] = {W.x64, W.r64, W.z64, W.q64, W.p64, W.x32, W.r32, W.z32, W.q32, W.p32,
                  W.b, W.parts, W.scal, W.sc};
    for (void *p : ps)
        if (p)
            hipFree(p);
    if (W.hsc)
        hipHostFree(W.hsc);
    if (W.have_ev)
        for (auto &e : W.ev)
            hipEventDestroy(e);
    W = Work();
}

static int ensure_work(void)
{
    if (W.n == g.n && W.dim == g.dim && W.sc)
        return LPMX_OK;
    lpmx_solver_free();
    const int64_t m = (int64_t)g.n * g.dim;
    double **d64[] = {&W.x64, &W.r64, &W.z64, &W.q64, &W.p64, &W.b};
    for (double **p : d64)
        HIPCHK(hipMalloc(p, sizeof(double) * m));
    float **d32[] = {&W.x32, &W.r32, &W.z32, &W.q32, &W.p32};
    for (float **p : d32)
        HIPCHK(hipMalloc(p, sizeof(float) * m));
    const int nbk = nblocks(g.n, 256) + 8;
    HIPCHK(hipMalloc(&W.parts, sizeof(double) * 2 * nbk));
    HIPCHK(hipMalloc(&W.scal, sizeof(double) * 8));
    HIPCHK(hipMalloc(&W.sc, sizeof(CgS)));
    HIPCHK(hipHostMalloc(&W.hsc, sizeof(CgS)));
    for (auto &e : W.ev)
        HIPCHK(hipEventCreate(&e));
    W.have_ev = true;
    W.n = g.n, W.dim = g.dim;
    return LPMX_OK;
}

// Start a CG on right-hand side b: a new solve (RESTART) or, for the mixed form, a
// residual replacement that keeps p and q; the device scalars get the stopping
// threshold rtol2 |r|^2 + atol2 of the new residual.
template <int D, typename TX, typename TM, bool RESTART>
static int cg_begin(const TM *Minv, TX *x, TX *r, TX *z, TX *q, TX *p, const double *b, double rtol2, double atol2,
                    int maxit)
{
    const int n = g.n;
    const int G = nblocks(n, 256);
    if (g_precond == 1)
    {
        hipLaunchKernelGGL((k_cg_init_np<D, TX, RESTART>), dim3(G), dim3(256), 0, g_stream, n, b, r, x, p, q);
        LAUNCHCHK();
        RETCHK(amg_apply(r, z, nullptr));
        hipLaunchKernelGGL((k_cg_dots<D, TX>), dim3(G), dim3(256), 0, g_stream, n, r, z, W.parts, nullptr);
    }
    else
        hipLaunchKernelGGL((k_cg_init<D, TX, TM, RESTART>), dim3(G), dim3(256), 0, g_stream, n, b, r, x, p, q, z,
                           Minv, W.parts);
    if (RESTART)
        hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(256), 0, g_stream, W.parts, G, W.sc, rtol2, atol2, maxit);
    else
        hipLaunchKernelGGL(k_cg_resume, dim3(1), dim3(256), 0, g_stream, W.parts, G, W.sc, rtol2, atol2);
    LAUNCHCHK();
    return LPMX_OK;
}

// Iterate until the device flag is set; W.hsc gets the final scalars.  it0 is the
// iteration count at entry (for the profiling filter and the runaway guard).
template <int D, typename TV, typename TX, typename TM>
static int cg_iterate(const TV *Kv, const TM *Minv, TX *x, TX *r, TX *z, TX *q, TX *p, int maxit)
{
    const int n = g.n;
    const int G = nblocks(n, 256);
    HIPCHK(hipMemcpyAsync(W.hsc, W.sc, sizeof(CgS), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    const int it0 = W.hsc->it;
    if (W.hsc->done)
        return LPMX_OK;
    // Iterations queued per host check.  An AMG-preconditioned solve takes tens of
    // iterations of several milliseconds each: its first batch is about as long as the
    // last solve, then two at a time, so hardly any V-cycle is queued after the stop.
    int batch = g_precond == 1 ? std::max(1, g_amg_last_iters - 1) : n > 200000 ? 16 : 48;
    int t = 0, ev_used = 0;
    for (;;)
    {
        const int it_before = W.hsc->it;
        for (int k = 0; k < batch; k++, t++)
        {
            const bool timed = g_prof && ev_used < 64;
            if (timed)
                HIPCHK(hipEventRecord(W.ev[2 * ev_used], g_stream));
            hipLaunchKernelGGL((k_cg_spmv<D, TV, TX, spmv_unroll<TV>()>), dim3(G), dim3(256), 0, g_stream, n, g.sptr,
                               g.col, Kv, z, p, q, W.parts, W.sc);
            if (timed)
            {
                HIPCHK(hipEventRecord(W.ev[2 * ev_used + 1], g_stream));
                ev_used++;
            }
            hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, g_stream, W.parts, G, W.sc);
            if (g_precond == 1)
            {
                hipLaunchKernelGGL((k_cg_update_np<D, TX>), dim3(G), dim3(256), 0, g_stream, n, x, p, r, q, W.sc);
                lpmx_amg_prof_tag(k);
                RETCHK(amg_apply(r, z, &W.sc->done));
                hipLaunchKernelGGL((k_cg_dots<D, TX>), dim3(G), dim3(256), 0, g_stream, n, r, z, W.parts, W.sc);
            }
            else
                hipLaunchKernelGGL((k_cg_update<D, TX, TM>), dim3(G), dim3(256), 0, g_stream, n, x, p, r, q, z, Minv,
                                   W.parts, W.sc);
            hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(256), 0, g_stream, W.parts, G, W.sc);
        }
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(W.hsc, W.sc, sizeof(CgS), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        if (g_precond == 1)
        {
            RETCHK(lpmx_amg_prof_flush(W.hsc->it - it_before)); // V-cycles of iterations that ran
            batch = 2;
            if (getenv("LPMX_CG_TRACE"))
                fprintf(stderr, "[lpmx] AMG-PCG it %d rr %.6e rz %.6e pq %.6e alpha %.6e beta %.6e done %d\n",
                        W.hsc->it, W.hsc->rr, W.hsc->rz, W.hsc->pq, W.hsc->alpha, W.hsc->beta, W.hsc->done);
        }
        if (W.hsc->done)
            break;
        if (t > (int64_t)maxit + 2 * batch) // the device flag must have been set by now
            return lpmx_fail(LPMX_ERR_STATE, "CG loop ran past its cap without a stop flag");
    }
    if (g_precond == 1)
        g_amg_last_iters = std::max(1, W.hsc->it - it0);
    if (g_prof)
        for (int e = 0; e < ev_used; e++)
        {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, W.ev[2 * e], W.ev[2 * e + 1]));
            // a launch skipped because the solve had stopped takes ~no time; skip it
            if (e < W.hsc->it - it0)
                g_prof_ms += ms, g_prof_n++;
        }
    return LPMX_OK;
}

// A complete PCG solve on the prepared buffers.
template <int D, typename TV, typename TX, typename TM>
static int pcg_loop(const TV *Kv, const TM *Minv, TX *x, TX *r, TX *z, TX *q, TX *p, const double *b, double rtol2,
                    double atol2, int maxit, int *iters, double *rr_out, int *status)
{
    RETCHK((cg_begin<D, TX, TM, true>(Minv, x, r, z, q, p, b, rtol2, atol2, maxit)));
    RETCHK((cg_iterate<D, TV, TX, TM>(Kv, Minv, x, r, z, q, p, maxit)));
    *iters = W.hsc->it;
    *rr_out = W.hsc->rr;
    *status = W.hsc->done;
    return LPMX_OK;
}

template <int D>
static int solve_fp64(const double *b, double *x, double rtol2, double atol2, int maxit, int *iters, double *rr,
                      int *status)
{
    return pcg_loop<D, double, double, double>(g.K, g.Minv, x, W.r64, W.z64, W.q64, W.p64, b, rtol2, atol2, maxit,
                                               iters, rr, status);
}

// Mixed form: one FP32 block-Jacobi CG whose residual is replaced by the FP64 one
// every time the FP32 recurrence has dropped by the factor sqrt(seg_r2) (the
// "reliable update" form of iterative refinement): x32 is folded into the FP64
// solution, r = b - K x is formed with the FP64 matrix, and the CG continues from
// it with its search direction kept.  The outer test is FP64 throughout.
template <int D>
static int solve_mixed(const double *b, double *x, double rtol2, double atol2, double seg_r2, int max_replace,
                       int maxit, int *iters, double *rr, int *status, int *passes)
{
    const int n = g.n;
    const int64_t m = (int64_t)n * D;
    const int G = nblocks(n, 256);
    double bb = 0.0;
    RETCHK(lpmx_dev_sumsq(b, m, &bb));
    const double thr = rtol2 * bb + atol2;
    HIPCHK(hipMemsetAsync(x, 0, sizeof(double) * m, g_stream));
    HIPCHK(hipMemcpyAsync(W.b, b, sizeof(double) * m, hipMemcpyDeviceToDevice, g_stream));
    *iters = 0, *rr = bb, *status = 1, *passes = 0;
    if (bb <= thr)
        return LPMX_OK;
    // segment target: seg_r2 of the current residual, never below a quarter of the
    // outer threshold (the FP64 residual then lands under it with margin)
    RETCHK((cg_begin<D, float, float, true>(g.Minv32, W.x32, W.r32, W.z32, W.q32, W.p32, W.b, seg_r2, 0.25 * thr,
                                            maxit)));
    double cur = bb;
    int pass = 0, st = 2;
    for (;;)
    {
        RETCHK((cg_iterate<D, float, float, float>(g.K32, g.Minv32, W.x32, W.r32, W.z32, W.q32, W.p32, maxit)));
        st = W.hsc->done;
        pass++;
        hipLaunchKernelGGL((k_add_f32<D>), dim3(nblocks(m)), dim3(LPMX_BLOCK), 0, g_stream, n, x, W.x32);
        hipLaunchKernelGGL((k_residual64<D>), dim3(G), dim3(256), 0, g_stream, n, g.sptr, g.col, g.K, x, b,
                           W.b, W.parts);
        hipLaunchKernelGGL(k_parts_total, dim3(1), dim3(256), 0, g_stream, W.parts, G, W.scal);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(&cur, W.scal, sizeof(double), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        if (cur <= thr)
        {
            st = 1;
            break;
        }
        if (st == 3 || W.hsc->it >= maxit || pass > max_replace)
        {
            st = st == 3 ? 3 : 2;
            break;
        }
        RETCHK((cg_begin<D, float, float, false>(g.Minv32, W.x32, W.r32, W.z32, W.q32, W.p32, W.b, seg_r2, 0.25 * thr,
                                                 maxit)));
    }
    *iters = W.hsc->it;
    *rr = cur;
    *status = st;
    *passes = pass;
    return LPMX_OK;
}

static int g_last_passes = 0;
int lpmx_last_ir_passes(void) { return g_last_passes; }

__global__ void k_to_f32(int64_t m, const double *__restrict__ a, float *__restrict__ b)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m)
        b[i] = (float)a[i];
}

__global__ void k_to_f64(int64_t m, const float *__restrict__ a, double *__restrict__ b)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m)
        b[i] = (double)a[i];
}

// x = B b with one AMG V-cycle and no Krylov iteration (preconditioner 2: the Newton
// loop's fixed-point map becomes u + B R(u), accelerated by Anderson mixing, instead
// of u + K^-1 R(u) with an inner CG per iteration).  FP32 cycle in the mixed form.
static int vcycle_only(int precision, const double *b, double *x)
{
    const int64_t m = (int64_t)g.n * g.dim;
    if (precision == LPMX_MIXED)
    {
        hipLaunchKernelGGL(k_to_f32, dim3(nblocks(m)), dim3(LPMX_BLOCK), 0, g_stream, m, b, W.r32);
        LAUNCHCHK();
        RETCHK(lpmx_amg_apply_f32(W.r32, W.z32, nullptr));
        hipLaunchKernelGGL(k_to_f64, dim3(nblocks(m)), dim3(LPMX_BLOCK), 0, g_stream, m, W.z32, x);
        LAUNCHCHK();
        return LPMX_OK;
    }
    return lpmx_amg_apply_f64(b, x, nullptr);
}

// K x = b on the current (constraint-eliminated) matrix: block-Jacobi inverse, then
// PCG (FP64 or mixed).  b and x are DOF vectors; status: 1 converged, 2 cap,
// 3 breakdown.
int lpmx_solve_into(int precision, double rtol2, double atol2, int maxit, double inner_rtol2, int ir_maxit,
                    const double *b, double *x, int *iters, double *rr, int *status)
{
    RETCHK(lpmx_require_pattern());
    RETCHK(ensure_work());
    RETCHK(lpmx_bj_inverse());
    if (g_precond >= 1 && g_amg_epoch != lpmx_tangent_epoch())
    {
        // a new tangent: keep the hierarchy if the pattern and constraints are those
        // it was built for
        if (!lpmx_amg_current())
            RETCHK(lpmx_amg_build());
        g_amg_epoch = lpmx_tangent_epoch();
    }
    int it = 0, st = 0, passes = 1;
    double r2 = 0;
    if (g_precond == 2)
    {
        RETCHK(vcycle_only(precision, b, x));
        it = 1, st = 1;
    }
    else if (precision == LPMX_MIXED)
    {
        if (g.dim == 3)
            RETCHK(solve_mixed<3>(b, x, rtol2, atol2, inner_rtol2, ir_maxit, maxit, &it, &r2, &st, &passes));
        else
            RETCHK(solve_mixed<2>(b, x, rtol2, atol2, inner_rtol2, ir_maxit, maxit, &it, &r2, &st, &passes));
    }
    else
    {
        if (g.dim == 3)
            RETCHK(solve_fp64<3>(b, x, rtol2, atol2, maxit, &it, &r2, &st));
        else
            RETCHK(solve_fp64<2>(b, x, rtol2, atol2, maxit, &it, &r2, &st));
    }
    g_last_passes = passes;
    if (iters)
        *iters = it;
    if (rr)
        *rr = r2;
    if (status)
        *status = st;
    return LPMX_OK;
}

// Full solve of the Newton correction K disp = residual, then xyz += disp.
extern "C" int lpmx_solve_ex(int precision, double rtol2, double atol2, int maxit, double inner_rtol2,
                             int ir_maxit, int *iters, double *rr, int *status)
{
    int it = 0, st = 0;
    double r2 = 0;
    RETCHK(lpmx_solve_into(precision, rtol2, atol2, maxit, inner_rtol2, ir_maxit, g.res, g.disp, &it, &r2, &st));
    hipLaunchKernelGGL(k_apply_disp, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g.n, g.dim, g.x, g.disp);
    LAUNCHCHK();
    if (iters)
        *iters = it;
    if (rr)
        *rr = r2;
    if (status)
        *status = st;
    return LPMX_OK;
}

extern "C" int lpmx_solve(int precision, double rtol2, double atol2, int maxit, int *iters, double *rr)
{
    int st = 0;
    RETCHK(lpmx_solve_ex(precision, rtol2, atol2, maxit, 1e-2, 100, iters, rr, &st));
    HIPCHK(hipStreamSynchronize(g_stream));
    if (st == 3)
        return lpmx_fail(LPMX_ERR_NOCONV, "CG breakdown (p.Kp <= 0)");
    if (st != 1)
        return lpmx_fail(LPMX_ERR_NOCONV, "CG reached its iteration cap");
    return LPMX_OK;
}

// ------------------------------------------------------------------ measuring ----
extern "C" void lpmx_profile_spmv(int on)
{
    g_prof = on;
    g_prof_ms = 0.0;
    g_prof_n = 0;
    lpmx_amg_prof_enable(on != 0);
}

extern "C" void lpmx_profile_spmv_read(double *total_ms, int64_t *launches)
{
    *total_ms = g_prof_ms;
    *launches = g_prof_n;
}

extern "C" int lpmx_spmv(int precision, const double *x, double *y)
{
    RETCHK(lpmx_require_pattern());
    const int64_t m = (int64_t)g.n * g.dim;
    double *dx = nullptr, *dy = nullptr;
    HIPCHK(hipMalloc(&dx, sizeof(double) * m));
    HIPCHK(hipMalloc(&dy, sizeof(double) * m));
    HIPCHK(hipMemcpyAsync(dx, x, sizeof(double) * m, hipMemcpyHostToDevice, g_stream));
    const int G = nblocks(g.n, 256);
    if (g.dim == 3)
    {
        if (precision == LPMX_MIXED)
            hipLaunchKernelGGL((k_spmv_plain<3, float>), dim3(G), dim3(256), 0, g_stream, g.n, g.sptr, g.col, g.K32, dx, dy);
        else
            hipLaunchKernelGGL((k_spmv_plain<3, double>), dim3(G), dim3(256), 0, g_stream, g.n, g.sptr, g.col, g.K, dx, dy);
    }
    else
    {
        if (precision == LPMX_MIXED)
            hipLaunchKernelGGL((k_spmv_plain<2, float>), dim3(G), dim3(256), 0, g_stream, g.n, g.sptr, g.col, g.K32, dx, dy);
        else
            hipLaunchKernelGGL((k_spmv_plain<2, double>), dim3(G), dim3(256), 0, g_stream, g.n, g.sptr, g.col, g.K, dx, dy);
    }
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(y, dy, sizeof(double) * m, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(dx), hipFree(dy);
    return LPMX_OK;
}

// Algorithmic HBM bytes of one k_cg_spmv launch: every stored block and its column
// index once, the slice pointers, z read once (its gathers are re-reads from cache),
// p and q read and written once.
static double spmv_bytes(bool f32)
{
    const double sv = f32 ? 4.0 : 8.0;
    const int D = g.dim;
    return (double)g.nnzb * (D * D * sv + 4) + 4.0 * (g.nslices + 1) + 5.0 * g.n * D * sv;
}

extern "C" int lpmx_spmv_bytes(int precision, double *bytes)
{
    RETCHK(lpmx_require_pattern());
    *bytes = spmv_bytes(precision == LPMX_MIXED);
    return LPMX_OK;
}

// Deterministic vector in [0.5, 1.5) with no structure a stiffness matrix could
// annihilate (it is the timing input of the SpMV benchmark).
template <typename TX>
__global__ void k_fill_pattern(int64_t m, TX *v, unsigned salt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m)
    {
        unsigned h = (unsigned)i * 2654435761u ^ salt;
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        v[i] = (TX)(0.5 + (double)(h >> 8) * (1.0 / 16777216.0));
    }
}

// Times the CG SpMV kernel itself (the solve's dominant kernel) on the current
// matrix: reps launches on the library stream, bracketed by HIP events.  z and
// p_old hold fixed non-null vectors, so p.Kp stays positive and no launch takes
// the early exit; a set `done` flag after the run is reported as an error.
extern "C" int lpmx_bench_spmv(int precision, int reps, double *avg_ms, double *bytes)
{
    RETCHK(lpmx_require_pattern());
    RETCHK(ensure_work());
    const int n = g.n, D = g.dim, G = nblocks(n, 256);
    CgS h;
    memset(&h, 0, sizeof(h));
    h.rz = 1.0, h.beta = 0.5, h.maxit = 1 << 30;
    HIPCHK(hipMemcpyAsync(W.sc, &h, sizeof(CgS), hipMemcpyHostToDevice, g_stream));
    const int64_t m = (int64_t)n * D;
    const bool f32 = precision == LPMX_MIXED;
    if (reps <= 0)
        return lpmx_fail(LPMX_ERR_ARG, "reps must be positive");
    const int FG = (int)((m + 255) / 256);
    if (f32)
    {
        hipLaunchKernelGGL(k_fill_pattern<float>, dim3(FG), dim3(256), 0, g_stream, m, W.z32, 0x1234u);
        hipLaunchKernelGGL(k_fill_pattern<float>, dim3(FG), dim3(256), 0, g_stream, m, W.p32, 0x9e37u);
        hipLaunchKernelGGL(k_fill_pattern<float>, dim3(FG), dim3(256), 0, g_stream, m, W.q32, 0x51edu);
    }
    else
    {
        hipLaunchKernelGGL(k_fill_pattern<double>, dim3(FG), dim3(256), 0, g_stream, m, W.z64, 0x1234u);
        hipLaunchKernelGGL(k_fill_pattern<double>, dim3(FG), dim3(256), 0, g_stream, m, W.p64, 0x9e37u);
        hipLaunchKernelGGL(k_fill_pattern<double>, dim3(FG), dim3(256), 0, g_stream, m, W.q64, 0x51edu);
    }
    LAUNCHCHK();
    hipEvent_t e0 = W.ev[0], e1 = W.ev[1];
    for (int pass = 0; pass < 2; pass++) // warm-up pass, then the timed pass
    {
        if (pass == 1)
            HIPCHK(hipEventRecord(e0, g_stream));
        for (int r = 0; r < (pass ? reps : 3); r++)
        {
            if (D == 3 && f32)
                hipLaunchKernelGGL((k_cg_spmv<3, float, float, spmv_unroll<float>()>), dim3(G), dim3(256), 0, g_stream, n,
                                   g.sptr, g.col, g.K32, W.z32, W.p32, W.q32, W.parts, W.sc);
            else if (D == 3)
                hipLaunchKernelGGL((k_cg_spmv<3, double, double, spmv_unroll<double>()>), dim3(G), dim3(256), 0, g_stream,
                                   n, g.sptr, g.col, g.K, W.z64, W.p64, W.q64, W.parts, W.sc);
            else if (f32)
                hipLaunchKernelGGL((k_cg_spmv<2, float, float, spmv_unroll<float>()>), dim3(G), dim3(256), 0, g_stream, n,
                                   g.sptr, g.col, g.K32, W.z32, W.p32, W.q32, W.parts, W.sc);
            else
                hipLaunchKernelGGL((k_cg_spmv<2, double, double, spmv_unroll<double>()>), dim3(G), dim3(256), 0, g_stream,
                                   n, g.sptr, g.col, g.K, W.z64, W.p64, W.q64, W.parts, W.sc);
        }
        LAUNCHCHK();
    }
    HIPCHK(hipEventRecord(e1, g_stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    CgS after;
    HIPCHK(hipMemcpy(&after, W.sc, sizeof(CgS), hipMemcpyDeviceToHost));
    if (after.done)
        return lpmx_fail(LPMX_ERR_STATE, "SpMV benchmark: the solver's stop flag was set (%d)", after.done);
    *avg_ms = ms / reps;
    *bytes = spmv_bytes(f32);
    return LPMX_OK;
}
