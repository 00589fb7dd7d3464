// Anderson acceleration of the modified-Newton iteration of a load step.
//
// The reference's Newton loop (src/lpmc_project.c:424-447) is the fixed-point
// iteration u <- u + d(u) with d(u) = K^-1 R(u), K the elastic tangent assembled once
// per step.  In plastic flow the elastic tangent over-estimates the stiffness by
// E/H and that iteration contracts by nearly 1 per step.  Anderson's method (type
// II, Walker & Ni, SIAM J. Numer. Anal. 49 (2011) 1715-1735) keeps the last m
// differences dU_j = u_j+1 - u_j, dF_j = d_j+1 - d_j and moves to
//     u_k+1 = u_k + d_k - sum_j gamma_j (dU_j + dF_j),
//     gamma = argmin | d_k - sum_j gamma_j dF_j |_2,
// which has the same fixed points (R(u) = 0) as the plain iteration.  Constrained
// DOFs have d = 0 in every iterate, so every difference is 0 there and the
// prescribed displacements are untouched.
//
// The same mixing serves the fixed-point iteration of the adjoint solve on a DOF
// vector (lpmx_aa_*).
//
// Device: the history lives in DOF vectors (n*dim doubles); the 2m dot products of
// an iteration are one pass with per-block partials and a fixed-order total, so the
// result is deterministic.  Host: the m x m least-squares system, by Cholesky of the
// Tikhonov-regularised Gram matrix.
#include <math.h>
#include <string.h>

#include "common.h"

#define AA_MAX 32

// dC_j = dU_j + dF_j (the iterate and correction differences, summed: the mix uses
// only their sum) and dF_j (the Gram matrix and right-hand side use the corrections),
// stored in FP64, or in FP32 for the mixed-precision Newton loop (the iterate, the
// corrections and the dot products stay FP64).
template <typename T>
struct AaPtrsT
{
    T *dC[AA_MAX], *dF[AA_MAX];
};
typedef AaPtrsT<double> AaPtrs;
typedef AaPtrsT<float> AaPtrs32;

struct AaGamma
{
    double g[AA_MAX];
};

// One Anderson history over DOF vectors of n*dim entries.  The iterate itself may
// be strided: entry (i, c) of the iterate is at xs*i + c (xs = 3 for positions,
// xs = dim for a DOF vector).
struct AaHist
{
    int n = 0, dim = 0, depth = 0, cnt = 0, next = 0, have_prev = 0, f32 = 0;
    double *u_prev = nullptr, *f_prev = nullptr, *parts = nullptr, *out = nullptr;
    AaPtrs P{};
    AaPtrs32 P32{};
    double G[AA_MAX][AA_MAX];
};

static AaHist A; // the Newton loop's history

// work-groups of the dot pass: three per CU (256 CUs).  At the kernel's 177 VGPRs
// (two waves per SIMD) that is 1.5 resident rounds; two per CU (one round) took 0.082
// against 0.0995 ms at 1M, but its other grouping of the partial sums moved the
// iterates enough that one step of the 216k perturbed-parameter history at Newton
// 1e-10 stalled at its cap, so the grouping stays.
static const int AA_GRID = 768;

static void aa_release(AaHist *a)
{
    for (int j = 0; j < AA_MAX; j++)
    {
        void *ps[] = {a->P.dC[j], a->P.dF[j], a->P32.dC[j], a->P32.dF[j]};
        for (void *p : ps)
            if (p)
                hipFree(p);
    }
    void *ps[] = {a->u_prev, a->f_prev, a->parts, a->out};
    for (void *p : ps)
        if (p)
            hipFree(p);
    *a = AaHist();
}

void lpmx_accel_free(void) { aa_release(&A); }

// u = x - d (the iterate the update started from); with a previous iterate the new
// differences go to slot s; then (u, d) become the previous iterate.
template <typename T>
__global__ void k_aa_push(int n, int dim, int xs, const double *__restrict__ x, const double *__restrict__ d,
                          double *__restrict__ u_prev, double *__restrict__ f_prev, T *__restrict__ dC,
                          T *__restrict__ dF, int have_prev)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * dim)
        return;
    const int64_t i = t / dim;
    const int c = (int)(t - i * dim);
    const double u = x[xs * i + c] - d[t];
    if (have_prev)
    {
        const double du = u - u_prev[t], df = d[t] - f_prev[t];
        dC[t] = (T)(du + df);
        dF[t] = (T)df;
    }
    u_prev[t] = u;
    f_prev[t] = d[t];
}

// Partials of dF_s . dF_j and dF_j . d for the cnt active slots: parts[b][2*AA_MAX].
// Every slot is read: the host points the slots from cnt on at slot s (cached after
// its first read), so the AA_MAX loads of an element are issued together.  With the
// loads under `j < cnt` each one was waited for before the next was issued.
template <typename T>
__global__ __launch_bounds__(256) void k_aa_dots(int64_t m, AaPtrsT<T> P, int cnt, int s,
                                                 const double *__restrict__ d, double *parts)
{
    double g[AA_MAX], h[AA_MAX];
#pragma unroll
    for (int j = 0; j < AA_MAX; j++)
        g[j] = 0.0, h[j] = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m; t += (int64_t)gridDim.x * blockDim.x)
    {
        const double fs = (double)P.dF[s][t], dt = d[t];
        double f[AA_MAX];
#pragma unroll
        for (int j = 0; j < AA_MAX; j++)
            f[j] = (double)P.dF[j][t];
#pragma unroll
        for (int j = 0; j < AA_MAX; j++)
        {
            g[j] += fs * f[j];
            h[j] += f[j] * dt;
        }
    }
    // wave sums of all 2 cnt values first, then one exchange through LDS: the
    // work-group totals in wave order (fixed, so deterministic)
    __shared__ double wsum[4][2 * AA_MAX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < AA_MAX; j++)
        if (j < cnt)
        {
            const double a = wave_sum(g[j]), b = wave_sum(h[j]);
            if (lane == 0)
                wsum[w][j] = a, wsum[w][AA_MAX + j] = b;
        }
    __syncthreads();
    const int q = threadIdx.x;
    if (q < 2 * AA_MAX && (q % AA_MAX) < cnt)
    {
        double t = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++)
            t += wsum[k][q];
        parts[(int64_t)blockIdx.x * 2 * AA_MAX + q] = t;
    }
}

// One work-group per dot product (q = blockIdx.x): the partials in a fixed order.
__global__ __launch_bounds__(256) void k_aa_dots_total(const double *parts, int np, int cnt, double *out)
{
    __shared__ double sh[4];
    const int q = blockIdx.x;
    if ((q % AA_MAX) >= cnt)
        return;
    double s = 0.0;
    for (int k = threadIdx.x; k < np; k += blockDim.x)
        s += parts[(int64_t)k * 2 * AA_MAX + q];
    s = block_sum(s, sh);
    if (threadIdx.x == 0)
        out[q] = s;
}

// x -= sum_j gamma_j dC_j.  Every slot is read, as in k_aa_dots: the host points the
// slots from cnt on at slot s, with gamma_j = +0 there, so their terms are +-0 and
// leave the sum's bits as they are, and the AA_MAX loads of an element are issued
// together.
template <typename T>
__global__ void k_aa_mix(int n, int dim, int xs, double *__restrict__ x, AaPtrsT<T> P, AaGamma gam)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * dim)
        return;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < AA_MAX; j++)
        s += gam.g[j] * (double)P.dC[j][t];
    const int64_t i = t / dim;
    const int c = (int)(t - i * dim);
    x[xs * i + c] -= s;
}

// x = u_prev + f_prev: the plain iterate of the last update.
__global__ void k_aa_plain(int n, int dim, int xs, double *__restrict__ x, const double *__restrict__ u_prev,
                           const double *__restrict__ f_prev)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * dim)
        return;
    const int64_t i = t / dim;
    const int c = (int)(t - i * dim);
    x[xs * i + c] = u_prev[t] + f_prev[t];
}

// Start a new history (a new load step, or a re-solve after bond breakage: the
// tangent, and with it the map, has changed).
static int aa_begin(AaHist *a, int depth, int f32 = 0)
{
    if (depth < 0 || depth > AA_MAX)
        return lpmx_fail(LPMX_ERR_ARG, "acceleration depth %d outside 0..%d", depth, AA_MAX);
    if (a->n != g.n || a->dim != g.dim)
        aa_release(a);
    const int64_t m = (int64_t)g.n * g.dim;
    for (int j = 0; j < depth; j++)
    {
        if (f32)
        {
            if (!a->P32.dC[j])
                HIPCHK(hipMalloc(&a->P32.dC[j], sizeof(float) * m));
            if (!a->P32.dF[j])
                HIPCHK(hipMalloc(&a->P32.dF[j], sizeof(float) * m));
        }
        else
        {
            if (!a->P.dC[j])
                HIPCHK(hipMalloc(&a->P.dC[j], sizeof(double) * m));
            if (!a->P.dF[j])
                HIPCHK(hipMalloc(&a->P.dF[j], sizeof(double) * m));
        }
    }
    a->f32 = f32;
    if (depth > 0 && !a->u_prev)
    {
        HIPCHK(hipMalloc(&a->u_prev, sizeof(double) * m));
        HIPCHK(hipMalloc(&a->f_prev, sizeof(double) * m));
        HIPCHK(hipMalloc(&a->parts, sizeof(double) * AA_GRID * 2 * AA_MAX));
        HIPCHK(hipMalloc(&a->out, sizeof(double) * 2 * AA_MAX));
    }
    a->n = g.n, a->dim = g.dim;
    a->depth = depth;
    a->cnt = a->next = a->have_prev = 0;
    return LPMX_OK;
}

static void aa_forget(AaHist *a) { a->cnt = a->next = 0; }

static int aa_revert(AaHist *a, double *x, int xs)
{
    if (a->depth == 0 || !a->have_prev)
        return LPMX_OK;
    hipLaunchKernelGGL(k_aa_plain, dim3(nblocks((int64_t)g.n * g.dim)), dim3(LPMX_BLOCK), 0, g_stream, g.n, g.dim, xs,
                       x, a->u_prev, a->f_prev);
    LAUNCHCHK();
    aa_forget(a);
    return LPMX_OK;
}

// Cholesky solve of (G + lam I) x = h for the cnt x cnt system; false if not PD.
static bool small_spd_solve(int cnt, const double G[AA_MAX][AA_MAX], const double *h, double *x)
{
    double L[AA_MAX][AA_MAX];
    double dmax = 0.0;
    for (int j = 0; j < cnt; j++)
        dmax = fmax(dmax, G[j][j]);
    if (!(dmax > 0.0) || !isfinite(dmax))
        return false;
    const double lam = 1e-12 * dmax;
    for (int i = 0; i < cnt; i++)
        for (int j = 0; j <= i; j++)
        {
            double s = G[i][j] + (i == j ? lam : 0.0);
            for (int k = 0; k < j; k++)
                s -= L[i][k] * L[j][k];
            if (i == j)
            {
                if (!(s > 0.0))
                    return false;
                L[i][i] = sqrt(s);
            }
            else
                L[i][j] = s / L[j][j];
        }
    double y[AA_MAX];
    for (int i = 0; i < cnt; i++)
    {
        double s = h[i];
        for (int k = 0; k < i; k++)
            s -= L[i][k] * y[k];
        y[i] = s / L[i][i];
    }
    for (int i = cnt - 1; i >= 0; i--)
    {
        double s = y[i];
        for (int k = i + 1; k < cnt; k++)
            s -= L[k][i] * x[k];
        x[i] = s / L[i][i];
        if (!isfinite(x[i]))
            return false;
    }
    return true;
}

// After an update has set x = u_k + d_k: record the iterate and, with a history,
// move x to the accelerated iterate.  *mixed = 1 when it did.
static int aa_step(AaHist *a, double *x, int xs, const double *d, int *mixed)
{
    *mixed = 0;
    if (a->depth == 0)
        return LPMX_OK;
    const int n = g.n, dim = g.dim;
    const int64_t m = (int64_t)n * dim;
    const int nb = nblocks(m);
    const int s = a->next;
    if (a->f32)
        hipLaunchKernelGGL(k_aa_push<float>, dim3(nb), dim3(LPMX_BLOCK), 0, g_stream, n, dim, xs, x, d, a->u_prev,
                           a->f_prev, a->P32.dC[s], a->P32.dF[s], a->have_prev);
    else
        hipLaunchKernelGGL(k_aa_push<double>, dim3(nb), dim3(LPMX_BLOCK), 0, g_stream, n, dim, xs, x, d, a->u_prev,
                           a->f_prev, a->P.dC[s], a->P.dF[s], a->have_prev);
    LAUNCHCHK();
    if (!a->have_prev)
    {
        a->have_prev = 1;
        return LPMX_OK;
    }
    if (a->cnt < a->depth)
        a->cnt++;
    a->next = (a->next + 1) % a->depth;
    const int cnt = a->cnt;
    // the unused slots read slot s (written just now, so finite and cached)
    AaPtrs32 Q32 = a->P32;
    AaPtrs Q64 = a->P;
    for (int j = cnt; j < AA_MAX; j++)
    {
        Q32.dF[j] = a->P32.dF[s], Q32.dC[j] = a->P32.dC[s];
        Q64.dF[j] = a->P.dF[s], Q64.dC[j] = a->P.dC[s];
    }
    if (a->f32)
        hipLaunchKernelGGL(k_aa_dots<float>, dim3(AA_GRID), dim3(256), 0, g_stream, m, Q32, cnt, s, d, a->parts);
    else
        hipLaunchKernelGGL(k_aa_dots<double>, dim3(AA_GRID), dim3(256), 0, g_stream, m, Q64, cnt, s, d, a->parts);
    hipLaunchKernelGGL(k_aa_dots_total, dim3(2 * AA_MAX), dim3(256), 0, g_stream, a->parts, AA_GRID, cnt, a->out);
    LAUNCHCHK();
    double o[2 * AA_MAX];
    HIPCHK(hipMemcpyAsync(o, a->out, sizeof(o), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    for (int j = 0; j < cnt; j++)
        a->G[s][j] = a->G[j][s] = o[j];
    AaGamma gam;
    for (int j = 0; j < AA_MAX; j++)
        gam.g[j] = 0.0;
    if (!small_spd_solve(cnt, a->G, o + AA_MAX, gam.g))
    {
        // degenerate history: keep the plain iterate, start the differences afresh
        aa_forget(a);
        return LPMX_OK;
    }
    if (a->f32)
        hipLaunchKernelGGL(k_aa_mix<float>, dim3(nb), dim3(LPMX_BLOCK), 0, g_stream, n, dim, xs, x, Q32, gam);
    else
        hipLaunchKernelGGL(k_aa_mix<double>, dim3(nb), dim3(LPMX_BLOCK), 0, g_stream, n, dim, xs, x, Q64, gam);
    LAUNCHCHK();
    *mixed = 1;
    return LPMX_OK;
}

// --------------------------------------------------------- the Newton loop's -----
// precision LPMX_MIXED keeps the differences in FP32
int lpmx_accel_begin(int depth, int precision) { return aa_begin(&A, depth, precision == LPMX_MIXED); }
// Keep the history's differences and drop only its last point: the next step records
// a new point without a difference across it (a re-solve of the same load step after
// bond breakage, whose map differs from the last one by the broken bonds).
void lpmx_accel_rebase(void) { A.have_prev = 0; }
void lpmx_accel_forget(void) { aa_forget(&A); }
// Revert xyz to the plain iterate of the last solve and forget the history.
int lpmx_accel_revert(void) { return aa_revert(&A, g.x, 3); }
// After a solve has set xyz = u_k + d_k (d_k in g.disp).
int lpmx_accel_step(int *mixed) { return aa_step(&A, g.x, 3, g.disp, mixed); }

// ------------------------------------------------------------ other histories ----
AaHist *lpmx_aa_create(void) { return new AaHist(); }
void lpmx_aa_destroy(AaHist *a)
{
    if (a)
    {
        aa_release(a);
        delete a;
    }
}
int lpmx_aa_begin(AaHist *a, int depth) { return aa_begin(a, depth); }
int lpmx_aa_step(AaHist *a, double *x, const double *d, int *mixed) { return aa_step(a, x, g.dim, d, mixed); }
