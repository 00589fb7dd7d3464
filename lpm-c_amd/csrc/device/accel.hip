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
// Device: the history lives in DOF vectors (n*dim doubles); the 2m dot products of
// an iteration are one pass with per-block partials and a fixed-order total, so the
// result is deterministic.  Host: the m x m least-squares system, by Cholesky of the
// Tikhonov-regularised Gram matrix.
#include <math.h>
#include <string.h>

#include "common.h"

#define AA_MAX 8

struct AaPtrs
{
    double *dU[AA_MAX], *dF[AA_MAX];
};

static struct
{
    int n = 0, dim = 0, depth = 0, cnt = 0, next = 0, have_prev = 0;
    double *u_prev = nullptr, *f_prev = nullptr, *parts = nullptr, *out = nullptr;
    AaPtrs P{};
    double G[AA_MAX][AA_MAX];
} A;

static const int AA_GRID = 1024;

void lpmx_accel_free(void)
{
    for (int j = 0; j < AA_MAX; j++)
    {
        if (A.P.dU[j])
            hipFree(A.P.dU[j]);
        if (A.P.dF[j])
            hipFree(A.P.dF[j]);
    }
    void *ps[] = {A.u_prev, A.f_prev, A.parts, A.out};
    for (void *p : ps)
        if (p)
            hipFree(p);
    memset(&A.P, 0, sizeof(A.P));
    A.u_prev = A.f_prev = A.parts = A.out = nullptr;
    A.n = A.dim = A.depth = A.cnt = A.next = A.have_prev = 0;
}

// u = xyz - d (the iterate the solve started from); with a previous iterate the new
// differences go to slot s; then (u, d) become the previous iterate.
__global__ void k_aa_push(int n, int dim, const double *__restrict__ xyz, const double *__restrict__ d,
                          double *__restrict__ u_prev, double *__restrict__ f_prev, double *__restrict__ dU,
                          double *__restrict__ dF, int have_prev)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * dim)
        return;
    const int64_t i = t / dim;
    const int c = (int)(t - i * dim);
    const double u = xyz[3 * i + c] - d[t];
    if (have_prev)
    {
        dU[t] = u - u_prev[t];
        dF[t] = d[t] - f_prev[t];
    }
    u_prev[t] = u;
    f_prev[t] = d[t];
}

// Partials of dF_s . dF_j and dF_j . d for the cnt active slots: parts[b][2*AA_MAX].
__global__ __launch_bounds__(256) void k_aa_dots(int64_t m, AaPtrs P, int cnt, int s, const double *__restrict__ d,
                                                 double *parts)
{
    __shared__ double sh[4];
    double g[AA_MAX], h[AA_MAX];
#pragma unroll
    for (int j = 0; j < AA_MAX; j++)
        g[j] = 0.0, h[j] = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m; t += (int64_t)gridDim.x * blockDim.x)
    {
        const double fs = P.dF[s][t], dt = d[t];
#pragma unroll
        for (int j = 0; j < AA_MAX; j++)
            if (j < cnt)
            {
                const double fj = P.dF[j][t];
                g[j] += fs * fj;
                h[j] += fj * dt;
            }
    }
    for (int j = 0; j < cnt; j++)
    {
        const double a = block_sum(g[j], sh);
        const double b = block_sum(h[j], sh);
        if (threadIdx.x == 0)
            parts[(int64_t)blockIdx.x * 2 * AA_MAX + j] = a, parts[(int64_t)blockIdx.x * 2 * AA_MAX + AA_MAX + j] = b;
    }
}

__global__ __launch_bounds__(256) void k_aa_dots_total(const double *parts, int np, int cnt, double *out)
{
    __shared__ double sh[4];
    for (int q = 0; q < 2 * AA_MAX; q++)
    {
        if ((q % AA_MAX) >= cnt)
            continue;
        double s = 0.0;
        for (int k = threadIdx.x; k < np; k += blockDim.x)
            s += parts[(int64_t)k * 2 * AA_MAX + q];
        s = block_sum(s, sh);
        if (threadIdx.x == 0)
            out[q] = s;
    }
}

// xyz -= sum_j gamma_j (dU_j + dF_j)
__global__ void k_aa_mix(int n, int dim, double *__restrict__ xyz, AaPtrs P, int cnt, double g0, double g1, double g2,
                         double g3, double g4, double g5, double g6, double g7)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * dim)
        return;
    const double gam[AA_MAX] = {g0, g1, g2, g3, g4, g5, g6, g7};
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < AA_MAX; j++)
        if (j < cnt)
            s += gam[j] * (P.dU[j][t] + P.dF[j][t]);
    const int64_t i = t / dim;
    const int c = (int)(t - i * dim);
    xyz[3 * i + c] -= s;
}

// xyz = u_prev + f_prev: the plain modified-Newton iterate of the last solve.
__global__ void k_aa_plain(int n, int dim, double *__restrict__ xyz, const double *__restrict__ u_prev,
                           const double *__restrict__ f_prev)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * dim)
        return;
    const int64_t i = t / dim;
    const int c = (int)(t - i * dim);
    xyz[3 * i + c] = u_prev[t] + f_prev[t];
}

// Start a new history (a new load step, or a re-solve after bond breakage: the
// tangent, and with it the map, has changed).
int lpmx_accel_begin(int depth)
{
    if (depth < 0 || depth > AA_MAX)
        return lpmx_fail(LPMX_ERR_ARG, "acceleration depth %d outside 0..%d", depth, AA_MAX);
    if (A.n != g.n || A.dim != g.dim)
        lpmx_accel_free();
    const int64_t m = (int64_t)g.n * g.dim;
    for (int j = 0; j < depth; j++)
    {
        if (!A.P.dU[j])
            HIPCHK(hipMalloc(&A.P.dU[j], sizeof(double) * m));
        if (!A.P.dF[j])
            HIPCHK(hipMalloc(&A.P.dF[j], sizeof(double) * m));
    }
    if (depth > 0 && !A.u_prev)
    {
        HIPCHK(hipMalloc(&A.u_prev, sizeof(double) * m));
        HIPCHK(hipMalloc(&A.f_prev, sizeof(double) * m));
        HIPCHK(hipMalloc(&A.parts, sizeof(double) * AA_GRID * 2 * AA_MAX));
        HIPCHK(hipMalloc(&A.out, sizeof(double) * 2 * AA_MAX));
    }
    A.n = g.n, A.dim = g.dim;
    A.depth = depth;
    A.cnt = A.next = A.have_prev = 0;
    return LPMX_OK;
}

// Forget the differences (keeps the last iterate as the reference of the next one).
void lpmx_accel_forget(void) { A.cnt = A.next = 0; }

// Revert xyz to the plain iterate of the last solve and forget the history.
int lpmx_accel_revert(void)
{
    if (A.depth == 0 || !A.have_prev)
        return LPMX_OK;
    hipLaunchKernelGGL(k_aa_plain, dim3(nblocks((int64_t)g.n * g.dim)), dim3(LPMX_BLOCK), 0, g_stream, g.n, g.dim, g.x,
                       A.u_prev, A.f_prev);
    LAUNCHCHK();
    lpmx_accel_forget();
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

// After a solve has set xyz = u_k + d_k (d_k in g.disp): record the iterate and, with
// a history, move xyz to the accelerated iterate.  *mixed = 1 when it did.
int lpmx_accel_step(int *mixed)
{
    *mixed = 0;
    if (A.depth == 0)
        return LPMX_OK;
    const int n = g.n, dim = g.dim;
    const int64_t m = (int64_t)n * dim;
    const int nb = nblocks(m);
    const int s = A.next;
    hipLaunchKernelGGL(k_aa_push, dim3(nb), dim3(LPMX_BLOCK), 0, g_stream, n, dim, g.x, g.disp, A.u_prev, A.f_prev,
                       A.P.dU[s], A.P.dF[s], A.have_prev);
    LAUNCHCHK();
    if (!A.have_prev)
    {
        A.have_prev = 1;
        return LPMX_OK;
    }
    if (A.cnt < A.depth)
        A.cnt++;
    A.next = (A.next + 1) % A.depth;
    const int cnt = A.cnt;
    hipLaunchKernelGGL(k_aa_dots, dim3(AA_GRID), dim3(256), 0, g_stream, m, A.P, cnt, s, g.disp, A.parts);
    hipLaunchKernelGGL(k_aa_dots_total, dim3(1), dim3(256), 0, g_stream, A.parts, AA_GRID, cnt, A.out);
    LAUNCHCHK();
    double o[2 * AA_MAX];
    HIPCHK(hipMemcpyAsync(o, A.out, sizeof(o), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    for (int j = 0; j < cnt; j++)
        A.G[s][j] = A.G[j][s] = o[j];
    double gam[AA_MAX] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!small_spd_solve(cnt, A.G, o + AA_MAX, gam))
    {
        // degenerate history: keep the plain iterate, start the differences afresh
        lpmx_accel_forget();
        return LPMX_OK;
    }
    hipLaunchKernelGGL(k_aa_mix, dim3(nb), dim3(LPMX_BLOCK), 0, g_stream, n, dim, g.x, A.P, cnt, gam[0], gam[1],
                       gam[2], gam[3], gam[4], gam[5], gam[6], gam[7]);
    LAUNCHCHK();
    *mixed = 1;
    return LPMX_OK;
}
