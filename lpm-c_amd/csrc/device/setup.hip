// One-time setup and output post-processing on the device: the lattice points of
// createCuboid (reference src/initialization.c:6-560) and the least-squares strain
// of computeStrain (reference src/lpm_basic.c:127-249).  Compiled without
// contraction: the positions come out bit for bit as the reference's arithmetic
// gives them.
#include <math.h>

#include <vector>

#include "bond_common.h"

// ------------------------------------------------------------- lattice points ----
// createCuboid generates a candidate grid over [-a, a] (a the box diagonal) in rows
// (layer k, row j) of points along x, rotates every point and keeps those inside
// the closed box, in generation order.  Row r of the grid (r = (k-1) rows + (j-1))
// is one thread: the first pass counts its kept points, the second writes them at
// the row's offset, so the order is the reference's.
struct Grid
{
    int lattice, dim, pfr, rows, layers;
    double hx, hy, hz, radius, bt[6], box[6], R[9];
};

__device__ __forceinline__ int row_length(const Grid &G, int k, int j)
{
    switch (G.lattice)
    {
    case 0:
    case 2:
        return G.pfr;
    case 1:
        return j % 2 == 1 ? G.pfr : G.pfr - 1;
    case 3:
        return ((k % 2 == 1) == (j % 2 == 1)) ? G.pfr : G.pfr - 1;
    default:
        return k % 2 == 1 ? G.pfr : (j <= G.rows - 1 ? G.pfr - 1 : 0);
    }
}

// Candidate i (1-based) of row (k, j), before rotation (initialization.c per lattice).
__device__ __forceinline__ void candidate(const Grid &G, int k, int j, int i, double *t)
{
    const double *bt = G.bt;
    switch (G.lattice)
    {
    case 0:
        t[0] = bt[0] + G.hx * (i - 1) + G.radius, t[1] = bt[2] + G.hy * (j - 1) + G.radius, t[2] = 0.0;
        break;
    case 1:
        t[1] = bt[2] + G.hy * (j - 1), t[2] = 0.0;
        t[0] = j % 2 == 1 ? bt[0] + G.hx * (i - 1) : bt[0] + G.hx * (2 * i - 1) / 2.0;
        break;
    case 2:
        t[0] = G.box[0] + G.hx * (i - 1), t[1] = bt[2] + G.hy * (j - 1), t[2] = bt[4] + G.hz * (k - 1);
        break;
    case 3:
        t[2] = bt[4] + G.hz / 2.0 * (k - 1), t[1] = bt[2] + G.hy / 2.0 * (j - 1);
        t[0] = ((k % 2 == 1) == (j % 2 == 1)) ? bt[0] + G.hx * (i - 1) : bt[0] + G.hx * (i - 1) + G.hx / 2.0;
        break;
    default:
        t[2] = bt[4] + G.hz / 2.0 * (k - 1);
        if (k % 2 == 1)
            t[0] = bt[0] + G.hx * (i - 1), t[1] = bt[2] + G.hy * (j - 1);
        else
            t[0] = bt[0] + G.hx * (i - 1) + G.hx / 2.0, t[1] = bt[2] + G.hy * (j - 1) + G.hy / 2.0;
        break;
    }
}

__device__ __forceinline__ bool keep(const Grid &G, const double *t, double *p)
{
    for (int r = 0; r < 3; r++)
        p[r] = G.R[3 * r] * t[0] + G.R[3 * r + 1] * t[1] + G.R[3 * r + 2] * t[2];
    if (!(p[0] >= G.box[0] && p[0] <= G.box[1] && p[1] >= G.box[2] && p[1] <= G.box[3]))
        return false;
    if (G.dim == 3 && !(p[2] >= G.box[4] && p[2] <= G.box[5]))
        return false;
    return true;
}

__global__ void k_lattice_rows(Grid G, int nrows, int *cnt, const int *off, double *xyz, int cap)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows)
        return;
    const int k = r / G.rows + 1, j = r % G.rows + 1, len = row_length(G, k, j);
    int c = 0;
    int64_t o = off ? off[r] : 0;
    for (int i = 1; i <= len; i++)
    {
        double t[3], p[3];
        candidate(G, k, j, i, t);
        if (!keep(G, t, p))
            continue;
        if (off && o + c < cap)
        {
            xyz[3 * (o + c)] = p[0];
            xyz[3 * (o + c) + 1] = p[1];
            xyz[3 * (o + c) + 2] = G.dim == 3 ? p[2] : 0.0;
        }
        c++;
    }
    if (!off)
        cnt[r] = c;
}

extern "C" int lpmx_lattice_points(const lpmx_lattice_spec *s, double *xyz_host, int cap, int *count)
{
    RETCHK(lpmx_device_init());
    if (!s || !count || s->lattice < 0 || s->lattice > 4 || s->pfr < 1 || s->rows < 1 ||
        (s->lattice >= 2 && s->layers < 1))
        return lpmx_fail(LPMX_ERR_ARG, "lattice description");
    Grid G;
    G.lattice = s->lattice, G.dim = s->dim, G.pfr = s->pfr, G.rows = s->rows;
    G.layers = s->lattice >= 2 ? s->layers : 1;
    G.hx = s->hx, G.hy = s->hy, G.hz = s->hz, G.radius = s->radius;
    for (int c = 0; c < 6; c++)
        G.bt[c] = s->bt[c], G.box[c] = s->box[c];
    for (int c = 0; c < 9; c++)
        G.R[c] = s->R[c];
    const int64_t nrows64 = (int64_t)G.layers * G.rows;
    if (nrows64 >= 0x7fffffff)
        return lpmx_fail(LPMX_ERR_CAPACITY, "candidate grid of %lld rows", (long long)nrows64);
    const int nrows = (int)nrows64;
    int *cnt = nullptr, *off = nullptr;
    double *xyz = nullptr;
    HIPCHK(hipMalloc(&cnt, sizeof(int) * (nrows + 1)));
    HIPCHK(hipMalloc(&off, sizeof(int) * (nrows + 1)));
    hipLaunchKernelGGL(k_lattice_rows, dim3(nblocks(nrows)), dim3(LPMX_BLOCK), 0, g_stream, G, nrows, cnt, nullptr,
                       nullptr, 0);
    LAUNCHCHK();
    int total = 0;
    HIPCHK(hipMemsetAsync(cnt + nrows, 0, sizeof(int), g_stream));
    RETCHK(lpmx_dev_scan_int(cnt, off, nrows + 1, &total));
    *count = total;
    int rc = LPMX_OK;
    if (xyz_host && cap > 0)
    {
        if (cap < total)
            rc = lpmx_fail(LPMX_ERR_ARG, "position buffer of %d points for %d", cap, total);
        else
        {
            HIPCHK(hipMalloc(&xyz, sizeof(double) * 3 * ((int64_t)total + 1)));
            hipLaunchKernelGGL(k_lattice_rows, dim3(nblocks(nrows)), dim3(LPMX_BLOCK), 0, g_stream, G, nrows, cnt,
                               off, xyz, total);
            LAUNCHCHK();
            HIPCHK(hipMemcpyAsync(xyz_host, xyz, sizeof(double) * 3 * (int64_t)total, hipMemcpyDeviceToHost,
                                  g_stream));
        }
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(cnt), hipFree(off);
    if (xyz)
        hipFree(xyz);
    return rc;
}

// --------------------------------------------------------------- strain -----------
// computeStrain (lpm_basic.c:127-249): per particle, the strain that best fits the
// bond stretches dL/L0 in weighted least squares (weight 0.1 on the first shell,
// 0.9 on the second), from the normal equations (6 x 6 in 3D, 3 x 3 in 2D) solved by
// elimination with partial pivoting, in the role LAPACKE_dgesv plays there.  When
// the system is exactly singular dgesv leaves the right-hand side unsolved, and the
// reference then stores that right-hand side as the strain (its copy of the previous
// particle's strain is overwritten by it), except at particle 0, which keeps what it
// had.  Each particle is independent: one thread per particle.
__device__ int gauss_solve_dev(int n, double *A, double *b)
{
    for (int k = 0; k < n; k++)
    {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i * n + k]) > fabs(A[p * n + k]))
                p = i;
        if (A[p * n + k] == 0.0)
            return k + 1;
        if (p != k)
        {
            for (int j = 0; j < n; j++)
            {
                const double t = A[k * n + j];
                A[k * n + j] = A[p * n + j];
                A[p * n + j] = t;
            }
            const double t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
        for (int i = k + 1; i < n; i++)
        {
            const double f = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; j++)
                A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; k--)
    {
        double s = b[k];
        for (int j = k + 1; j < n; j++)
            s -= A[k * n + j] * b[j];
        b[k] = s / A[k * n + k];
    }
    return 0;
}

// strain: [n][6] in the reference's component order xx yy zz yz xz xy
template <int D>
__global__ void k_strain(DV v, double *strain)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    constexpr int NS = 3 * (D - 1);
    double r4[D][D][D][D], r2[D][D];
    for (int a = 0; a < D; a++)
        for (int c = 0; c < D; c++)
        {
            r2[a][c] = 0.0;
            for (int m = 0; m < D; m++)
                for (int l = 0; l < D; l++)
                    r4[a][c][m][l] = 0.0;
        }
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int s = v.nsign[b];
        if (s != 0 && s != 1)
            continue;
        const double w = s == 0 ? 0.1 : 0.9, L = v.L0[b], dl = v.dL[b];
        const int q = v.nbr[b];
        double u[D]; // x_q - x_i, the bond in the initial configuration
        for (int c = 0; c < D; c++)
            u[c] = v.x0[3 * q + c] - v.x0[3 * i + c];
        // the reference's left-to-right products and quotients (lpm_basic.c:148-175)
        for (int a = 0; a < D; a++)
            for (int c = 0; c < D; c++)
                for (int m = 0; m < D; m++)
                    for (int l = 0; l < D; l++)
                        r4[a][c][m][l] += w * u[a] / L * u[c] / L * u[m] / L * u[l] / L;
        for (int a = 0; a < D; a++)
            for (int c = 0; c < D; c++)
                r2[a][c] += w * dl / L * u[a] / L * u[c] / L;
    }
    double A[NS * NS], bb[NS], rhs[NS];
    int ii = 0;
    for (int j = 0; j < D; j++)
        for (int k = 0; k < D; k++)
            for (int m = 0; m < D; m++)
                for (int l = 0; l < D; l++)
                    if (m <= l && j <= k)
                        A[ii++] = r4[m][l][j][k];
    ii = 0;
    for (int j = 0; j < D; j++)
        for (int k = 0; k < D; k++)
            if (j <= k)
                rhs[ii] = bb[ii] = r2[j][k], ii++;
    if (gauss_solve_dev(NS, A, bb) > 0)
    {
        if (i == 0)
            return;
        for (int c = 0; c < NS; c++)
            bb[c] = rhs[c];
    }
    double *s = strain + 6 * (int64_t)i;
    if (D == 2)
        s[0] = bb[0], s[5] = bb[1], s[1] = bb[2];
    else
        s[0] = bb[0], s[5] = bb[1], s[4] = bb[2], s[1] = bb[3], s[3] = bb[4], s[2] = bb[5];
}

extern "C" int lpmx_compute_strain(double *strain_io)
{
    RETCHK(lpmx_require_model());
    if (!strain_io)
        return lpmx_fail(LPMX_ERR_ARG, "strain buffer");
    double *d = nullptr;
    const size_t bytes = sizeof(double) * 6 * (size_t)g.n;
    HIPCHK(hipMalloc(&d, bytes));
    HIPCHK(hipMemcpyAsync(d, strain_io, bytes, hipMemcpyHostToDevice, g_stream));
    if (g.dim == 3)
        hipLaunchKernelGGL(k_strain<3>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, d);
    else
        hipLaunchKernelGGL(k_strain<2>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(strain_io, d, bytes, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(g_stream);
    hipFree(d);
    HIPCHK(e);
    return LPMX_OK;
}
