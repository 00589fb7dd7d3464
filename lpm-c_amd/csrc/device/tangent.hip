// Tangent stiffness on the device: analytic elastic Jacobian rows, the reference's
// symmetrisation into the sliced-ELL block matrix, displacement-BC elimination,
// block-Jacobi inverse, and the reference's upper-CSR view of the matrix.
#include <math.h>

#include "common.h"

static double *t_A = nullptr, *t_B = nullptr; // shell sums of the elastic law
static int *t_kstart = nullptr;               // K_pointer[i][1] (0-based starts), n+1
static double t_nd = 0.0;                     // diagonal value of the constrained DOFs
static bool t_nd_valid = false;               // computed since the last tangent assembly
// version of the device K, bumped by every kernel sequence that writes it (assembly,
// import, elimination): the block-Jacobi inverse is formed once per version
static int t_kver = 0, t_bj_kver = -1;

// Shell sums A_s = sum dL, B_s = sum Tv dL of every particle under the elastic law
// (constitutive.c:241-261): dL = (|x_i - x_j| - L0 - dLp[0]) * damage_broken.
__global__ void k_elastic_sums(DV v, double *A, double *B)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    const double xi = v.x[3 * i], yi = v.x[3 * i + 1], zi = v.x[3 * i + 2];
    for (int a = 0; a < v.nb0[i]; a++)
    {
        int64_t b = (int64_t)a * v.n + i;
        int j = v.nbr[b];
        double dx = xi - v.x[3 * j], dy = yi - v.x[3 * j + 1], dz = zi - v.x[3 * j + 2];
        double dis = sqrt(dx * dx + dy * dy + dz * dz);
        double dl = dis - v.L0[b];
        dl -= v.dLp0[b];
        dl *= v.db[b];
        if (v.nsign[b] == 0)
            a0 += dl, b0 += v.Tv[b] * dl;
        else
            a1 += dl, b1 += v.Tv[b] * dl;
    }
    A[i] = a0, A[(int64_t)v.n + i] = a1;
    B[i] = b0, B[(int64_t)v.n + i] = b1;
}

__device__ __forceinline__ int conn_find(const int *sc, int nc, int q)
{
    int lo = 0, hi = nc - 1;
    while (lo <= hi)
    {
        int mid = (lo + hi) >> 1;
        int c = sc[mid];
        if (c == q)
            return mid;
        if (c < q)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

// Row i of the Jacobian J_im = dPin_i/dx_m of the elastic law, m in conn[i]; one
// wavefront per row.  With n = n_ij, g_ik = db_ik n_ik, P = I - n n^T:
//   J_ii += F P/d + 2 Kn db^2 n n^T,            J_ij -= the same
//   J_ii += n (sum_k w_k g_ik)^T,  J_ik -= w_k n g_ik^T,  w_k = db (Tv_ik + Tv_ij)/2
//   J_ij += n (sum_l w_l g_jl)^T,  J_il -= w_l n g_jl^T,  w_l = db (Tv_jl + Tv_ij)/2
// (k over i's shell-s bonds, l over j's shell-s bonds).  Every LDS block entry has
// one writer per phase and the phases run in bond order, so the sums are
// deterministic.
template <int D>
__global__ __launch_bounds__(64) void k_jacobian(DV v, const double *A, const double *B)
{
    constexpr int D2 = D * D;
    __shared__ double J[LPMX_MAXCONN][D2];
    __shared__ int sc[LPMX_MAXCONN];
    const int i = blockIdx.x, lane = threadIdx.x;
    const int n = v.n, nc = v.nb_conn[i];
    if (lane < nc)
        sc[lane] = v.conn[(int64_t)lane * n + i];
    for (int e = lane; e < LPMX_MAXCONN * D2; e += 64)
        (&J[0][0])[e] = 0.0;
    __syncthreads();
    const int pi = v.dslot[i];
    const double xi[3] = {v.x[3 * i], v.x[3 * i + 1], v.x[3 * i + 2]};
    const int nbi = v.nb0[i];
    for (int a = 0; a < nbi; a++)
    {
        const int64_t b = (int64_t)a * n + i;
        const double db = v.db[b];
        if (db == 0.0)
            continue;
        const int j = v.nbr[b], s = v.nsign[b];
        const double Kn = v.Kn[b], Tvij = v.Tv[b];
        const double xj[3] = {v.x[3 * j], v.x[3 * j + 1], v.x[3 * j + 2]};
        double nv[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        const double dij = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        for (int c = 0; c < 3; c++)
            nv[c] /= dij;
        double dl = (dij - v.L0[b] - v.dLp0[b]) * db;
        const double F = db * (2.0 * Kn * dl + 0.5 * (B[(int64_t)s * n + i] + B[(int64_t)s * n + j]) +
                               0.5 * Tvij * (A[(int64_t)s * n + i] + A[(int64_t)s * n + j]));
        const int pj = conn_find(sc, nc, j);
        if (pj < 0 || pi < 0 || pi >= nc)
            continue; // cannot happen for a conn built from these lists
        // geometric and direct stretch terms
        if (lane < D2)
        {
            const int r = lane / D, c = lane % D;
            double gval = F * ((r == c ? 1.0 : 0.0) - nv[r] * nv[c]) / dij + 2.0 * Kn * db * db * nv[r] * nv[c];
            J[pi][lane] += gval;
            J[pj][lane] -= gval;
        }
        __syncthreads();
        // i's own shell-s bonds
        double u[3] = {0, 0, 0};
        if (lane < nbi)
        {
            const int64_t be = (int64_t)lane * n + i;
            const double dbe = v.db[be];
            if (v.nsign[be] == s && dbe != 0.0)
            {
                const int k = v.nbr[be];
                double gk[3] = {xi[0] - v.x[3 * k], xi[1] - v.x[3 * k + 1], xi[2] - v.x[3 * k + 2]};
                const double dik = sqrt(gk[0] * gk[0] + gk[1] * gk[1] + gk[2] * gk[2]);
                for (int c = 0; c < 3; c++)
                    gk[c] = dbe * gk[c] / dik;
                const double w = 0.5 * db * (v.Tv[be] + Tvij);
                const int pk = conn_find(sc, nc, k);
                if (pk >= 0)
                    for (int r = 0; r < D; r++)
                    for (int c = 0; c < D; c++)
                        J[pk][r * D + c] -= w * nv[r] * gk[c];
                for (int c = 0; c < 3; c++)
                    u[c] = w * gk[c];
            }
        }
        for (int c = 0; c < 3; c++)
            u[c] = wave_sum(u[c]);
        __syncthreads();
        if (lane < D2)
            J[pi][lane] += nv[lane / D] * u[lane % D];
        __syncthreads();
        // j's shell-s bonds
        double q[3] = {0, 0, 0};
        const int nbj = v.nb0[j];
        if (lane < nbj)
        {
            const int64_t be = (int64_t)lane * n + j;
            const double dbe = v.db[be];
            if (v.nsign[be] == s && dbe != 0.0)
            {
                const int l = v.nbr[be];
                double gl[3] = {xj[0] - v.x[3 * l], xj[1] - v.x[3 * l + 1], xj[2] - v.x[3 * l + 2]};
                const double djl = sqrt(gl[0] * gl[0] + gl[1] * gl[1] + gl[2] * gl[2]);
                for (int c = 0; c < 3; c++)
                    gl[c] = dbe * gl[c] / djl;
                const double w = 0.5 * db * (v.Tv[be] + Tvij);
                const int pl = conn_find(sc, nc, l);
                if (pl >= 0)
                    for (int r = 0; r < D; r++)
                    for (int c = 0; c < D; c++)
                        J[pl][r * D + c] -= w * nv[r] * gl[c];
                for (int c = 0; c < 3; c++)
                    q[c] = w * gl[c];
            }
        }
        for (int c = 0; c < 3; c++)
            q[c] = wave_sum(q[c]);
        __syncthreads();
        if (lane < D2)
            J[pj][lane] += nv[lane / D] * q[lane % D];
        __syncthreads();
    }
    const int64_t base = v.rowptr[i] * D2;
    for (int e = lane; e < nc * D2; e += 64)
        v.J[base + e] = (&J[0][0])[e];
}

// K(i,m) = 0.5 (J_im + J_mi^T) for m != i; K(i,i) = the upper triangle of J_ii
// mirrored (stiffness.c:441-483).  One 64-thread group per slice; lane = row.
template <int D>
__global__ __launch_bounds__(64) void k_symmetrize(DV v)
{
    constexpr int D2 = D * D;
    const int s = blockIdx.x, l = threadIdx.x, i = s * LPMX_SLICE + l;
    if (i >= v.n)
        return;
    const int w = v.sptr[s + 1] - v.sptr[s], nc = v.nb_conn[i];
    const int64_t jrow = v.rowptr[i];
    for (int k = 0; k < w; k++)
    {
        const int64_t slot = v.sptr[s] + k;
        double out[D2];
        if (k >= nc)
        {
            for (int e = 0; e < D2; e++)
                out[e] = 0.0;
        }
        else
        {
            const int m = v.col[slot * LPMX_SLICE + l];
            const double *Jim = v.J + (jrow + k) * D2;
            if (m == i)
            {
                for (int r = 0; r < D; r++)
                    for (int c = 0; c < D; c++)
                        out[r * D + c] = r <= c ? Jim[r * D + c] : Jim[c * D + r];
            }
            else
            {
                const int tp = v.tpos[slot * LPMX_SLICE + l];
                const int kt = tp < 0 ? -1 : tp / LPMX_SLICE - v.sptr[m / LPMX_SLICE];
                const double *Jmi = v.J + (v.rowptr[m] + (kt < 0 ? 0 : kt)) * D2;
                for (int r = 0; r < D; r++)
                    for (int c = 0; c < D; c++)
                        out[r * D + c] = 0.5 * Jim[r * D + c] + (kt < 0 ? 0.0 : 0.5 * Jmi[c * D + r]);
            }
        }
        for (int e = 0; e < D2; e++)
        {
            const int64_t o = (slot * D2 + e) * LPMX_SLICE + l;
            v.K[o] = out[e];
            v.K32[o] = (float)out[e];
        }
    }
}

static int ensure_tangent_buffers(void)
{
    if (!t_A)
    {
        HIPCHK(hipMalloc(&t_A, sizeof(double) * 2 * (int64_t)g.n));
        HIPCHK(hipMalloc(&t_B, sizeof(double) * 2 * (int64_t)g.n));
    }
    return LPMX_OK;
}

void lpmx_tangent_free(void)
{
    if (t_A)
        hipFree(t_A), hipFree(t_B);
    if (t_kstart)
        hipFree(t_kstart);
    t_A = t_B = nullptr;
    t_kstart = nullptr;
    t_nd_valid = false;
}

static int t_epoch = 0;
int lpmx_tangent_epoch(void) { return t_epoch; }

extern "C" int lpmx_tangent(void)
{
    RETCHK(lpmx_require_pattern());
    RETCHK(ensure_tangent_buffers());
    t_nd_valid = false;
    t_epoch++;
    t_kver++;
    hipLaunchKernelGGL(k_elastic_sums, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, t_A, t_B);
    LAUNCHCHK();
    if (g.dim == 3)
    {
        hipLaunchKernelGGL(k_jacobian<3>, dim3(g.n), dim3(64), 0, g_stream, g, t_A, t_B);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_symmetrize<3>, dim3(g.nslices), dim3(64), 0, g_stream, g);
    }
    else
    {
        hipLaunchKernelGGL(k_jacobian<2>, dim3(g.n), dim3(64), 0, g_stream, g, t_A, t_B);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_symmetrize<2>, dim3(g.nslices), dim3(64), 0, g_stream, g);
    }
    LAUNCHCHK();
    return LPMX_OK;
}

// -------------------------------------------------------- BC elimination ---------
__device__ __forceinline__ bool dof_fixed(const DV &v, int i, int c)
{
    const int64_t k = (int64_t)i * v.dim + c;
    return v.dispBC[k] == 0 || v.fix[k] == 0;
}

// Zero the rows of constrained DOFs of row i and the columns of constrained DOFs
// of the column particle, in every block of row i; the diagonal entry of a
// constrained DOF is left for k_bc_diag (boundary.c:159-281).
template <int D>
__global__ __launch_bounds__(64) void k_bc_zero(DV v)
{
    constexpr int D2 = D * D;
    const int s = blockIdx.x, l = threadIdx.x, i = s * LPMX_SLICE + l;
    if (i >= v.n)
        return;
    bool ri[D];
    bool anyi = false;
    for (int c = 0; c < D; c++)
        ri[c] = dof_fixed(v, i, c), anyi |= ri[c];
    const int w = v.nb_conn[i];
    for (int k = 0; k < w; k++)
    {
        const int64_t slot = v.sptr[s] + k;
        const int m = v.col[slot * LPMX_SLICE + l];
        bool cm[D];
        bool any = anyi;
        for (int c = 0; c < D; c++)
            cm[c] = dof_fixed(v, m, c), any |= cm[c];
        if (!any)
            continue;
        for (int r = 0; r < D; r++)
            for (int c = 0; c < D; c++)
                if ((ri[r] || cm[c]) && !(m == i && r == c))
                {
                    const int64_t o = (slot * D2 + r * D + c) * LPMX_SLICE + l;
                    v.K[o] = 0.0;
                    v.K32[o] = 0.0f;
                }
    }
}

template <int D>
__global__ void k_diag_sq(DV v, double *sq)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int s = i / LPMX_SLICE, l = i % LPMX_SLICE;
    const int64_t slot = v.sptr[s] + v.dslot[i];
    for (int c = 0; c < D; c++)
    {
        double d = v.K[(slot * D * D + c * D + c) * LPMX_SLICE + l];
        sq[(int64_t)i * D + c] = d;
    }
}

template <int D>
__global__ void k_bc_diag(DV v, double nd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int s = i / LPMX_SLICE, l = i % LPMX_SLICE;
    const int64_t slot = v.sptr[s] + v.dslot[i];
    for (int c = 0; c < D; c++)
        if (dof_fixed(v, i, c))
        {
            const int64_t o = (slot * D * D + c * D + c) * LPMX_SLICE + l;
            v.K[o] = nd;
            v.K32[o] = (float)nd;
            v.res[(int64_t)i * D + c] = 0.0;
        }
}

// fixed_diag = 0: the reference's rule, |diag(K)|_2 of the current matrix (which
// already holds the previous call's value on constrained DOFs, so it grows by about
// sqrt(#constrained DOFs) per call).  fixed_diag = 1: the value of the first call
// after the tangent assembly, kept until the next assembly.  Constrained rows and
// columns are decoupled, so the value never changes the solution.
extern "C" int lpmx_bc_eliminate_ex(int fixed_diag)
{
    RETCHK(lpmx_require_pattern());
    const int n = g.n, D = g.dim;
    double *sq = lpmx_scratch(sizeof(double) * (int64_t)D * n + 64);
    if (!sq)
        return lpmx_fail(LPMX_ERR_HIP, "scratch allocation failed");
    // the current diagonal first (the reference takes its norm before zeroing)
    if (D == 3)
        hipLaunchKernelGGL(k_diag_sq<3>, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, sq);
    else
        hipLaunchKernelGGL(k_diag_sq<2>, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, sq);
    LAUNCHCHK();
    double nd = t_nd;
    if (!(fixed_diag && t_nd_valid))
    {
        double ss = 0;
        RETCHK(lpmx_dev_sumsq(sq, (int64_t)D * n, &ss));
        nd = sqrt(ss);
        t_nd = nd;
        t_nd_valid = true;
    }
    t_kver++;
    if (D == 3)
    {
        hipLaunchKernelGGL(k_bc_zero<3>, dim3(g.nslices), dim3(64), 0, g_stream, g);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_bc_diag<3>, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, nd);
    }
    else
    {
        hipLaunchKernelGGL(k_bc_zero<2>, dim3(g.nslices), dim3(64), 0, g_stream, g);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_bc_diag<2>, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, nd);
    }
    LAUNCHCHK();
    return LPMX_OK;
}

extern "C" int lpmx_bc_eliminate(void) { return lpmx_bc_eliminate_ex(0); }

template <int D>
__global__ void k_bc_res(DV v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    for (int c = 0; c < D; c++)
        if (dof_fixed(v, i, c))
            v.res[(int64_t)i * D + c] = 0.0;
}

// A repeated elimination with the fixed diagonal on an unchanged tangent and
// constraint set leaves K as it is: only the residual of the constrained DOFs is
// zeroed (the fused loop's Newton iterations after the first of a tangent).
int lpmx_bc_eliminate_repeat(void)
{
    RETCHK(lpmx_require_pattern());
    if (!t_nd_valid)
        return lpmx_bc_eliminate_ex(1);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_bc_res<3>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g);
    else
        hipLaunchKernelGGL(k_bc_res<2>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g);
    LAUNCHCHK();
    return LPMX_OK;
}

// ------------------------------------------------------ block-Jacobi inverse -----
template <int D>
__global__ void k_bj_inverse(DV v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int s = i / LPMX_SLICE, l = i % LPMX_SLICE;
    const int64_t slot = v.sptr[s] + v.dslot[i];
    double a[D * D], inv[D * D];
    for (int e = 0; e < D * D; e++)
        a[e] = v.K[(slot * D * D + e) * LPMX_SLICE + l];
    if constexpr (D == 3)
    {
        inv[0] = a[4] * a[8] - a[5] * a[7];
        inv[1] = a[2] * a[7] - a[1] * a[8];
        inv[2] = a[1] * a[5] - a[2] * a[4];
        inv[3] = a[5] * a[6] - a[3] * a[8];
        inv[4] = a[0] * a[8] - a[2] * a[6];
        inv[5] = a[2] * a[3] - a[0] * a[5];
        inv[6] = a[3] * a[7] - a[4] * a[6];
        inv[7] = a[1] * a[6] - a[0] * a[7];
        inv[8] = a[0] * a[4] - a[1] * a[3];
        double det = a[0] * inv[0] + a[1] * inv[3] + a[2] * inv[6];
        double id = det != 0.0 ? 1.0 / det : 0.0;
        for (int e = 0; e < 9; e++)
            inv[e] *= id;
    }
    else
    {
        double det = a[0] * a[3] - a[1] * a[2];
        double id = det != 0.0 ? 1.0 / det : 0.0;
        inv[0] = a[3] * id, inv[1] = -a[1] * id, inv[2] = -a[2] * id, inv[3] = a[0] * id;
    }
    for (int e = 0; e < D * D; e++)
    {
        v.Minv[(int64_t)e * v.n + i] = inv[e];
        v.Minv32[(int64_t)e * v.n + i] = (float)inv[e];
    }
}

int lpmx_bj_inverse(void)
{
    if (!g.Minv)
    {
        HIPCHK(hipMalloc(&g.Minv, sizeof(double) * g.d2 * (int64_t)g.n));
        HIPCHK(hipMalloc(&g.Minv32, sizeof(float) * g.d2 * (int64_t)g.n));
        t_bj_kver = -1;
    }
    if (t_bj_kver == t_kver)
        return LPMX_OK;
    t_bj_kver = t_kver;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_bj_inverse<3>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g);
    else
        hipLaunchKernelGGL(k_bj_inverse<2>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g);
    LAUNCHCHK();
    return LPMX_OK;
}

// ------------------------------------------------- upper CSR (reference view) ----
__global__ void k_row_len_k(const int *kcnt, int *len, int n, int d2, int sub)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        len[i] = d2 * kcnt[i] - sub;
    else if (i == n)
        len[i] = 0;
}

static int ensure_kstart(int *total)
{
    const int n = g.n;
    if (!t_kstart)
        HIPCHK(hipMalloc(&t_kstart, sizeof(int) * ((int64_t)n + 1)));
    int *len = nullptr;
    HIPCHK(hipMalloc(&len, sizeof(int) * ((int64_t)n + 1)));
    hipLaunchKernelGGL(k_row_len_k, dim3(nblocks(n + 1)), dim3(LPMX_BLOCK), 0, g_stream, g.kcnt, len, n, g.d2,
                       g.dim == 2 ? 1 : 3);
    LAUNCHCHK();
    int rc = lpmx_dev_scan_int(len, t_kstart, n + 1, total);
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(len);
    return rc;
}

// Position in the reference's upper CSR of entry (row r of particle i, column c
// of the num-th block of i with column index >= i; num 0 is the diagonal block),
// stiffness.c:441-505: row 3i+r starts at start + r*(3k) - r*(r-1)/2 ... written
// out per row below.
template <int D>
__device__ __forceinline__ int64_t csr_pos(int64_t start, int k, int r, int num, int c)
{
    if (D == 3)
    {
        if (num == 0)
            return r == 0 ? start + c : (r == 1 ? start + 3 * k + (c - 1) : start + 6 * k - 1);
        return r == 0 ? start + 3 * num + c : (r == 1 ? start + 3 * k + 3 * num - 1 + c : start + 6 * k + 3 * num - 3 + c);
    }
    if (num == 0)
        return r == 0 ? start + c : start + 2 * k;
    return r == 0 ? start + 2 * num + c : start + 2 * k + 2 * num - 1 + c;
}

template <int D>
__global__ void k_export_csr(DV v, const int *kstart, int *IK, int *JK, double *Kc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int s = i / LPMX_SLICE, l = i % LPMX_SLICE;
    const int64_t start = kstart[i];
    const int k = v.kcnt[i];
    int num = 0;
    for (int q = 0; q < v.nb_conn[i]; q++)
    {
        const int m = v.conn[(int64_t)q * v.n + i];
        if (m < i)
            continue;
        const int64_t slot = v.sptr[s] + q;
        for (int r = 0; r < D; r++)
            for (int c = (m == i ? r : 0); c < D; c++)
            {
                int64_t p = csr_pos<D>(start, k, r, num, c);
                if (JK)
                    JK[p] = D * m + c + 1;
                if (Kc)
                    Kc[p] = v.K[(slot * D * D + r * D + c) * LPMX_SLICE + l];
            }
        num++;
    }
    if (IK)
    {
        IK[D * i] = (int)start + 1;
        IK[D * i + 1] = (int)(start + D * k + 1);
        if (D == 3)
            IK[D * i + 2] = (int)(start + 2 * D * k);
        if (i == v.n - 1)
            IK[D * v.n] = kstart[v.n] + 1;
    }
}

template <int D>
__global__ void k_import_csr(DV v, const int *kstart, const double *Kc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int s = i / LPMX_SLICE, l = i % LPMX_SLICE;
    const int64_t start = kstart[i];
    const int k = v.kcnt[i];
    int num = 0;
    for (int q = 0; q < v.nb_conn[i]; q++)
    {
        const int m = v.conn[(int64_t)q * v.n + i];
        if (m < i)
            continue;
        const int64_t slot = v.sptr[s] + q;
        const int tp = v.tpos[slot * LPMX_SLICE + l];
        for (int r = 0; r < D; r++)
            for (int c = (m == i ? r : 0); c < D; c++)
            {
                const double val = Kc[csr_pos<D>(start, k, r, num, c)];
                const int64_t o = (slot * D * D + r * D + c) * LPMX_SLICE + l;
                v.K[o] = val;
                v.K32[o] = (float)val;
                if (m == i)
                {
                    const int64_t o2 = (slot * D * D + c * D + r) * LPMX_SLICE + l;
                    v.K[o2] = val;
                    v.K32[o2] = (float)val;
                }
                else
                {
                    // transposed block (m, i): element (c, r)
                    const int64_t tslot = tp / LPMX_SLICE, tl = tp % LPMX_SLICE;
                    const int64_t o2 = (tslot * D * D + c * D + r) * LPMX_SLICE + tl;
                    v.K[o2] = val;
                    v.K32[o2] = (float)val;
                }
            }
        num++;
    }
}

extern "C" int lpmx_export_upper_csr(int *IK, int *JK, double *K)
{
    RETCHK(lpmx_require_pattern());
    int total = 0;
    RETCHK(ensure_kstart(&total));
    const int n = g.n, D = g.dim;
    int *dIK = nullptr, *dJK = nullptr;
    double *dK = nullptr;
    if (IK)
        HIPCHK(hipMalloc(&dIK, sizeof(int) * ((int64_t)D * n + 1)));
    if (JK)
        HIPCHK(hipMalloc(&dJK, sizeof(int) * (int64_t)total));
    if (K)
        HIPCHK(hipMalloc(&dK, sizeof(double) * (int64_t)total));
    if (D == 3)
        hipLaunchKernelGGL(k_export_csr<3>, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, t_kstart, dIK, dJK, dK);
    else
        hipLaunchKernelGGL(k_export_csr<2>, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, t_kstart, dIK, dJK, dK);
    LAUNCHCHK();
    if (IK)
        HIPCHK(hipMemcpyAsync(IK, dIK, sizeof(int) * ((int64_t)D * n + 1), hipMemcpyDeviceToHost, g_stream));
    if (JK)
        HIPCHK(hipMemcpyAsync(JK, dJK, sizeof(int) * (int64_t)total, hipMemcpyDeviceToHost, g_stream));
    if (K)
        HIPCHK(hipMemcpyAsync(K, dK, sizeof(double) * (int64_t)total, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    if (dIK)
        hipFree(dIK);
    if (dJK)
        hipFree(dJK);
    if (dK)
        hipFree(dK);
    return LPMX_OK;
}

extern "C" int lpmx_import_upper_csr(const double *K)
{
    t_epoch++;
    t_kver++;
    RETCHK(lpmx_require_pattern());
    if (!K)
        return lpmx_fail(LPMX_ERR_ARG, "null K");
    int total = 0;
    RETCHK(ensure_kstart(&total));
    double *dK = nullptr;
    HIPCHK(hipMalloc(&dK, sizeof(double) * (int64_t)total));
    HIPCHK(hipMemcpyAsync(dK, K, sizeof(double) * (int64_t)total, hipMemcpyHostToDevice, g_stream));
    if (g.dim == 3)
        hipLaunchKernelGGL(k_import_csr<3>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, t_kstart, dK);
    else
        hipLaunchKernelGGL(k_import_csr<2>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, t_kstart, dK);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(dK);
    return LPMX_OK;
}
