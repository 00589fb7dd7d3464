// Tangent stiffness on the device: analytic elastic Jacobian rows symmetrised as the
// reference does, straight into the sliced-ELL block matrix, displacement-BC elimination,
// block-Jacobi inverse, and the reference's upper-CSR view of the matrix.
#include <math.h>

#include "common.h"

static double *t_S = nullptr;   // per-particle shell sums of the elastic law, TS_NPLANE planes of n
static int *t_kstart = nullptr; // K_pointer[i][1] (0-based starts), n+1
static double t_nd = 0.0;       // diagonal value of the constrained DOFs
static bool t_nd_valid = false; // computed since the last tangent assembly
// version of the device K, bumped by every kernel sequence that writes it (assembly,
// import, elimination): the block-Jacobi inverse is formed once per version
static int t_kver = 0, t_bj_kver = -1;

// Planes of t_S: A_s = sum dL, B_s = sum Tv dL (s = 0, 1), V_s = sum db n and
// W_s = sum db Tv n (3 components each) over the particle's shell-s bonds.
#define TS_A 0
#define TS_B 2
#define TS_V 4
#define TS_W 10
#define TS_NPLANE 16

// Shell sums of every particle under the elastic law (constitutive.c:241-261):
// dL = (|x_i - x_j| - L0 - dLp[0]) * damage_broken, n = (x_i - x_j)/|x_i - x_j|.
__global__ void k_tangent_sums(DV v, double *S)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int64_t n = v.n;
    double a[2] = {0, 0}, bb[2] = {0, 0}, V[2][3] = {{0, 0, 0}, {0, 0, 0}}, W[2][3] = {{0, 0, 0}, {0, 0, 0}};
    const double xi = v.x[3 * i], yi = v.x[3 * i + 1], zi = v.x[3 * i + 2];
    for (int e = 0; e < v.nb0[i]; e++)
    {
        const int64_t b = (int64_t)e * n + i;
        const int j = v.nbr[b], s = v.nsign[b];
        const double dx = xi - v.x[3 * j], dy = yi - v.x[3 * j + 1], dz = zi - v.x[3 * j + 2];
        const double dis = sqrt(dx * dx + dy * dy + dz * dz);
        const double db = v.db[b], tv = v.Tv[b];
        double dl = dis - v.L0[b];
        dl -= v.dLp0[b];
        dl *= db;
        const double nv[3] = {dx / dis, dy / dis, dz / dis};
        for (int q = 0; q < 2; q++)
            if (s == q)
            {
                a[q] += dl, bb[q] += tv * dl;
                for (int c = 0; c < 3; c++)
                    V[q][c] += db * nv[c], W[q][c] += db * tv * nv[c];
            }
    }
    for (int q = 0; q < 2; q++)
    {
        S[(TS_A + q) * n + i] = a[q];
        S[(TS_B + q) * n + i] = bb[q];
        for (int c = 0; c < 3; c++)
            S[(TS_V + 3 * q + c) * n + i] = V[q][c], S[(TS_W + 3 * q + c) * n + i] = W[q][c];
    }
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

// Row i of the symmetrised tangent, written straight into the sliced-ELL K and K32:
// K(i,m) = 0.5 (J_im + J_mi^T) for m != i, the upper triangle of J_ii mirrored for
// m = i (stiffness.c:441-483), J = dPin/dx the Jacobian of the elastic law.  With
// n = n_ij, g_ij = db_ij n_ij, F_ij = db_ij [2Kn dL_ij + (B_i + B_j)/2 + Tv_ij (A_i + A_j)/2]
// (sums over the bond's shell s) and G_ij = F_ij/d (I - n n^T) + 2 Kn db_ij^2 n n^T,
// the terms of bond (i,j) are
//   J_ii += G_ij + db_ij/2 n (W_i + Tv_ij V_i)^T            (stretch and i's shell)
//   J_ij += -G_ij + db_ij/2 n (W_j + Tv_ij V_j)^T - 1/2 (Tv_ij V_i + W_i) g_ij^T
//   J_il -= db_ij/2 (Tv_jl + Tv_ij) n g_jl^T,   l in j's shell-s bonds (two hops)
// with V = sum db n, W = sum db Tv n over the particle's shell-s bonds.  A two-hop
// path i -> j -> l contributes to K(i,l) through J_il and, reversed, through J_li:
// 0.5 (alpha + beta) n_ij n_jl^T with alpha = -db_ij (Tv_jl + Tv_ij) db_jl / 2 and
// beta = -db_lj (Tv_ji + Tv_lj) db_ji / 2; for l = i it is J_ii's alpha term.
// One 64-lane work-group per row.  Lane a handles bond a; every bond's off-diagonal
// block lands on a distinct conn slot, and the two-hop rounds (bond a of i, lane t =
// bond t of j) land on distinct slots within a round and are added in bond order:
// the sums are deterministic.  Rows are mapped so that each XCD assembles a
// contiguous range, so the partial cache lines of a slice meet in one L2.
template <int D>
__global__ __launch_bounds__(64) void k_tangent_rows(DV v, const double *S)
{
    constexpr int D2 = D * D;
    __shared__ double acc[LPMX_MAXCONN][D2];
    __shared__ double dg[LPMX_MAXNB][D2];
    __shared__ double bn[LPMX_MAXNB][3], bx[LPMX_MAXNB][3], bdb[LPMX_MAXNB], btv[LPMX_MAXNB], bdbr[LPMX_MAXNB],
        btvr[LPMX_MAXNB];
    __shared__ int sc[LPMX_MAXCONN], bj[LPMX_MAXNB], bs[LPMX_MAXNB];
    const int i = xcd_block(), lane = threadIdx.x;
    const int64_t n = v.n;
    const int nc = v.nb_conn[i], nbi = v.nb0[i];
    if (lane < nc)
        sc[lane] = v.conn[(int64_t)lane * n + i];
    for (int e = lane; e < LPMX_MAXCONN * D2; e += 64)
        (&acc[0][0])[e] = 0.0;
    __syncthreads();
    const double xi[3] = {v.x[3 * i], v.x[3 * i + 1], v.x[3 * i + 2]};
    // ---- bonds of i, one per lane
    if (lane < nbi)
    {
        const int64_t b = (int64_t)lane * n + i;
        const int j = v.nbr[b], s = v.nsign[b];
        const double db = v.db[b], tv = v.Tv[b], kn = v.Kn[b];
        const int64_t br = (int64_t)v.rev[b] * n + j;
        const double dbr = v.db[br], tvr = v.Tv[br], knr = v.Kn[br];
        const double xj[3] = {v.x[3 * j], v.x[3 * j + 1], v.x[3 * j + 2]};
        double nv[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        const double d = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        for (int c = 0; c < 3; c++)
            nv[c] /= d;
        const double Ai = S[(TS_A + s) * n + i], Bi = S[(TS_B + s) * n + i];
        const double Aj = S[(TS_A + s) * n + j], Bj = S[(TS_B + s) * n + j];
        const double dl = (d - v.L0[b] - v.dLp0[b]) * db, dlr = (d - v.L0[br] - v.dLp0[br]) * dbr;
        const double F = db * (2.0 * kn * dl + 0.5 * (Bi + Bj) + 0.5 * tv * (Ai + Aj));
        const double Fr = dbr * (2.0 * knr * dlr + 0.5 * (Bj + Bi) + 0.5 * tvr * (Aj + Ai));
        double Vi[3], Wi[3], Vj[3], Wj[3];
        for (int c = 0; c < 3; c++)
        {
            Vi[c] = S[(TS_V + 3 * s + c) * n + i], Wi[c] = S[(TS_W + 3 * s + c) * n + i];
            Vj[c] = S[(TS_V + 3 * s + c) * n + j], Wj[c] = S[(TS_W + 3 * s + c) * n + j];
        }
        const double fd = F / d, fdr = Fr / d, k2 = 2.0 * kn * db * db, k2r = 2.0 * knr * dbr * dbr;
        const double w2 = 0.5 * db * dbr * (tvr + tv);
        const int pj = conn_find(sc, nc, j);
        for (int r = 0; r < D; r++)
            for (int c = 0; c < D; c++)
            {
                const double nn = nv[r] * nv[c], del = r == c ? 1.0 : 0.0;
                // J_ij of this bond, and J_ji's transpose (n_ji = -n)
                const double jij = -(fd * (del - nn) + k2 * nn) + 0.5 * db * nv[r] * (Wj[c] + tv * Vj[c]) -
                                   0.5 * (tv * Vi[r] + Wi[r]) * db * nv[c];
                const double jjiT = -(fdr * (del - nn) + k2r * nn) - 0.5 * dbr * (Wi[r] + tvr * Vi[r]) * nv[c] +
                                    0.5 * dbr * nv[r] * (tvr * Vj[c] + Wj[c]);
                if (pj >= 0)
                    acc[pj][r * D + c] = 0.5 * (jij + jjiT);
                dg[lane][r * D + c] =
                    fd * (del - nn) + k2 * nn + 0.5 * db * nv[r] * (Wi[c] + tv * Vi[c]) + w2 * nn;
            }
        bj[lane] = j, bs[lane] = s;
        for (int c = 0; c < 3; c++)
            bn[lane][c] = nv[c], bx[lane][c] = xj[c];
        bdb[lane] = db, btv[lane] = tv, bdbr[lane] = dbr, btvr[lane] = tvr;
    }
    __syncthreads();
    const int pii = v.dslot[i];
    if (lane < D2)
    {
        double t = 0.0;
        for (int a = 0; a < nbi; a++)
            t += dg[a][lane];
        acc[pii][lane] += t;
    }
    __syncthreads();
    // ---- two-hop paths i -> j -> l: groups of nb lanes take consecutive bonds of i
    const int nbm = v.nb, ng = 64 / nbm, grp = lane / nbm, t = lane % nbm;
    for (int a0 = 0; a0 < nbi; a0 += ng)
    {
        const int a = a0 + grp;
        double gam = 0.0, njl[3] = {0, 0, 0};
        int pl = -1;
        if (grp < ng && a < nbi)
        {
            const int j = bj[a], s = bs[a];
            if (t < v.nb0[j])
            {
                const int64_t bt = (int64_t)t * n + j;
                const int l = v.nbr[bt];
                if (v.nsign[bt] == s && l != i)
                {
                    const double dbjl = v.db[bt], tvjl = v.Tv[bt];
                    const int64_t br = (int64_t)v.rev[bt] * n + l;
                    const double dblj = v.db[br], tvlj = v.Tv[br];
                    double u[3] = {bx[a][0] - v.x[3 * l], bx[a][1] - v.x[3 * l + 1], bx[a][2] - v.x[3 * l + 2]};
                    const double djl = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                    for (int c = 0; c < 3; c++)
                        njl[c] = u[c] / djl;
                    const double alpha = -0.5 * bdb[a] * (tvjl + btv[a]) * dbjl;
                    const double beta = -0.5 * dblj * (btvr[a] + tvlj) * bdbr[a];
                    gam = 0.5 * (alpha + beta);
                    pl = conn_find(sc, nc, l);
                }
            }
        }
        // groups add in bond order; within a group the slots are distinct
        for (int q = 0; q < ng; q++)
        {
            if (grp == q && pl >= 0)
            {
                const int a = a0 + q;
                for (int r = 0; r < D; r++)
                    for (int c = 0; c < D; c++)
                        acc[pl][r * D + c] += gam * bn[a][r] * njl[c];
            }
            __syncthreads();
        }
    }
    // ---- the row's blocks in slice order; padding slots hold zero blocks
    const int si = i / LPMX_SLICE, li = i % LPMX_SLICE;
    const int w = v.sptr[si + 1] - v.sptr[si];
    for (int k = lane; k < w; k += 64)
    {
        const int64_t slot = v.sptr[si] + k;
        for (int r = 0; r < D; r++)
            for (int c = 0; c < D; c++)
            {
                double val = 0.0;
                if (k < nc)
                    val = (k == pii && r > c) ? acc[k][c * D + r] : acc[k][r * D + c];
                const int64_t o = (slot * D2 + r * D + c) * LPMX_SLICE + li;
                v.K[o] = val;
                v.K32[o] = (float)val;
            }
    }
}

static int ensure_tangent_buffers(void)
{
    if (!t_S)
        HIPCHK(hipMalloc(&t_S, sizeof(double) * TS_NPLANE * (int64_t)g.n));
    return LPMX_OK;
}

void lpmx_tangent_free(void)
{
    if (t_S)
        hipFree(t_S);
    if (t_kstart)
        hipFree(t_kstart);
    t_S = nullptr;
    t_kstart = nullptr;
    t_nd_valid = false;
}

static int t_epoch = 0;
int lpmx_tangent_epoch(void) { return t_epoch; }

extern "C" int lpmx_tangent(void)
{
    RETCHK(lpmx_require_pattern());
    if (g.nb > 64 || g.nb > LPMX_MAXNB || g.cm > LPMX_MAXCONN)
        return lpmx_fail(LPMX_ERR_CAPACITY, "tangent rows handle at most %d bonds and %d conn entries", LPMX_MAXNB,
                         LPMX_MAXCONN);
    RETCHK(ensure_tangent_buffers());
    t_nd_valid = false;
    t_epoch++;
    t_kver++;
    hipLaunchKernelGGL(k_tangent_sums, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, t_S);
    LAUNCHCHK();
    if (g.dim == 3)
        hipLaunchKernelGGL(k_tangent_rows<3>, dim3(g.n), dim3(64), 0, g_stream, g, t_S);
    else
        hipLaunchKernelGGL(k_tangent_rows<2>, dim3(g.n), dim3(64), 0, g_stream, g, t_S);
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
