// Tangent stiffness on the device: analytic elastic Jacobian rows symmetrised as the
// reference does, straight into the sliced-ELL block matrix, displacement-BC elimination,
// block-Jacobi inverse, and the reference's upper-CSR view of the matrix.
#include <math.h>

#include "common.h"

static double *t_S = nullptr;   // per-particle shell sums of the elastic law, [n][TS_NPLANE]
static void *t_rec = nullptr;   // bond records, [n][nb] TBond then [n][nb] TBondOwn
static int64_t t_rec_bytes = 0;
static uint16_t *t_pstart = nullptr; // two-hop paths per conn slot: starts [n][LPMX_MAXCONN + 1]
static uint16_t *t_plist = nullptr;  // and the paths (a << 5 | t) of every row, [n][nb * nb]
static int t_paths_version = -1;     // pattern version the path lists belong to
static int64_t t_paths_cap = 0;      // entries of the one allocation holding both
static int *t_kstart = nullptr; // K_pointer[i][1] (0-based starts), n+1
static double t_nd = 0.0;       // diagonal value of the constrained DOFs
static bool t_nd_valid = false; // computed since the last tangent assembly
// version of the device K, bumped by every kernel sequence that writes it (assembly,
// import, elimination): the block-Jacobi inverse is formed once per version
static int t_kver = 0, t_bj_kver = -1;

// Entries of a row of t_S: A_s = sum dL, B_s = sum Tv dL (s = 0, 1), V_s = sum db n
// and W_s = sum db Tv n (3 components each) over the particle's shell-s bonds.
#define TS_A 0
#define TS_B 2
#define TS_V 4
#define TS_W 10
#define TS_NPLANE 16

// A bond (p, t) as the row kernel reads it: partner l (-1 for an empty slot), shell s,
// damage factors and Tv of the bond and of its reverse (l, rev), the unit vector
// (x_p - x_l) / |x_p - x_l|: one 64-byte line.  TBondOwn: Kn and dL of the bond and
// of its reverse, and the distance (read for the row particle's own bonds only).
struct TBond
{
    int l, s;
    double db, tv, dbr, tvr, n[3];
};
struct TBondOwn
{
    double kn, dl, knr, dlr, d, pad;
};

// Shell sums and bond records of every particle under the elastic law
// (constitutive.c:241-261): dL = (|x_i - x_j| - L0 - dLp[0]) * damage_broken,
// n = (x_i - x_j)/|x_i - x_j|.  The records turn the slot planes into rows, so a
// neighbour's bonds are one contiguous read for the row kernel.
__global__ void k_tangent_sums(DV v, double *__restrict__ S, TBond *__restrict__ rec, TBondOwn *__restrict__ own)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int64_t n = v.n;
    double a[2] = {0, 0}, bb[2] = {0, 0}, V[2][3] = {{0, 0, 0}, {0, 0, 0}}, W[2][3] = {{0, 0, 0}, {0, 0, 0}};
    const double xi = v.x[3 * i], yi = v.x[3 * i + 1], zi = v.x[3 * i + 2];
    TBond *ri = rec + (int64_t)i * v.nb;
    TBondOwn *oi = own + (int64_t)i * v.nb;
    const int nbi = v.nb0[i];
    for (int e = 0; e < nbi; e++)
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
        const int64_t br = (int64_t)v.rev[b] * n + j;
        const double dbr = v.db[br];
        TBond r;
        r.l = j, r.s = s, r.db = db, r.tv = tv, r.dbr = dbr, r.tvr = v.Tv[br];
        r.n[0] = nv[0], r.n[1] = nv[1], r.n[2] = nv[2];
        ri[e] = r;
        TBondOwn o;
        o.kn = v.Kn[b], o.dl = dl, o.knr = v.Kn[br], o.dlr = (dis - v.L0[br] - v.dLp0[br]) * dbr, o.d = dis, o.pad = 0;
        oi[e] = o;
    }
    for (int e = nbi; e < v.nb; e++)
        ri[e].l = -1;
    double *Si = S + (int64_t)i * TS_NPLANE;
    for (int q = 0; q < 2; q++)
    {
        Si[TS_A + q] = a[q];
        Si[TS_B + q] = bb[q];
        for (int c = 0; c < 3; c++)
            Si[TS_V + 3 * q + c] = V[q][c], Si[TS_W + 3 * q + c] = W[q][c];
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

// Two-hop paths of every row, grouped by the conn slot they land on: path (a, t) is
// i -> j = nbr(i, a) -> l = nbr(j, t) with the same shell on both bonds and l != i;
// it adds to K(i, l).  Per slot the paths keep the order of (a, t); the lists
// depend on the neighbour lists and the pattern only, so they are built once per
// pattern.  One thread per row.
__global__ void k_tangent_paths(DV v, uint16_t *__restrict__ pstart, uint16_t *__restrict__ plist)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int64_t n = v.n;
    const int nc = v.nb_conn[i], cap = v.nb * v.nb;
    uint16_t cnt[LPMX_MAXCONN + 1];
    for (int k = 0; k <= nc; k++)
        cnt[k] = 0;
    uint16_t *ps = pstart + (int64_t)i * (LPMX_MAXCONN + 1);
    uint16_t *pl = plist + (int64_t)i * cap;
    for (int pass = 0; pass < 2; pass++)
    {
        for (int a = 0; a < v.nb0[i]; a++)
        {
            const int64_t b = (int64_t)a * n + i;
            const int j = v.nbr[b], s = v.nsign[b];
            for (int t = 0; t < v.nb0[j]; t++)
            {
                const int64_t bt = (int64_t)t * n + j;
                const int l = v.nbr[bt];
                if (v.nsign[bt] != s || l == i || l < 0)
                    continue;
                int lo = 0, hi = nc - 1, k = -1;
                while (lo <= hi)
                {
                    const int mid = (lo + hi) >> 1, c = v.conn[(int64_t)mid * n + i];
                    if (c == l)
                    {
                        k = mid;
                        break;
                    }
                    if (c < l)
                        lo = mid + 1;
                    else
                        hi = mid - 1;
                }
                if (k < 0)
                    continue;
                if (pass == 0)
                    cnt[k]++;
                else
                    pl[cnt[k]++] = (uint16_t)(a << 5 | t);
            }
        }
        if (pass == 0)
        {
            int acc = 0;
            for (int k = 0; k < nc; k++)
            {
                const int c = cnt[k];
                ps[k] = (uint16_t)acc;
                cnt[k] = (uint16_t)acc;
                acc += c;
            }
            ps[nc] = (uint16_t)acc;
        }
    }
}

#define TR_ROWS 8 // rows per work-group: one wavefront each, 8 consecutive lanes of a slice

// Rows of the symmetrised tangent, written straight into the sliced-ELL K and K32:
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
// One wavefront per row, TR_ROWS rows per work-group.  Lane a handles bond a; every
// bond's off-diagonal block lands on a distinct conn slot, and the two-hop rounds
// (bond a of i, lane t = bond t of j, read from j's contiguous records) land on
// distinct slots within a round and are added in bond order: the sums are
// deterministic.  The rows of a work-group are 8 consecutive lanes of one slice, so
// the work-group stores every slot entry of its rows as one 64-byte segment.  Row
// groups are mapped so that each XCD assembles a contiguous range of slices.
template <int D>
__global__ __launch_bounds__(64 * TR_ROWS) void k_tangent_rows(DV v, const double *__restrict__ S,
                                                               const TBond *__restrict__ rec,
                                                               const TBondOwn *__restrict__ own,
                                                               const uint16_t *__restrict__ pstart,
                                                               const uint16_t *__restrict__ plist)
{
    constexpr int D2 = D * D;
    __shared__ double acc[TR_ROWS][LPMX_MAXCONN][D2];
    __shared__ double dg[TR_ROWS][LPMX_MAXNB][D2];
    __shared__ double bn[TR_ROWS][LPMX_MAXNB][3], bdb[TR_ROWS][LPMX_MAXNB], btv[TR_ROWS][LPMX_MAXNB],
        bdbr[TR_ROWS][LPMX_MAXNB], btvr[TR_ROWS][LPMX_MAXNB];
    __shared__ int sc[TR_ROWS][LPMX_MAXCONN], bj[TR_ROWS][LPMX_MAXNB], bs[TR_ROWS][LPMX_MAXNB];
    __shared__ int rnc[TR_ROWS], rpii[TR_ROWS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = xcd_block() * TR_ROWS, i = row0 + w;
    const int64_t n = v.n;
    const bool live = i < v.n;
    const int si = row0 / LPMX_SLICE, li0 = row0 % LPMX_SLICE;
    const int64_t slot0 = v.sptr[si];
    const int width = v.sptr[si + 1] - (int)slot0;
    const int nc = live ? v.nb_conn[i] : 0, nbi = live ? v.nb0[i] : 0;
    // the rows' columns, 8 consecutive lanes of the slice per slot
    for (int t = threadIdx.x; t < width * TR_ROWS; t += 64 * TR_ROWS)
    {
        const int r8 = t % TR_ROWS, k = t / TR_ROWS;
        sc[r8][k] = v.col[(slot0 + k) * LPMX_SLICE + li0 + r8];
    }
    if (lane == 0)
        rnc[w] = nc, rpii[w] = live ? v.dslot[i] : 0;
    for (int e = lane; e < LPMX_MAXCONN * D2; e += 64)
        (&acc[w][0][0])[e] = 0.0;
    __syncthreads();
    // ---- bonds of i: D lanes per bond, one row r of the bond's blocks each
    for (int u = lane; u < D * nbi; u += 64)
    {
        const int ab = u / D, r = u % D;
        const TBond rb = rec[(int64_t)i * v.nb + ab];
        const TBondOwn ob = own[(int64_t)i * v.nb + ab];
        const int j = rb.l, s = rb.s;
        const double db = rb.db, tv = rb.tv, kn = ob.kn, dbr = rb.dbr, tvr = rb.tvr, knr = ob.knr;
        const double nv[3] = {rb.n[0], rb.n[1], rb.n[2]};
        const double d = ob.d;
        const double *Si = S + (int64_t)i * TS_NPLANE, *Sj = S + (int64_t)j * TS_NPLANE;
        const double Ai = Si[TS_A + s], Bi = Si[TS_B + s];
        const double Aj = Sj[TS_A + s], Bj = Sj[TS_B + s];
        const double dl = ob.dl, dlr = ob.dlr;
        const double F = db * (2.0 * kn * dl + 0.5 * (Bi + Bj) + 0.5 * tv * (Ai + Aj));
        const double Fr = dbr * (2.0 * knr * dlr + 0.5 * (Bj + Bi) + 0.5 * tvr * (Aj + Ai));
        double Vj[3], Wj[3];
        for (int c = 0; c < 3; c++)
            Vj[c] = Sj[TS_V + 3 * s + c], Wj[c] = Sj[TS_W + 3 * s + c];
        const double Vir = Si[TS_V + 3 * s + r], Wir = Si[TS_W + 3 * s + r];
        const double fd = F / d, fdr = Fr / d, k2 = 2.0 * kn * db * db, k2r = 2.0 * knr * dbr * dbr;
        const double w2 = 0.5 * db * dbr * (tvr + tv);
        const int pj = conn_find(sc[w], nc, j);
        for (int c = 0; c < D; c++)
        {
            const double nn = nv[r] * nv[c], del = r == c ? 1.0 : 0.0;
            const double Vic = Si[TS_V + 3 * s + c], Wic = Si[TS_W + 3 * s + c];
            // J_ij of this bond, and J_ji's transpose (n_ji = -n)
            const double jij = -(fd * (del - nn) + k2 * nn) + 0.5 * db * nv[r] * (Wj[c] + tv * Vj[c]) -
                               0.5 * (tv * Vir + Wir) * db * nv[c];
            const double jjiT = -(fdr * (del - nn) + k2r * nn) - 0.5 * dbr * (Wir + tvr * Vir) * nv[c] +
                                0.5 * dbr * nv[r] * (tvr * Vj[c] + Wj[c]);
            if (pj >= 0)
                acc[w][pj][r * D + c] = 0.5 * (jij + jjiT);
            dg[w][ab][r * D + c] = fd * (del - nn) + k2 * nn + 0.5 * db * nv[r] * (Wic + tv * Vic) + w2 * nn;
        }
        if (r == 0)
        {
            bj[w][ab] = j, bs[w][ab] = s;
            for (int c = 0; c < 3; c++)
                bn[w][ab][c] = nv[c];
            bdb[w][ab] = db, btv[w][ab] = tv, bdbr[w][ab] = dbr, btvr[w][ab] = tvr;
        }
    }
    __syncthreads();
    const int pii = rpii[w];
    if (lane < D2 && live)
    {
        double t = 0.0;
        for (int a = 0; a < nbi; a++)
            t += dg[w][a][lane];
        acc[w][pii][lane] += t;
    }
    __syncthreads();
    // ---- two-hop paths i -> j -> l: lane k sums the paths that land on slot k, in
    // their listed order, from j's contiguous records
    if (lane < nc)
    {
        double sum[D2];
        for (int e = 0; e < D2; e++)
            sum[e] = 0.0;
        const uint16_t *ps = pstart + (int64_t)i * (LPMX_MAXCONN + 1);
        const uint16_t *pl = plist + (int64_t)i * v.nb * v.nb;
        const int p1 = ps[lane + 1];
        for (int p = ps[lane]; p < p1; p++)
        {
            const int code = pl[p], a = code >> 5;
            const TBond *rt = rec + (int64_t)bj[w][a] * v.nb + (code & 31);
            const double alpha = -0.5 * bdb[w][a] * (rt->tv + btv[w][a]) * rt->db;
            const double beta = -0.5 * rt->dbr * (btvr[w][a] + rt->tvr) * bdbr[w][a];
            const double gam = 0.5 * (alpha + beta);
            for (int r = 0; r < D; r++)
                for (int c = 0; c < D; c++)
                    sum[r * D + c] += gam * bn[w][a][r] * rt->n[c];
        }
        for (int e = 0; e < D2; e++)
            acc[w][lane][e] += sum[e];
    }
    __syncthreads();
    // ---- the rows' blocks in slice order, 8 lanes of the slice per store; padding
    // slots hold zero blocks
    for (int u = threadIdx.x; u < width * D2 * TR_ROWS; u += 64 * TR_ROWS)
    {
        const int r8 = u % TR_ROWS, ke = u / TR_ROWS, k = ke / D2, e = ke % D2;
        if (row0 + r8 >= v.n)
            continue;
        const int r = e / D, c = e % D;
        double val = 0.0;
        if (k < rnc[r8])
            val = (k == rpii[r8] && r > c) ? acc[r8][k][c * D + r] : acc[r8][k][e];
        const int64_t o = ((slot0 + k) * D2 + e) * LPMX_SLICE + li0 + r8;
        v.K[o] = val;
        v.K32[o] = (float)val;
    }
}

// The reference's tangent is its upper CSR (stiffness.c:441-483); the solvers read
// both triangles here, so every block below the diagonal is made the exact
// transpose of its mirror above it (the row kernel's lower blocks agree with those
// only to round-off).  One thread per row.  The lower blocks written and the upper
// blocks read are disjoint, so K is passed as a read-only view of the upper blocks
// and a write-only view of the lower ones: a block's D*D loads are in flight together
// and need not wait for the previous block's stores.  K32 holds the FP32 rounding
// of K at every write, so the lower FP32 block is rounded from the copied values.
template <int D>
__global__ void k_tangent_mirror(int n, const int *__restrict__ nb_conn, const int *__restrict__ sptr,
                                 const int *__restrict__ col, const int *__restrict__ tpos,
                                 const double *__restrict__ Kup, double *__restrict__ Klo,
                                 float *__restrict__ K32lo)
{
    constexpr int D2 = D * D;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int s = i / LPMX_SLICE, l = i % LPMX_SLICE;
    const int nc = nb_conn[i];
    const int64_t s0 = sptr[s];
    for (int k = 0; k < nc; k++)
    {
        const int64_t e0 = (s0 + k) * LPMX_SLICE + l;
        if (col[e0] >= i)
            continue;
        const int64_t tp = tpos[e0];
        if (tp < 0)
            continue;
        const int64_t st = tp / LPMX_SLICE, lt = tp % LPMX_SLICE;
        double blk[D2];
#pragma unroll
        for (int r = 0; r < D; r++)
#pragma unroll
            for (int c = 0; c < D; c++)
                blk[r * D + c] = Kup[(st * D2 + c * D + r) * LPMX_SLICE + lt];
#pragma unroll
        for (int e = 0; e < D2; e++)
        {
            const int64_t o = ((s0 + k) * D2 + e) * LPMX_SLICE + l;
            Klo[o] = blk[e];
            K32lo[o] = (float)blk[e];
        }
    }
}

static int ensure_tangent_buffers(void)
{
    if (!t_S)
        HIPCHK(hipMalloc(&t_S, sizeof(double) * TS_NPLANE * (int64_t)g.n));
    const int64_t need = ((int64_t)sizeof(TBond) + (int64_t)sizeof(TBondOwn)) * g.nb * (int64_t)g.n;
    if (need > t_rec_bytes)
    {
        if (t_rec)
            hipFree(t_rec);
        t_rec = nullptr;
        t_rec_bytes = 0;
        HIPCHK(hipMalloc(&t_rec, need));
        t_rec_bytes = need;
    }
    const int64_t pneed = (int64_t)(LPMX_MAXCONN + 1 + g.nb * g.nb) * g.n;
    if (pneed > t_paths_cap)
    {
        if (t_pstart)
            hipFree(t_pstart);
        t_pstart = t_plist = nullptr;
        t_paths_cap = 0;
        HIPCHK(hipMalloc(&t_pstart, sizeof(uint16_t) * pneed));
        t_paths_cap = pneed;
        t_paths_version = -1;
    }
    t_plist = t_pstart + (int64_t)(LPMX_MAXCONN + 1) * g.n; // the paths follow this model's starts
    if (t_paths_version != lpmx_pattern_version())
    {
        hipLaunchKernelGGL(k_tangent_paths, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, t_pstart, t_plist);
        LAUNCHCHK();
        t_paths_version = lpmx_pattern_version();
    }
    return LPMX_OK;
}

void lpmx_tangent_free(void)
{
    if (t_S)
        hipFree(t_S);
    if (t_rec)
        hipFree(t_rec);
    if (t_pstart)
        hipFree(t_pstart);
    if (t_kstart)
        hipFree(t_kstart);
    t_S = nullptr;
    t_rec = nullptr;
    t_rec_bytes = 0;
    t_pstart = t_plist = nullptr;
    t_paths_cap = 0;
    t_paths_version = -1;
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
    TBond *rec = (TBond *)t_rec;
    TBondOwn *own = (TBondOwn *)(rec + (int64_t)g.nb * g.n);
    hipLaunchKernelGGL(k_tangent_sums, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, t_S, rec, own);
    LAUNCHCHK();
    const dim3 G(nblocks(g.n, TR_ROWS)), T(64 * TR_ROWS);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_tangent_rows<3>, G, T, 0, g_stream, g, t_S, rec, own, t_pstart, t_plist);
    else
        hipLaunchKernelGGL(k_tangent_rows<2>, G, T, 0, g_stream, g, t_S, rec, own, t_pstart, t_plist);
    LAUNCHCHK();
    if (g.dim == 3)
        hipLaunchKernelGGL(k_tangent_mirror<3>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g.n, g.nb_conn,
                           g.sptr, g.col, g.tpos, g.K, g.K, g.K32);
    else
        hipLaunchKernelGGL(k_tangent_mirror<2>, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g.n, g.nb_conn,
                           g.sptr, g.col, g.tpos, g.K, g.K, g.K32);
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
