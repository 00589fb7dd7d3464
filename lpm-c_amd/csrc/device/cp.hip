// Crystal plasticity of the reference (plmode 1, src/constitutive.c:866-1396): per
// particle, trial resolved shear stresses on the Schmid tensors, an active-set search
// with an inner Newton iteration on the slips (dense LU with partial pivoting on the
// nslip x nslip slip Jacobian), the plastic bond stretch, then the same final force
// pass as the J2 law.  One thread per particle; every per-particle array, the slip
// Jacobian included, is a plane [entry][particle], so the 64 lanes of a wavefront
// touch 64 consecutive particles' entry e at once.
//
// The reference factorises the full nslip x nslip slip Jacobian in place
// (LAPACKE_dgesv), with zeros in every row and column of an inactive system but its
// unit diagonal.  Partial pivoting never picks such a row (its entries in the active
// columns are zero), its multipliers are zero and its right-hand side entry is
// zero, so every elimination step on it subtracts zero products: the active
// systems' solution is the LU solve of the active submatrix, rows and columns kept
// in ascending system order, with the same operations in the same order.  This
// kernel factorises only that submatrix (na x na, na active systems) and sums the
// hardening matrix over the active systems only; the results are the same values,
// and the work per inner iteration drops from nslip^3 to na^3.
//
// Compiled without floating-point contraction, in the reference's evaluation order.
#include <math.h>
#include <string.h>

#include "common.h"

#define BI(j, i) ((int64_t)(j) * v.n + (i))
#define CPI(e, i) ((int64_t)(e) * v.n + (i))
#define CP_TOLITER 1e-4 // the reference's TOLITER (include/lpm.h:41)
#define CP_MAXSMALL 20  // the reference's MAXSMALL (include/lpm.h:38)

// pow() of the hardening and rate terms: the exponents 1, 2 and 0 that the
// linear-hardening, linear-rate settings produce are exact, as in a correctly
// rounded pow; other exponents go to the device pow.
__device__ __forceinline__ double cp_pow(double x, double y)
{
    if (y == 1.0)
        return x;
    if (y == 2.0)
        return x * x;
    if (y == 0.0)
        return 1.0;
    return pow(x, y);
}

__device__ __forceinline__ double proj(double cx, double cy, double cz, const double *t)
{
    return cx * cx * t[0] + cy * cy * t[1] + cz * cz * t[2] + cy * cz * t[3] + cx * cz * t[4] + cx * cy * t[5];
}

__device__ __forceinline__ double cp_opp(const DV &v, int i, int j)
{
    if (v.nbcur[i] == v.nb)
        return 0.5;
    const int o = v.opp[BI(j, i)];
    if (o < 0)
        return 1.0;
    return v.db[BI(o, i)] <= LPMX_EPS ? 1.0 : 0.5;
}

// computeCab (constitutive.c:1864-1917): C_ab of every particle at the current
// geometry (distance, cs), C[(a*ns+b)][i].
__global__ void k_cp_cab(DV v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int ns = v.nslip, nbi = v.nb0[i];
    for (int a = 0; a < ns; a++)
        for (int b = 0; b < ns; b++)
        {
            const double *tb = v.schmid + 6 * b, *ta = v.schmid + 6 * a;
            double Ls[2] = {0, 0};
            for (int j = 0; j < nbi; j++)
            {
                const int64_t q = BI(j, i);
                Ls[v.nsign[q]] += v.Tv[q] * v.dist[q] * proj(v.csx[q], v.csy[q], v.csz[q], tb);
            }
            double c = 0.;
            for (int j = 0; j < nbi; j++)
            {
                const int64_t q = BI(j, i);
                const double cx = v.csx[q], cy = v.csy[q], cz = v.csz[q];
                const double dF = -2.0 * v.Kn[q] * v.dist[q] * proj(cx, cy, cz, tb) - 2.0 * Ls[v.nsign[q]];
                c += -cp_opp(v, i, j) / v.volume * v.dist[q] * dF * proj(cx, cy, cz, ta);
            }
            v.cp_Cab[CPI(a * ns + b, i)] = c;
        }
}

// Local stress of particle i from elastic stretches xdl (per slot) and the shell
// sums xa, xb (constitutive.c:963-998).
__device__ void cp_stress(const DV &v, int i, const double *xdl, const double *xa, const double *xb, double *sl)
{
    for (int c = 0; c < 6; c++)
        sl[c] = 0.0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int s = v.nsign[b];
        double Fij = 2.0 * v.Kn[b] * xdl[j] + xb[s] + v.Tv[b] * xa[s];
        Fij *= v.dw[b];
        const double of = cp_opp(v, i, j);
        const double cx = v.csx[b], cy = v.csy[b], cz = v.csz[b], L0 = v.L0[b];
        sl[0] += of / v.volume * L0 * Fij * cx * cx;
        sl[1] += of / v.volume * L0 * Fij * cy * cy;
        sl[2] += of / v.volume * L0 * Fij * cz * cz;
        sl[3] += of / v.volume * L0 * Fij * cy * cz;
        sl[4] += of / v.volume * L0 * Fij * cx * cz;
        sl[5] += of / v.volume * L0 * Fij * cx * cy;
    }
}

// Solve D x = r in place for particle i (D an ns x ns matrix in planes, entry
// (a, b) at plane a*ns + b): Gaussian elimination with partial pivoting (first
// largest |d_ik| on or below the diagonal), whole-row swaps, multipliers stored
// below the diagonal.  Returns false on an exact zero pivot.
__device__ bool cp_lu_solve(const DV &v, int i, int ns, double *r)
{
    int piv[LPMX_MAXSLIP];
    double *D = v.cp_D;
    for (int k = 0; k < ns; k++)
    {
        int p = k;
        double amax = fabs(D[CPI(k * ns + k, i)]);
        for (int r2 = k + 1; r2 < ns; r2++)
        {
            const double t = fabs(D[CPI(r2 * ns + k, i)]);
            if (t > amax)
                amax = t, p = r2;
        }
        piv[k] = p;
        if (D[CPI(p * ns + k, i)] == 0.0)
            return false;
        if (p != k)
            for (int j = 0; j < ns; j++)
            {
                const double t = D[CPI(k * ns + j, i)];
                D[CPI(k * ns + j, i)] = D[CPI(p * ns + j, i)];
                D[CPI(p * ns + j, i)] = t;
            }
        const double dkk = D[CPI(k * ns + k, i)];
        for (int r2 = k + 1; r2 < ns; r2++)
        {
            const double l = D[CPI(r2 * ns + k, i)] / dkk;
            D[CPI(r2 * ns + k, i)] = l;
            // four columns at a time, so their loads are in flight together (one
            // memory round trip per four entries instead of per entry)
            for (int j0 = k + 1; j0 < ns; j0 += 4)
            {
                double x[4], y[4];
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (j0 + t < ns)
                        x[t] = D[CPI(r2 * ns + j0 + t, i)], y[t] = D[CPI(k * ns + j0 + t, i)];
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (j0 + t < ns)
                        D[CPI(r2 * ns + j0 + t, i)] = x[t] - l * y[t];
            }
        }
    }
    for (int k = 0; k < ns; k++)
        if (piv[k] != k)
        {
            const double t = r[k];
            r[k] = r[piv[k]];
            r[piv[k]] = t;
        }
    for (int a = 0; a < ns; a++)
        for (int j = 0; j < a; j++)
            r[a] -= D[CPI(a * ns + j, i)] * r[j];
    for (int a = ns - 1; a >= 0; a--)
    {
        for (int j = a + 1; j < ns; j++)
            r[a] -= D[CPI(a * ns + j, i)] * r[j];
        r[a] /= D[CPI(a * ns + a, i)];
    }
    return true;
}

// BLAS dnrm2 (scaled sum of squares), as the reference's cblas_dnrm2.
__device__ double cp_nrm2(int n, const double *x)
{
    double scale = 0.0, ssq = 1.0;
    for (int k = 0; k < n; k++)
        if (x[k] != 0.0)
        {
            const double ax = fabs(x[k]);
            if (scale < ax)
            {
                ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
                scale = ax;
            }
            else
                ssq += (ax / scale) * (ax / scale);
        }
    return scale * sqrt(ssq);
}

// One stretch pass of particle i with plastic stretch xp (plane) into dL, cs and
// the shell sums.
__device__ void cp_stretch(const DV &v, int i, const double *xp, double *a, double *bt)
{
    a[0] = 0, bt[0] = 0;
    a[1] = 0, bt[1] = 0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b];
        const double dx = v.x[3 * i] - v.x[3 * q], dy = v.x[3 * i + 1] - v.x[3 * q + 1],
                     dz = v.x[3 * i + 2] - v.x[3 * q + 2];
        const double dis = sqrt(dx * dx + dy * dy + dz * dz);
        double dl = dis - v.L0[b];
        dl -= xp[b];
        dl *= v.db[b];
        v.dL[b] = dl;
        a[v.nsign[b]] += dl;
        bt[v.nsign[b]] += v.Tv[b] * dl;
        v.csx[b] = (v.x[3 * i] - v.x[3 * q]) / dis;
        v.csy[b] = (v.x[3 * i + 1] - v.x[3 * q + 1]) / dis;
        v.csz[b] = (v.x[3 * i + 2] - v.x[3 * q + 2]) / dis;
    }
}

// The return map of each particle (constitutive.c:944-1320), then its second
// stretch pass with the updated plastic stretch and slot 2 of its state.  dLp2
// holds dLp0 + ddLp meanwhile, then damage_broken times it.
__global__ void k_cp_particle(DV v, int first, int count)
{
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= first + count)
        return;
    const int ns = v.nslip, nbi = v.nb0[i];
    double a[2], bt[2];
    cp_stretch(v, i, v.dLp0, a, bt);
    double dl[LPMX_MAXNB], xdl[LPMX_MAXNB], sl[6];
    for (int j = 0; j < nbi; j++)
        dl[j] = v.dL[BI(j, i)];
    cp_stress(v, i, dl, a, bt, sl);
    double rss[LPMX_MAXSLIP], yf[LPMX_MAXSLIP], xgy[LPMX_MAXSLIP], gam[LPMX_MAXSLIP], r[LPMX_MAXSLIP],
        rr[LPMX_MAXSLIP], dgy[LPMX_MAXSLIP];
    int Jact[LPMX_MAXSLIP], act[LPMX_MAXSLIP];
    double ract[LPMX_MAXSLIP];
    double tmax = 0.0;
    for (int s = 0; s < ns; s++)
    {
        const double *t = v.schmid + 6 * s;
        rss[s] = sl[0] * t[0] + sl[1] * t[1] + sl[2] * t[2] + sl[3] * t[3] + sl[4] * t[4] + sl[5] * t[5];
        yf[s] = rss[s] - v.cp_gy0[CPI(s, i)];
        if (yf[s] > tmax)
            tmax = yf[s];
        gam[s] = 0.0, dgy[s] = 0.0, Jact[s] = 0;
    }
    double dA = 0.0;
    if (tmax <= LPMX_EPS)
    {
        for (int j = 0; j < nbi; j++)
            v.ddLp[BI(j, i)] = 0.0;
    }
    else
    {
        v.pl_flag[i] = 1;
        const double A0 = v.cp_A0[i], tsd = v.cp_taus - v.cp_tau0;
        int nouter = 0;
        for (;;)
        {
            nouter++;
            double nr = 1.0;
            int na = 0; // the active systems of this pass, ascending
            for (int s = 0; s < ns; s++)
            {
                gam[s] = 0.0, r[s] = 0.0, rr[s] = 0.0;
                if (Jact[s] == 1)
                    act[na++] = s;
            }
            int ninner = 0;
            do
            {
                ninner++;
                // sums over the active systems: an inactive one adds Jact * ... = 0
                double dpl[6] = {0, 0, 0, 0, 0, 0};
                for (int a = 0; a < na; a++)
                {
                    const int s = act[a];
                    for (int c = 0; c < 6; c++)
                        dpl[c] += gam[s] * v.schmid[6 * s + c];
                }
                double xa[2] = {0, 0}, xb[2] = {0, 0};
                for (int j = 0; j < nbi; j++)
                {
                    const int64_t b = BI(j, i);
                    const double cx = v.csx[b], cy = v.csy[b], cz = v.csz[b];
                    double d = v.L0[b] * (dpl[0] * cx * cx + dpl[1] * cy * cy + dpl[2] * cz * cz + dpl[3] * cy * cz +
                                          dpl[4] * cx * cz + dpl[5] * cx * cy);
                    d *= v.db[b];
                    v.ddLp[b] = d;
                    xdl[j] = dl[j] - d;
                    xa[v.nsign[b]] += xdl[j];
                    xb[v.nsign[b]] += v.Tv[b] * xdl[j];
                }
                cp_stress(v, i, xdl, xa, xb, sl);
                dA = 0.0;
                for (int s = 0; s < ns; s++)
                    dA += gam[s];
                const double xA = A0 + dA;
                const double h = v.cp_h0 / cp_pow(cosh(v.cp_h0 * xA / tsd), 2.0);
                const double hp = -2.0 * v.cp_h0 * v.cp_h0 / tsd * tanh(v.cp_h0 * xA / tsd) * h;
                for (int s = 0; s < ns; s++)
                    dgy[s] = 0.0;
                for (int a = 0; a < na; a++)
                {
                    const int s = act[a];
                    double t1 = 0.0;
                    for (int c = 0; c < na; c++)
                    {
                        const int u = act[c];
                        const double hab = s == u ? h : v.cp_q * h;
                        t1 += hab * gam[u];
                    }
                    dgy[s] = t1;
                }
                for (int s = 0; s < ns; s++)
                    xgy[s] = v.cp_gy0[CPI(s, i)] + dgy[s];
                for (int s = 0; s < ns; s++)
                {
                    const double t1 = cp_pow(1. + gam[s] * v.cp_eta / v.dtime, 1. / v.cp_p);
                    const double *t = v.schmid + 6 * s;
                    rss[s] = sl[0] * t[0] + sl[1] * t[1] + sl[2] * t[2] + sl[3] * t[3] + sl[4] * t[4] + sl[5] * t[5];
                    r[s] = Jact[s] * (rss[s] - xgy[s] * t1);
                    rr[s] = r[s];
                }
                // the active submatrix of the slip Jacobian; the hardening sum over
                // the active systems (the inactive ones add Jact * hd = 0).  By
                // position c of system d = act[c] (act ascending, so d = m exactly when
                // c = a and d = u exactly when c = b): hd(m, u, d) is t0[c] but at c = a
                // and c = b.  t0 is a register array (unrolled, compile-time indices):
                // the sum does not chase act[c] and gam[d] through scratch memory.
                double t0[LPMX_MAXSLIP];
#pragma unroll
                for (int c = 0; c < LPMX_MAXSLIP; c++)
                    t0[c] = c < na ? v.cp_q * hp * gam[act[c]] : 0.0;
                for (int a = 0; a < na; a++)
                {
                    const int m = act[a];
                    const double t1 = xgy[m] * (v.cp_eta / v.cp_p / v.dtime *
                                                cp_pow(1. + v.cp_eta * gam[m] / v.dtime, (1. - v.cp_p) / v.cp_p));
                    const double pw = cp_pow(1. + v.cp_eta * gam[m] / v.dtime, (1. / v.cp_p));
                    for (int b = 0; b < na; b++)
                    {
                        const int u = act[b];
                        // hd at d = m (h + hp gam_m when m = u, else hp gam_m) and at
                        // d = u != m (q (h + hp gam_u)); the same expressions as the
                        // reference's four cases
                        const double hm = m == u ? h + hp * gam[m] : hp * gam[m], hu = v.cp_q * (h + hp * gam[u]);
                        double hs = 0.0;
#pragma unroll
                        for (int c = 0; c < LPMX_MAXSLIP; c++)
                            if (c < na)
                                hs += c == a ? hm : (c == b ? hu : t0[c]);
                        const double t2 = hs * pw;
                        const double cab = v.cp_Cab[CPI(m * ns + u, i)];
                        v.cp_D[CPI(a * na + b, i)] = m == u ? cab + t1 + t2 : cab + t2;
                    }
                    ract[a] = rr[m];
                }
                if (!cp_lu_solve(v, i, na, ract))
                {
                    atomicAdd(v.cp_fail, 1);
                    goto done;
                }
                for (int a = 0; a < na; a++)
                    rr[act[a]] = ract[a];
                for (int s = 0; s < ns; s++)
                    gam[s] += Jact[s] * rr[s];
                nr = cp_nrm2(ns, r);
            } while (nr > CP_TOLITER && ninner < CP_MAXSMALL);
            int minI = -1;
            double minY = 0.0;
            for (int s = 0; s < ns; s++)
            {
                yf[s] = rss[s] - xgy[s];
                if (Jact[s] == 1 && gam[s] <= 0.0 && yf[s] < minY)
                    minY = yf[s], minI = s;
            }
            if (minI != -1)
            {
                Jact[minI] = 0;
                continue;
            }
            int maxI = -1;
            double maxY = 0.0;
            for (int s = 0; s < ns; s++)
                if (Jact[s] == 0 && yf[s] > 0.0 && yf[s] > maxY)
                    maxY = yf[s], maxI = s;
            if (maxI != -1)
                Jact[maxI] = 1;
            if (maxI != -1 && nouter < v.cp_maxloop)
                continue;
            break;
        }
    }
done:
    for (int s = 0; s < ns; s++)
    {
        const double das = Jact[s] * gam[s];
        v.cp_Jact[CPI(s, i)] = Jact[s];
        v.cp_RSS[CPI(s, i)] = rss[s];
        v.cp_dgy[CPI(s, i)] = dgy[s];
        v.cp_dAs[CPI(s, i)] = das;
        v.cp_gy2[CPI(s, i)] = v.cp_gy0[CPI(s, i)] + dgy[s];
        v.cp_As2[CPI(s, i)] = v.cp_As0[CPI(s, i)] + das;
    }
    v.cp_dA[i] = dA;
    v.cp_A2[i] = v.cp_A0[i] + dA;
    // plastic stretch of the step, second stretch pass, slot 2
    for (int j = 0; j < v.nb; j++)
    {
        const int64_t b = BI(j, i);
        double xp = v.dLp0[b];
        if (j < nbi)
            xp += v.ddLp[b];
        v.dLp2[b] = xp;
    }
    cp_stretch(v, i, v.dLp2, a, bt);
    for (int j = 0; j < v.nb; j++)
    {
        const int64_t b = BI(j, i);
        v.dLp2[b] = v.db[b] * v.dLp2[b];
    }
    v.A[i] = a[0], v.A[(int64_t)v.n + i] = a[1];
    v.Bt[i] = bt[0], v.Bt[(int64_t)v.n + i] = bt[1];
}

static double *cp_alloc(int64_t count)
{
    double *p = nullptr;
    if (hipMalloc(&p, sizeof(double) * count) != hipSuccess)
        return nullptr;
    return p;
}

void lpmx_cp_free(void)
{
    void *ps[] = {g.schmid, g.cp_gy0, g.cp_gy1, g.cp_gy2, g.cp_As0, g.cp_As1, g.cp_As2, g.cp_A0, g.cp_A1,
                  g.cp_A2, g.cp_dA, g.cp_dgy, g.cp_dAs, g.cp_RSS, g.cp_Jact, g.cp_Cab, g.cp_D, g.cp_fail};
    for (void *p : ps)
        if (p)
            hipFree(p);
    g.schmid = g.cp_gy0 = g.cp_gy1 = g.cp_gy2 = g.cp_As0 = g.cp_As1 = g.cp_As2 = nullptr;
    g.cp_A0 = g.cp_A1 = g.cp_A2 = g.cp_dA = g.cp_dgy = g.cp_dAs = g.cp_RSS = g.cp_Cab = g.cp_D = nullptr;
    g.cp_Jact = g.cp_fail = nullptr;
    g.nslip = 0;
}

__global__ void k_cp_init(DV v, double tau0)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t m = (int64_t)v.nslip * v.n;
    if (t < m)
    {
        v.cp_gy0[t] = v.cp_gy1[t] = v.cp_gy2[t] = tau0;
        v.cp_As0[t] = v.cp_As1[t] = v.cp_As2[t] = 0.0;
        v.cp_dgy[t] = v.cp_dAs[t] = v.cp_RSS[t] = 0.0;
        v.cp_Jact[t] = 0;
    }
    if (t < v.n)
        v.cp_A0[t] = v.cp_A1[t] = v.cp_A2[t] = v.cp_dA[t] = 0.0;
}

extern "C" int lpmx_set_slip_systems(int nslip, const double *schmid, const lpmx_cp_params *p)
{
    RETCHK(lpmx_require_model());
    if (nslip <= 0 || nslip > LPMX_MAXSLIP || !schmid || !p)
        return lpmx_fail(LPMX_ERR_ARG, "slip systems: nslip %d outside 1..%d or null input", nslip, LPMX_MAXSLIP);
    if (!(p->dtime > 0.0) || !(p->cp_p > 0.0) || p->taus == p->tau0)
        return lpmx_fail(LPMX_ERR_ARG, "slip systems: dtime and p must be positive and taus != tau0");
    lpmx_cp_free();
    const int64_t n = g.n, m = (int64_t)nslip * n;
    double **planes[] = {&g.cp_gy0, &g.cp_gy1, &g.cp_gy2, &g.cp_As0, &g.cp_As1, &g.cp_As2, &g.cp_dgy, &g.cp_dAs,
                         &g.cp_RSS};
    for (double **q : planes)
        if (!(*q = cp_alloc(m)))
            return lpmx_fail(LPMX_ERR_HIP, "slip-system state allocation");
    double **vecs[] = {&g.cp_A0, &g.cp_A1, &g.cp_A2, &g.cp_dA};
    for (double **q : vecs)
        if (!(*q = cp_alloc(n)))
            return lpmx_fail(LPMX_ERR_HIP, "slip-system state allocation");
    if (!(g.cp_Cab = cp_alloc(m * nslip)) || !(g.cp_D = cp_alloc(m * nslip)) ||
        !(g.schmid = cp_alloc(6 * nslip)))
        return lpmx_fail(LPMX_ERR_HIP, "slip-system matrix allocation");
    HIPCHK(hipMalloc(&g.cp_Jact, sizeof(int) * m));
    HIPCHK(hipMalloc(&g.cp_fail, sizeof(int)));
    HIPCHK(hipMemsetAsync(g.cp_fail, 0, sizeof(int), g_stream));
    HIPCHK(hipMemsetAsync(g.cp_Cab, 0, sizeof(double) * m * nslip, g_stream));
    HIPCHK(hipMemcpyAsync(g.schmid, schmid, sizeof(double) * 6 * nslip, hipMemcpyHostToDevice, g_stream));
    g.nslip = nslip;
    g.cp_tau0 = p->tau0, g.cp_taus = p->taus, g.cp_h0 = p->h0, g.cp_q = p->q;
    g.cp_eta = p->eta, g.cp_p = p->cp_p, g.dtime = p->dtime, g.cp_maxloop = p->maxloop;
    hipLaunchKernelGGL(k_cp_init, dim3(nblocks(m > n ? m : n)), dim3(LPMX_BLOCK), 0, g_stream, g, p->tau0);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    return LPMX_OK;
}

extern "C" int lpmx_compute_cab(void)
{
    RETCHK(lpmx_require_model());
    if (g.nslip <= 0)
        return lpmx_fail(LPMX_ERR_STATE, "computeCab before the slip systems are set");
    hipLaunchKernelGGL(k_cp_cab, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g);
    LAUNCHCHK();
    return LPMX_OK;
}

// Return maps of all particles (plmode 1); the caller then runs the J2 force pass.
int lpmx_cp_particles(void)
{
    if (g.nslip <= 0)
        return lpmx_fail(LPMX_ERR_STATE, "plmode 1 before the slip systems are set");
    hipLaunchKernelGGL(k_cp_particle, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, 0, g.n);
    LAUNCHCHK();
    return LPMX_OK;
}

extern "C" int lpmx_cp_failures(int *count)
{
    RETCHK(lpmx_require_model());
    *count = 0;
    if (g.cp_fail)
    {
        HIPCHK(hipMemcpyAsync(count, g.cp_fail, sizeof(int), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    return LPMX_OK;
}
