/*
 * Oracle: rate-dependent crystal plasticity of the reference (plmode 1, Miehe-type
 * active-set return map with saturating hardening) for FCC lattices.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define O_EPS 1e-6
#define O_TOLITER 1e-4
#define O_MAXSMALL 20

static inline int64_t B(const orc_model *m, int i, int j) { return (int64_t)i * m->nbmax + j; }

/* FCC slip systems (reference src/initialization.c:565-638): 4 {111} planes, on each
 * 6 <110> directions (both senses of the 3 directions), in the reference's order,
 * which decides ties in the active-set search.  Plane normals and directions are
 * rotated by R (row-major 3x3, y = R x), then schmid[s] = (n0 d0, n1 d1, n2 d2,
 * n1 d2 + n2 d1, n0 d2 + n2 d0, n0 d1 + n1 d0).  Returns the number of systems. */
int orc_slip_systems_fcc(const double *R, double *schmid)
{
    static const int pl[4][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1}};
    static const int dr[4][6][3] = {
        {{1, -1, 0}, {-1, 1, 0}, {-1, 0, 1}, {1, 0, -1}, {0, 1, -1}, {0, -1, 1}},
        {{1, 0, 1}, {-1, 0, -1}, {-1, -1, 0}, {1, 1, 0}, {0, 1, -1}, {0, -1, 1}},
        {{-1, 0, 1}, {1, 0, -1}, {0, -1, -1}, {0, 1, 1}, {1, 1, 0}, {-1, -1, 0}},
        {{-1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {-1, 0, -1}, {0, -1, -1}, {0, 1, 1}}};
    const double a = 1 / sqrt(3), b = 1 / sqrt(2);
    int s = 0;
    for (int p = 0; p < 4; p++)
    {
        double nl[3], n[3];
        for (int c = 0; c < 3; c++)
            nl[c] = pl[p][c] * a;
        for (int r = 0; r < 3; r++)
            n[r] = R[3 * r] * nl[0] + R[3 * r + 1] * nl[1] + R[3 * r + 2] * nl[2];
        for (int k = 0; k < 6; k++, s++)
        {
            double dl[3], d[3];
            for (int c = 0; c < 3; c++)
                dl[c] = dr[p][k][c] * b;
            for (int r = 0; r < 3; r++)
                d[r] = R[3 * r] * dl[0] + R[3 * r + 1] * dl[1] + R[3 * r + 2] * dl[2];
            double *t = schmid + 6 * s;
            t[0] = n[0] * d[0];
            t[1] = n[1] * d[1];
            t[2] = n[2] * d[2];
            t[3] = n[1] * d[2] + n[2] * d[1];
            t[4] = n[0] * d[2] + n[2] * d[0];
            t[5] = n[0] * d[1] + n[1] * d[0];
        }
    }
    return s;
}

/* Schmid projection of the bond direction: cs.T.cs with engineering shear terms. */
static inline double proj(const double *cs, const double *t)
{
    return cs[0] * cs[0] * t[0] + cs[1] * cs[1] * t[1] + cs[2] * cs[2] * t[2] + cs[1] * cs[2] * t[3] +
           cs[0] * cs[2] * t[4] + cs[0] * cs[1] * t[5];
}

static double opp_flag(const orc_model *m, int i, int j)
{
    if (m->nb[i] == m->nbmax)
        return 0.5;
    const double *cj = m->cs0 + 3 * B(m, i, j);
    for (int k = 0; k < m->nb_initial[i]; k++)
    {
        const double *ck = m->cs0 + 3 * B(m, i, k);
        if (fabs(ck[0] + cj[0]) < O_EPS && fabs(ck[1] + cj[1]) < O_EPS && fabs(ck[2] + cj[2]) < O_EPS)
            return m->dmg_broken[B(m, i, k)] <= O_EPS ? 1.0 : 0.5;
    }
    return 1.0;
}

/* computeCab (reference src/constitutive.c:1864-1917): per particle, the slip
 * stiffness C_ab = -dRSS_a/dgamma_b of the bond model at the current geometry. */
void orc_compute_cab(orc_model *m)
{
    const int ns = m->nslip;
    for (int i = 0; i < m->n; i++)
        for (int a = 0; a < ns; a++)
            for (int bb = 0; bb < ns; bb++)
            {
                double Ls[2] = {0, 0};
                for (int j = 0; j < m->nb_initial[i]; j++)
                {
                    int64_t b = B(m, i, j);
                    Ls[m->nsign[b]] += m->Tv[b] * m->dist[b] * proj(m->cs + 3 * b, m->schmid + 6 * bb);
                }
                double *c = m->cp_Cab + ((int64_t)i * ns + a) * ns + bb;
                *c = 0.;
                for (int j = 0; j < m->nb_initial[i]; j++)
                {
                    int64_t b = B(m, i, j);
                    const double *cs = m->cs + 3 * b;
                    double dF = -2.0 * m->Kn[b] * m->dist[b] * proj(cs, m->schmid + 6 * bb) - 2.0 * Ls[m->nsign[b]];
                    double of = opp_flag(m, i, j);
                    *c += -of / m->volume * m->dist[b] * dF * proj(cs, m->schmid + 6 * a);
                }
            }
}

/* Dense solve a x = b (n x n, row-major), Gaussian elimination with partial pivoting
 * (first largest |a_ik| below the diagonal), whole-row swaps; on return a holds the
 * unit-lower multipliers below and U on and above the diagonal of the permuted
 * matrix (LAPACK dgetrf's layout), b holds x.  Returns k+1 if U(k,k) is 0, else 0. */
int orc_lu_solve(int n, double *a, double *b)
{
    int info = 0;
    int piv[64];
    for (int k = 0; k < n; k++)
    {
        int p = k;
        double amax = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(a[i * n + k]) > amax)
                amax = fabs(a[i * n + k]), p = i;
        piv[k] = p;
        if (a[p * n + k] == 0.0)
        {
            if (!info)
                info = k + 1;
            continue;
        }
        if (p != k)
            for (int j = 0; j < n; j++)
            {
                double t = a[k * n + j];
                a[k * n + j] = a[p * n + j];
                a[p * n + j] = t;
            }
        for (int i = k + 1; i < n; i++)
        {
            a[i * n + k] /= a[k * n + k];
            const double l = a[i * n + k];
            for (int j = k + 1; j < n; j++)
                a[i * n + j] -= l * a[k * n + j];
        }
    }
    if (info)
        return info;
    for (int k = 0; k < n; k++)
        if (piv[k] != k)
        {
            double t = b[k];
            b[k] = b[piv[k]];
            b[piv[k]] = t;
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++)
            b[i] -= a[i * n + j] * b[j];
    for (int i = n - 1; i >= 0; i--)
    {
        for (int j = i + 1; j < n; j++)
            b[i] -= a[i * n + j] * b[j];
        b[i] /= a[i * n + i];
    }
    return 0;
}

/* BLAS dnrm2 with the scaled sum of squares (the reference's cblas_dnrm2). */
static double nrm2(int n, const double *x)
{
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++)
        if (x[i] != 0.0)
        {
            double ax = fabs(x[i]);
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

/* Bond stretch pass of particle i with plastic stretch xdLp (constitutive.c:911-933). */
static void stretch_pass(orc_model *m, int i, const double *xdLp)
{
    double *A = m->dLtot + 2 * i, *Bt = m->TdLtot + 2 * i;
    A[0] = 0, Bt[0] = 0;
    A[1] = 0, Bt[1] = 0;
    for (int j = 0; j < m->nb_initial[i]; j++)
    {
        int64_t b = B(m, i, j);
        int q = m->nbr[b], s = m->nsign[b];
        double dx = m->x[3 * i] - m->x[3 * q], dy = m->x[3 * i + 1] - m->x[3 * q + 1],
               dz = m->x[3 * i + 2] - m->x[3 * q + 2];
        double dis = sqrt(dx * dx + dy * dy + dz * dz);
        m->dL[b] = dis - m->L0[b];
        m->dL[b] -= xdLp[j];
        m->dL[b] *= m->dmg_broken[b];
        A[s] += m->dL[b];
        Bt[s] += m->Tv[b] * m->dL[b];
        for (int c = 0; c < 3; c++)
            m->cs[3 * b + c] = (m->x[3 * i + c] - m->x[3 * q + c]) / dis;
    }
}

/* Local stress of particle i from stretches xdL and shell sums xA, xB. */
static void local_stress(const orc_model *m, int i, const double *xdL, const double *xA, const double *xB,
                         double *sl)
{
    for (int c = 0; c < 6; c++)
        sl[c] = 0.0;
    for (int j = 0; j < m->nb_initial[i]; j++)
    {
        int64_t b = B(m, i, j);
        int s = m->nsign[b];
        double Fij = 2.0 * m->Kn[b] * xdL[j] + xB[s] + m->Tv[b] * xA[s];
        Fij *= m->dmg_w[b];
        double of = opp_flag(m, i, j);
        const double cx = m->cs[3 * b], cy = m->cs[3 * b + 1], cz = m->cs[3 * b + 2], L0 = m->L0[b];
        sl[0] += of / m->volume * L0 * Fij * cx * cx;
        sl[1] += of / m->volume * L0 * Fij * cy * cy;
        sl[2] += of / m->volume * L0 * Fij * cz * cz;
        sl[3] += of / m->volume * L0 * Fij * cy * cz;
        sl[4] += of / m->volume * L0 * Fij * cx * cz;
        sl[5] += of / m->volume * L0 * Fij * cx * cy;
    }
}

/* Resolved shear stresses of particle i after plastic slips gamma on every system,
 * from the stretches and sums left in the model by a stretch pass (the stress path
 * of the return map's inner iteration, constitutive.c:1044-1109); for checks. */
void orc_cp_rss(orc_model *m, int i, const double *gamma, double *rss)
{
    const int ns = m->nslip, nbi = m->nb_initial[i];
    double dpl[6] = {0, 0, 0, 0, 0, 0}, xdL[64], sl[6], xA[2] = {0, 0}, xB[2] = {0, 0};
    for (int s = 0; s < ns; s++)
        for (int c = 0; c < 6; c++)
            dpl[c] += gamma[s] * m->schmid[6 * s + c];
    for (int j = 0; j < nbi; j++)
    {
        int64_t b = B(m, i, j);
        const double *cs = m->cs + 3 * b;
        double d = m->L0[b] * (dpl[0] * cs[0] * cs[0] + dpl[1] * cs[1] * cs[1] + dpl[2] * cs[2] * cs[2] +
                               dpl[3] * cs[1] * cs[2] + dpl[4] * cs[0] * cs[2] + dpl[5] * cs[0] * cs[1]);
        d *= m->dmg_broken[b];
        xdL[j] = m->dL[b] - d;
        xA[m->nsign[b]] += xdL[j];
        xB[m->nsign[b]] += m->Tv[b] * xdL[j];
    }
    local_stress(m, i, xdL, xA, xB, sl);
    for (int a = 0; a < ns; a++)
    {
        const double *t = m->schmid + 6 * a;
        rss[a] = sl[0] * t[0] + sl[1] * t[1] + sl[2] * t[2] + sl[3] * t[3] + sl[4] * t[4] + sl[5] * t[5];
    }
}

/* The return map of particle i (constitutive.c:944-1320) on the trial stretches in
 * m->dL / dLtot / TdLtot: writes ddLp, cp_dA, cp_dgy, cp_dA_single, cp_Jact,
 * cp_RSS and pl_flag of i. */
static void cp_return_map(orc_model *m, int i)
{
    const int ns = m->nslip, nbi = m->nb_initial[i];
    const double *sch = m->schmid;
    int *Jact = m->cp_Jact + (int64_t)i * ns;
    double *RSS = m->cp_RSS + (int64_t)i * ns, *dgy = m->cp_dgy + (int64_t)i * ns;
    double *dAs = m->cp_dA_single + (int64_t)i * ns;
    const double *Cab = m->cp_Cab + (int64_t)i * ns * ns;
    double dL[64], xdL[64], sl[6], yf[64], xgy[64];
    for (int j = 0; j < nbi; j++)
        dL[j] = m->dL[B(m, i, j)];
    local_stress(m, i, dL, m->dLtot + 2 * i, m->TdLtot + 2 * i, sl);
    double tmax = 0.0;
    for (int a = 0; a < ns; a++)
    {
        const double *t = sch + 6 * a;
        RSS[a] = sl[0] * t[0] + sl[1] * t[1] + sl[2] * t[2] + sl[3] * t[3] + sl[4] * t[4] + sl[5] * t[5];
        yf[a] = RSS[a] - m->cp_gy[((int64_t)i * ns + a) * 3 + 0];
        if (yf[a] > tmax)
            tmax = yf[a];
    }
    double gamma[64];
    memset(gamma, 0, sizeof(gamma));
    if (tmax <= O_EPS)
    {
        for (int j = 0; j < nbi; j++)
            m->ddLp[B(m, i, j)] = 0.0;
        m->cp_dA[i] = 0.0;
        for (int a = 0; a < ns; a++)
            Jact[a] = 0, dgy[a] = 0.0, dAs[a] = 0.0;
        return;
    }
    m->pl_flag[i] = 1;
    for (int a = 0; a < ns; a++)
        Jact[a] = 0;
    const double A0 = m->cp_A[3 * i + 0];
    const double tsd = m->cp_taus - m->cp_tau0;
    double *D = malloc(sizeof(double) * ns * ns);
    double r[64], rr[64];
    int nouter = 0;
    for (;;)
    {
        nouter++;
        double nr = 1.0;
        memset(gamma, 0, sizeof(double) * ns);
        memset(r, 0, sizeof(double) * ns);
        memset(rr, 0, sizeof(double) * ns);
        memset(D, 0, sizeof(double) * ns * ns);
        int ninner = 0;
        do
        {
            ninner++;
            double dpl[6] = {0, 0, 0, 0, 0, 0};
            for (int s = 0; s < ns; s++)
                for (int c = 0; c < 6; c++)
                    dpl[c] += Jact[s] * gamma[s] * sch[6 * s + c];
            double xA[2] = {0, 0}, xB[2] = {0, 0};
            for (int j = 0; j < nbi; j++)
            {
                int64_t b = B(m, i, j);
                const double *cs = m->cs + 3 * b;
                double d = m->L0[b] * (dpl[0] * cs[0] * cs[0] + dpl[1] * cs[1] * cs[1] + dpl[2] * cs[2] * cs[2] +
                                       dpl[3] * cs[1] * cs[2] + dpl[4] * cs[0] * cs[2] + dpl[5] * cs[0] * cs[1]);
                d *= m->dmg_broken[b];
                m->ddLp[b] = d;
                xdL[j] = dL[j] - d;
                xA[m->nsign[b]] += xdL[j];
                xB[m->nsign[b]] += m->Tv[b] * xdL[j];
            }
            local_stress(m, i, xdL, xA, xB, sl);
            double dA = 0.0;
            for (int s = 0; s < ns; s++)
                dA += gamma[s];
            m->cp_dA[i] = dA;
            const double xA_ = A0 + dA;
            const double h = m->cp_h0 / pow(cosh(m->cp_h0 * xA_ / tsd), 2.0);
            const double hp = -2.0 * m->cp_h0 * m->cp_h0 / tsd * tanh(m->cp_h0 * xA_ / tsd) * h;
            for (int a = 0; a < ns; a++)
            {
                double t1 = 0.0;
                for (int bb = 0; bb < ns; bb++)
                {
                    double hab = a == bb ? h : m->cp_q * h;
                    t1 += Jact[bb] * hab * gamma[bb];
                }
                dgy[a] = Jact[a] * t1;
                xgy[a] = m->cp_gy[((int64_t)i * ns + a) * 3 + 0] + dgy[a];
            }
            for (int a = 0; a < ns; a++)
            {
                const double t1 = pow(1. + gamma[a] * m->cp_eta / m->dtime, 1. / m->cp_p);
                const double *t = sch + 6 * a;
                RSS[a] = sl[0] * t[0] + sl[1] * t[1] + sl[2] * t[2] + sl[3] * t[3] + sl[4] * t[4] + sl[5] * t[5];
                r[a] = Jact[a] * (RSS[a] - xgy[a] * t1);
                rr[a] = r[a];
            }
            for (int a = 0; a < ns; a++)
                for (int bb = 0; bb < ns; bb++)
                {
                    if (Jact[a] == 1 && Jact[bb] == 1)
                    {
                        double hs = 0.0;
                        for (int d = 0; d < ns; d++)
                        {
                            double hd;
                            if (a == d && bb == d)
                                hd = h + hp * gamma[d];
                            else if (a == d && bb != d)
                                hd = hp * gamma[d];
                            else if (a != d && bb == d)
                                hd = m->cp_q * (h + hp * gamma[d]);
                            else
                                hd = m->cp_q * hp * gamma[d];
                            hs += Jact[d] * hd;
                        }
                        const double t1 = xgy[a] * (m->cp_eta / m->cp_p / m->dtime *
                                                    pow(1. + m->cp_eta * gamma[a] / m->dtime, (1. - m->cp_p) / m->cp_p));
                        const double t2 = hs * pow(1. + m->cp_eta * gamma[a] / m->dtime, (1. / m->cp_p));
                        if (a == bb)
                            D[a * ns + bb] = Cab[a * ns + bb] + t1 + t2;
                        else
                            D[a * ns + bb] = Cab[a * ns + bb] + t2;
                    }
                    else if (a == bb)
                        D[a * ns + bb] = 1.0;
                }
            /* the reference solves in place: D keeps the factors, entries not
             * rewritten above carry them into the next inner iteration */
            if (orc_lu_solve(ns, D, rr))
            {
                m->cp_fail++;
                free(D);
                return;
            }
            for (int a = 0; a < ns; a++)
                gamma[a] += Jact[a] * rr[a];
            nr = nrm2(ns, r);
        } while (nr > O_TOLITER && ninner < O_MAXSMALL);
        int minI = -1;
        double minY = 0.0;
        for (int a = 0; a < ns; a++)
        {
            yf[a] = RSS[a] - xgy[a];
            if (Jact[a] == 1 && gamma[a] <= 0.0 && yf[a] < minY)
                minY = yf[a], minI = a;
        }
        if (minI != -1)
        {
            Jact[minI] = 0;
            continue;
        }
        int maxI = -1;
        double maxY = 0.0;
        for (int a = 0; a < ns; a++)
            if (Jact[a] == 0 && yf[a] > 0.0 && yf[a] > maxY)
                maxY = yf[a], maxI = a;
        if (maxI != -1)
            Jact[maxI] = 1;
        if (maxI != -1 && nouter < m->cp_maxloop)
            continue;
        break;
    }
    for (int a = 0; a < ns; a++)
        dAs[a] = Jact[a] * gamma[a];
    free(D);
}

/* computeBondForceCPMiehe (reference src/constitutive.c:866-1396) for particle ii:
 * the list {ii} + unbroken partners, trial stretches from slot 0, return maps of
 * list members not yet done in this sweep (state_v), second stretch pass with the
 * updated plastic stretch, force of ii from the half-averaged stretch, slot 2 of
 * ii's state. */
void orc_bond_force_cp(orc_model *m, int ii)
{
    const int nbm = m->nbmax, ns = m->nslip;
    int lst[64];
    int nl = 1;
    lst[0] = ii;
    for (int k = 0; k < nbm; k++)
    {
        int64_t b = B(m, ii, k);
        if (m->dmg_broken[b] > O_EPS && m->nbr[b] != -1)
            lst[nl++] = m->nbr[b];
    }
    double *xdLp = malloc(sizeof(double) * nl * nbm);
    for (int k = 0; k < nl; k++)
        for (int j = 0; j < nbm; j++)
            xdLp[k * nbm + j] = m->dLp[3 * B(m, lst[k], j) + 0];
    for (int k = 0; k < nl; k++)
        stretch_pass(m, lst[k], xdLp + k * nbm);
    for (int k = 0; k < nl; k++)
    {
        int i = lst[k];
        if (!m->state_v[i])
        {
            m->state_v[i] = 1;
            cp_return_map(m, i);
        }
        for (int j = 0; j < m->nb_initial[i]; j++)
            xdLp[k * nbm + j] += m->ddLp[B(m, i, j)];
    }
    for (int k = 0; k < nl; k++)
        stretch_pass(m, lst[k], xdLp + k * nbm);
    int i = ii;
    double *P = m->Pin + 3 * i;
    P[0] = 0.0, P[1] = 0.0, P[2] = 0.0;
    for (int j = 0; j < m->nb_initial[i]; j++)
    {
        int64_t b = B(m, i, j);
        int q = m->nbr[b], s = m->nsign[b];
        for (int jj = 0; jj < nbm; jj++)
            if (m->nbr[B(m, q, jj)] == i)
                m->dL_ave[b] = 0.5 * (m->dL[b] + m->dL[B(m, q, jj)]);
        m->F[b] = 2.0 * m->Kn[b] * m->dL_ave[b] + 0.5 * (m->TdLtot[2 * i + s] + m->TdLtot[2 * q + s]) +
                  0.5 * m->Tv[b] * (m->dLtot[2 * i + s] + m->dLtot[2 * q + s]);
        m->F[b] *= m->dmg_w[b];
        P[0] += m->cs[3 * b + 0] * m->F[b];
        P[1] += m->cs[3 * b + 1] * m->F[b];
        P[2] += m->cs[3 * b + 2] * m->F[b];
    }
    for (int j = 0; j < nbm; j++)
        m->dLp[3 * B(m, i, j) + 2] = m->dmg_broken[B(m, i, j)] * xdLp[j];
    for (int a = 0; a < ns; a++)
    {
        int64_t o = (int64_t)i * ns + a;
        m->cp_gy[3 * o + 2] = m->cp_gy[3 * o + 0] + m->cp_dgy[o];
        m->cp_A_single[3 * o + 2] = m->cp_A_single[3 * o + 0] + m->cp_dA_single[o];
    }
    m->cp_A[3 * i + 2] = m->cp_A[3 * i + 0] + m->cp_dA[i];
    free(xdLp);
}
