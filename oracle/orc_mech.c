/*
 * Oracle: bond constants, constitutive laws, stress, tangent stiffness, residual
 * and displacement boundary conditions.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define O_EPS 1e-6
#define O_NDIM 3

static inline int64_t B(const orc_model *m, int i, int j) { return (int64_t)i * m->nbmax + j; }

/* per-particle modulus factor of the extension (1 when absent) */
static inline double emod(const orc_model *m, int i) { return m->emod ? m->emod[i] : 1.0; }

/* calcKnTv (reference src/stiffness.c:11-268): bond constants from C11/C12/C44 of
 * the particle types.  The 3x3 (2x2 for hex) mapping times Ce is written out; only
 * SC averages the constants over both end types.  With the per-particle modulus
 * extension each end's constants are scaled by its factor before averaging. */
void orc_kntv(orc_model *m)
{
    double *kt = malloc(sizeof(double) * 3 * m->ntype);
    for (int t = 0; t < m->ntype; t++)
    {
        const double c11 = m->Ce[3 * t], c12 = m->Ce[3 * t + 1], c44 = m->Ce[3 * t + 2];
        const double r = m->radius;
        if (m->lattice == 0)
        {
            kt[3 * t + 0] = 1 / 2.0 * c11 + -1 / 2.0 * c12 + 0.0 * c44;
            kt[3 * t + 1] = 0.0 * c11 + 0.0 * c12 + 1 / 2.0 * c44;
            kt[3 * t + 2] = 0.0 * c11 + 1.0 / 12.0 * c12 + -1.0 / 12.0 * c44;
        }
        else if (m->lattice == 1)
        {
            kt[3 * t + 0] = sqrt(3.0) / 12.0 * c11 + -sqrt(3.0) / 12.0 * c12;
            kt[3 * t + 1] = -sqrt(3.0) / 144.0 * c11 + sqrt(3.0) / 48.0 * c12;
            kt[3 * t + 2] = 0.0;
        }
        else if (m->lattice == 2)
        {
            kt[3 * t + 0] = r * (1 * c11 + -1 * c12 + -1 * c44);
            kt[3 * t + 1] = r * (0 * c11 + 0 * c12 + 1 * c44);
            kt[3 * t + 2] = r * (0 * c11 + 1.0 / 18.0 * c12 + -1.0 / 18.0 * c44);
        }
        else if (m->lattice == 3)
        {
            kt[3 * t + 0] = r * (0 * c11 + 0 * c12 + sqrt(2.0) * c44);
            kt[3 * t + 1] = r * (sqrt(2.0) / 4.0 * c11 + -sqrt(2.0) / 4.0 * c12 + -sqrt(2.0) / 4.0 * c44);
            kt[3 * t + 2] = r * (0 * c11 + sqrt(2.0) / 24.0 * c12 + -sqrt(2.0) / 24.0 * c44);
        }
        else
        {
            kt[3 * t + 0] = r * (0. * c11 + 0. * c12 + sqrt(3.0) * c44);
            kt[3 * t + 1] = r * (1. / sqrt(3.0) * c11 + -1. / sqrt(3.0) * c12 + 0. * c44);
            kt[3 * t + 2] = r * (0. * c11 + sqrt(3.0) / 14.0 * c12 + sqrt(3.0) / 14.0 * c44);
        }
    }
    for (int i = 0; i < m->n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int s = m->nsign[b], ti = m->type[i], tj = m->type[m->nbr[b]];
            double si = emod(m, i), sj = emod(m, m->nbr[b]);
            if (m->lattice == 1)
            {
                m->Kn[b] = si * kt[3 * ti + 0];
                m->Tv[b] = si * kt[3 * ti + 1];
            }
            else if (m->lattice == 2)
            {
                m->Kn[b] = 0.5 * (si * kt[3 * ti + s] + sj * kt[3 * tj + s]);
                m->Tv[b] = 0.5 * (si * kt[3 * ti + 2] + sj * kt[3 * tj + 2]);
            }
            else
            {
                m->Kn[b] = si * kt[3 * ti + s];
                m->Tv[b] = si * kt[3 * ti + 2];
            }
        }
    free(kt);
}

static inline double dist_ij(const double *x, int i, int j)
{
    double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* computedL (reference src/lpm_basic.c:252-291). */
void orc_computedL(orc_model *m)
{
    for (int i = 0; i < m->n; i++)
    {
        double *A = m->dLtot + 2 * i, *Bt = m->TdLtot + 2 * i;
        A[0] = A[1] = Bt[0] = Bt[1] = 0;
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int q = m->nbr[b], s = m->nsign[b];
            double dis = dist_ij(m->x, i, q);
            if (m->dist)
                m->dist[b] = dis;
            m->dL[b] = dis - m->L0[b];
            m->dL[b] -= m->dLp[3 * b + 0];
            A[s] += m->dL[b];
            Bt[s] += m->Tv[b] * m->dL[b];
            for (int c = 0; c < 3; c++)
                m->cs[3 * b + c] = (m->x[3 * i + c] - m->x[3 * q + c]) / dis;
        }
    }
}

/* Particle list {ii} followed by its unbroken bond partners
 * (constitutive.c:230-237, 468-475, 169-176). */
static int particle_list(const orc_model *m, int ii, int *lst)
{
    int s = 1;
    lst[0] = ii;
    for (int k = 1; k < m->nbmax + 1; k++)
    {
        int64_t b = B(m, ii, k - 1);
        if (m->dmg_broken[b] > O_EPS && m->nbr[b] != -1)
            lst[s++] = m->nbr[b];
    }
    return s;
}

/* Bond stretch pass over particle i with plastic stretch xdLp (constitutive.c:495-515). */
static void stretch_pass(orc_model *m, int i, const double *xdLp)
{
    double *A = m->dLtot + 2 * i, *Bt = m->TdLtot + 2 * i;
    A[0] = 0, Bt[0] = 0;
    A[1] = 0, Bt[1] = 0;
    for (int j = 0; j < m->nb_initial[i]; j++)
    {
        int64_t b = B(m, i, j);
        int q = m->nbr[b], s = m->nsign[b];
        double dis = dist_ij(m->x, i, q);
        m->dL[b] = dis - m->L0[b];
        m->dL[b] -= xdLp[j];
        m->dL[b] *= m->dmg_broken[b];
        A[s] += m->dL[b];
        Bt[s] += m->Tv[b] * m->dL[b];
        for (int c = 0; c < 3; c++)
            m->cs[3 * b + c] = (m->x[3 * i + c] - m->x[3 * q + c]) / dis;
    }
}

/* computeBondForceElastic (reference src/constitutive.c:228-283). */
void orc_bond_force_elastic(orc_model *m, int ii)
{
    int lst[64];
    double xd[64];
    int nl = particle_list(m, ii, lst);
    for (int k = 0; k < nl; k++)
    {
        int i = lst[k];
        for (int j = 0; j < m->nbmax; j++)
            xd[j] = m->dLp[3 * B(m, i, j) + 0];
        stretch_pass(m, i, xd);
    }
    int i = ii;
    double *P = m->Pin + 3 * i;
    P[0] = 0.0, P[1] = 0.0, P[2] = 0.0;
    for (int j = 0; j < m->nb_initial[i]; j++)
    {
        int64_t b = B(m, i, j);
        int q = m->nbr[b], s = m->nsign[b];
        double f = 2.0 * m->Kn[b] * m->dL[b] + 0.5 * (m->TdLtot[2 * i + s] + m->TdLtot[2 * q + s]) +
                   0.5 * m->Tv[b] * (m->dLtot[2 * i + s] + m->dLtot[2 * q + s]);
        f *= m->dmg_broken[b];
        m->F[b] = f;
        P[0] += m->cs[3 * b + 0] * m->F[b];
        P[1] += m->cs[3 * b + 1] * m->F[b];
        P[2] += m->cs[3 * b + 2] * m->F[b];
    }
}

/* Opposite-bond factor of the stress sum (constitutive.c:540-559, lpm_basic.c:71-90). */
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

/* computeBondForceJ2mixedLinear3D (reference src/constitutive.c:466-686): trial
 * stress from the particle's own sums, radial return with linear mixed hardening,
 * plastic bond stretch, second stretch pass, final force from the half-averaged
 * elastic stretch.  Side effects on the shared scratch follow the reference. */
void orc_bond_force_j2(orc_model *m, int ii)
{
    int lst[64];
    int nl = particle_list(m, ii, lst);
    const int nbm = m->nbmax;
    double *xdLp = malloc(sizeof(double) * nl * nbm);
    double xbeta[64][6], xalpha[64], xdlam[64];
    for (int k = 0; k < nl; k++)
    {
        int i = lst[k];
        for (int j = 0; j < nbm; j++)
            xdLp[k * nbm + j] = m->dLp[3 * B(m, i, j) + 0];
        for (int j = 0; j < 6; j++)
            xbeta[k][j] = m->beta[(6 * i + j) * 3 + 0];
        xalpha[k] = m->alpha[3 * i + 0];
        xdlam[k] = 0.0;
    }
    for (int k = 0; k < nl; k++)
        stretch_pass(m, lst[k], xdLp + k * nbm);

    for (int k = 0; k < nl; k++)
    {
        double sl[6] = {0, 0, 0, 0, 0, 0}, dep[6] = {0, 0, 0, 0, 0, 0};
        int i = lst[k];
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int s = m->nsign[b];
            double Fij = 2.0 * m->Kn[b] * m->dL[b] + m->TdLtot[2 * i + s] + m->Tv[b] * m->dLtot[2 * i + s];
            Fij *= m->dmg_w[b];
            double of = opp_flag(m, i, j);
            const double cx = m->cs[3 * b], cy = m->cs[3 * b + 1], cz = m->cs[3 * b + 2];
            sl[0] += of / m->volume * m->L0[b] * Fij * cx * cx;
            sl[1] += of / m->volume * m->L0[b] * Fij * cy * cy;
            sl[2] += of / m->volume * m->L0[b] * Fij * cz * cz;
            sl[3] += of / m->volume * m->L0[b] * Fij * cy * cz;
            sl[4] += of / m->volume * m->L0[b] * Fij * cx * cz;
            sl[5] += of / m->volume * m->L0[b] * Fij * cx * cy;
        }
        double tmp = 1.0 / 3.0 * (sl[0] + sl[1] + sl[2]);
        for (int j = 0; j < 3; j++)
            sl[j] -= tmp;
        for (int j = 0; j < 6; j++)
            sl[j] -= xbeta[k][j];
        double seq = 0.0;
        for (int j = 0; j < 6; j++)
        {
            if (j < 3)
                seq += sl[j] * sl[j];
            else
                seq += 2.0 * sl[j] * sl[j];
        }
        seq = sqrt(3.0 / 2.0 * seq);
        double yf = seq - (m->sigmay[i] + (1.0 - m->J2_xi) * m->J2_H * xalpha[k]);
        if (yf > 0.0)
        {
            m->pl_flag[i] = 1;
            double c44 = m->Ce[3 * m->type[i] + 2] * emod(m, i);
            xdlam[k] = yf / (3 * c44 + m->J2_H);
        }
        xalpha[k] += xdlam[k];
        for (int j = 0; j < 6; j++)
            if (fabs(seq) > O_EPS)
            {
                dep[j] = xdlam[k] * 1.5 * sl[j] / seq;
                xbeta[k][j] += 2. / 3. * m->J2_xi * m->J2_H * dep[j];
            }
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            const double cx = m->cs[3 * b], cy = m->cs[3 * b + 1], cz = m->cs[3 * b + 2];
            m->ddLp[b] = m->L0[b] * (dep[0] * cx * cx + dep[1] * cy * cy + dep[2] * cz * cz +
                                     2 * dep[3] * cy * cz + 2 * dep[4] * cx * cz + 2 * dep[5] * cx * cy);
            m->ddLp[b] *= m->dmg_broken[b];
            xdLp[k * nbm + j] += m->ddLp[b];
        }
    }
    for (int k = 0; k < nl; k++)
        stretch_pass(m, lst[k], xdLp + k * nbm);

    int i = ii;
    for (int c = 0; c < 6; c++)
        m->stress[6 * i + c] = 0.0;
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
    for (int j = 0; j < 6; j++)
        m->beta[(6 * i + j) * 3 + 2] = xbeta[0][j];
    m->alpha[3 * i + 2] = xalpha[0];
    m->dlambda[i] = xdlam[0];
    free(xdLp);
}

/* The energy-based J2 return map of one particle i of a list (reference
 * src/constitutive.c:336-398): trial distortional energy from the particle's own
 * stretches and shell sums (the first stretch pass), equivalent stress, yield test
 * against the slot-0 state, bisection for the multiplier on [0, 1] to TOLITER, then
 * the updated equivalent plastic strain and back stress and the plastic stretch
 * increment of each of the first nb[i] bonds, added to xdLp.  Reads only particle i's
 * arrays and its slot-0 state. */
#define ORC_TOLITER 1e-4 /* the reference's TOLITER (include/lpm.h:41) */
static void j2energy_update(orc_model *m, int i, int t, double *xdLp, double *xalpha, double *xbeta_eq,
                            double *xdlam)
{
    int nb1 = 0;
    for (int c = 0; c < m->nb1max; c++)
        if (m->nbr1[(int64_t)i * m->nb1max + c] != -1)
            nb1++;
    int nb2 = m->nb[i] - nb1;
    double U_d = 0.0, J2_k = 0.0, J2_sigma = 0.0;
    double J2_V = m->volume * m->nb[i] / m->nbmax;
    for (int j = 0; j < m->nb[i]; j++)
    {
        int64_t b = B(m, i, j);
        if (m->nsign[b] == 0)
        {
            U_d += 0.5 * m->Kn[b] * (m->dL[b] - m->dLtot[2 * i + 0] / nb1) * (m->dL[b] - m->dLtot[2 * i + 0] / nb1);
            J2_k = m->Kn[b];
        }
        else if (m->nsign[b] == 1)
            U_d += 0.5 * m->Kn[b] * (m->dL[b] - m->dLtot[2 * i + 1] / nb2) * (m->dL[b] - m->dLtot[2 * i + 1] / nb2);
    }
    const double c44 = m->Ce[3 * m->type[i] + 2] * emod(m, i);
    J2_sigma = sqrt(6.0 * c44 * U_d / J2_V);
    const double beq0 = m->beta_eq[3 * i + 0], al0 = m->alpha[3 * i + 0], H = m->J2_H, xi = m->J2_xi;
    double dlambda = 0.0;
    double yield_func = fabs(t * J2_sigma - beq0) - (m->sigmay[i] + (1.0 - xi) * H * al0);
    if (yield_func > 0.0)
    {
        m->pl_flag[i] = 1;
        double a = 0.0, bb = 1.0, ya = yield_func;
        while ((bb - a) > ORC_TOLITER)
        {
            dlambda = (a + bb) / 2.0;
            yield_func = fabs(t * J2_sigma / (1.0 + 0.5 * dlambda) -
                              (beq0 + t * xi * H * dlambda * J2_sigma * sqrt(m->radius / J2_k / c44) / 6.0 /
                                          (1.0 + 0.5 * dlambda))) -
                         (m->sigmay[i] + (1.0 - xi) * H * al0 +
                          (1.0 - xi) * H * dlambda * J2_sigma * sqrt(m->radius / J2_k / c44) / 6.0 /
                              (1.0 + 0.5 * dlambda));
            if (yield_func * ya < 0.0)
                bb = dlambda;
            else
            {
                a = dlambda;
                ya = yield_func;
            }
        }
    }
    *xdlam = dlambda;
    *xalpha += dlambda / (1.0 + 0.5 * dlambda) * sqrt(m->radius / 6.0 / J2_k * U_d / J2_V);
    *xbeta_eq = (*xbeta_eq + t * xi * H * dlambda / (1.0 + 0.5 * dlambda) * sqrt(m->radius / 6.0 / J2_k * U_d / J2_V));
    double f_d = 0.0;
    for (int j = 0; j < m->nb[i]; j++)
    {
        int64_t b = B(m, i, j);
        int s = m->nsign[b];
        if (s == 0)
            f_d = 2.0 * m->Kn[b] / (1 + dlambda / 2.0) * (m->dL[b] - m->dLtot[2 * i + s] / nb1);
        if (s == 1)
            f_d = 2.0 * m->Kn[b] / (1 + dlambda / 2.0) * (m->dL[b] - m->dLtot[2 * i + s] / nb2);
        m->ddLp[b] = dlambda * f_d / (4.0 * m->Kn[b]);
        m->ddLp[b] *= m->dmg_broken[b];
        xdLp[j] += m->ddLp[b];
    }
}

/* Final force of particle ii from the post-update stretches of itself and its
 * partners (reference src/constitutive.c:422-440): half-averaged elastic stretch,
 * shell sums of both ends, times 1 - damage_D of slot 0. */
static void j2energy_force(orc_model *m, int ii)
{
    const int i = ii, nbm = m->nbmax;
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
        m->F[b] *= (1.0 - m->dmgD[2 * b + 0]);
        P[0] += m->cs[3 * b + 0] * m->F[b];
        P[1] += m->cs[3 * b + 1] * m->F[b];
        P[2] += m->cs[3 * b + 2] * m->F[b];
    }
}

/* Slot 2 of particle ii (reference src/constitutive.c:442-447). */
static void j2energy_store(orc_model *m, int ii, const double *xdLp, double xalpha, double xbeta_eq, double xdlam)
{
    for (int j = 0; j < m->nbmax; j++)
        m->dLp[3 * B(m, ii, j) + 2] = m->dmg_broken[B(m, ii, j)] * xdLp[j];
    m->beta_eq[3 * ii + 2] = xbeta_eq;
    m->alpha[3 * ii + 2] = xalpha;
    m->dlambda[ii] = xdlam;
}

/* computeBondForceJ2energyReturnMap (reference src/constitutive.c:286-463), one
 * particle ii in the reference's order: its list {ii} and unbroken partners from
 * slot 0, first stretch pass of the list, the update of every list member, second
 * stretch pass, the force of ii, slot 2 of ii.  The shared scratch (dL, the shell
 * sums, the cosines, ddLp) is written for the whole list, as in the reference. */
static void j2energy_particle(orc_model *m, int ii, int t)
{
    int lst[64];
    int nl = particle_list(m, ii, lst);
    const int nbm = m->nbmax;
    double *xdLp = malloc(sizeof(double) * nl * nbm);
    double xalpha[64], xbeta_eq[64], xdlam[64];
    for (int k = 0; k < nl; k++)
    {
        int i = lst[k];
        for (int j = 0; j < nbm; j++)
            xdLp[k * nbm + j] = m->dLp[3 * B(m, i, j) + 0];
        xbeta_eq[k] = m->beta_eq[3 * i + 0];
        xalpha[k] = m->alpha[3 * i + 0];
        xdlam[k] = 0.0;
    }
    for (int k = 0; k < nl; k++)
        stretch_pass(m, lst[k], xdLp + k * nbm);
    for (int k = 0; k < nl; k++)
        j2energy_update(m, lst[k], t, xdLp + k * nbm, &xalpha[k], &xbeta_eq[k], &xdlam[k]);
    for (int k = 0; k < nl; k++)
        stretch_pass(m, lst[k], xdLp + k * nbm);
    j2energy_force(m, ii);
    j2energy_store(m, ii, xdLp, xalpha[0], xbeta_eq[0], xdlam[0]);
    free(xdLp);
}

void orc_bond_force_j2energy(orc_model *m, int t, int jacobi)
{
    const int n = m->n, nbm = m->nbmax;
    if (!jacobi)
        for (int ii = 0; ii < n; ii++)
            j2energy_particle(m, ii, t);
    else
    {
        double *xdLp = malloc(sizeof(double) * (size_t)n * nbm), *xa = malloc(sizeof(double) * n),
               *xb = malloc(sizeof(double) * n), *xl = malloc(sizeof(double) * n);
        for (int i = 0; i < n; i++)
        {
            for (int j = 0; j < nbm; j++)
                xdLp[(int64_t)i * nbm + j] = m->dLp[3 * B(m, i, j) + 0];
            xb[i] = m->beta_eq[3 * i + 0];
            xa[i] = m->alpha[3 * i + 0];
            xl[i] = 0.0;
        }
        for (int i = 0; i < n; i++)
            stretch_pass(m, i, xdLp + (int64_t)i * nbm);
        for (int i = 0; i < n; i++)
            j2energy_update(m, i, t, xdLp + (int64_t)i * nbm, &xa[i], &xb[i], &xl[i]);
        for (int i = 0; i < n; i++)
            stretch_pass(m, i, xdLp + (int64_t)i * nbm);
        for (int i = 0; i < n; i++)
        {
            j2energy_force(m, i);
            j2energy_store(m, i, xdLp + (int64_t)i * nbm, xa[i], xb[i], xl[i]);
        }
        free(xdLp), free(xa), free(xb), free(xl);
    }
    orc_compute_stress(m);
    orc_switch_state(m, 2);
}

/* The reference's isotropic hardening law SY (include/lpm.h:50), MPa. */
static inline double orc_SY(double x) { return 620.0 + 3300.0 * (1.0 - exp(-0.4 * x)); }

/* The nonlinear-isotropic J2 return map of one particle i (reference
 * src/constitutive.c:726-812): trial bond forces and stress of i from its first-pass
 * stretches, deviator minus the back stress, bisection of the multiplier on [0, 1] to
 * TOLITER, then the multiplier, the accumulated plastic strain, the back stress and
 * the plastic stretch of slot 0 advanced in place.  C is the kinematic modulus J2_C. */
static void j2nl_update(orc_model *m, int i, double C)
{
    const double J2_V = m->volume * m->nb[i] / m->nbmax;
    double sl[6] = {0, 0, 0, 0, 0, 0}, dpl[6];
    for (int j = 0; j < m->nb_initial[i]; j++)
    {
        int64_t b = B(m, i, j);
        int s = m->nsign[b];
        m->F[b] = 2.0 * m->Kn[b] * m->dL[b] + m->TdLtot[2 * i + s] + m->Tv[b] * m->dLtot[2 * i + s];
        m->F[b] *= m->dmg_broken[b];
        const double cx = m->cs[3 * b], cy = m->cs[3 * b + 1], cz = m->cs[3 * b + 2];
        sl[0] += 0.5 / J2_V * m->L0[b] * m->F[b] * cx * cx;
        sl[1] += 0.5 / J2_V * m->L0[b] * m->F[b] * cy * cy;
        sl[2] += 0.5 / J2_V * m->L0[b] * m->F[b] * cz * cz;
        sl[3] += 0.5 / J2_V * m->L0[b] * m->F[b] * cy * cz;
        sl[4] += 0.5 / J2_V * m->L0[b] * m->F[b] * cx * cz;
        sl[5] += 0.5 / J2_V * m->L0[b] * m->F[b] * cx * cy;
    }
    double temp = 1.0 / 3.0 * (sl[0] + sl[1] + sl[2]);
    for (int c = 0; c < 3; c++)
        sl[c] -= temp;
    for (int c = 0; c < 6; c++)
        sl[c] -= m->beta[(6 * i + c) * 3 + 0];
    double seq = 0.0;
    for (int c = 0; c < 6; c++)
    {
        if (c < 3)
            seq += sl[c] * sl[c];
        else
            seq += 2.0 * sl[c] * sl[c];
    }
    seq = sqrt(3.0 / 2.0 * seq);
    const double c44 = m->Ce[3 * m->type[i] + 2] * emod(m, i), al0 = m->alpha[3 * i + 0];
    double dlambda = 0.0;
    double yield_func = seq - orc_SY(al0);
    if (yield_func > 0.0)
    {
        double a = 0.0, bb = 1.0, ya = yield_func;
        while ((bb - a) > ORC_TOLITER)
        {
            dlambda = (a + bb) / 2.0;
            yield_func = seq - 1.5 * dlambda * (2.0 * c44 + C) - orc_SY(al0 + dlambda);
            if (yield_func * ya < 0.0)
                bb = dlambda;
            else
            {
                a = dlambda;
                ya = yield_func;
            }
        }
    }
    m->dlambda[i] = dlambda;
    m->alpha[3 * i + 0] += m->dlambda[i];
    for (int c = 0; c < 6; c++)
    {
        dpl[c] = dlambda * 1.5 * sl[c] / seq;
        m->beta[(6 * i + c) * 3 + 0] += C * dpl[c];
    }
    for (int j = 0; j < m->nb_initial[i]; j++)
    {
        int64_t b = B(m, i, j);
        const double cx = m->cs[3 * b], cy = m->cs[3 * b + 1], cz = m->cs[3 * b + 2];
        m->ddLp[b] = m->L0[b] * (dpl[0] * cx * cx + dpl[1] * cy * cy + dpl[2] * cz * cz + 2 * dpl[3] * cy * cz +
                                 2 * dpl[4] * cx * cz + 2 * dpl[5] * cx * cy);
        m->dLp[3 * b + 0] += m->ddLp[b];
    }
}

/* Stretch pass of particle i with its current slot-0 plastic stretch. */
static void j2nl_stretch(orc_model *m, int i)
{
    double xd[64];
    for (int j = 0; j < m->nbmax; j++)
        xd[j] = m->dLp[3 * B(m, i, j) + 0];
    stretch_pass(m, i, xd);
}

/* Final force of particle ii (reference src/constitutive.c:834-856). */
static void j2nl_force(orc_model *m, int ii)
{
    const int i = ii;
    for (int c = 0; c < 6; c++)
        m->stress[6 * i + c] = 0.0;
    double *P = m->Pin + 3 * i;
    P[0] = 0.0, P[1] = 0.0, P[2] = 0.0;
    for (int j = 0; j < m->nb_initial[i]; j++)
    {
        int64_t b = B(m, i, j);
        int q = m->nbr[b], s = m->nsign[b];
        for (int jj = 0; jj < m->nbmax; jj++)
            if (m->nbr[B(m, q, jj)] == i)
                m->dL_ave[b] = 0.5 * (m->dL[b] + m->dL[B(m, q, jj)]);
        m->F[b] = 2.0 * m->Kn[b] * m->dL_ave[b] + 0.5 * (m->TdLtot[2 * i + s] + m->TdLtot[2 * q + s]) +
                  0.5 * m->Tv[b] * (m->dLtot[2 * i + s] + m->dLtot[2 * q + s]);
        m->F[b] *= m->dmg_w[b];
        P[0] += m->cs[3 * b + 0] * m->F[b];
        P[1] += m->cs[3 * b + 1] * m->F[b];
        P[2] += m->cs[3 * b + 2] * m->F[b];
    }
}

void orc_bond_force_j2nonlinear(orc_model *m, double C, int jacobi)
{
    const int n = m->n;
    if (!jacobi)
        /* computeBondForceJ2nonlinearIso (reference src/constitutive.c:689-863) per
         * particle ii in order: the list's stretches from the current slot 0, every
         * list member's update in place, the list's stretches again, the force of ii */
        for (int ii = 0; ii < n; ii++)
        {
            int lst[64];
            int nl = particle_list(m, ii, lst);
            for (int k = 0; k < nl; k++)
                j2nl_stretch(m, lst[k]);
            for (int k = 0; k < nl; k++)
                j2nl_update(m, lst[k], C);
            for (int k = 0; k < nl; k++)
                j2nl_stretch(m, lst[k]);
            j2nl_force(m, ii);
        }
    else
    {
        for (int i = 0; i < n; i++)
            j2nl_stretch(m, i);
        for (int i = 0; i < n; i++)
            j2nl_update(m, i, C);
        for (int i = 0; i < n; i++)
            j2nl_stretch(m, i);
        for (int i = 0; i < n; i++)
            j2nl_force(m, i);
    }
    orc_compute_stress(m);
    orc_switch_state(m, 2);
}

/* computeBondForceIncrementalUpdating (reference src/constitutive.c:167-225): the
 * elastic predictor with stretch increments from xyz_temp; Pin uses the direction
 * cosines left by the previous force evaluation. */
void orc_bond_force_incremental(orc_model *m, int ii)
{
    int lst[64];
    int nl = particle_list(m, ii, lst);
    for (int k = 0; k < nl; k++)
    {
        int i = lst[k];
        double *A = m->ddLtot + 2 * i, *Bt = m->TddLtot + 2 * i;
        A[0] = 0, Bt[0] = 0;
        A[1] = 0, Bt[1] = 0;
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int q = m->nbr[b], s = m->nsign[b];
            double dis0 = dist_ij(m->xt, i, q);
            double dis1 = dist_ij(m->x, i, q);
            m->ddL[b] = m->dmg_broken[b] * (dis1 - dis0);
            A[s] += m->ddL[b];
            Bt[s] += m->Tv[b] * m->ddL[b];
        }
    }
    int i = ii;
    double *P = m->Pin + 3 * i;
    P[0] = 0.0, P[1] = 0.0, P[2] = 0.0;
    for (int j = 0; j < m->nb_initial[i]; j++)
    {
        int64_t b = B(m, i, j);
        int q = m->nbr[b], s = m->nsign[b];
        m->F[b] = m->Ftemp[b] + 2.0 * m->Kn[b] * m->ddL[b] + 0.5 * (m->TddLtot[2 * i + s] + m->TddLtot[2 * q + s]) +
                  0.5 * m->Tv[b] * (m->ddLtot[2 * i + s] + m->ddLtot[2 * q + s]);
        m->F[b] *= m->dmg_broken[b];
        P[0] += m->cs[3 * b + 0] * m->F[b];
        P[1] += m->cs[3 * b + 1] * m->F[b];
        P[2] += m->cs[3 * b + 2] * m->F[b];
    }
}

/* computeStress (reference src/lpm_basic.c:53-125). */
void orc_compute_stress(orc_model *m)
{
    for (int i = 0; i < m->n; i++)
    {
        double *st = m->stress + 6 * i;
        for (int c = 0; c < 6; c++)
            st[c] = 0.0;
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            double of = opp_flag(m, i, j);
            const double cx = m->cs[3 * b], cy = m->cs[3 * b + 1], cz = m->cs[3 * b + 2];
            st[0] += of / m->volume * m->L0[b] * m->F[b] * cx * cx;
            st[1] += of / m->volume * m->L0[b] * m->F[b] * cy * cy;
            st[2] += of / m->volume * m->L0[b] * m->F[b] * cz * cz;
            st[3] += of / m->volume * m->L0[b] * m->F[b] * cy * cz;
            st[4] += of / m->volume * m->L0[b] * m->F[b] * cx * cz;
            st[5] += of / m->volume * m->L0[b] * m->F[b] * cx * cy;
        }
        m->seq[i] = 0.0;
        for (int j = 0; j < 6; j++)
        {
            if (j < 3)
                m->seq[i] += st[j] * st[j];
            else
                m->seq[i] += 2.0 * st[j] * st[j];
        }
        m->seq[i] = sqrt(3.0 / 2.0 * m->seq[i]);
        m->sm[i] = 1.0 / 3.0 * (st[0] + st[1] + st[2]);
        if (m->seq[i] > O_EPS)
            m->triax[i] = m->sm[i] / m->seq[i];
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            const double cx = m->cs[3 * b], cy = m->cs[3 * b + 1], cz = m->cs[3 * b + 2];
            m->bond_stress[b] = (st[0] * cx * cx + st[1] * cy * cy + st[2] * cz * cz + 2 * st[3] * cy * cz +
                                 2 * st[4] * cx * cz + 2 * st[5] * cx * cy);
        }
    }
}

/* Gaussian elimination with partial pivoting (the first largest pivot in a column)
 * on the dense n x n row-major system A x = b: the algorithm of LAPACK's dgesv, which
 * computeStrain calls.  Returns 0, or the 1-based column of an exactly zero pivot,
 * in which case b is left as given (dgesv does not solve a singular system). */
static int orc_gauss(int n, double *A, double *b)
{
    double b0[6];
    for (int k = 0; k < n; k++)
        b0[k] = b[k];
    for (int k = 0; k < n; k++)
    {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i * n + k]) > fabs(A[p * n + k]))
                p = i;
        if (A[p * n + k] == 0.0)
        {
            for (int c = 0; c < n; c++)
                b[c] = b0[c];
            return k + 1;
        }
        if (p != k)
        {
            for (int j = 0; j < n; j++)
            {
                double t = A[k * n + j];
                A[k * n + j] = A[p * n + j];
                A[p * n + j] = t;
            }
            double t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
        for (int i = k + 1; i < n; i++)
        {
            double f = A[i * n + k] / A[k * n + k];
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

/* computeStrain (reference src/lpm_basic.c:127-249): per particle, the fourth- and
 * second-order bond sums with weights 0.1 (first shell) and 0.9 (second shell), the
 * normal equations of the weighted least-squares strain in the reference's packing
 * (:196-217), a solve, and the reference's component order (:236-246).  A singular
 * system stores its unsolved right-hand side, except at particle 0, which keeps its
 * previous strain (:228-234 followed by :236-246).  strain: [n][6], in and out. */
void orc_compute_strain(const orc_model *m, double *strain)
{
    const int d = m->dim, ns = 3 * (d - 1);
    for (int i = 0; i < m->n; i++)
    {
        double r4[3][3][3][3] = {{{{0.0}}}}, r2[3][3] = {{0.0}};
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            const int s = m->nsign[b];
            if (s != 0 && s != 1)
                continue;
            const double w = s == 0 ? 0.1 : 0.9, L = m->L0[b];
            const double *xq = m->x0 + 3 * m->nbr[b], *xi = m->x0 + 3 * i;
            for (int k = 0; k < d; k++)
                for (int n = 0; n < d; n++)
                    for (int a = 0; a < d; a++)
                        for (int l = 0; l < d; l++)
                            r4[k][n][a][l] += w * (xq[k] - xi[k]) / L * (xq[n] - xi[n]) / L * (xq[a] - xi[a]) / L *
                                              (xq[l] - xi[l]) / L;
            for (int k = 0; k < d; k++)
                for (int n = 0; n < d; n++)
                    r2[k][n] += w * m->dL[b] / L * (xq[k] - xi[k]) / L * (xq[n] - xi[n]) / L;
        }
        double A[36], rhs[6];
        int ii = 0;
        for (int j = 0; j < d; j++)
            for (int k = 0; k < d; k++)
                for (int a = 0; a < d; a++)
                    for (int l = 0; l < d; l++)
                        if (a <= l && j <= k)
                            A[ii++] = r4[a][l][j][k];
        ii = 0;
        for (int j = 0; j < d; j++)
            for (int k = 0; k < d; k++)
                if (j <= k)
                    rhs[ii++] = r2[j][k];
        if (orc_gauss(ns, A, rhs) > 0 && i == 0)
            continue;
        double *st = strain + 6 * i;
        if (d == 2)
            st[0] = rhs[0], st[5] = rhs[1], st[1] = rhs[2];
        else
            st[0] = rhs[0], st[5] = rhs[1], st[4] = rhs[2], st[1] = rhs[3], st[3] = rhs[4], st[2] = rhs[5];
    }
}

/* switchStateV (reference src/constitutive.c:10-85); crystal-plasticity slots are
 * not part of the oracle's laws. */
void orc_switch_state(orc_model *m, int flag)
{
    for (int i = 0; i < m->n; i++)
    {
        if (flag == 0 || flag == 1)
        {
            int d = flag == 0 ? 0 : 1, s = flag == 0 ? 1 : 0;
            for (int j = 0; j < m->nbmax; j++)
            {
                int64_t b = B(m, i, j);
                m->dLp[3 * b + d] = m->dLp[3 * b + s];
                m->dmgD[2 * b + d] = m->dmgD[2 * b + s];
            }
            for (int j = 0; j < 6; j++)
                m->beta[(6 * i + j) * 3 + d] = m->beta[(6 * i + j) * 3 + s];
            m->alpha[3 * i + d] = m->alpha[3 * i + s];
            m->beta_eq[3 * i + d] = m->beta_eq[3 * i + s];
            m->dloc[2 * i + d] = m->dloc[2 * i + s];
            m->dnl[2 * i + d] = m->dnl[2 * i + s];
            if (m->cp_gy)
            {
                for (int a = 0; a < m->nslip; a++)
                {
                    int64_t o = (int64_t)i * m->nslip + a;
                    m->cp_gy[3 * o + d] = m->cp_gy[3 * o + s];
                    m->cp_A_single[3 * o + d] = m->cp_A_single[3 * o + s];
                }
                m->cp_A[3 * i + d] = m->cp_A[3 * i + s];
            }
        }
        else
        {
            for (int j = 0; j < m->nbmax; j++)
            {
                int64_t b = B(m, i, j);
                m->dLp[3 * b + 0] = m->dLp[3 * b + 2];
            }
            for (int j = 0; j < 6; j++)
                m->beta[(6 * i + j) * 3 + 0] = m->beta[(6 * i + j) * 3 + 2];
            m->alpha[3 * i + 0] = m->alpha[3 * i + 2];
            m->beta_eq[3 * i + 0] = m->beta_eq[3 * i + 2];
            if (m->cp_gy)
            {
                for (int a = 0; a < m->nslip; a++)
                {
                    int64_t o = (int64_t)i * m->nslip + a;
                    m->cp_gy[3 * o + 0] = m->cp_gy[3 * o + 2];
                    m->cp_A_single[3 * o + 0] = m->cp_A_single[3 * o + 2];
                }
                m->cp_A[3 * i + 0] = m->cp_A[3 * i + 2];
            }
        }
    }
}

/* computeBondForceGeneral (reference src/constitutive.c:88-146): serial sweep in
 * particle order, then stress and switchStateV(2). */
void orc_bond_force_general(orc_model *m, int mode)
{
    /* plmode 3 and 5 (with their own stress and switch): load indicator +1; plmode 5 in
     * the reference's serial order, or in its one-update-per-particle form */
    if (mode == 3)
    {
        orc_bond_force_j2energy(m, 1, 0);
        return;
    }
    if (mode == 5)
    {
        orc_bond_force_j2nonlinear(m, m->J2_C, m->j2nl_jacobi);
        return;
    }
    if (mode == 1)
        memset(m->state_v, 0, sizeof(int) * m->n);
    for (int i = 0; i < m->n; i++)
    {
        if (mode == 0)
            orc_bond_force_j2(m, i);
        else if (mode == 1)
            orc_bond_force_cp(m, i);
        else if (mode == 4)
            orc_bond_force_incremental(m, i);
        else if (mode == 6)
            orc_bond_force_elastic(m, i);
    }
    orc_compute_stress(m);
    orc_switch_state(m, 2);
}

/* Elastic internal force of every particle at positions x (no side effects on the
 * model): the function whose Jacobian the tangent is. */
void orc_internal_force(orc_model *m, const double *x, double *pin)
{
    int n = m->n;
    double *A = calloc(2 * (size_t)n, sizeof(double)), *Bt = calloc(2 * (size_t)n, sizeof(double));
    double *dl = malloc(sizeof(double) * (size_t)n * m->nbmax);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int q = m->nbr[b], s = m->nsign[b];
            double dis = dist_ij(x, i, q);
            dl[b] = dis - m->L0[b];
            dl[b] -= m->dLp[3 * b];
            dl[b] *= m->dmg_broken[b];
            A[2 * i + s] += dl[b];
            Bt[2 * i + s] += m->Tv[b] * dl[b];
        }
    for (int i = 0; i < n; i++)
    {
        pin[3 * i] = pin[3 * i + 1] = pin[3 * i + 2] = 0;
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int q = m->nbr[b], s = m->nsign[b];
            double dis = dist_ij(x, i, q);
            double f = 2.0 * m->Kn[b] * dl[b] + 0.5 * (Bt[2 * i + s] + Bt[2 * q + s]) +
                       0.5 * m->Tv[b] * (A[2 * i + s] + A[2 * q + s]);
            f *= m->dmg_broken[b];
            for (int c = 0; c < 3; c++)
                pin[3 * i + c] += (x[3 * i + c] - x[3 * q + c]) / dis * f;
        }
    }
    free(A);
    free(Bt);
    free(dl);
}

/* ---------------------------------------------------------------------------------
 * Tangent assembly into the reference's upper CSR (stiffness.c:431-515 in 3D,
 * :318-380 in 2D).  Kb[r*dim+s] is the block dPin_i[r]/dx_m[s].
 * -------------------------------------------------------------------------------- */
static int64_t row_start(const orc_model *m, int i) { return m->kptr[2 * i + 1]; }
static int kcount(const orc_model *m, int i) { return m->kptr[2 * i]; }

static void assemble_block(orc_model *m, int i, int jj, int num1, const double *Kb)
{
    const int d = m->dim;
    int64_t p = row_start(m, i);
    int k = kcount(m, i);
    if (d == 3)
    {
        if (jj == i)
        {
            m->K[p] += Kb[0];
            m->K[p + 1] += Kb[1];
            m->K[p + 2] += Kb[2];
            m->K[p + 3 * k] += Kb[4];
            m->K[p + 3 * k + 1] += Kb[5];
            m->K[p + 6 * k - 1] += Kb[8];
        }
        else if (jj > i)
        {
            for (int s = 0; s < 3; s++)
            {
                m->K[p + 3 * num1 + s] += 0.5 * Kb[0 * 3 + s];
                m->K[p + 3 * k + 3 * num1 - 1 + s] += 0.5 * Kb[1 * 3 + s];
                m->K[p + 6 * k + 3 * num1 - 3 + s] += 0.5 * Kb[2 * 3 + s];
            }
        }
        else
        {
            int num2 = 0;
            int64_t pj = row_start(m, jj);
            int kj = kcount(m, jj);
            for (int c = 0; c < m->nb_conn[jj]; c++)
            {
                int cc = m->conn[(int64_t)jj * m->connmax + c];
                if (cc <= jj)
                    continue;
                num2++;
                if (cc == i)
                    for (int a = 0; a < 3; a++)
                    {
                        m->K[pj + 3 * num2 + a] += 0.5 * Kb[a * 3 + 0];
                        m->K[pj + 3 * kj + 3 * num2 - 1 + a] += 0.5 * Kb[a * 3 + 1];
                        m->K[pj + 6 * kj + 3 * num2 - 3 + a] += 0.5 * Kb[a * 3 + 2];
                    }
            }
        }
        if (jj == i)
        {
            m->JK[p] = 3 * (jj + 1) - 2;
            m->JK[p + 1] = 3 * (jj + 1) - 1;
            m->JK[p + 2] = 3 * (jj + 1);
            m->JK[p + 3 * k] = 3 * (jj + 1) - 1;
            m->JK[p + 3 * k + 1] = 3 * (jj + 1);
            m->JK[p + 6 * k - 1] = 3 * (jj + 1);
        }
        else if (jj > i)
            for (int s = 0; s < 3; s++)
            {
                m->JK[p + 3 * num1 + s] = 3 * (jj + 1) - 2 + s;
                m->JK[p + 3 * k + 3 * num1 - 1 + s] = 3 * (jj + 1) - 2 + s;
                m->JK[p + 6 * k + 3 * num1 - 3 + s] = 3 * (jj + 1) - 2 + s;
            }
    }
    else
    {
        if (jj == i)
        {
            m->K[p] += Kb[0];
            m->K[p + 1] += Kb[1];
            m->K[p + 2 * k] += Kb[3];
        }
        else if (jj > i)
        {
            m->K[p + 2 * num1] += 0.5 * Kb[0];
            m->K[p + 2 * num1 + 1] += 0.5 * Kb[1];
            m->K[p + 2 * k + 2 * num1 - 1] += 0.5 * Kb[2];
            m->K[p + 2 * k + 2 * num1] += 0.5 * Kb[3];
        }
        else
        {
            int num2 = 0;
            int64_t pj = row_start(m, jj);
            int kj = kcount(m, jj);
            for (int c = 0; c < m->nb_conn[jj]; c++)
            {
                int cc = m->conn[(int64_t)jj * m->connmax + c];
                if (cc <= jj)
                    continue;
                num2++;
                if (cc == i)
                {
                    m->K[pj + 2 * num2] += 0.5 * Kb[0];
                    m->K[pj + 2 * num2 + 1] += 0.5 * Kb[2];
                    m->K[pj + 2 * kj + 2 * num2 - 1] += 0.5 * Kb[1];
                    m->K[pj + 2 * kj + 2 * num2] += 0.5 * Kb[3];
                }
            }
        }
        if (jj == i)
        {
            m->JK[p] = 2 * (jj + 1) - 1;
            m->JK[p + 1] = 2 * (jj + 1);
            m->JK[p + 2 * k] = 2 * (jj + 1);
        }
        else if (jj > i)
        {
            m->JK[p + 2 * num1] = 2 * (jj + 1) - 1;
            m->JK[p + 2 * num1 + 1] = 2 * (jj + 1);
            m->JK[p + 2 * k + 2 * num1 - 1] = 2 * (jj + 1) - 1;
            m->JK[p + 2 * k + 2 * num1] = 2 * (jj + 1);
        }
    }
}

static void finish_ik(orc_model *m)
{
    const int d = m->dim;
    for (int i = 0; i < m->n; i++)
    {
        int64_t p = row_start(m, i);
        int k = kcount(m, i);
        m->IK[d * i] = (int)(p + 1);
        m->IK[d * i + 1] = (int)(p + d * k + 1);
        if (d == 3)
            m->IK[d * i + 2] = (int)(p + 2 * d * k);
    }
    m->IK[d * m->n] = (int)(m->kptr[2 * m->n + 1] + 1);
}

/* calcStiffness{3,2}DFiniteDifference (reference src/stiffness.c:271-516): forward
 * differences of Pin_i under a perturbation EPS*radius of each conn member, with
 * the elastic law, assembled as 0.5(A + A^T) into the upper CSR.  The perturbed
 * evaluations leave their side effects (Pin, cs, dL) exactly as the reference. */
void orc_tangent_fd(orc_model *m)
{
    const int d = m->dim;
    memset(m->K, 0, sizeof(double) * m->nK);
    double *Kl = malloc(sizeof(double) * 3 * m->connmax * 3);
    for (int i = 0; i < m->n; i++)
    {
        int nc = m->nb_conn[i];
        orc_bond_force_elastic(m, i);
        double f0[3] = {m->Pin[3 * i], m->Pin[3 * i + 1], m->Pin[3 * i + 2]};
        for (int jid = 0; jid < nc; jid++)
        {
            int mm = m->conn[(int64_t)i * m->connmax + jid];
            for (int r = 0; r < d; r++)
            {
                double xt = m->x[3 * mm + r];
                m->x[3 * mm + r] = xt + O_EPS * m->radius;
                orc_bond_force_elastic(m, i);
                m->x[3 * mm + r] = xt;
                for (int s = 0; s < d; s++)
                    Kl[d * nc * r + d * jid + s] = (m->Pin[3 * i + s] - f0[s]) / O_EPS / m->radius;
            }
        }
        int num1 = 0;
        for (int jid = 0; jid < nc; jid++)
        {
            int jj = m->conn[(int64_t)i * m->connmax + jid];
            double Kb[9];
            for (int r = 0; r < d; r++)
                for (int s = 0; s < d; s++)
                    Kb[r * d + s] = Kl[d * nc * s + d * jid + r];
            if (jj > i)
                num1++;
            assemble_block(m, i, jj, num1, Kb);
        }
    }
    finish_ik(m);
    free(Kl);
}

/* Position of particle q in the sorted conn row of i, or -1. */
static int conn_pos(const orc_model *m, int i, int q)
{
    const int *c = m->conn + (int64_t)i * m->connmax;
    int lo = 0, hi = m->nb_conn[i] - 1;
    while (lo <= hi)
    {
        int mid = (lo + hi) >> 1;
        if (c[mid] == q)
            return mid;
        if (c[mid] < q)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

/* Analytic Jacobian J_im = dPin_i/dx_m of the elastic law (the law the reference
 * differentiates numerically), one particle row at a time.  With g_ik = db_ik n_ik,
 * F_ij = db_ij [2Kn dL_ij + (B_is + B_js)/2 + Tv_ij (A_is + A_js)/2]:
 *   J_ii += F P/d + 2Kn db^2 n n^T,  J_ij -= the same         (P = I - n n^T)
 *   J_ii += w n g_ik^T, J_ik -= w n g_ik^T,  w = db_ij (Tv_ik + Tv_ij)/2, k in N_s(i)
 *   J_ij += w n g_jl^T, J_il -= w n g_jl^T,  w = db_ij (Tv_jl + Tv_ij)/2, l in N_s(j)
 * Jrow holds nb_conn[i] blocks of dim*dim in conn order. */
static void orc_jacobian_row(orc_model *m, int i, double *Jrow)
{
    const int d = m->dim;
    const int nc = m->nb_conn[i];
    memset(Jrow, 0, sizeof(double) * nc * d * d);
    /* bond stretches and sums of every particle in the 2-hop neighbourhood */
    for (int a = 0; a < m->nb_initial[i]; a++)
    {
        int64_t b = B(m, i, a);
        double db = m->dmg_broken[b];
        if (db == 0.0)
            continue;
        int j = m->nbr[b], s = m->nsign[b];
        double n[3], dij = dist_ij(m->x, i, j);
        for (int c = 0; c < 3; c++)
            n[c] = (m->x[3 * i + c] - m->x[3 * j + c]) / dij;
        /* sums A, B of i and j over shell s, with the current plastic stretch */
        double Ai = 0, Bi = 0, Aj = 0, Bj = 0, dl_ij = 0;
        for (int e = 0; e < m->nb_initial[i]; e++)
        {
            int64_t be = B(m, i, e);
            double dl = (dist_ij(m->x, i, m->nbr[be]) - m->L0[be] - m->dLp[3 * be]) * m->dmg_broken[be];
            if (e == a)
                dl_ij = dl;
            if (m->nsign[be] == s)
                Ai += dl, Bi += m->Tv[be] * dl;
        }
        for (int e = 0; e < m->nb_initial[j]; e++)
        {
            int64_t be = B(m, j, e);
            if (m->nsign[be] != s)
                continue;
            double dl = (dist_ij(m->x, j, m->nbr[be]) - m->L0[be] - m->dLp[3 * be]) * m->dmg_broken[be];
            Aj += dl, Bj += m->Tv[be] * dl;
        }
        double F = db * (2.0 * m->Kn[b] * dl_ij + 0.5 * (Bi + Bj) + 0.5 * m->Tv[b] * (Ai + Aj));
        int pi = conn_pos(m, i, i), pj = conn_pos(m, i, j);
        double G[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                G[3 * r + c] = F * ((r == c) - n[r] * n[c]) / dij + 2.0 * m->Kn[b] * db * db * n[r] * n[c];
        for (int r = 0; r < d; r++)
            for (int c = 0; c < d; c++)
            {
                Jrow[pi * d * d + r * d + c] += G[3 * r + c];
                Jrow[pj * d * d + r * d + c] -= G[3 * r + c];
            }
        for (int e = 0; e < m->nb_initial[i]; e++)
        {
            int64_t be = B(m, i, e);
            if (m->nsign[be] != s || m->dmg_broken[be] == 0.0)
                continue;
            int k = m->nbr[be];
            double dik = dist_ij(m->x, i, k), g[3];
            for (int c = 0; c < 3; c++)
                g[c] = m->dmg_broken[be] * (m->x[3 * i + c] - m->x[3 * k + c]) / dik;
            double w = 0.5 * db * (m->Tv[be] + m->Tv[b]);
            int pk = conn_pos(m, i, k);
            for (int r = 0; r < d; r++)
                for (int c = 0; c < d; c++)
                {
                    Jrow[pi * d * d + r * d + c] += w * n[r] * g[c];
                    Jrow[pk * d * d + r * d + c] -= w * n[r] * g[c];
                }
        }
        for (int e = 0; e < m->nb_initial[j]; e++)
        {
            int64_t be = B(m, j, e);
            if (m->nsign[be] != s || m->dmg_broken[be] == 0.0)
                continue;
            int l = m->nbr[be];
            double djl = dist_ij(m->x, j, l), g[3];
            for (int c = 0; c < 3; c++)
                g[c] = m->dmg_broken[be] * (m->x[3 * j + c] - m->x[3 * l + c]) / djl;
            double w = 0.5 * db * (m->Tv[be] + m->Tv[b]);
            int pl = conn_pos(m, i, l);
            for (int r = 0; r < d; r++)
                for (int c = 0; c < d; c++)
                {
                    Jrow[pj * d * d + r * d + c] += w * n[r] * g[c];
                    Jrow[pl * d * d + r * d + c] -= w * n[r] * g[c];
                }
        }
    }
}

/* The analytic counterpart of orc_tangent_fd: same pattern and symmetrisation,
 * exact derivative instead of forward differences, no side effects. */
void orc_tangent_analytic(orc_model *m)
{
    const int d = m->dim;
    memset(m->K, 0, sizeof(double) * m->nK);
    double *J = malloc(sizeof(double) * m->connmax * 9);
    for (int i = 0; i < m->n; i++)
    {
        orc_jacobian_row(m, i, J);
        int num1 = 0;
        for (int jid = 0; jid < m->nb_conn[i]; jid++)
        {
            int jj = m->conn[(int64_t)i * m->connmax + jid];
            if (jj > i)
                num1++;
            assemble_block(m, i, jj, num1, J + jid * d * d);
        }
    }
    finish_ik(m);
    free(J);
}

/* updateRR (reference src/stiffness.c:519-534). */
void orc_update_rr(orc_model *m)
{
    int ii = 0;
    const int d = m->dim;
    for (int i = 0; i < m->n; i++)
        for (int k = 0; k < d; k++)
        {
            m->residual[d * i + k] = m->dispBC[d * i + k] * (m->Pex[d * i + k] - m->Pin[3 * i + k]);
            if (m->dispBC[d * i + k] == 0)
                m->reaction[ii++] = m->Pin[3 * i + k];
        }
    m->nreaction = ii;
}

/* cblas_dnrm2 as the plain square root of the ordered sum of squares.  MKL's
 * internal scaling and blocking are closed source; the norms only feed tolerance
 * comparisons. */
double orc_nrm2(int n, const double *x)
{
    double s = 0.0;
    for (int i = 0; i < n; i++)
        s += x[i] * x[i];
    return sqrt(s);
}

/* setDispBC (reference src/boundary.c:12-45). */
void orc_set_disp_bc(orc_model *m, int nbd, const orc_dispbc *bc)
{
    const int d = m->dim;
    for (int i = 0; i < m->n; i++)
        for (int k = 0; k < nbd; k++)
            if (m->type[i] == bc[k].type)
            {
                int c = bc[k].flag == 'x' ? 0 : bc[k].flag == 'y' ? 1 : bc[k].flag == 'z' ? 2 : -1;
                if (c < 0)
                    continue;
                m->x[3 * i + c] += bc[k].step;
                m->dispBC[d * i + c] = 0;
            }
}

/* setForceBC (reference src/boundary.c:48-70). */
void orc_set_force_bc(orc_model *m, int nbf, const orc_forcebc *bc)
{
    const int d = m->dim;
    for (int k = 0; k < nbf; k++)
    {
        int cnt = 0;
        for (int i = 0; i < m->n; i++)
            if (m->type[i] == bc[k].type)
                cnt++;
        for (int i = 0; i < m->n; i++)
            if (m->type[i] == bc[k].type)
            {
                m->Pex[d * i] += bc[k].step[0] / cnt;
                m->Pex[d * i + 1] += bc[k].step[1] / cnt;
                if (d == 3)
                    m->Pex[d * i + 2] += bc[k].step[2] / cnt;
            }
    }
}

/* setDispBC_stiffnessUpdate{3,2}D (reference src/boundary.c:72-281): for every DOF
 * with dispBC_index==0 or fix_index==0, zero its row and column in the upper CSR,
 * put the 2-norm of the current diagonal on its diagonal, zero its residual. */
void orc_bc_stiffness_update(orc_model *m)
{
    const int d = m->dim;
    double *diag = malloc(sizeof(double) * d * m->n);
    for (int i = 0; i < m->n; i++)
    {
        int64_t p = row_start(m, i);
        int k = kcount(m, i);
        diag[d * i] = m->K[p];
        diag[d * i + 1] = m->K[p + d * k];
        if (d == 3)
            diag[d * i + 2] = m->K[p + 2 * d * k - 1];
    }
    double nd = orc_nrm2(d * m->n, diag);
    free(diag);
    orc_bc_stiffness_update_nd(m, nd);
}

/* The same elimination with a given diagonal value for the constrained DOFs. */
void orc_bc_stiffness_update_nd(orc_model *m, double nd)
{
    const int d = m->dim;
    for (int i = 0; i < m->n; i++)
    {
        int64_t p = row_start(m, i);
        int k = kcount(m, i);
        int64_t rs[3], re[3]; /* [start, end] of each of the d rows of particle i */
        rs[0] = p;
        rs[1] = p + d * k;
        rs[2] = p + 2 * d * k - 1;
        re[0] = rs[1] - 1;
        re[1] = d == 3 ? rs[2] - 1 : m->kptr[2 * (i + 1) + 1] - 1;
        re[2] = m->kptr[2 * (i + 1) + 1] - 1;
        for (int c = 0; c < d; c++)
        {
            if (!(m->dispBC[d * i + c] == 0 || m->fixidx[d * i + c] == 0))
                continue;
            /* column c inside the diagonal block (rows above c), then row c */
            if (d == 3)
            {
                if (c == 1)
                    m->K[p + 1] = 0.0;
                if (c == 2)
                {
                    m->K[p + 2] = 0.0;
                    m->K[p + d * k + 1] = 0.0;
                }
            }
            else if (c == 1)
                m->K[p + 1] = 0.0;
            for (int64_t kk = rs[c] + 1; kk <= re[c]; kk++)
                m->K[kk] = 0.0;
            m->K[rs[c]] = nd;
            /* column entries in the rows of lower-index conn members */
            for (int jc = 0; jc < m->nb_conn[i]; jc++)
            {
                int q = m->conn[(int64_t)i * m->connmax + jc];
                if (q >= i)
                    continue;
                int num2 = 0;
                int64_t pq = row_start(m, q);
                int kq = kcount(m, q);
                for (int c2 = 0; c2 < m->nb_conn[q]; c2++)
                {
                    int cc = m->conn[(int64_t)q * m->connmax + c2];
                    if (cc <= q)
                        continue;
                    num2++;
                    if (cc == i)
                    {
                        if (d == 3)
                        {
                            m->K[pq + 3 * num2 + c] = 0.0;
                            m->K[pq + 3 * kq + 3 * num2 - 1 + c] = 0.0;
                            m->K[pq + 6 * kq + 3 * num2 - 3 + c] = 0.0;
                        }
                        else
                        {
                            m->K[pq + 2 * num2 + c] = 0.0;
                            m->K[pq + 2 * kq + 2 * num2 - 1 + c] = 0.0;
                        }
                    }
                }
            }
            m->residual[d * i + c] = 0.0;
        }
    }
}
