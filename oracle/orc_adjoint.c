/*
 * Oracle: reverse mode (vector-Jacobian products) of one J2 force evaluation, and the
 * discrete adjoint of a load history built on it.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference has no derivatives; this is the checker of the device adjoint.  The
 * forward function differentiated is the J2 law of orc_bond_force_j2 (reference
 * src/constitutive.c:466-686): per particle, a trial stretch pass, the radial
 * return with linear mixed hardening, the plastic bond stretch and a second stretch
 * pass; per bond, the force from the half-averaged stretch and both ends' sums;
 * Pin_i = sum_j cs_ij F_ij.  Every particle's part is a pure function of the
 * positions and its slot-0 state, so the forward values are recomputed here per
 * particle.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define O_EPS 1e-6

static inline int64_t B(const orc_model *m, int i, int j) { return (int64_t)i * m->nbmax + j; }

/* per-type bond constants of orc_kntv (reference src/stiffness.c:11-268), kt[3t+c] */
static void kt_table(const orc_model *m, double *kt)
{
    for (int t = 0; t < m->ntype; t++)
    {
        const double c11 = m->Ce[3 * t], c12 = m->Ce[3 * t + 1], c44 = m->Ce[3 * t + 2], r = m->radius;
        double *k = kt + 3 * t;
        if (m->lattice == 0)
            k[0] = 0.5 * c11 - 0.5 * c12, k[1] = 0.5 * c44, k[2] = c12 / 12.0 - c44 / 12.0;
        else if (m->lattice == 1)
            k[0] = sqrt(3.0) / 12.0 * (c11 - c12), k[1] = -sqrt(3.0) / 144.0 * c11 + sqrt(3.0) / 48.0 * c12, k[2] = 0;
        else if (m->lattice == 2)
            k[0] = r * (c11 - c12 - c44), k[1] = r * c44, k[2] = r * (c12 - c44) / 18.0;
        else if (m->lattice == 3)
            k[0] = r * sqrt(2.0) * c44, k[1] = r * sqrt(2.0) / 4.0 * (c11 - c12 - c44), k[2] = r * sqrt(2.0) / 24.0 * (c12 - c44);
        else
            k[0] = r * sqrt(3.0) * c44, k[1] = r / sqrt(3.0) * (c11 - c12), k[2] = r * sqrt(3.0) / 14.0 * (c12 + c44);
    }
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

static inline void cc6(const double *c, double *o)
{
    o[0] = c[0] * c[0], o[1] = c[1] * c[1], o[2] = c[2] * c[2];
    o[3] = c[1] * c[2], o[4] = c[0] * c[2], o[5] = c[0] * c[1];
}

/* adjoint of cc6: cb[k] += sum_c ob[c] d o_c / d c_k */
static inline void cc6_bar(const double *c, const double *ob, double *cb)
{
    cb[0] += 2 * c[0] * ob[0] + c[2] * ob[4] + c[1] * ob[5];
    cb[1] += 2 * c[1] * ob[1] + c[2] * ob[3] + c[0] * ob[5];
    cb[2] += 2 * c[2] * ob[2] + c[1] * ob[3] + c[0] * ob[4];
}

/* Forward values of one particle's J2 evaluation that the reverse sweep needs. */
typedef struct
{
    double dis[64], cs[64][3], dL1[64], dL2[64], Ftr[64], A1[2], B1[2], A2[2], B2[2];
    double sdev[6], seq, yf, dlam, dep[6], c44;
    int plastic;
} part_fwd;

static void fwd_particle(const orc_model *m, int i, part_fwd *p)
{
    const int nbi = m->nb_initial[i];
    memset(p, 0, sizeof(*p));
    for (int j = 0; j < nbi; j++)
    {
        int64_t b = B(m, i, j);
        int q = m->nbr[b], s = m->nsign[b];
        double d[3];
        for (int c = 0; c < 3; c++)
            d[c] = m->x[3 * i + c] - m->x[3 * q + c];
        p->dis[j] = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int c = 0; c < 3; c++)
            p->cs[j][c] = d[c] / p->dis[j];
        p->dL1[j] = (p->dis[j] - m->L0[b] - m->dLp[3 * b + 0]) * m->dmg_broken[b];
        p->A1[s] += p->dL1[j];
        p->B1[s] += m->Tv[b] * p->dL1[j];
    }
    double sl[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nbi; j++)
    {
        int64_t b = B(m, i, j);
        int s = m->nsign[b];
        p->Ftr[j] = (2.0 * m->Kn[b] * p->dL1[j] + p->B1[s] + m->Tv[b] * p->A1[s]) * m->dmg_w[b];
        double cc[6];
        cc6(p->cs[j], cc);
        const double w = opp_flag(m, i, j) / m->volume * m->L0[b] * p->Ftr[j];
        for (int c = 0; c < 6; c++)
            sl[c] += w * cc[c];
    }
    const double mean = (sl[0] + sl[1] + sl[2]) / 3.0;
    for (int c = 0; c < 6; c++)
        p->sdev[c] = (c < 3 ? sl[c] - mean : sl[c]) - m->beta[(6 * i + c) * 3 + 0];
    double q2 = 0.0;
    for (int c = 0; c < 6; c++)
        q2 += (c < 3 ? 1.0 : 2.0) * p->sdev[c] * p->sdev[c];
    p->seq = sqrt(1.5 * q2);
    const double H = m->J2_H, xi = m->J2_xi;
    p->c44 = m->Ce[3 * m->type[i] + 2] * (m->emod ? m->emod[i] : 1.0);
    p->yf = p->seq - (m->sigmay[i] + (1.0 - xi) * H * m->alpha[3 * i + 0]);
    p->plastic = p->yf > 0.0;
    p->dlam = p->plastic ? p->yf / (3 * p->c44 + H) : 0.0;
    for (int c = 0; c < 6; c++)
        p->dep[c] = fabs(p->seq) > O_EPS ? p->dlam * 1.5 * p->sdev[c] / p->seq : 0.0;
    for (int j = 0; j < nbi; j++)
    {
        int64_t b = B(m, i, j);
        int s = m->nsign[b];
        double cc[6];
        cc6(p->cs[j], cc);
        double dd = 0.0;
        for (int c = 0; c < 6; c++)
            dd += (c < 3 ? 1.0 : 2.0) * p->dep[c] * cc[c];
        const double ddLp = m->L0[b] * dd * m->dmg_broken[b];
        p->dL2[j] = (p->dis[j] - m->L0[b] - (m->dLp[3 * b + 0] + ddLp)) * m->dmg_broken[b];
        p->A2[s] += p->dL2[j];
        p->B2[s] += m->Tv[b] * p->dL2[j];
    }
}

void orc_eval_j2(orc_model *m)
{
    for (int i = 0; i < m->n; i++)
        orc_bond_force_j2(m, i);
}

static int rev_slot(const orc_model *m, int i, int j)
{
    const int q = m->nbr[B(m, i, j)];
    for (int jj = 0; jj < m->nbmax; jj++)
        if (m->nbr[B(m, q, jj)] == i)
            return jj;
    return -1;
}

void orc_vjp_j2(orc_model *m, const double *Pb, const double *dLp2b, const double *beta2b, const double *alpha2b,
                double *xb, double *dLp0b, double *beta0b, double *alpha0b, double *eb, double *syb, double *Hb)
{
    const int n = m->n, nbm = m->nbmax;
    part_fwd *P = malloc(sizeof(part_fwd) * n);
    for (int i = 0; i < n; i++)
        fwd_particle(m, i, P + i);
    double *Fb = calloc((size_t)n * nbm, sizeof(double));     /* adjoint of F_ij */
    double *gb = calloc((size_t)n * nbm * 3, sizeof(double)); /* adjoint of x_i - x_q per bond */
    double *Knb = calloc((size_t)n * nbm, sizeof(double)), *Tvb = calloc((size_t)n * nbm, sizeof(double));
    double *c44b = calloc(n, sizeof(double));
    double Hacc = 0.0;
    memset(xb, 0, sizeof(double) * 3 * n);
    memset(dLp0b, 0, sizeof(double) * n * nbm);
    memset(beta0b, 0, sizeof(double) * 6 * n);
    memset(alpha0b, 0, sizeof(double) * n);
    memset(eb, 0, sizeof(double) * n);
    memset(syb, 0, sizeof(double) * n);

    /* force pass: Pin_i = sum_j cs_ij F_ij */
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int q = m->nbr[b], s = m->nsign[b], r = rev_slot(m, i, j);
            const part_fwd *pi = P + i, *pq = P + q;
            const double dla = r >= 0 ? 0.5 * (pi->dL2[j] + pq->dL2[r]) : m->dL_ave[b];
            const double F = (2.0 * m->Kn[b] * dla + 0.5 * (pi->B2[s] + pq->B2[s]) +
                              0.5 * m->Tv[b] * (pi->A2[s] + pq->A2[s])) * m->dmg_w[b];
            const double *cs = pi->cs[j];
            double csb[3];
            Fb[b] = 0.0;
            for (int c = 0; c < 3; c++)
                Fb[b] += Pb[3 * i + c] * cs[c], csb[c] = Pb[3 * i + c] * F;
            const double dot = cs[0] * csb[0] + cs[1] * csb[1] + cs[2] * csb[2];
            for (int c = 0; c < 3; c++)
                gb[3 * b + c] += (csb[c] - dot * cs[c]) / pi->dis[j];
            Knb[b] += Fb[b] * m->dmg_w[b] * 2.0 * dla;
            Tvb[b] += Fb[b] * m->dmg_w[b] * 0.5 * (pi->A2[s] + pq->A2[s]);
        }

    /* per particle: adjoints of its second-pass stretches and sums, gathered over its
     * own bonds and the reverse bonds, then the reverse of its return map */
    for (int i = 0; i < n; i++)
    {
        const part_fwd *p = P + i;
        const int nbi = m->nb_initial[i];
        double dL2b[64], A2b[2] = {0, 0}, B2b[2] = {0, 0};
        for (int j = 0; j < nbi; j++)
        {
            int64_t b = B(m, i, j);
            int q = m->nbr[b], s = m->nsign[b], r = rev_slot(m, i, j);
            double t = Fb[b] * m->dmg_w[b];
            dL2b[j] = r >= 0 ? t * m->Kn[b] : 0.0;
            A2b[s] += t * 0.5 * m->Tv[b];
            B2b[s] += t * 0.5;
            if (r >= 0)
            {
                int64_t br = B(m, q, r);
                double tr = Fb[br] * m->dmg_w[br];
                dL2b[j] += tr * m->Kn[br];
                A2b[s] += tr * 0.5 * m->Tv[br];
                B2b[s] += tr * 0.5;
            }
        }
        /* second pass and the plastic stretch of slot 2 */
        double depb[6] = {0, 0, 0, 0, 0, 0}, disb[64], csbar[64][3];
        for (int j = 0; j < nbi; j++)
        {
            int64_t b = B(m, i, j);
            int s = m->nsign[b];
            const double db = m->dmg_broken[b];
            double tot = dL2b[j] + A2b[s] + B2b[s] * m->Tv[b];
            Tvb[b] += B2b[s] * p->dL2[j];
            disb[j] = tot * db;
            double xpb = -tot * db + dLp2b[B(m, i, j)] * db; /* xp = dLp0 + ddLp */
            dLp0b[B(m, i, j)] += xpb;
            /* ddLp = L0 db sum_c w_c dep_c cc_c */
            double cc[6], ccb[6];
            cc6(p->cs[j], cc);
            for (int c = 0; c < 6; c++)
            {
                const double wc = c < 3 ? 1.0 : 2.0;
                depb[c] += xpb * m->L0[b] * db * wc * cc[c];
                ccb[c] = xpb * m->L0[b] * db * wc * p->dep[c];
            }
            csbar[j][0] = csbar[j][1] = csbar[j][2] = 0.0;
            cc6_bar(p->cs[j], ccb, csbar[j]);
        }
        for (int j = nbi; j < nbm; j++)
            dLp0b[B(m, i, j)] += dLp2b[B(m, i, j)] * m->dmg_broken[B(m, i, j)];
        /* state of slot 2 */
        const double H = m->J2_H, xi = m->J2_xi;
        double dlamb = alpha2b[i];
        alpha0b[i] += alpha2b[i];
        for (int c = 0; c < 6; c++)
        {
            beta0b[6 * i + c] += beta2b[6 * i + c];
            depb[c] += 2. / 3. * xi * H * beta2b[6 * i + c];
            Hacc += 2. / 3. * xi * p->dep[c] * beta2b[6 * i + c];
        }
        /* dep = dlam 1.5 sdev / seq */
        double sdevb[6] = {0, 0, 0, 0, 0, 0}, seqb = 0.0;
        if (fabs(p->seq) > O_EPS)
            for (int c = 0; c < 6; c++)
            {
                dlamb += depb[c] * 1.5 * p->sdev[c] / p->seq;
                sdevb[c] += depb[c] * 1.5 * p->dlam / p->seq;
                seqb -= depb[c] * 1.5 * p->dlam * p->sdev[c] / (p->seq * p->seq);
            }
        if (p->plastic)
        {
            const double den = 3 * p->c44 + H;
            const double yfb = dlamb / den;
            c44b[i] -= dlamb * p->yf * 3.0 / (den * den);
            Hacc -= dlamb * p->yf / (den * den);
            seqb += yfb;
            syb[i] -= yfb;
            Hacc -= yfb * (1.0 - xi) * m->alpha[3 * i + 0];
            alpha0b[i] -= yfb * (1.0 - xi) * H;
        }
        if (p->seq > 0.0)
            for (int c = 0; c < 6; c++)
                sdevb[c] += seqb * 1.5 * (c < 3 ? 1.0 : 2.0) * p->sdev[c] / p->seq;
        double slb[6];
        const double mb = (sdevb[0] + sdevb[1] + sdevb[2]) / 3.0;
        for (int c = 0; c < 6; c++)
        {
            beta0b[6 * i + c] -= sdevb[c];
            slb[c] = c < 3 ? sdevb[c] - mb : sdevb[c];
        }
        /* trial stress and trial force */
        double A1b[2] = {0, 0}, B1b[2] = {0, 0}, dL1b[64];
        for (int j = 0; j < nbi; j++)
        {
            int64_t b = B(m, i, j);
            int s = m->nsign[b];
            double cc[6], ccb[6];
            cc6(p->cs[j], cc);
            const double w = opp_flag(m, i, j) / m->volume * m->L0[b];
            double Ftrb = 0.0;
            for (int c = 0; c < 6; c++)
                Ftrb += slb[c] * w * cc[c], ccb[c] = slb[c] * w * p->Ftr[j];
            cc6_bar(p->cs[j], ccb, csbar[j]);
            const double t = Ftrb * m->dmg_w[b];
            dL1b[j] = t * 2.0 * m->Kn[b];
            B1b[s] += t;
            A1b[s] += t * m->Tv[b];
            Knb[b] += t * 2.0 * p->dL1[j];
            Tvb[b] += t * p->A1[s];
        }
        for (int j = 0; j < nbi; j++)
        {
            int64_t b = B(m, i, j);
            int s = m->nsign[b];
            double tot = dL1b[j] + A1b[s] + B1b[s] * m->Tv[b];
            Tvb[b] += B1b[s] * p->dL1[j];
            disb[j] += tot * m->dmg_broken[b];
            dLp0b[b] -= tot * m->dmg_broken[b];
            /* dis = |d|, cs = d / dis */
            const double *cs = p->cs[j];
            const double dot = cs[0] * csbar[j][0] + cs[1] * csbar[j][1] + cs[2] * csbar[j][2];
            for (int c = 0; c < 3; c++)
                gb[3 * b + c] += disb[j] * cs[c] + (csbar[j][c] - dot * cs[c]) / p->dis[j];
        }
    }
    /* positions from the bond vectors; moduli from the bond constants */
    double *kt = malloc(sizeof(double) * 3 * m->ntype);
    kt_table(m, kt);
    for (int i = 0; i < n; i++)
    {
        const int ti = m->type[i];
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int q = m->nbr[b], s = m->nsign[b], r = rev_slot(m, i, j);
            for (int c = 0; c < 3; c++)
                xb[3 * i + c] += gb[3 * b + c];
            if (r >= 0)
                for (int c = 0; c < 3; c++)
                    xb[3 * i + c] -= gb[3 * B(m, q, r) + c];
            const double kn = m->lattice == 1 ? kt[3 * ti] : kt[3 * ti + s];
            const double tv = m->lattice == 1 ? kt[3 * ti + 1] : kt[3 * ti + 2];
            if (m->lattice == 2)
            {
                eb[i] += 0.5 * (Knb[b] * kn + Tvb[b] * tv);
                if (r >= 0)
                    eb[i] += 0.5 * (Knb[B(m, q, r)] * kn + Tvb[B(m, q, r)] * tv);
            }
            else
                eb[i] += Knb[b] * kn + Tvb[b] * tv;
        }
        eb[i] += c44b[i] * m->Ce[3 * ti + 2];
    }
    *Hb = Hacc;
    free(kt), free(P), free(Fb), free(gb), free(Knb), free(Tvb), free(c44b);
}

/* ---------------------------------------------------------------- history adjoint --- */
/* Gaussian elimination with partial pivoting, row swaps applied to b as they happen;
 * a (n x n, row-major) is destroyed, b receives the solution.  Returns 1 if singular. */
static int dense_solve(int n, double *a, double *b)
{
    for (int k = 0; k < n; k++)
    {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(a[(size_t)i * n + k]) > fabs(a[(size_t)p * n + k]))
                p = i;
        if (a[(size_t)p * n + k] == 0.0)
            return 1;
        if (p != k)
        {
            for (int j = 0; j < n; j++)
            {
                double t = a[(size_t)k * n + j];
                a[(size_t)k * n + j] = a[(size_t)p * n + j], a[(size_t)p * n + j] = t;
            }
            double t = b[k];
            b[k] = b[p], b[p] = t;
        }
        for (int i = k + 1; i < n; i++)
        {
            const double l = a[(size_t)i * n + k] / a[(size_t)k * n + k];
            if (l == 0.0)
                continue;
            for (int j = k + 1; j < n; j++)
                a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
            b[i] -= l * b[k];
        }
    }
    for (int k = n - 1; k >= 0; k--)
    {
        double t = b[k];
        for (int j = k + 1; j < n; j++)
            t -= a[(size_t)k * n + j] * b[j];
        b[k] = t / a[(size_t)k * n + k];
    }
    return 0;
}

int orc_adjoint_run(orc_model *m, const orc_solver *s, int nsteps, const int *nbd, const orc_dispbc *dbc,
                    const int *nbf, const orc_forcebc *fbc, int maxbc, char axis, int t1, const double *target,
                    double *J, double *reactions, double *g_emod, double *g_sigmay, double *g_H)
{
    const int n = m->n, nbm = m->nbmax, d = m->dim, nd = d * n;
    const int ax = axis == 'x' ? 0 : (axis == 'y' ? 1 : 2);
    orc_solver sf = *s;
    sf.skip_damage = 1;
    /* per step: converged positions, incoming state, free-DOF mask */
    double *X = malloc(sizeof(double) * 3 * n * nsteps);
    double *Sp = malloc(sizeof(double) * (size_t)n * nbm * nsteps), *Sb = malloc(sizeof(double) * 6 * n * nsteps),
           *Sa = malloc(sizeof(double) * n * nsteps);
    int *Fr = malloc(sizeof(int) * nd * nsteps);
    int fail = 0;
    *J = 0.0;
    for (int st = 0; st < nsteps && !fail; st++)
    {
        for (int b = 0; b < n * nbm; b++)
            Sp[(size_t)st * n * nbm + b] = m->dLp[3 * b + 0];
        for (int k = 0; k < 6 * n; k++)
            Sb[(size_t)st * 6 * n + k] = m->beta[3 * k + 0];
        for (int i = 0; i < n; i++)
            Sa[(size_t)st * n + i] = m->alpha[3 * i + 0];
        orc_stepout o;
        if (orc_run_steps(m, &sf, 1, nbd + st, dbc + (size_t)st * maxbc, nbf + st, fbc + (size_t)st * maxbc, maxbc,
                          axis, t1, &o))
            fail = st + 1;
        memcpy(X + (size_t)st * 3 * n, m->x, sizeof(double) * 3 * n);
        for (int k = 0; k < nd; k++)
            Fr[(size_t)st * nd + k] = m->dispBC[k] && m->fixidx[k];
        reactions[st] = o.reaction;
        *J += (o.reaction - target[st]) * (o.reaction - target[st]);
    }
    if (fail)
    {
        free(X), free(Sp), free(Sb), free(Sa), free(Fr);
        return fail;
    }
    /* backward sweep */
    double *sp = calloc((size_t)n * nbm, sizeof(double)), *sb = calloc(6 * n, sizeof(double)),
           *sa = calloc(n, sizeof(double));
    double *Pb = malloc(sizeof(double) * 3 * n), *xb = malloc(sizeof(double) * 3 * n);
    double *o1 = malloc(sizeof(double) * n * nbm), *o2 = malloc(sizeof(double) * 6 * n), *o3 = malloc(sizeof(double) * n);
    double *eb = malloc(sizeof(double) * n), *yb = malloc(sizeof(double) * n);
    double *zp = calloc((size_t)n * nbm, sizeof(double)), *zb = calloc(6 * n, sizeof(double)), *za = calloc(n, sizeof(double));
    int *fi = malloc(sizeof(int) * nd);
    double *A = NULL, *bv = NULL;
    memset(g_emod, 0, sizeof(double) * n);
    memset(g_sigmay, 0, sizeof(double) * n);
    *g_H = 0.0;
    for (int st = nsteps - 1; st >= 0; st--)
    {
        memcpy(m->x, X + (size_t)st * 3 * n, sizeof(double) * 3 * n);
        for (int b = 0; b < n * nbm; b++)
            m->dLp[3 * b + 0] = Sp[(size_t)st * n * nbm + b];
        for (int k = 0; k < 6 * n; k++)
            m->beta[3 * k + 0] = Sb[(size_t)st * 6 * n + k];
        for (int i = 0; i < n; i++)
            m->alpha[3 * i + 0] = Sa[(size_t)st * n + i];
        orc_eval_j2(m);
        int nf = 0;
        for (int k = 0; k < nd; k++)
            if (Fr[(size_t)st * nd + k])
                fi[nf++] = k;
        const double jr = 2.0 * (reactions[st] - target[st]);
        /* R = -sum_{type t1} Pin_axis */
        memset(Pb, 0, sizeof(double) * 3 * n);
        for (int i = 0; i < n; i++)
            if (m->type[i] == t1)
                Pb[3 * i + ax] = -jr;
        double Hb;
        orc_vjp_j2(m, Pb, sp, sb, sa, xb, o1, o2, o3, eb, yb, &Hb);
        A = realloc(A, sizeof(double) * (size_t)nf * nf);
        bv = realloc(bv, sizeof(double) * nf);
        for (int r = 0; r < nf; r++)
            bv[r] = xb[3 * (fi[r] / d) + fi[r] % d];
        /* A[r][c] = d(P_f[c]) / d(u_f[r]) = column c of the free-DOF Jacobian, transposed */
        for (int c = 0; c < nf; c++)
        {
            memset(Pb, 0, sizeof(double) * 3 * n);
            Pb[3 * (fi[c] / d) + fi[c] % d] = 1.0;
            double h2;
            orc_vjp_j2(m, Pb, zp, zb, za, xb, o1, o2, o3, eb, yb, &h2);
            for (int r = 0; r < nf; r++)
                A[(size_t)r * nf + c] = xb[3 * (fi[r] / d) + fi[r] % d];
        }
        if (dense_solve(nf, A, bv))
        {
            fail = -1;
            break;
        }
        memset(Pb, 0, sizeof(double) * 3 * n);
        for (int i = 0; i < n; i++)
            if (m->type[i] == t1)
                Pb[3 * i + ax] = -jr;
        for (int r = 0; r < nf; r++)
            Pb[3 * (fi[r] / d) + fi[r] % d] -= bv[r];
        orc_vjp_j2(m, Pb, sp, sb, sa, xb, o1, o2, o3, eb, yb, &Hb);
        memcpy(sp, o1, sizeof(double) * n * nbm);
        memcpy(sb, o2, sizeof(double) * 6 * n);
        memcpy(sa, o3, sizeof(double) * n);
        for (int i = 0; i < n; i++)
            g_emod[i] += eb[i], g_sigmay[i] += yb[i];
        *g_H += Hb;
    }
    free(X), free(Sp), free(Sb), free(Sa), free(Fr), free(sp), free(sb), free(sa), free(Pb), free(xb);
    free(o1), free(o2), free(o3), free(eb), free(yb), free(zp), free(zb), free(za), free(fi), free(A), free(bv);
    return fail;
}
