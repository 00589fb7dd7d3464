/*
 * Oracle: linear solve, damage, crack bookkeeping, outputs and the load-step loop.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define O_EPS 1e-6
#define O_PI 3.14159265358979323846

static inline int64_t B(const orc_model *m, int i, int j) { return (int64_t)i * m->nbmax + j; }

/* y = K x for the symmetric matrix stored as the reference's 1-based upper CSR
 * (the operation of mkl_sparse_d_mv with SPARSE_MATRIX_TYPE_SYMMETRIC/UPPER,
 * src/solver.c:195-206,243). */
void orc_spmv_upper(const orc_model *m, const double *x, double *y)
{
    const int nr = m->dim * m->n;
    memset(y, 0, sizeof(double) * nr);
    for (int r = 0; r < nr; r++)
    {
        double acc = 0.0;
        for (int p = m->IK[r] - 1; p < m->IK[r + 1] - 1; p++)
        {
            int c = m->JK[p] - 1;
            acc += m->K[p] * x[c];
            if (c != r)
                y[c] += m->K[p] * x[r];
        }
        y[r] += acc;
    }
}

/* Parallel variant for the CPU timing baseline: rows are split into thread bands,
 * each thread scatters into its own buffer, buffers are summed at the end. */
static void spmv_upper_par(const orc_model *m, const double *x, double *y, double *scratch, int nt)
{
    const int nr = m->dim * m->n;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        double *yt = scratch + (size_t)t * nr;
        memset(yt, 0, sizeof(double) * nr);
#pragma omp for schedule(static)
        for (int r = 0; r < nr; r++)
        {
            double acc = 0.0;
            for (int p = m->IK[r] - 1; p < m->IK[r + 1] - 1; p++)
            {
                int c = m->JK[p] - 1;
                acc += m->K[p] * x[c];
                if (c != r)
                    yt[c] += m->K[p] * x[r];
            }
            yt[r] += acc;
        }
#pragma omp for schedule(static)
        for (int r = 0; r < nr; r++)
        {
            double s = 0.0;
            for (int u = 0; u < nt; u++)
                s += scratch[(size_t)u * nr + r];
            y[r] = s;
        }
    }
#else
    (void)scratch, (void)nt;
    orc_spmv_upper(m, x, y);
#endif
}

/* solverCG (reference src/solver.c:188-270): unpreconditioned conjugate gradients
 * from disp = 0, stopping when |r|^2 <= rtol2*|r0|^2 + atol2 (MKL dcg's residual
 * test with dpar[0]=rtol2, dpar[1]=atol2 on squared norms) or after maxit
 * iterations; then xyz += disp, also on failure.  Returns the iteration count,
 * negative if maxit was reached. */
int orc_solve_cg(orc_model *m, const orc_solver *s, double *final_rr)
{
    const int nr = m->dim * m->n;
    const int nt = s->nthreads > 1 ? s->nthreads : 1;
    double *r = malloc(sizeof(double) * nr), *p = malloc(sizeof(double) * nr), *q = malloc(sizeof(double) * nr);
    double *scratch = nt > 1 ? malloc(sizeof(double) * (size_t)nr * nt) : NULL;
    double *x = m->disp;
    memset(x, 0, sizeof(double) * nr);
    memcpy(r, m->residual, sizeof(double) * nr);
    memcpy(p, r, sizeof(double) * nr);
    double rr = 0.0;
    for (int k = 0; k < nr; k++)
        rr += r[k] * r[k];
    const double thr = s->rtol2 * rr + s->atol2;
    int it = 0, conv = rr <= thr;
    while (!conv && it < s->maxit)
    {
        if (nt > 1)
            spmv_upper_par(m, p, q, scratch, nt);
        else
            orc_spmv_upper(m, p, q);
        double pq = 0.0;
        for (int k = 0; k < nr; k++)
            pq += p[k] * q[k];
        double alpha = rr / pq, rrn = 0.0;
        for (int k = 0; k < nr; k++)
        {
            x[k] += alpha * p[k];
            r[k] -= alpha * q[k];
            rrn += r[k] * r[k];
        }
        it++;
        if (rrn <= thr)
        {
            rr = rrn;
            conv = 1;
            break;
        }
        double beta = rrn / rr;
        rr = rrn;
        for (int k = 0; k < nr; k++)
            p[k] = r[k] + beta * p[k];
    }
    if (final_rr)
        *final_rr = rr;
    for (int i = 0; i < m->n; i++)
        for (int c = 0; c < m->dim; c++)
            m->x[3 * i + c] += x[m->dim * i + c];
    free(r), free(p), free(q), free(scratch);
    return conv ? it : -it;
}

/* CPU timing baseline of the reference's linear solve (solverCG, src/solver.c:188-270:
 * unpreconditioned CG, the rate-setting part of a load step) on host cores.  The
 * symmetric upper CSR is expanded once to full CSR (untimed), as a threaded
 * symmetric SpMV would otherwise need per-thread scatter buffers; then `iters` CG
 * iterations from disp = 0 on m->residual run with the SpMV, dots and updates
 * spread over `nthreads` OpenMP threads.  Returns the wall seconds of those
 * iterations (negative on allocation failure); *rr_out gets the last |r|^2.
 * Timing only: the summation order is not the specification's. */
double orc_cg_baseline(const orc_model *m, int nthreads, int iters, double *rr_out)
{
    const int nr = m->dim * m->n;
    const int nt = nthreads > 0 ? nthreads : 1;
    int64_t *rp = calloc((size_t)nr + 1, sizeof(int64_t));
    if (!rp)
        return -1.0;
    for (int r = 0; r < nr; r++)
        for (int p = m->IK[r] - 1; p < m->IK[r + 1] - 1; p++)
        {
            int c = m->JK[p] - 1;
            rp[r + 1]++;
            if (c != r)
                rp[c + 1]++;
        }
    for (int r = 0; r < nr; r++)
        rp[r + 1] += rp[r];
    int *ci = malloc(sizeof(int) * (size_t)rp[nr]);
    double *va = malloc(sizeof(double) * (size_t)rp[nr]);
    int64_t *fill = malloc(sizeof(int64_t) * (size_t)nr);
    double *x = malloc(sizeof(double) * nr), *r = malloc(sizeof(double) * nr), *p = malloc(sizeof(double) * nr),
           *q = malloc(sizeof(double) * nr);
    if (!ci || !va || !fill || !x || !r || !p || !q)
    {
        free(rp), free(ci), free(va), free(fill), free(x), free(r), free(p), free(q);
        return -1.0;
    }
    memcpy(fill, rp, sizeof(int64_t) * nr);
    for (int row = 0; row < nr; row++)
        for (int k = m->IK[row] - 1; k < m->IK[row + 1] - 1; k++)
        {
            int c = m->JK[k] - 1;
            ci[fill[row]] = c, va[fill[row]++] = m->K[k];
            if (c != row)
                ci[fill[c]] = row, va[fill[c]++] = m->K[k];
        }
    double rr = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : rr) schedule(static)
    for (int k = 0; k < nr; k++)
    {
        x[k] = 0.0, r[k] = m->residual[k], p[k] = r[k];
        rr += r[k] * r[k];
    }
#ifdef _OPENMP
    const double t0 = omp_get_wtime();
#endif
    for (int it = 0; it < iters; it++)
    {
        double pq = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : pq) schedule(static)
        for (int row = 0; row < nr; row++)
        {
            double acc = 0.0;
            for (int64_t k = rp[row]; k < rp[row + 1]; k++)
                acc += va[k] * p[ci[k]];
            q[row] = acc;
            pq += p[row] * acc;
        }
        const double alpha = rr / pq;
        double rrn = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : rrn) schedule(static)
        for (int k = 0; k < nr; k++)
        {
            x[k] += alpha * p[k];
            r[k] -= alpha * q[k];
            rrn += r[k] * r[k];
        }
        const double beta = rrn / rr;
        rr = rrn;
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int k = 0; k < nr; k++)
            p[k] = r[k] + beta * p[k];
    }
    double secs = -1.0;
#ifdef _OPENMP
    secs = omp_get_wtime() - t0;
#endif
    if (rr_out)
        *rr_out = rr;
    free(rp), free(ci), free(va), free(fill), free(x), free(r), free(p), free(q);
    return secs;
}

/* Broken-bond file of the damage update being run (NULL: none); the damage
 * functions write their lines where the reference writes them. */
static FILE *o_fp = NULL;

/* updateDuctileDamagePwiseNonlocal (reference src/constitutive.c:1757-1862):
 * Gaussian-weighted average of dlambda*(1 + A*triaxiality) over all particles
 * within 3L in the initial configuration (index order), bond breakage when either
 * end exceeds the threshold, bond damage max(D_i, D_j), w = 1 - D. */
int orc_damage_pwise_nonlocal(orc_model *m)
{
    const double L = m->damage_L;
    for (int i = 0; i < m->n; i++)
    {
        double *dn = m->dnl + 2 * i;
        if (dn[0] > m->damage_threshold)
        {
            if (dn[0] > 1.0)
                dn[0] = 1.0;
            continue;
        }
        double Dd = 0, A = 0;
        for (int j = 0; j < m->n; j++)
        {
            double dx = m->x0[3 * j] - m->x0[3 * i], dy = m->x0[3 * j + 1] - m->x0[3 * i + 1],
                   dz = m->x0[3 * j + 2] - m->x0[3 * i + 2];
            double dis = sqrt(dx * dx + dy * dy + dz * dz);
            if (dis < 3 * L)
            {
                double loc = 0;
                double f = (1.0 + m->damagec_A * m->triax[j]);
                if (f > 0.0)
                    loc = m->dlambda[j] * (1.0 + m->damagec_A * m->triax[j]);
                double phi = 1.0 / L / sqrt(2 * O_PI) * exp(-0.5 * dis * dis / L / L);
                Dd += loc * phi * m->volume;
                A += phi * m->volume;
            }
        }
        if (Dd > 0.0)
            dn[0] += 1.0 / A * Dd;
    }
    int k = 0;
    for (int i = 0; i < m->n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int q = m->nbr[b];
            if (m->dnl[2 * i] > m->damage_threshold || m->dnl[2 * q] > m->damage_threshold)
                if (fabs(m->dmg_broken[b]) > O_EPS)
                {
                    m->dmg_broken[b] = 0.0;
                    m->dmgD[2 * b] = 1.0;
                    if (o_fp)
                        fprintf(o_fp, "%d %d \n", i, q);
                    k++;
                }
        }
    for (int i = 0; i < m->n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            int q = m->nbr[b];
            if (fabs(m->dmg_broken[b]) > O_EPS)
            {
                double a = m->dnl[2 * i], c = m->dnl[2 * q];
                m->dmgD[2 * b] = a < c ? c : a;
            }
            m->dmg_w[b] = 1.0 - m->dmgD[2 * b];
        }
    return k;
}

/* The shell sort of updateBrittleDamage (reference src/constitutive.c:1486-1508):
 * gaps k/2, k/4, ..., 1, insertion with strict '>' so that ties keep their gap
 * order.  order[] is permuted alongside the keys; returns k. */
int orc_shell_sort_select(int k, const double *bstrain, int nbreak, int *order)
{
    double *key = malloc(sizeof(double) * (k > 0 ? k : 1));
    for (int i = 0; i < k; i++)
        key[i] = bstrain[i], order[i] = i;
    for (int r = k / 2; r >= 1; r = r / 2)
        for (int i = r; i < k; ++i)
        {
            double tb = key[i];
            int to = order[i];
            int j = i - r;
            while (j >= 0 && key[j] > tb)
            {
                key[j + r] = key[j];
                order[j + r] = order[j];
                j = j - r;
            }
            key[j + r] = tb;
            order[j + r] = to;
        }
    free(key);
    (void)nbreak;
    return k;
}

/* updateDuctileDamageBwiseNonlocal (reference src/constitutive.c:1529-1605): the
 * nonlocal damage of every particle below the threshold grows by the Gaussian-
 * weighted mean (DAM_PHI of the initial bond length, include/lpm.h:51) of the local
 * rates (1 + A triaxiality) dlambda of itself and its list; then every directed bond
 * whose two ends average above the threshold breaks, and the weights follow.  Each
 * update reads only multipliers and triaxialities, and each break only its two ends,
 * so the particle order does not matter.  Returns the directed entries broken. */
int orc_damage_bwise_nonlocal(orc_model *m)
{
    const int n = m->n, nbm = m->nbmax;
    const double L = m->damage_L, V = m->volume;
    for (int i = 0; i < n; i++)
    {
        if (m->dnl[2 * i + 0] > m->damage_threshold)
        {
            if (m->dnl[2 * i + 0] > 1.0)
                m->dnl[2 * i + 0] = 1.0;
            continue;
        }
        double DdotLocal = 0;
        double damage_f = (1.0 + m->damagec_A * m->triax[i]);
        if (damage_f > 0.0)
            DdotLocal = m->dlambda[i] * (1.0 + m->damagec_A * m->triax[i]);
        double Ddot = DdotLocal * V;
        double A = V;
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            const int64_t b = (int64_t)i * nbm + j;
            const int q = m->nbr[b];
            DdotLocal = 0;
            damage_f = (1.0 + m->damagec_A * m->triax[q]);
            if (damage_f > 0.0)
                DdotLocal = m->dlambda[q] * (1.0 + m->damagec_A * m->triax[q]);
            const double x = m->L0[b];
            const double phi = 1.0 / L / sqrt(2 * O_PI) * exp(-0.5 * x * x / L / L);
            Ddot += DdotLocal * phi * V;
            A += phi * V;
        }
        if (Ddot > 0.0)
            m->dnl[2 * i + 0] += 1.0 / A * Ddot;
    }
    int k = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            const int64_t b = (int64_t)i * nbm + j;
            const int q = m->nbr[b];
            if (0.5 * (m->dnl[2 * i + 0] + m->dnl[2 * q + 0]) > m->damage_threshold && fabs(m->dmg_broken[b]) > O_EPS)
            {
                m->dmg_broken[b] = 0.0;
                m->dmgD[2 * b + 0] = 1.0;
                if (o_fp)
                    fprintf(o_fp, "%d %d \n", i, q);
                k++;
            }
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            const int64_t b = (int64_t)i * nbm + j;
            const int q = m->nbr[b];
            if (fabs(m->dmg_broken[b]) > O_EPS)
                m->dmgD[2 * b + 0] = 0.5 * (m->dnl[2 * i + 0] + m->dnl[2 * q + 0]);
            m->dmg_w[b] = 1.0 - m->dmgD[2 * b + 0];
        }
    return k;
}

/* updateDuctileDamagePwiseLocal (reference src/constitutive.c:1698-1755): local
 * damage from the plastic multiplier; a particle passing the threshold (and not yet
 * at 1) is set to 1 and all its bonds and their reverse entries break; then the
 * bond damage is the larger of its two ends.  Each particle's decision reads only
 * its own state: `order` 1 (reverse) gives the state of 0 (the reference's order).
 * Returns the particles that failed. */
int orc_damage_pwise_local(orc_model *m, int order)
{
    const int n = m->n, nbm = m->nbmax;
    int k = 0;
    for (int t = 0; t < n; t++)
    {
        const int i = order ? n - 1 - t : t;
        double damage_f = (1.0 + m->damagec_A * m->triax[i]);
        if (damage_f > 0.0 && m->dloc[2 * i + 0] <= m->damage_threshold)
            m->dloc[2 * i + 0] += damage_f * m->dlambda[i];
        if (m->dloc[2 * i + 0] > m->damage_threshold && fabs(m->dloc[2 * i + 0] - 1.0) > O_EPS)
        {
            m->dloc[2 * i + 0] = 1.0;
            for (int j = 0; j < m->nb_initial[i]; j++)
            {
                const int64_t b = (int64_t)i * nbm + j;
                const int q = m->nbr[b];
                m->dmg_broken[b] = 0.0;
                for (int jj = 0; jj < nbm; jj++)
                    if (m->nbr[(int64_t)q * nbm + jj] == i)
                        m->dmg_broken[(int64_t)q * nbm + jj] = 0.0;
            }
            if (o_fp)
                fprintf(o_fp, "%d \n", i);
            k++;
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            const int64_t b = (int64_t)i * nbm + j;
            const int q = m->nbr[b];
            m->dmgD[2 * b + 0] = m->dloc[2 * i + 0] < m->dloc[2 * q + 0] ? m->dloc[2 * q + 0] : m->dloc[2 * i + 0];
            m->dmg_w[b] = 1.0 - m->dmgD[2 * b + 0];
        }
    return k;
}

/* updateDuctileDamageBwiseLocal (reference src/constitutive.c:1607-1695): local
 * damage accumulated from the plastic multiplier, bond damage as the mean of the two
 * ends, a bond (and its reverse entry) broken when that mean passes the threshold,
 * then the weights.  `order` = 0 visits the particles in the reference's order,
 * 1 in reverse: the decision on a bond depends only on its two ends, so both give the
 * same state (the tests check it).  Returns the number of bonds broken (each pair
 * once). */
int orc_damage_bwise_local(orc_model *m, int order)
{
    const int n = m->n, nbm = m->nbmax;
    for (int i = 0; i < n; i++)
    {
        double damage_f = (1.0 + m->damagec_A * m->triax[i]);
        if (damage_f > 0.0 && m->dloc[2 * i + 0] <= m->damage_threshold)
            m->dloc[2 * i + 0] += damage_f * m->dlambda[i];
        else if (m->dloc[2 * i + 0] > 1.0)
            m->dloc[2 * i + 0] = 1.0;
    }
    int k = 0;
    for (int t = 0; t < n; t++)
    {
        const int i = order ? n - 1 - t : t;
        m->nb[i] = m->nb_initial[i];
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            const int64_t b = (int64_t)i * nbm + j;
            const int q = m->nbr[b];
            m->dmgD[2 * b + 0] = 0.5 * (m->dloc[2 * i + 0] + m->dloc[2 * q + 0]);
            if (m->dmgD[2 * b + 0] > m->damage_threshold && m->dmg_broken[b] > O_EPS)
            {
                m->dmgD[2 * b + 0] = 1.0;
                m->dmg_broken[b] = 0.0;
                for (int jj = 0; jj < nbm; jj++)
                    if (m->nbr[(int64_t)q * nbm + jj] == i)
                    {
                        m->dmgD[2 * ((int64_t)q * nbm + jj) + 0] = 1.0;
                        m->dmg_broken[(int64_t)q * nbm + jj] = 0.0;
                    }
                if (o_fp)
                    fprintf(o_fp, "%d %d \n", i, q);
                k++;
            }
            if (m->dmg_broken[b] <= O_EPS)
                m->nb[i] -= 1;
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            const int64_t b = (int64_t)i * nbm + j;
            const int q = m->nbr[b];
            if (fabs(m->dmg_broken[b]) < O_EPS || m->nb[i] == 0 || m->nb[q] == 0)
                m->dmgD[2 * b + 0] = 1.0;
            m->dmg_w[b] = 1.0 - m->dmgD[2 * b + 0];
        }
    return k;
}

/* updateBrittleDamage (reference src/constitutive.c:1437-1526): bonds with
 * dL/L0 >= critical (per-bond critical stretch with the extension), collected in
 * (i, slot) order; all break if there are at most nbreak, otherwise the nbreak
 * largest after the reference's shell sort.  The reference's fixed 400-entry
 * candidate buffer is replaced by an unbounded one.  Broken (i, j) pairs are
 * written to broken_i/broken_j (up to cap); returns the candidate count k. */
int orc_damage_brittle(orc_model *m, int *bi, int *bj, int cap)
{
    int k = 0, capc = 1024;
    int *ci = malloc(sizeof(int) * capc), *cj = malloc(sizeof(int) * capc);
    double *cb = malloc(sizeof(double) * capc);
    for (int i = 0; i < m->n; i++)
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            double crit = m->crit ? m->crit[b] : m->critical_bstrain;
            double ave = m->dL[b] / m->L0[b];
            if (ave >= crit)
            {
                if (k == capc)
                {
                    capc *= 2;
                    ci = realloc(ci, sizeof(int) * capc);
                    cj = realloc(cj, sizeof(int) * capc);
                    cb = realloc(cb, sizeof(double) * capc);
                }
                ci[k] = i, cj[k] = j, cb[k] = ave;
                k++;
            }
        }
    int nb_out = 0;
    if (k > 0 && k <= m->nbreak)
    {
        for (int t = 0; t < k; t++)
        {
            int64_t b = B(m, ci[t], cj[t]);
            m->dmgD[2 * b] = 1.0;
            m->dmg_w[b] = 0.0;
            m->dmg_broken[b] = 0.0;
            if (nb_out < cap)
                bi[nb_out] = ci[t], bj[nb_out] = m->nbr[b], nb_out++;
            if (o_fp)
                fprintf(o_fp, "%d %d \n", (int)(b / m->nbmax), m->nbr[b]);
        }
    }
    else if (k > m->nbreak)
    {
        int *ord = malloc(sizeof(int) * k);
        orc_shell_sort_select(k, cb, m->nbreak, ord);
        for (int t = k - m->nbreak; t < k; t++)
        {
            int64_t b = B(m, ci[ord[t]], cj[ord[t]]);
            m->dmgD[2 * b] = 1.0;
            m->dmg_w[b] = 0.0;
            m->dmg_broken[b] = 0.0;
            if (nb_out < cap)
                bi[nb_out] = ci[ord[t]], bj[nb_out] = m->nbr[b], nb_out++;
            if (o_fp)
                fprintf(o_fp, "%d %d \n", (int)(b / m->nbmax), m->nbr[b]);
        }
        free(ord);
    }
    free(ci), free(cj), free(cb);
    return k;
}

/* updateCrack (reference src/constitutive.c:1399-1434): live-bond count, visual
 * damage, F *= damage_w (a second time after the constitutive law), Pin from F,
 * particles without bonds fixed. */
void orc_update_crack(orc_model *m)
{
    const int d = m->dim;
    for (int i = 0; i < m->n; i++)
    {
        m->nb[i] = m->nb_initial[i];
        m->damage_visual[i] = 0.0;
        double *P = m->Pin + 3 * i;
        P[0] = 0.0, P[1] = 0.0, P[2] = 0.0;
        for (int j = 0; j < m->nb_initial[i]; j++)
        {
            int64_t b = B(m, i, j);
            if (m->dmg_broken[b] <= O_EPS)
                m->nb[i] -= 1;
            m->damage_visual[i] += m->dmg_broken[b];
            m->F[b] *= m->dmg_w[b];
            P[0] += m->cs[3 * b + 0] * m->F[b];
            P[1] += m->cs[3 * b + 1] * m->F[b];
            P[2] += m->cs[3 * b + 2] * m->F[b];
        }
        if (m->nb[i] < 1)
        {
            m->fixidx[d * i] = 0;
            m->fixidx[d * i + 1] = 0;
            if (d == 3)
                m->fixidx[d * i + 2] = 0;
        }
        m->damage_visual[i] = 1 - m->damage_visual[i] / m->nb_initial[i];
    }
}

/* writeReaction / writeDisp values (reference src/data_handler.c:278-335). */
double orc_reaction(const orc_model *m, char axis, int t1)
{
    /* an axis character other than x, y, z adds nothing (data_handler.c:320-327) */
    int c = axis == 'x' ? 0 : axis == 'y' ? 1 : axis == 'z' ? 2 : -1;
    double force = 0.;
    if (c < 0)
        return force;
    for (int i = 0; i < m->n; i++)
        if (m->type[i] == t1)
            for (int j = 0; j < m->nb_initial[i]; j++)
            {
                int64_t b = B(m, i, j);
                force += -m->cs[3 * b + c] * m->F[b];
            }
    return force;
}

double orc_mean_disp(const orc_model *m, char axis, int t1)
{
    int c = axis == 'x' ? 0 : axis == 'y' ? 1 : 2;
    int num = 0;
    double p = 0, p0 = 0;
    for (int i = 0; i < m->n; i++)
        if (m->type[i] == t1)
        {
            p += m->x[3 * i + c];
            p0 += m->x0[3 * i + c];
            num++;
        }
    return (p - p0) / num;
}

/* The diagonal value the last elimination put on a constrained DOF (-1 if none). */
static double first_bc_diag(const orc_model *m)
{
    const int d = m->dim;
    for (int i = 0; i < m->n; i++)
        for (int c = 0; c < d; c++)
            if (m->dispBC[d * i + c] == 0 || m->fixidx[d * i + c] == 0)
            {
                int64_t p = m->kptr[2 * i + 1];
                int k = m->kptr[2 * i];
                int64_t pos = c == 0 ? p : (c == 1 ? p + d * k : p + 2 * d * k - 1);
                return m->K[pos];
            }
    return -1.0;
}

/* Reference-format files written by orc_run_steps when set (NULL: none): the
 * broken-bond file of updateBrittleDamage / updateDuctileDamagePwiseNonlocal, a
 * "TIMESTEP" line per call and one "i neighbour" line per newly broken bond in the
 * order the reference breaks them (constitutive.c:1440-1442, 1481, 1519, 1760-1763,
 * 1838), and writeReaction / writeDisp of (axis, type) after every Newton loop of a
 * step, re-solves after breakage included (data_handler.c:278-335, called at
 * lpmc_project.c:494-501).  Step numbers are the step index + 1 + step_offset. */
static char o_bond[1024], o_reac[1024], o_disp[1024];
static char o_axis = 'z';
static int o_type = 1, o_off = 0;

void orc_set_output(const char *bond_file, const char *reaction_file, const char *disp_file, char axis, int t1,
                    int step_offset)
{
    snprintf(o_bond, sizeof(o_bond), "%s", bond_file ? bond_file : "");
    snprintf(o_reac, sizeof(o_reac), "%s", reaction_file ? reaction_file : "");
    snprintf(o_disp, sizeof(o_disp), "%s", disp_file ? disp_file : "");
    o_axis = axis, o_type = t1, o_off = step_offset;
}

static int damage_general(orc_model *m, int tstep)
{
    /* updateDamageGeneral (reference src/constitutive.c:149-164) */
    o_fp = o_bond[0] ? fopen(o_bond, "a+") : NULL;
    if (o_fp)
        fprintf(o_fp, "TIMESTEP %d\n", tstep);
    int k = 0;
    if (m->plmode == 0)
        k = orc_damage_pwise_nonlocal(m);
    else if (m->plmode == 5)
        k = orc_damage_bwise_local(m, 0);
    else if (m->plmode == 6)
        k = orc_damage_brittle(m, NULL, NULL, 0);
    if (o_fp)
        fclose(o_fp);
    o_fp = NULL;
    return k;
}

static void write_outputs(const orc_model *m, int tstep)
{
    if (o_reac[0])
    {
        FILE *fp = fopen(o_reac, "a+");
        if (fp)
        {
            fprintf(fp, "%d %8.8e \n", tstep, orc_reaction(m, o_axis, o_type));
            fclose(fp);
        }
    }
    if (o_disp[0])
    {
        FILE *fp = fopen(o_disp, "a+");
        if (fp)
        {
            fprintf(fp, "%d %8.8e\n", tstep, orc_mean_disp(m, o_axis, o_type));
            fclose(fp);
        }
    }
}

/* The load-step loop of the reference driver (src/lpmc_project.c:382-546) with the
 * solver tolerances as parameters.  computeStrain (output only) is not run.
 * Returns 0, or the index+1 of the first step whose Newton loop hit the cap. */
/* Anderson mixing of the modified-Newton iterates u <- u + d(u) (Walker & Ni 2011,
 * type II): the last `depth` differences of iterates and corrections, the
 * least-squares weights from the normal equations (regularised by 1e-12 of their
 * largest diagonal), u <- u + d - sum_j g_j (dU_j + dF_j).  An accelerated iterate
 * whose residual grew more than 100-fold is replaced by the plain one; when the
 * residual has not halved over the last `depth` iterations (a stale history, as
 * particles switch between elastic and plastic) the differences start afresh. */
typedef struct
{
    int depth, cnt, next, have, nd;
    double *dU, *dF, *up, *fp, G[8][8];
} orc_aa;

static void aa_begin(orc_aa *a, int depth, int nd)
{
    if (a->nd != nd || a->depth != depth)
    {
        free(a->dU), free(a->dF), free(a->up), free(a->fp);
        a->dU = malloc(sizeof(double) * (size_t)nd * (depth > 0 ? depth : 1));
        a->dF = malloc(sizeof(double) * (size_t)nd * (depth > 0 ? depth : 1));
        a->up = malloc(sizeof(double) * nd), a->fp = malloc(sizeof(double) * nd);
    }
    a->depth = depth, a->nd = nd, a->cnt = 0, a->next = 0, a->have = 0;
}

/* x holds u + d after the solve, m->disp holds d; returns 1 when it mixed */
static int aa_step(orc_aa *a, orc_model *m)
{
    const int d = m->dim, nd = a->nd;
    if (a->depth <= 0)
        return 0;
    const int s = a->next;
    double *dU = a->dU + (size_t)s * nd, *dF = a->dF + (size_t)s * nd;
    for (int t = 0; t < nd; t++)
    {
        const int i = t / d, c = t % d;
        const double f = m->disp[t], u = m->x[3 * i + c] - f;
        if (a->have)
            dU[t] = u - a->up[t], dF[t] = f - a->fp[t];
        a->up[t] = u, a->fp[t] = f;
    }
    if (!a->have)
    {
        a->have = 1;
        return 0;
    }
    if (a->cnt < a->depth)
        a->cnt++;
    a->next = (a->next + 1) % a->depth;
    const int k = a->cnt;
    double h[8], gmat[64], dmax = 0.0;
    for (int j = 0; j < k; j++)
    {
        const double *Fj = a->dF + (size_t)j * nd;
        double g = 0.0, hh = 0.0;
        for (int t = 0; t < nd; t++)
            g += dF[t] * Fj[t], hh += Fj[t] * m->disp[t];
        a->G[s][j] = a->G[j][s] = g;
        h[j] = hh;
    }
    for (int j = 0; j < k; j++)
        dmax = a->G[j][j] > dmax ? a->G[j][j] : dmax;
    if (!(dmax > 0.0))
    {
        a->cnt = a->next = 0;
        return 0;
    }
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++)
            gmat[i * k + j] = a->G[i][j] + (i == j ? 1e-12 * dmax : 0.0);
    if (orc_lu_solve(k, gmat, h))
    {
        a->cnt = a->next = 0;
        return 0;
    }
    for (int t = 0; t < nd; t++)
    {
        double sum = 0.0;
        for (int j = 0; j < k; j++)
            sum += h[j] * (a->dU[(size_t)j * nd + t] + a->dF[(size_t)j * nd + t]);
        const int i = t / d, c = t % d;
        m->x[3 * i + c] -= sum;
    }
    return 1;
}

#define AA_RING 64

int orc_run_steps(orc_model *m, const orc_solver *s, int nsteps,
                  const int *nbd, const orc_dispbc *dbc, const int *nbf, const orc_forcebc *fbc,
                  int maxbc, char out_axis, int out_type, orc_stepout *out)
{
    const int d = m->dim, n = m->n;
    int first_fail = 0;
    orc_aa aa;
    memset(&aa, 0, sizeof(aa));
    for (int st = 0; st < nsteps; st++)
    {
        orc_stepout o;
        memset(&o, 0, sizeof(o));
        memcpy(m->xt, m->x, sizeof(double) * 3 * n);
        memcpy(m->Ftemp, m->F, sizeof(double) * (size_t)n * m->nbmax);
        memcpy(m->Pex_temp, m->Pex, sizeof(double) * d * n);
        if (s->tangent == 0)
            orc_tangent_fd(m);
        else
            orc_tangent_analytic(m);
        orc_set_disp_bc(m, nbd[st], dbc + (size_t)st * maxbc);
        orc_set_force_bc(m, nbf[st], fbc + (size_t)st * maxbc);
        orc_bond_force_general(m, 4);
        for (;;)
        {
            orc_update_rr(m);
            double nres = orc_nrm2(d * n, m->residual);
            double nrea = orc_nrm2(m->nreaction, m->reaction);
            double tolm = nres > nrea ? nres : nrea;
            int ni = 0;
            double nd = -1.0;
            aa_begin(&aa, s->accel > 8 ? 8 : s->accel, d * n);
            double nprev = nres;
            double ring[AA_RING];
            int since = 0;
            ring[0] = nres;
            while (nres > s->newton_tol * tolm && ni < s->newton_maxit)
            {
                orc_switch_state(m, 0);
                if (s->fixed_diag && nd >= 0.0)
                    orc_bc_stiffness_update_nd(m, nd);
                else
                {
                    orc_bc_stiffness_update(m);
                    nd = first_bc_diag(m);
                }
                int it = orc_solve_cg(m, s, NULL);
                o.cg_iters += it < 0 ? -it : it;
                const int mixed = aa_step(&aa, m);
                orc_bond_force_general(m, m->plmode);
                orc_update_rr(m);
                ni++;
                nres = orc_nrm2(d * n, m->residual);
                if (mixed && !(nres <= 1e2 * nprev))
                {
                    /* back to the plain iterate, history restarted */
                    for (int t = 0; t < d * n; t++)
                        m->x[3 * (t / d) + t % d] = aa.up[t] + aa.fp[t];
                    aa.cnt = aa.next = 0;
                    orc_switch_state(m, 0);
                    orc_bond_force_general(m, m->plmode);
                    orc_update_rr(m);
                    nres = orc_nrm2(d * n, m->residual);
                }
                nprev = nres;
                /* stagnation: less than a halving over the last `depth` iterations
                 * starts the differences afresh */
                since++;
                if (aa.depth > 0 && since > aa.depth && nres > 0.5 * ring[(since - aa.depth) % AA_RING])
                {
                    aa.cnt = aa.next = 0;
                    since = 0;
                }
                ring[since % AA_RING] = nres;
            }
            if (ni >= s->newton_maxit && !first_fail)
                first_fail = st + 1;
            o.newton_iters += ni;
            o.res_norm = nres;
            o.tol_mult = tolm;
            int broken = s->skip_damage ? 0 : damage_general(m, st + 1 + o_off);
            orc_update_crack(m);
            orc_switch_state(m, 1);
            write_outputs(m, st + 1 + o_off);
            o.broken += broken;
            o.reaction = orc_reaction(m, out_axis, out_type);
            o.disp = orc_mean_disp(m, out_axis, out_type);
            if (broken > 0)
            {
                o.resolves++;
                if (s->tangent == 0)
                    orc_tangent_fd(m);
                else
                    orc_tangent_analytic(m);
                continue;
            }
            break;
        }
        if (out)
            out[st] = o;
    }
    free(aa.dU), free(aa.dF), free(aa.up), free(aa.fp);
    return first_fail;
}
