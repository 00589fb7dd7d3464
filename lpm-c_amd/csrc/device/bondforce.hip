// Constitutive laws, stress, state slots, residual, boundary conditions, damage,
// crack bookkeeping and outputs on the device.  One thread per particle, bond data
// in slot planes, so a wavefront reads 64 consecutive particles' slot j at once.
//
// The reference evaluates each particle's law on a list of the particle and its
// unbroken partners, recomputing the partners' sums and return maps every time
// (src/constitutive.c:466-686).  Each particle's result is a pure function of the
// positions and the slot-0 state, so here every particle runs its return map
// once (k_j2_particle) and the final forces gather the partners' results
// (k_j2_force).  Arithmetic follows the reference's evaluation order; this file is
// compiled without floating-point contraction.
#include <math.h>
#include <string.h>

#include <vector>

#include "common.h"

#include "bond_common.h"

// ------------------------------------------------------------ elastic law -------
__global__ void k_elastic_pass(DV v, int first, int count)
{
    int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= first + count)
        return;
    double a0, a1, b0, b1;
    stretch_pass(v, i, v.dLp0, a0, a1, b0, b1);
    v.A[i] = a0, v.A[(int64_t)v.n + i] = a1;
    v.Bt[i] = b0, v.Bt[(int64_t)v.n + i] = b1;
}

// computeBondForceElastic (constitutive.c:263-279): F from the half sums of both
// ends, times damage_broken; Pin = sum cs F.
__global__ void k_elastic_force(DV v, int first, int count)
{
    int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= first + count)
        return;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b], s = v.nsign[b];
        double f = 2.0 * v.Kn[b] * v.dL[b] + 0.5 * (v.Bt[(int64_t)s * v.n + i] + v.Bt[(int64_t)s * v.n + q]) +
                   0.5 * v.Tv[b] * (v.A[(int64_t)s * v.n + i] + v.A[(int64_t)s * v.n + q]);
        f *= v.db[b];
        v.F[b] = f;
        p0 += v.csx[b] * f;
        p1 += v.csy[b] * f;
        p2 += v.csz[b] * f;
    }
    v.Pin[3 * i] = p0, v.Pin[3 * i + 1] = p1, v.Pin[3 * i + 2] = p2;
}

// computedL (lpm_basic.c:252-291): no damage factor, writes distance as well.
__global__ void k_computedL(DV v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b];
        const double dis = dist_xyz(v.x, i, q);
        v.dist[b] = dis;
        double dl = dis - v.L0[b];
        dl -= v.dLp0[b];
        v.dL[b] = dl;
        if (v.nsign[b] == 0)
            a0 += dl, b0 += v.Tv[b] * dl;
        else
            a1 += dl, b1 += v.Tv[b] * dl;
        v.csx[b] = (v.x[3 * i] - v.x[3 * q]) / dis;
        v.csy[b] = (v.x[3 * i + 1] - v.x[3 * q + 1]) / dis;
        v.csz[b] = (v.x[3 * i + 2] - v.x[3 * q + 2]) / dis;
    }
    v.A[i] = a0, v.A[(int64_t)v.n + i] = a1;
    v.Bt[i] = b0, v.Bt[(int64_t)v.n + i] = b1;
}

// -------------------------------------------------------------- J2 law ---------
// Per particle: trial stress from its own sums, radial return with linear mixed
// hardening, plastic bond stretch, second stretch pass (constitutive.c:495-645,
// 669-675).  Slot 2 of the state receives the updated values.
__global__ void k_j2_particle(DV v, int first, int count)
{
    int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= first + count)
        return;
    const int n = v.n, nb0 = v.nb0[i];
    double a0, a1, b0, b1;
    stretch_pass(v, i, v.dLp0, a0, a1, b0, b1);
    double sl[6] = {0, 0, 0, 0, 0, 0}, dep[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nb0; j++)
    {
        const int64_t b = BI(j, i);
        const int s = v.nsign[b];
        double Fij = 2.0 * v.Kn[b] * v.dL[b] + (s == 0 ? b0 : b1) + v.Tv[b] * (s == 0 ? a0 : a1);
        Fij *= v.dw[b];
        const double of = opp_flag(v, i, j);
        const double cx = v.csx[b], cy = v.csy[b], cz = v.csz[b], L0 = v.L0[b];
        sl[0] += of / v.volume * L0 * Fij * cx * cx;
        sl[1] += of / v.volume * L0 * Fij * cy * cy;
        sl[2] += of / v.volume * L0 * Fij * cz * cz;
        sl[3] += of / v.volume * L0 * Fij * cy * cz;
        sl[4] += of / v.volume * L0 * Fij * cx * cz;
        sl[5] += of / v.volume * L0 * Fij * cx * cy;
    }
    const double tmp = 1.0 / 3.0 * (sl[0] + sl[1] + sl[2]);
    for (int c = 0; c < 3; c++)
        sl[c] -= tmp;
    double beta[6];
    for (int c = 0; c < 6; c++)
    {
        beta[c] = v.beta0[(int64_t)c * n + i];
        sl[c] -= beta[c];
    }
    double seq = 0.0;
    for (int c = 0; c < 6; c++)
    {
        if (c < 3)
            seq += sl[c] * sl[c];
        else
            seq += 2.0 * sl[c] * sl[c];
    }
    seq = sqrt(3.0 / 2.0 * seq);
    double alpha = v.alpha0[i], dlam = 0.0;
    const double yf = seq - (v.sigmay[i] + (1.0 - v.xi) * v.H * alpha);
    if (yf > 0.0)
    {
        v.pl_flag[i] = 1;
        dlam = yf / (3 * v.c44[i] + v.H);
    }
    alpha += dlam;
    for (int c = 0; c < 6; c++)
        if (fabs(seq) > LPMX_EPS)
        {
            dep[c] = dlam * 1.5 * sl[c] / seq;
            beta[c] += 2. / 3. * v.xi * v.H * dep[c];
        }
    // plastic bond stretch; dLp2 holds the updated plastic stretch meanwhile
    for (int j = 0; j < v.nb; j++)
    {
        const int64_t b = BI(j, i);
        double xp = v.dLp0[b];
        if (j < nb0)
        {
            const double cx = v.csx[b], cy = v.csy[b], cz = v.csz[b];
            double d = v.L0[b] * (dep[0] * cx * cx + dep[1] * cy * cy + dep[2] * cz * cz + 2 * dep[3] * cy * cz +
                                  2 * dep[4] * cx * cz + 2 * dep[5] * cx * cy);
            d *= v.db[b];
            v.ddLp[b] = d;
            xp += d;
        }
        v.dLp2[b] = xp;
    }
    stretch_pass(v, i, v.dLp2, a0, a1, b0, b1);
    for (int j = 0; j < v.nb; j++)
    {
        const int64_t b = BI(j, i);
        v.dLp2[b] = v.db[b] * v.dLp2[b];
    }
    v.A[i] = a0, v.A[(int64_t)n + i] = a1;
    v.Bt[i] = b0, v.Bt[(int64_t)n + i] = b1;
    for (int c = 0; c < 6; c++)
        v.beta2[(int64_t)c * n + i] = beta[c];
    v.alpha2[i] = alpha;
    v.dlam[i] = dlam;
}

// k_j2_particle with the bond loops unrolled to NBM list slots and every array a
// separate restrict-qualified argument: the loads of all bonds can then be issued
// ahead of the stores (the generic kernel's pointers may alias, which serialises the
// dependent index and position loads of one bond after another).  Same arithmetic in
// the same order, so the same bits.  The kernel is bound by memory latency, so it is
// laid out for occupancy (125 VGPRs, four waves per SIMD):
// - registers keep across the three bond loops only each bond's neighbour index,
//   elastic stretch and shell bit; the distance and direction cosines are recomputed
//   from the neighbour's position in the two later loops (the same expressions, so
//   the same values; the positions come back from cache), which halves the
//   registers against keeping them (234 VGPRs, two waves per SIMD: 41% slower);
// - L0 is read from HBM once and kept in LDS for the later loops, one column per
//   thread (no thread reads another's column, so no barrier), 18 KiB per work-group;
// - damage factors that are exactly 1.0 or +0.0 are kept as bits, so the later loops
//   and the opposite-bond factor do not read them again;
// - read-once streams are loaded and write-once streams stored nontemporally.
#define J2_WG 128
template <int NBM>
__global__ __launch_bounds__(J2_WG) void k_j2_particle_u(
    int n, int nbmax, double volume, double H, double xi, const int *__restrict__ nb0_,
    const int *__restrict__ nbcur, const int *__restrict__ nbr, const int *__restrict__ nsign,
    const int *__restrict__ opp, const double *__restrict__ X, const double *__restrict__ L0,
    const double *__restrict__ dLp0, const double *__restrict__ db, const double *__restrict__ dw,
    const double *__restrict__ Kn, const double *__restrict__ Tv, const double *__restrict__ beta0,
    const double *__restrict__ alpha0, const double *__restrict__ sigmay, const double *__restrict__ c44,
    double *__restrict__ dL, double *__restrict__ csx, double *__restrict__ csy, double *__restrict__ csz,
    double *__restrict__ ddLp, double *__restrict__ dLp2, double *__restrict__ A, double *__restrict__ Bt,
    double *__restrict__ beta2, double *__restrict__ alpha2, double *__restrict__ dlam_, int *__restrict__ pl_flag)
{
    static_assert(NBM <= 32, "the shell bits of a particle's bonds share one 32-bit word");
    __shared__ double sL0[NBM][J2_WG];
    const int tl = threadIdx.x;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int nb0 = nb0_[i];
    const double xi0 = X[3 * i], xi1 = X[3 * i + 1], xi2 = X[3 * i + 2];
    int qs[NBM];
    double rdl[NBM];
    unsigned sbit = 0, one = 0, zero = 0;
    double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
#pragma unroll
    for (int j = 0; j < NBM; j++)
        if (j < nb0)
        {
            const int64_t b = (int64_t)j * n + i;
            const int q = __builtin_nontemporal_load(&nbr[b]);
            qs[j] = q;
            const double dx = xi0 - X[3 * q], dy = xi1 - X[3 * q + 1], dz = xi2 - X[3 * q + 2];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            const double l0 = __builtin_nontemporal_load(&L0[b]), tv = Tv[b],
                         dbv = __builtin_nontemporal_load(&db[b]);
            sL0[j][tl] = l0;
            const unsigned long long u = (unsigned long long)__double_as_longlong(dbv);
            if (u == 0x3FF0000000000000ULL)
                one |= 1u << j;
            else if (u == 0ULL)
                zero |= 1u << j;
            double dl = d - l0;
            dl -= dLp0[b];
            dl *= dbv;
            rdl[j] = dl;
            if (__builtin_nontemporal_load(&nsign[b]) == 0)
                a0 += dl, b0 += tv * dl;
            else
            {
                a1 += dl, b1 += tv * dl;
                sbit |= 1u << j;
            }
            __builtin_nontemporal_store(dx / d, &csx[b]);
            __builtin_nontemporal_store(dy / d, &csy[b]);
            __builtin_nontemporal_store(dz / d, &csz[b]);
        }
    const bool full = nbcur[i] == nbmax;
    double sl[6] = {0, 0, 0, 0, 0, 0}, dep[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < NBM; j++)
        if (j < nb0)
        {
            const int64_t b = (int64_t)j * n + i;
            const int q = qs[j];
            const double dx = xi0 - X[3 * q], dy = xi1 - X[3 * q + 1], dz = xi2 - X[3 * q + 2];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            const bool s1 = (sbit >> j) & 1u;
            double Fij = 2.0 * __builtin_nontemporal_load(&Kn[b]) * rdl[j] + (s1 ? b1 : b0) + Tv[b] * (s1 ? a1 : a0);
            Fij *= __builtin_nontemporal_load(&dw[b]);
            double of = 0.5;
            if (!full)
            {
                const int o = opp[b];
                if (o < 0)
                    of = 1.0;
                else
                {
                    double dbo;
                    if ((one >> o) & 1u)
                        dbo = 1.0;
                    else if ((zero >> o) & 1u)
                        dbo = 0.0;
                    else
                        dbo = db[(int64_t)o * n + i];
                    of = dbo <= LPMX_EPS ? 1.0 : 0.5;
                }
            }
            const double cx = dx / d, cy = dy / d, cz = dz / d, l0 = sL0[j][tl];
            sl[0] += of / volume * l0 * Fij * cx * cx;
            sl[1] += of / volume * l0 * Fij * cy * cy;
            sl[2] += of / volume * l0 * Fij * cz * cz;
            sl[3] += of / volume * l0 * Fij * cy * cz;
            sl[4] += of / volume * l0 * Fij * cx * cz;
            sl[5] += of / volume * l0 * Fij * cx * cy;
        }
    const double tmp = 1.0 / 3.0 * (sl[0] + sl[1] + sl[2]);
    for (int c = 0; c < 3; c++)
        sl[c] -= tmp;
    double beta[6];
    for (int c = 0; c < 6; c++)
    {
        beta[c] = beta0[(int64_t)c * n + i];
        sl[c] -= beta[c];
    }
    double seq = 0.0;
    for (int c = 0; c < 6; c++)
    {
        if (c < 3)
            seq += sl[c] * sl[c];
        else
            seq += 2.0 * sl[c] * sl[c];
    }
    seq = sqrt(3.0 / 2.0 * seq);
    double alpha = alpha0[i], dlam = 0.0;
    const double yf = seq - (sigmay[i] + (1.0 - xi) * H * alpha);
    if (yf > 0.0)
    {
        pl_flag[i] = 1;
        dlam = yf / (3 * c44[i] + H);
    }
    alpha += dlam;
    for (int c = 0; c < 6; c++)
        if (fabs(seq) > LPMX_EPS)
        {
            dep[c] = dlam * 1.5 * sl[c] / seq;
            beta[c] += 2. / 3. * xi * H * dep[c];
        }
    a0 = 0, a1 = 0, b0 = 0, b1 = 0;
#pragma unroll
    for (int j = 0; j < NBM; j++)
        if (j < nbmax)
        {
            const int64_t b = (int64_t)j * n + i;
            double dbb;
            if (j < nb0 && ((one >> j) & 1u))
                dbb = 1.0;
            else if (j < nb0 && ((zero >> j) & 1u))
                dbb = 0.0;
            else
                dbb = db[b];
            double x_ = dLp0[b];
            if (j < nb0)
            {
                const int q = qs[j];
                const double dx = xi0 - X[3 * q], dy = xi1 - X[3 * q + 1], dz = xi2 - X[3 * q + 2];
                const double dd = sqrt(dx * dx + dy * dy + dz * dz);
                const double cx = dx / dd, cy = dy / dd, cz = dz / dd, l0 = sL0[j][tl];
                double d = l0 * (dep[0] * cx * cx + dep[1] * cy * cy + dep[2] * cz * cz + 2 * dep[3] * cy * cz +
                                 2 * dep[4] * cx * cz + 2 * dep[5] * cx * cy);
                d *= dbb;
                __builtin_nontemporal_store(d, &ddLp[b]);
                x_ += d;
                double dl = dd - l0;
                dl -= x_;
                dl *= dbb;
                __builtin_nontemporal_store(dl, &dL[b]);
                const double tv = Tv[b];
                if (((sbit >> j) & 1u) == 0)
                    a0 += dl, b0 += tv * dl;
                else
                    a1 += dl, b1 += tv * dl;
            }
            __builtin_nontemporal_store(dbb * x_, &dLp2[b]);
        }
    A[i] = a0, A[(int64_t)n + i] = a1;
    Bt[i] = b0, Bt[(int64_t)n + i] = b1;
    for (int c = 0; c < 6; c++)
        beta2[(int64_t)c * n + i] = beta[c];
    alpha2[i] = alpha;
    dlam_[i] = dlam;
}

// The J2 return maps of all particles: the unrolled kernel for lists of up to 18
// slots (SC, FCC), the generic one otherwise.
static void launch_j2_particles(void)
{
    const dim3 G(nblocks(g.n)), B(LPMX_BLOCK);
    if (g.nb <= 18)
        hipLaunchKernelGGL(k_j2_particle_u<18>, dim3(nblocks(g.n, J2_WG)), dim3(J2_WG), 0, g_stream, g.n, g.nb, g.volume, g.H, g.xi, g.nb0, g.nbcur,
                           g.nbr, g.nsign, g.opp, g.x, g.L0, g.dLp0, g.db, g.dw, g.Kn, g.Tv, g.beta0, g.alpha0,
                           g.sigmay, g.c44, g.dL, g.csx, g.csy, g.csz, g.ddLp, g.dLp2, g.A, g.Bt, g.beta2, g.alpha2,
                           g.dlam, g.pl_flag);
    else
        hipLaunchKernelGGL(k_j2_particle, G, B, 0, g_stream, g, 0, g.n);
}

// Final J2 force of particle i from the half-averaged elastic stretch of the bond
// and its reverse bond, times damage_w (constitutive.c:647-667).
__global__ void k_j2_force(DV v, int first, int count)
{
    int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= first + count)
        return;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b], s = v.nsign[b], r = v.rev[b];
        double dla = v.dLave[b];
        if (r >= 0)
            dla = 0.5 * (v.dL[b] + v.dL[BI(r, q)]);
        v.dLave[b] = dla;
        double f = 2.0 * v.Kn[b] * dla + 0.5 * (v.Bt[(int64_t)s * v.n + i] + v.Bt[(int64_t)s * v.n + q]) +
                   0.5 * v.Tv[b] * (v.A[(int64_t)s * v.n + i] + v.A[(int64_t)s * v.n + q]);
        f *= v.dw[b];
        __builtin_nontemporal_store(f, &v.F[b]);
        p0 += v.csx[b] * f;
        p1 += v.csy[b] * f;
        p2 += v.csz[b] * f;
    }
    __builtin_nontemporal_store(p0, &v.Pin[3 * i]);
    __builtin_nontemporal_store(p1, &v.Pin[3 * i + 1]);
    __builtin_nontemporal_store(p2, &v.Pin[3 * i + 2]);
}

// ------------------------------------------------------------ predictor ---------
// computeBondForceIncrementalUpdating (constitutive.c:167-225): stretch increments
// from xyz_temp, F = F_temp + increments, Pin with the existing cs.
__global__ void k_pred_pass(DV v, int first, int count)
{
    int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= first + count)
        return;
    double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b];
        const double d0 = dist_xyz(v.xt, i, q), d1 = dist_xyz(v.x, i, q);
        const double dd = v.db[b] * (d1 - d0);
        v.ddL[b] = dd;
        if (v.nsign[b] == 0)
            a0 += dd, b0 += v.Tv[b] * dd;
        else
            a1 += dd, b1 += v.Tv[b] * dd;
    }
    v.dA_[i] = a0, v.dA_[(int64_t)v.n + i] = a1;
    v.dB_[i] = b0, v.dB_[(int64_t)v.n + i] = b1;
}

__global__ void k_pred_force(DV v, int first, int count)
{
    int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= first + count)
        return;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b], s = v.nsign[b];
        double f = v.Ft[b] + 2.0 * v.Kn[b] * v.ddL[b] + 0.5 * (v.dB_[(int64_t)s * v.n + i] + v.dB_[(int64_t)s * v.n + q]) +
                   0.5 * v.Tv[b] * (v.dA_[(int64_t)s * v.n + i] + v.dA_[(int64_t)s * v.n + q]);
        f *= v.db[b];
        v.F[b] = f;
        p0 += v.csx[b] * f;
        p1 += v.csy[b] * f;
        p2 += v.csz[b] * f;
    }
    v.Pin[3 * i] = p0, v.Pin[3 * i + 1] = p1, v.Pin[3 * i + 2] = p2;
}

// -------------------------------------------------------------- stress ----------
// computeStress (lpm_basic.c:53-125).
__global__ void k_stress(DV v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int n = v.n;
    double st[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const double of = opp_flag(v, i, j);
        const double cx = v.csx[b], cy = v.csy[b], cz = v.csz[b], L0 = v.L0[b], f = v.F[b];
        st[0] += of / v.volume * L0 * f * cx * cx;
        st[1] += of / v.volume * L0 * f * cy * cy;
        st[2] += of / v.volume * L0 * f * cz * cz;
        st[3] += of / v.volume * L0 * f * cy * cz;
        st[4] += of / v.volume * L0 * f * cx * cz;
        st[5] += of / v.volume * L0 * f * cx * cy;
    }
    double seq = 0.0;
    for (int c = 0; c < 6; c++)
    {
        v.stress[(int64_t)c * n + i] = st[c];
        if (c < 3)
            seq += st[c] * st[c];
        else
            seq += 2.0 * st[c] * st[c];
    }
    seq = sqrt(3.0 / 2.0 * seq);
    const double sm = 1.0 / 3.0 * (st[0] + st[1] + st[2]);
    v.seq[i] = seq;
    v.sm[i] = sm;
    if (seq > LPMX_EPS)
        v.triax[i] = sm / seq;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const double cx = v.csx[b], cy = v.csy[b], cz = v.csz[b];
        v.bstress[b] = (st[0] * cx * cx + st[1] * cy * cy + st[2] * cz * cz + 2 * st[3] * cy * cz + 2 * st[4] * cx * cz +
                        2 * st[5] * cx * cy);
    }
}

// --------------------------------------------------------------- slots ----------
// switchStateV (constitutive.c:10-85) for the slots of the laws on this path.
__global__ void k_switch(DV v, int flag)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nbn = (int64_t)v.n * v.nb;
    if (t < nbn)
    {
        if (flag == 0)
            v.dLp0[t] = v.dLp1[t], v.D0[t] = v.D1[t];
        else if (flag == 1)
            v.dLp1[t] = v.dLp0[t], v.D1[t] = v.D0[t];
        else
            v.dLp0[t] = v.dLp2[t];
    }
    if (t < 6 * (int64_t)v.n)
    {
        if (flag == 0)
            v.beta0[t] = v.beta1[t];
        else if (flag == 1)
            v.beta1[t] = v.beta0[t];
        else
            v.beta0[t] = v.beta2[t];
    }
    if (t < v.n)
    {
        if (flag == 0)
        {
            v.alpha0[t] = v.alpha1[t], v.beq0[t] = v.beq1[t];
            v.dloc0[t] = v.dloc1[t], v.dnl0[t] = v.dnl1[t];
        }
        else if (flag == 1)
        {
            v.alpha1[t] = v.alpha0[t], v.beq1[t] = v.beq0[t];
            v.dloc1[t] = v.dloc0[t], v.dnl1[t] = v.dnl0[t];
        }
        else
            v.alpha0[t] = v.alpha2[t], v.beq0[t] = v.beq2[t];
        if (v.nslip > 0)
        {
            if (flag == 0)
                v.cp_A0[t] = v.cp_A1[t];
            else if (flag == 1)
                v.cp_A1[t] = v.cp_A0[t];
            else
                v.cp_A0[t] = v.cp_A2[t];
        }
    }
    if (t < (int64_t)v.nslip * v.n)
    {
        if (flag == 0)
            v.cp_gy0[t] = v.cp_gy1[t], v.cp_As0[t] = v.cp_As1[t];
        else if (flag == 1)
            v.cp_gy1[t] = v.cp_gy0[t], v.cp_As1[t] = v.cp_As0[t];
        else
            v.cp_gy0[t] = v.cp_gy2[t], v.cp_As0[t] = v.cp_As2[t];
    }
}

extern "C" int lpmx_switch_state(int flag)
{
    RETCHK(lpmx_require_model());
    if (flag < 0 || flag > 2)
        return lpmx_fail(LPMX_ERR_ARG, "switchStateV flag %d", flag);
    int w = g.nb > 6 ? g.nb : 6;
    if (g.nslip > w)
        w = g.nslip;
    const int64_t m = (int64_t)g.n * w;
    hipLaunchKernelGGL(k_switch, dim3(nblocks(m)), dim3(LPMX_BLOCK), 0, g_stream, g, flag);
    LAUNCHCHK();
    return LPMX_OK;
}

extern "C" int lpmx_compute_stress(void)
{
    RETCHK(lpmx_require_model());
    hipLaunchKernelGGL(k_stress, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g);
    LAUNCHCHK();
    return LPMX_OK;
}

extern "C" int lpmx_computedL(void)
{
    RETCHK(lpmx_require_model());
    hipLaunchKernelGGL(k_computedL, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g);
    LAUNCHCHK();
    return LPMX_OK;
}

// HIP-event timing of the k_j2_particle launches while on: each launch's pair is
// read at the next launch (or at the read), after its end event has completed.
static struct ForceProf
{
    bool on = false, have_ev = false, pending = false;
    hipEvent_t e0, e1;
    double ms = 0.0;
    int64_t n = 0;
} FP;

static int force_prof_harvest(void)
{
    if (!FP.pending)
        return LPMX_OK;
    HIPCHK(hipEventSynchronize(FP.e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, FP.e0, FP.e1));
    FP.ms += ms, FP.n++;
    FP.pending = false;
    return LPMX_OK;
}

extern "C" void lpmx_profile_force(int on)
{
    FP.on = on != 0, FP.pending = false, FP.ms = 0.0, FP.n = 0;
    if (FP.on && !FP.have_ev)
    {
        hipEventCreate(&FP.e0), hipEventCreate(&FP.e1);
        FP.have_ev = true;
    }
}

extern "C" int lpmx_profile_force_read(double *total_ms, int64_t *launches)
{
    RETCHK(force_prof_harvest());
    *total_ms = FP.ms;
    *launches = FP.n;
    return LPMX_OK;
}

// Algorithmic HBM bytes of one k_j2_particle launch: per particle its own position,
// counts, yield stress, hardening state and slot-0 back stress read, its shell sums,
// slot-2 state and flag written; per bond the partner index and sign, L0, slot-0
// plastic stretch, both damage factors, Kn, Tv read and dL, the direction cosines,
// ddLp and the slot-2 plastic stretch written (gathered partner positions are
// re-reads from cache); the slot-2 plastic stretch of every list slot.
extern "C" int lpmx_j2_particle_bytes(double *bytes)
{
    RETCHK(lpmx_require_model());
    std::vector<int> nb0(g.n);
    HIPCHK(hipMemcpyAsync(nb0.data(), g.nb0, sizeof(int) * (size_t)g.n, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    int64_t nbond = 0;
    for (int c : nb0)
        nbond += c;
    *bytes = (double)g.n * (4 + 4 + 24 + 24 + 48) + (double)g.n * (32 + 48 + 16 + 4) + (double)nbond * (4 + 4 + 6 * 8) +
             (double)nbond * (5 * 8) + (double)g.n * g.nb * 8;
    return LPMX_OK;
}

static int bond_force_range(int mode, int t, int first, int count)
{
    RETCHK(lpmx_require_model());
    if (first < 0 || count < 0 || first + count > g.n)
        return lpmx_fail(LPMX_ERR_ARG, "particle range [%d, %d) outside [0, %d)", first, first + count, g.n);
    if (count == 0)
        return LPMX_OK;
    const dim3 G(nblocks(count)), B(LPMX_BLOCK);
    if (mode == 3 || mode == 5)
    {
        // plmode 5 advances the slot-0 state of every particle in place: no subrange
        if (first != 0 || count != g.n)
            return lpmx_fail(LPMX_ERR_ARG, "constitutive mode %d evaluates every particle", mode);
        return mode == 3 ? lpmx_law3(t) : lpmx_law5();
    }
    if (mode == 0)
    {
        // the final force of a particle needs its partners' return maps too
        if (FP.on)
        {
            RETCHK(force_prof_harvest());
            HIPCHK(hipEventRecord(FP.e0, g_stream));
        }
        launch_j2_particles();
        LAUNCHCHK();
        if (FP.on)
        {
            HIPCHK(hipEventRecord(FP.e1, g_stream));
            FP.pending = true;
        }
        hipLaunchKernelGGL(k_j2_force, G, B, 0, g_stream, g, first, count);
    }
    else if (mode == 1)
    {
        RETCHK(lpmx_cp_particles());
        hipLaunchKernelGGL(k_j2_force, G, B, 0, g_stream, g, first, count);
    }
    else if (mode == 4)
    {
        hipLaunchKernelGGL(k_pred_pass, dim3(nblocks(g.n)), B, 0, g_stream, g, 0, g.n);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_pred_force, G, B, 0, g_stream, g, first, count);
    }
    else if (mode == 6)
    {
        hipLaunchKernelGGL(k_elastic_pass, dim3(nblocks(g.n)), B, 0, g_stream, g, 0, g.n);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_elastic_force, G, B, 0, g_stream, g, first, count);
    }
    else
        return lpmx_fail(LPMX_ERR_ARG, "constitutive mode %d is not a law of the reference (0, 1, 3, 4, 5, 6 are)",
                         mode);
    LAUNCHCHK();
    return LPMX_OK;
}

extern "C" int lpmx_bond_force_range(int mode, int first, int count) { return bond_force_range(mode, 1, first, count); }

extern "C" int lpmx_bond_force_range_ex(int mode, int t, int first, int count)
{
    return bond_force_range(mode, t, first, count);
}

int lpmx_eval_j2(void)
{
    const dim3 G(nblocks(g.n)), B(LPMX_BLOCK);
    launch_j2_particles();
    LAUNCHCHK();
    hipLaunchKernelGGL(k_j2_force, G, B, 0, g_stream, g, 0, g.n);
    LAUNCHCHK();
    return LPMX_OK;
}

extern "C" int lpmx_bond_force_ex(int mode, int t)
{
    RETCHK(bond_force_range(mode, t, 0, g.n));
    RETCHK(lpmx_compute_stress());
    return lpmx_switch_state(2);
}

extern "C" int lpmx_bond_force(int mode)
{
    return lpmx_bond_force_ex(mode, 1);
}

// ------------------------------------------------------------ residual ----------
// updateRR (stiffness.c:519-534): residual = dispBC_index (Pex - Pin); squared
// sums of the residual and of Pin on constrained DOFs for the norms.
__global__ void k_update_rr(DV v, double *sq_res, double *sq_rea)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < v.dim; k++)
    {
        const int64_t o = (int64_t)v.dim * i + k;
        const double r = v.dispBC[o] * (v.Pex[o] - v.Pin[3 * i + k]);
        v.res[o] = r;
        a += r * r;
        if (v.dispBC[o] == 0)
            b += v.Pin[3 * i + k] * v.Pin[3 * i + k];
    }
    sq_res[i] = a;
    sq_rea[i] = b;
}

__global__ void k_count_bc(const int *dispBC, int m, int *cnt)
{
    __shared__ int sh;
    if (threadIdx.x == 0)
        sh = 0;
    __syncthreads();
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < m && dispBC[t] == 0)
        atomicAdd(&sh, 1);
    __syncthreads();
    if (threadIdx.x == 0 && sh)
        atomicAdd(cnt, sh);
}

// per-particle squares of the residual and reaction norms (2n), sized for the model
static double *u_sq = nullptr;
static int *u_cnt = nullptr;
static int64_t u_cap = 0;

static void update_rr_free(void)
{
    if (u_sq)
        hipFree(u_sq), hipFree(u_cnt);
    u_sq = nullptr, u_cnt = nullptr, u_cap = 0;
}

extern "C" int lpmx_update_rr(double *nres, double *nrea, int *nreaction)
{
    RETCHK(lpmx_require_model());
    const int n = g.n;
    if (u_cap < n)
    {
        update_rr_free();
        HIPCHK(hipMalloc(&u_sq, sizeof(double) * 2 * (int64_t)n));
        HIPCHK(hipMalloc(&u_cnt, sizeof(int)));
        u_cap = n;
    }
    hipLaunchKernelGGL(k_update_rr, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, u_sq, u_sq + n);
    LAUNCHCHK();
    double a = 0, b = 0;
    if (nres)
    {
        RETCHK(lpmx_dev_sum(u_sq, n, &a));
        *nres = sqrt(a);
    }
    if (nrea)
    {
        RETCHK(lpmx_dev_sum(u_sq + n, n, &b));
        *nrea = sqrt(b);
    }
    if (nreaction)
    {
        HIPCHK(hipMemsetAsync(u_cnt, 0, sizeof(int), g_stream));
        hipLaunchKernelGGL(k_count_bc, dim3(nblocks((int64_t)n * g.dim)), dim3(LPMX_BLOCK), 0, g_stream, g.dispBC,
                           n * g.dim, u_cnt);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(nreaction, u_cnt, sizeof(int), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    return LPMX_OK;
}

__global__ void k_flag_bc(const int *dispBC, int m, int *flag)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < m)
        flag[t] = dispBC[t] == 0;
}

__global__ void k_pack_reaction(DV v, const int *pos, double *out)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.dim)
        return;
    if (v.dispBC[t] == 0)
    {
        const int i = (int)(t / v.dim), k = (int)(t % v.dim);
        out[pos[t]] = v.Pin[3 * i + k];
    }
}

// reaction_force in the reference's packed DOF order (stiffness.c:529-531).
extern "C" int lpmx_get_reaction_force(double *host, int count)
{
    RETCHK(lpmx_require_model());
    const int m = g.n * g.dim;
    int *flag, *pos;
    double *out;
    HIPCHK(hipMalloc(&flag, sizeof(int) * (m + 1)));
    HIPCHK(hipMalloc(&pos, sizeof(int) * (m + 1)));
    hipLaunchKernelGGL(k_flag_bc, dim3(nblocks(m)), dim3(LPMX_BLOCK), 0, g_stream, g.dispBC, m, flag);
    LAUNCHCHK();
    int tot = 0;
    RETCHK(lpmx_dev_scan_int(flag, pos, m, &tot));
    HIPCHK(hipMalloc(&out, sizeof(double) * (tot + 1)));
    hipLaunchKernelGGL(k_pack_reaction, dim3(nblocks(m)), dim3(LPMX_BLOCK), 0, g_stream, g, pos, out);
    LAUNCHCHK();
    if (count < tot)
        return lpmx_fail(LPMX_ERR_ARG, "reaction buffer %d < %d", count, tot);
    HIPCHK(hipMemcpyAsync(host, out, sizeof(double) * tot, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(flag), hipFree(pos), hipFree(out);
    return LPMX_OK;
}

// ------------------------------------------------------- boundary conditions -----
struct BCs
{
    int n;
    int type[LPMX_MAXBC];
    int comp[LPMX_MAXBC];
    double step[LPMX_MAXBC][3];
    double inv_count[LPMX_MAXBC];
};

// setDispBC (boundary.c:12-45).
__global__ void k_disp_bc(DV v, BCs bc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int t = v.type[i];
    for (int k = 0; k < bc.n; k++)
        if (t == bc.type[k] && bc.comp[k] >= 0)
        {
            v.x[3 * i + bc.comp[k]] += bc.step[k][0];
            v.dispBC[(int64_t)v.dim * i + bc.comp[k]] = 0;
        }
}

// setForceBC (boundary.c:48-70): step / count of each entry's type.
__global__ void k_force_bc(DV v, BCs bc, const int *cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int t = v.type[i];
    for (int k = 0; k < bc.n; k++)
        if (t == bc.type[k])
        {
            const int c = cnt[k];
            v.Pex[(int64_t)v.dim * i] += bc.step[k][0] / c;
            v.Pex[(int64_t)v.dim * i + 1] += bc.step[k][1] / c;
            if (v.dim == 3)
                v.Pex[(int64_t)v.dim * i + 2] += bc.step[k][2] / c;
        }
}

__global__ void k_type_count(const int *type, int n, BCs bc, int *cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    for (int k = 0; k < bc.n; k++)
        if (type[i] == bc.type[k])
            atomicAdd(&cnt[k], 1);
}

extern "C" int lpmx_set_disp_bc(int nbd, const lpmx_dispbc *e)
{
    RETCHK(lpmx_require_model());
    if (nbd < 0 || nbd > LPMX_MAXBC || (nbd && !e))
        return lpmx_fail(LPMX_ERR_ARG, "displacement BC count %d", nbd);
    if (nbd == 0)
        return LPMX_OK;
    BCs bc;
    memset(&bc, 0, sizeof(bc));
    bc.n = nbd;
    for (int k = 0; k < nbd; k++)
    {
        bc.type[k] = e[k].type;
        bc.comp[k] = e[k].flag == 'x' ? 0 : e[k].flag == 'y' ? 1 : e[k].flag == 'z' ? 2 : -1;
        if (bc.comp[k] >= g.dim)
            bc.comp[k] = -1;
        bc.step[k][0] = e[k].step;
    }
    hipLaunchKernelGGL(k_disp_bc, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, bc);
    LAUNCHCHK();
    return LPMX_OK;
}

extern "C" int lpmx_set_force_bc(int nbf, const lpmx_forcebc *e)
{
    RETCHK(lpmx_require_model());
    if (nbf < 0 || nbf > LPMX_MAXBC || (nbf && !e))
        return lpmx_fail(LPMX_ERR_ARG, "force BC count %d", nbf);
    if (nbf == 0)
        return LPMX_OK;
    BCs bc;
    memset(&bc, 0, sizeof(bc));
    bc.n = nbf;
    for (int k = 0; k < nbf; k++)
    {
        bc.type[k] = e[k].type;
        for (int c = 0; c < 3; c++)
            bc.step[k][c] = e[k].step[c];
    }
    int *cnt;
    HIPCHK(hipMalloc(&cnt, sizeof(int) * LPMX_MAXBC));
    HIPCHK(hipMemsetAsync(cnt, 0, sizeof(int) * LPMX_MAXBC, g_stream));
    hipLaunchKernelGGL(k_type_count, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g.type, g.n, bc, cnt);
    LAUNCHCHK();
    int hc[LPMX_MAXBC];
    HIPCHK(hipMemcpyAsync(hc, cnt, sizeof(int) * LPMX_MAXBC, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    for (int k = 0; k < nbf; k++)
        if (hc[k] == 0)
        {
            hipFree(cnt);
            return lpmx_fail(LPMX_ERR_ARG, "force BC on type %d, which no particle has", bc.type[k]);
        }
    hipLaunchKernelGGL(k_force_bc, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g, bc, cnt);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(cnt);
    return LPMX_OK;
}

// ------------------------------------------------------------------ damage -------
// Static nonlocal neighbourhood: every j with |X_j - X_i| < 3L in the initial
// configuration, in increasing index, and the Gaussian weight DAM_PHI(dis).
static int d_built_n = -1;
static double d_built_L = -1;
static int *d_cnt = nullptr;

struct Grid3
{
    double lo[3], h;
    int nc[3];
};

__device__ __forceinline__ int gcoord(double v, double lo, double h, int nc)
{
    int c = (int)floor((v - lo) / h);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void k_grid_count(const double *x, int n, Grid3 gr, int *cid, int *cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int a = gcoord(x[3 * i], gr.lo[0], gr.h, gr.nc[0]), b = gcoord(x[3 * i + 1], gr.lo[1], gr.h, gr.nc[1]),
        c = gcoord(x[3 * i + 2], gr.lo[2], gr.h, gr.nc[2]);
    int id = (c * gr.nc[1] + b) * gr.nc[0] + a;
    cid[i] = id;
    atomicAdd(&cnt[id], 1);
}

__global__ void k_grid_fill(const int *cid, int n, const int *start, int *fillc, int *list)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    list[start[cid[i]] + atomicAdd(&fillc[cid[i]], 1)] = i;
}

// pass 0: count; pass 1: write the sorted neighbourhood and weights
__global__ void k_dnl_list(DV v, Grid3 gr, const int *start, const int *list, double cut, int pass, int *cnt,
                           const int *ptr, int *idx, double *wt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const double *X = v.x0;
    int ca = gcoord(X[3 * i], gr.lo[0], gr.h, gr.nc[0]), cb = gcoord(X[3 * i + 1], gr.lo[1], gr.h, gr.nc[1]),
        cc = gcoord(X[3 * i + 2], gr.lo[2], gr.h, gr.nc[2]);
    int c = 0;
    int64_t base = pass ? ptr[i] : 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++)
            {
                int qa = ca + dx, qb = cb + dy, qc = cc + dz;
                if (qa < 0 || qb < 0 || qc < 0 || qa >= gr.nc[0] || qb >= gr.nc[1] || qc >= gr.nc[2])
                    continue;
                int id = (qc * gr.nc[1] + qb) * gr.nc[0] + qa;
                for (int s = start[id]; s < start[id + 1]; s++)
                {
                    int j = list[s];
                    double ddx = X[3 * j] - X[3 * i], ddy = X[3 * j + 1] - X[3 * i + 1], ddz = X[3 * j + 2] - X[3 * i + 2];
                    double dis = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                    if (!(dis < cut))
                        continue;
                    if (pass)
                    {
                        // insertion by index into this particle's segment
                        int p = c;
                        while (p > 0 && idx[base + p - 1] > j)
                        {
                            idx[base + p] = idx[base + p - 1];
                            wt[base + p] = wt[base + p - 1];
                            p--;
                        }
                        const double L = v.dmgL;
                        idx[base + p] = j;
                        wt[base + p] = 1.0 / L / sqrt(2 * 3.14159265358979323846) * exp(-0.5 * dis * dis / L / L);
                    }
                    c++;
                }
            }
    if (!pass)
        cnt[i] = c;
}

void lpmx_damage_free(void)
{
    update_rr_free();
    if (g.dptr)
        hipFree(g.dptr), hipFree(g.didx), hipFree(g.dwt);
    if (d_cnt)
        hipFree(d_cnt);
    d_cnt = nullptr;
    d_built_n = -1;
    d_built_L = -1;
}

static int build_damage_lists(void)
{
    if (d_built_n == g.n && d_built_L == g.dmgL && g.dptr)
        return LPMX_OK;
    if (g.dptr)
        hipFree(g.dptr), hipFree(g.didx), hipFree(g.dwt);
    g.dptr = nullptr;
    const int n = g.n;
    const double cut = 3 * g.dmgL;
    std::vector<double> hx(3 * (size_t)n);
    HIPCHK(hipMemcpyAsync(hx.data(), g.x0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    Grid3 gr;
    gr.h = cut * 1.0001;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++)
            lo[c] = fmin(lo[c], hx[3 * i + c]), hi[c] = fmax(hi[c], hx[3 * i + c]);
    int64_t ncell = 1;
    for (int c = 0; c < 3; c++)
    {
        gr.lo[c] = lo[c];
        gr.nc[c] = 1 + (int)floor((hi[c] - lo[c]) / gr.h);
        ncell *= gr.nc[c];
    }
    int *cid, *cnt, *start, *fillc, *list, *pc;
    HIPCHK(hipMalloc(&cid, sizeof(int) * n));
    HIPCHK(hipMalloc(&cnt, sizeof(int) * (ncell + 1)));
    HIPCHK(hipMalloc(&start, sizeof(int) * (ncell + 1)));
    HIPCHK(hipMalloc(&fillc, sizeof(int) * (ncell + 1)));
    HIPCHK(hipMalloc(&list, sizeof(int) * n));
    HIPCHK(hipMalloc(&pc, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&g.dptr, sizeof(int) * (n + 1)));
    HIPCHK(hipMemsetAsync(cnt, 0, sizeof(int) * (ncell + 1), g_stream));
    HIPCHK(hipMemsetAsync(fillc, 0, sizeof(int) * (ncell + 1), g_stream));
    HIPCHK(hipMemsetAsync(pc, 0, sizeof(int) * (n + 1), g_stream));
    hipLaunchKernelGGL(k_grid_count, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g.x0, n, gr, cid, cnt);
    LAUNCHCHK();
    int tot = 0;
    RETCHK(lpmx_dev_scan_int(cnt, start, (int)ncell + 1, &tot));
    hipLaunchKernelGGL(k_grid_fill, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, cid, n, start, fillc, list);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_dnl_list, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, gr, start, list, cut, 0, pc,
                       nullptr, nullptr, nullptr);
    LAUNCHCHK();
    int total = 0;
    RETCHK(lpmx_dev_scan_int(pc, g.dptr, n + 1, &total));
    HIPCHK(hipMalloc(&g.didx, sizeof(int) * (total + 1)));
    HIPCHK(hipMalloc(&g.dwt, sizeof(double) * (total + 1)));
    hipLaunchKernelGGL(k_dnl_list, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, gr, start, list, cut, 1, pc,
                       g.dptr, g.didx, g.dwt);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(cid), hipFree(cnt), hipFree(start), hipFree(fillc), hipFree(list), hipFree(pc);
    d_built_n = n;
    d_built_L = g.dmgL;
    return LPMX_OK;
}

// updateDuctileDamagePwiseNonlocal (constitutive.c:1798-1826).
__global__ void k_dnl(DV v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    double dn = v.dnl0[i];
    if (dn > v.dthr)
    {
        if (dn > 1.0)
            v.dnl0[i] = 1.0;
        return;
    }
    double Dd = 0, A = 0;
    for (int k = v.dptr[i]; k < v.dptr[i + 1]; k++)
    {
        const int j = v.didx[k];
        const double phi = v.dwt[k];
        double loc = 0;
        const double f = (1.0 + v.dA * v.triax[j]);
        if (f > 0.0)
            loc = v.dlam[j] * (1.0 + v.dA * v.triax[j]);
        Dd += loc * phi * v.volume;
        A += phi * v.volume;
    }
    if (Dd > 0.0)
        v.dnl0[i] = dn + 1.0 / A * Dd;
}

// bond breakage (constitutive.c:1828-1846)
__global__ void k_dmg_break(DV v, int *count)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.nb)
        return;
    const int i = (int)(t % v.n), j = (int)(t / v.n);
    if (j >= v.nb0[i])
        return;
    const int q = v.nbr[t];
    if (v.dnl0[i] > v.dthr || v.dnl0[q] > v.dthr)
        if (fabs(v.db[t]) > LPMX_EPS)
        {
            v.db[t] = 0.0;
            v.D0[t] = 1.0;
            atomicAdd(count, 1);
        }
}

// bond damage and intact weight (constitutive.c:1849-1858)
__global__ void k_dmg_w(DV v)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.nb)
        return;
    const int i = (int)(t % v.n), j = (int)(t / v.n);
    if (j >= v.nb0[i])
        return;
    if (fabs(v.db[t]) > LPMX_EPS)
    {
        const double a = v.dnl0[i], c = v.dnl0[v.nbr[t]];
        v.D0[t] = a < c ? c : a;
    }
    v.dw[t] = 1.0 - v.D0[t];
}

// Brittle law (constitutive.c:1437-1526): candidates dL/L0 >= critical in (i, slot)
// order; all break if at most nbreak, else the reference's shell sort picks the
// last nbreak.
__global__ void k_brittle_count(DV v, int *cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > v.n)
        return;
    if (i == v.n)
    {
        cnt[i] = 0;
        return;
    }
    int c = 0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const double crit = v.crit ? v.crit[b] : v.crit0;
        if (v.dL[b] / v.L0[b] >= crit)
            c++;
    }
    cnt[i] = c;
}

__global__ void k_brittle_fill(DV v, const int *off, int *ci, int *cj, double *cb)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    int p = off[i];
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const double crit = v.crit ? v.crit[b] : v.crit0;
        const double s = v.dL[b] / v.L0[b];
        if (s >= crit)
            ci[p] = i, cj[p] = j, cb[p] = s, p++;
    }
}

// The shell sort of the reference on k candidates (gaps k/2, k/4, ..., 1; strict
// '>' so equal keys keep the order the gaps leave them in).  One thread: the
// result must be the reference's permutation, ties included.
__global__ void k_shell_sort(int k, double *key, int *ci, int *cj)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    for (int r = k / 2; r >= 1; r = r / 2)
        for (int i = r; i < k; ++i)
        {
            const double tb = key[i];
            const int ti = ci[i], tj = cj[i];
            int j = i - r;
            while (j >= 0 && key[j] > tb)
            {
                key[j + r] = key[j];
                ci[j + r] = ci[j];
                cj[j + r] = cj[j];
                j = j - r;
            }
            key[j + r] = tb;
            ci[j + r] = ti;
            cj[j + r] = tj;
        }
}

__global__ void k_brittle_break(DV v, const int *ci, const int *cj, int from, int to, int *bi, int *bj)
{
    int t = from + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= to)
        return;
    const int i = ci[t], j = cj[t];
    const int64_t b = BI(j, i);
    v.D0[b] = 1.0;
    v.dw[b] = 0.0;
    v.db[b] = 0.0;
    bi[t - from] = i;
    bj[t - from] = v.nbr[b];
}

static int *b_list_i = nullptr, *b_list_j = nullptr;
static int b_list_n = 0, b_list_cap = 0;

extern "C" int lpmx_get_broken(int *bi, int *bj, int cap, int *count)
{
    if (count)
        *count = b_list_n;
    int c = b_list_n < cap ? b_list_n : cap;
    if (c > 0)
    {
        HIPCHK(hipMemcpyAsync(bi, b_list_i, sizeof(int) * c, hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipMemcpyAsync(bj, b_list_j, sizeof(int) * c, hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    return LPMX_OK;
}

static int brittle(int *broken)
{
    const int n = g.n;
    int *cnt, *off;
    HIPCHK(hipMalloc(&cnt, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&off, sizeof(int) * (n + 1)));
    hipLaunchKernelGGL(k_brittle_count, dim3(nblocks(n + 1)), dim3(LPMX_BLOCK), 0, g_stream, g, cnt);
    LAUNCHCHK();
    int k = 0;
    RETCHK(lpmx_dev_scan_int(cnt, off, n + 1, &k));
    b_list_n = 0;
    if (k > 0)
    {
        int *ci, *cj;
        double *cb;
        HIPCHK(hipMalloc(&ci, sizeof(int) * k));
        HIPCHK(hipMalloc(&cj, sizeof(int) * k));
        HIPCHK(hipMalloc(&cb, sizeof(double) * k));
        hipLaunchKernelGGL(k_brittle_fill, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, off, ci, cj, cb);
        LAUNCHCHK();
        int from = 0;
        if (k > g.nbreak)
        {
            hipLaunchKernelGGL(k_shell_sort, dim3(1), dim3(64), 0, g_stream, k, cb, ci, cj);
            LAUNCHCHK();
            from = k - g.nbreak;
        }
        const int nbk = k - from;
        if (nbk > b_list_cap)
        {
            if (b_list_i)
                hipFree(b_list_i), hipFree(b_list_j);
            HIPCHK(hipMalloc(&b_list_i, sizeof(int) * nbk));
            HIPCHK(hipMalloc(&b_list_j, sizeof(int) * nbk));
            b_list_cap = nbk;
        }
        hipLaunchKernelGGL(k_brittle_break, dim3(nblocks(nbk)), dim3(LPMX_BLOCK), 0, g_stream, g, ci, cj, from, k,
                           b_list_i, b_list_j);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(g_stream));
        hipFree(ci), hipFree(cj), hipFree(cb);
        b_list_n = nbk;
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(cnt), hipFree(off);
    *broken = k;
    return LPMX_OK;
}

extern "C" int lpmx_damage(int plmode, int *broken)
{
    RETCHK(lpmx_require_model());
    int b = 0;
    if (plmode == 0)
    {
        RETCHK(build_damage_lists());
        if (!d_cnt)
            HIPCHK(hipMalloc(&d_cnt, sizeof(int)));
        hipLaunchKernelGGL(k_dnl, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g);
        LAUNCHCHK();
        HIPCHK(hipMemsetAsync(d_cnt, 0, sizeof(int), g_stream));
        const int64_t nbn = (int64_t)g.n * g.nb;
        hipLaunchKernelGGL(k_dmg_break, dim3(nblocks(nbn)), dim3(LPMX_BLOCK), 0, g_stream, g, d_cnt);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_dmg_w, dim3(nblocks(nbn)), dim3(LPMX_BLOCK), 0, g_stream, g);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(&b, d_cnt, sizeof(int), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
    }
    else if (plmode == 5)
        RETCHK(lpmx_damage_variant(LPMX_DMG_BWISE_LOCAL, &b));
    else if (plmode == 6)
        RETCHK(brittle(&b));
    if (broken)
        *broken = b;
    return LPMX_OK;
}

extern "C" int lpmx_damage_law(int law, int *count)
{
    RETCHK(lpmx_require_model());
    if (law == LPMX_DMG_PWISE_NONLOCAL)
        return lpmx_damage(0, count);
    if (law == LPMX_DMG_BRITTLE)
        return lpmx_damage(6, count);
    return lpmx_damage_variant(law, count);
}

// ------------------------------------------------------------------- crack -------
// updateCrack (constitutive.c:1399-1434).
__global__ void k_crack(DV v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int nb0 = v.nb0[i];
    int nbc = nb0;
    double vis = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int j = 0; j < nb0; j++)
    {
        const int64_t b = BI(j, i);
        if (v.db[b] <= LPMX_EPS)
            nbc -= 1;
        vis += v.db[b];
        const double f = v.F[b] * v.dw[b];
        v.F[b] = f;
        p0 += v.csx[b] * f;
        p1 += v.csy[b] * f;
        p2 += v.csz[b] * f;
    }
    v.nbcur[i] = nbc;
    v.Pin[3 * i] = p0, v.Pin[3 * i + 1] = p1, v.Pin[3 * i + 2] = p2;
    if (nbc < 1)
        for (int c = 0; c < v.dim; c++)
            v.fix[(int64_t)v.dim * i + c] = 0;
    v.dvis[i] = 1 - vis / nb0;
}

extern "C" int lpmx_update_crack(void)
{
    RETCHK(lpmx_require_model());
    hipLaunchKernelGGL(k_crack, dim3(nblocks(g.n)), dim3(LPMX_BLOCK), 0, g_stream, g);
    LAUNCHCHK();
    return LPMX_OK;
}

// ----------------------------------------------------------------- outputs -------
// writeReaction / writeDisp (data_handler.c:278-335) per particle, summed on the
// device in a fixed order.
__global__ void k_reaction(DV v, int c, int t1, double *f, double *dx, double *cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    double s = 0.0, d = 0.0, k = 0.0;
    if (v.type[i] == t1)
    {
        const double *cs = c == 0 ? v.csx : (c == 1 ? v.csy : v.csz);
        for (int j = 0; j < v.nb0[i]; j++)
            s += -cs[BI(j, i)] * v.F[BI(j, i)];
        d = v.x[3 * i + c] - v.x0[3 * i + c];
        k = 1.0;
    }
    f[i] = s, dx[i] = d, cnt[i] = k;
}

extern "C" int lpmx_reaction(char axis, int t1, double *force, double *mean_disp)
{
    RETCHK(lpmx_require_model());
    const int n = g.n, c = axis == 'x' ? 0 : axis == 'y' ? 1 : 2;
    double *buf = lpmx_scratch(sizeof(double) * 3 * (int64_t)n);
    if (!buf)
        return lpmx_fail(LPMX_ERR_HIP, "scratch allocation failed");
    hipLaunchKernelGGL(k_reaction, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, c, t1, buf, buf + n, buf + 2 * n);
    LAUNCHCHK();
    double f = 0, d = 0, k = 0;
    RETCHK(lpmx_dev_sum(buf, n, &f));
    RETCHK(lpmx_dev_sum(buf + n, n, &d));
    RETCHK(lpmx_dev_sum(buf + 2 * n, n, &k));
    if (force)
        *force = f;
    if (mean_disp)
        *mean_disp = k > 0 ? d / k : 0.0;
    return LPMX_OK;
}
