// The remaining constitutive laws and ductile damage variants of the reference on the
// device: the energy-based J2 return map (plmode 3), the nonlinear-isotropic J2 law
// (plmode 5), and the point-wise local, bond-wise nonlocal and bond-wise local damage
// updates.  One thread per particle, bond data in slot planes; arithmetic in the
// reference's evaluation order (this file is compiled without contraction).
#include <math.h>

#include "bond_common.h"

#define TOLITER_REF 1e-4 // the reference's TOLITER (include/lpm.h:41): the bisection bracket

// ------------------------------------------------------ plmode 3 (energy J2) -----
// computeBondForceJ2energyReturnMap (constitutive.c:286-463).  The reference
// recomputes each particle's list from slot 0 and keeps the partners' updates in
// local copies, so a particle's update is a pure function of the positions and its
// slot-0 state: every particle's update once, then every force, gives the
// reference's values.  Per particle: the first stretch pass, the distortional
// energy of its first nb[i] bonds about each shell's mean stretch, the equivalent
// stress, the yield test against the slot-0 equivalent back stress and hardening
// variable, bisection of the multiplier on [0, 1] to TOLITER, the updated state and
// plastic stretches (slot 2), and the second stretch pass.
__global__ void k_j2e_particle(DV v, int t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int n = v.n, nb0 = v.nb0[i], nbc = v.nbcur[i];
    double a0, a1, b0, b1;
    stretch_pass(v, i, v.dLp0, a0, a1, b0, b1);
    int nb1 = 0; // first-shell neighbours of the initial list (neighbors1)
    for (int j = 0; j < nb0; j++)
        nb1 += v.nsign[BI(j, i)] == 0;
    const int nb2 = nbc - nb1;
    double U_d = 0.0, J2_k = 0.0;
    const double J2_V = v.volume * nbc / v.nb;
    for (int j = 0; j < nbc; j++)
    {
        const int64_t b = BI(j, i);
        if (v.nsign[b] == 0)
        {
            U_d += 0.5 * v.Kn[b] * (v.dL[b] - a0 / nb1) * (v.dL[b] - a0 / nb1);
            J2_k = v.Kn[b];
        }
        else if (v.nsign[b] == 1)
            U_d += 0.5 * v.Kn[b] * (v.dL[b] - a1 / nb2) * (v.dL[b] - a1 / nb2);
    }
    const double c44 = v.c44[i], H = v.H, xi = v.xi;
    const double J2_sigma = sqrt(6.0 * c44 * U_d / J2_V);
    const double beq0 = v.beq0[i], al0 = v.alpha0[i];
    double dlambda = 0.0;
    double yf = fabs(t * J2_sigma - beq0) - (v.sigmay[i] + (1.0 - xi) * H * al0);
    if (yf > 0.0)
    {
        v.pl_flag[i] = 1;
        double a = 0.0, bb = 1.0, ya = yf;
        while ((bb - a) > TOLITER_REF)
        {
            dlambda = (a + bb) / 2.0;
            yf = fabs(t * J2_sigma / (1.0 + 0.5 * dlambda) -
                      (beq0 + t * xi * H * dlambda * J2_sigma * sqrt(v.radius / J2_k / c44) / 6.0 /
                                  (1.0 + 0.5 * dlambda))) -
                 (v.sigmay[i] + (1.0 - xi) * H * al0 +
                  (1.0 - xi) * H * dlambda * J2_sigma * sqrt(v.radius / J2_k / c44) / 6.0 / (1.0 + 0.5 * dlambda));
            if (yf * ya < 0.0)
                bb = dlambda;
            else
            {
                a = dlambda;
                ya = yf;
            }
        }
    }
    const double alpha = al0 + dlambda / (1.0 + 0.5 * dlambda) * sqrt(v.radius / 6.0 / J2_k * U_d / J2_V);
    const double beq = (beq0 + t * xi * H * dlambda / (1.0 + 0.5 * dlambda) * sqrt(v.radius / 6.0 / J2_k * U_d / J2_V));
    for (int j = 0; j < v.nb; j++)
    {
        const int64_t b = BI(j, i);
        double xp = v.dLp0[b];
        if (j < nbc)
        {
            const int s = v.nsign[b];
            double f_d = 0.0;
            if (s == 0)
                f_d = 2.0 * v.Kn[b] / (1 + dlambda / 2.0) * (v.dL[b] - a0 / nb1);
            if (s == 1)
                f_d = 2.0 * v.Kn[b] / (1 + dlambda / 2.0) * (v.dL[b] - a1 / nb2);
            double d = dlambda * f_d / (4.0 * v.Kn[b]);
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
    v.alpha2[i] = alpha;
    v.beq2[i] = beq;
    v.dlam[i] = dlambda;
}

// Final force of a particle from the half-averaged post-update stretches and both
// ends' shell sums (constitutive.c:422-440 for plmode 3, :834-856 for plmode 5);
// plmode 3 scales by 1 - damage_D of slot 0, plmode 5 by damage_w.
__global__ void k_avg_force(DV v, int use_D0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
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
        f *= use_D0 ? (1.0 - v.D0[b]) : v.dw[b];
        v.F[b] = f;
        p0 += v.csx[b] * f;
        p1 += v.csy[b] * f;
        p2 += v.csz[b] * f;
    }
    v.Pin[3 * i] = p0, v.Pin[3 * i + 1] = p1, v.Pin[3 * i + 2] = p2;
}

// ----------------------------------------------- plmode 5 (nonlinear isotropic) ----
// The reference's hardening law SY (include/lpm.h:50), MPa.
__device__ __forceinline__ double sy_law(double x) { return 620.0 + 3300.0 * (1.0 - exp(-0.4 * x)); }

// Stretch pass of every particle with its current slot-0 plastic stretch.
__global__ void k_stretch_slot0(DV v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    double a0, a1, b0, b1;
    stretch_pass(v, i, v.dLp0, a0, a1, b0, b1);
    v.A[i] = a0, v.A[(int64_t)v.n + i] = a1;
    v.Bt[i] = b0, v.Bt[(int64_t)v.n + i] = b1;
}

// computeBondForceJ2nonlinearIso (constitutive.c:726-812), one update per particle:
// trial bond forces and stress from the particle's own sums, deviator minus the
// slot-0 back stress, bisection of the multiplier on [0, 1] to TOLITER against the
// hardening law, then the multiplier, hardening variable, back stress and plastic
// stretch of slot 0 advanced in place.  The reference advances a particle once per
// list it is in, in its particle order; this is the single update per particle the
// order-free reading gives.  Like the reference, the flow direction divides by the
// equivalent stress with no guard.  C is the kinematic modulus J2_C.
__global__ void k_j2nl_update(DV v, double C)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int n = v.n, nb0 = v.nb0[i];
    const double J2_V = v.volume * v.nbcur[i] / v.nb;
    double sl[6] = {0, 0, 0, 0, 0, 0}, dpl[6];
    for (int j = 0; j < nb0; j++)
    {
        const int64_t b = BI(j, i);
        const int s = v.nsign[b];
        double f = 2.0 * v.Kn[b] * v.dL[b] + v.Bt[(int64_t)s * n + i] + v.Tv[b] * v.A[(int64_t)s * n + i];
        f *= v.db[b];
        v.F[b] = f;
        const double cx = v.csx[b], cy = v.csy[b], cz = v.csz[b], L0 = v.L0[b];
        sl[0] += 0.5 / J2_V * L0 * f * cx * cx;
        sl[1] += 0.5 / J2_V * L0 * f * cy * cy;
        sl[2] += 0.5 / J2_V * L0 * f * cz * cz;
        sl[3] += 0.5 / J2_V * L0 * f * cy * cz;
        sl[4] += 0.5 / J2_V * L0 * f * cx * cz;
        sl[5] += 0.5 / J2_V * L0 * f * cx * cy;
    }
    const double temp = 1.0 / 3.0 * (sl[0] + sl[1] + sl[2]);
    for (int c = 0; c < 3; c++)
        sl[c] -= temp;
    for (int c = 0; c < 6; c++)
        sl[c] -= v.beta0[(int64_t)c * n + i];
    double seq = 0.0;
    for (int c = 0; c < 6; c++)
    {
        if (c < 3)
            seq += sl[c] * sl[c];
        else
            seq += 2.0 * sl[c] * sl[c];
    }
    seq = sqrt(3.0 / 2.0 * seq);
    const double c44 = v.c44[i], al0 = v.alpha0[i];
    double dlambda = 0.0;
    double yf = seq - sy_law(al0);
    if (yf > 0.0)
    {
        double a = 0.0, bb = 1.0, ya = yf;
        while ((bb - a) > TOLITER_REF)
        {
            dlambda = (a + bb) / 2.0;
            yf = seq - 1.5 * dlambda * (2.0 * c44 + C) - sy_law(al0 + dlambda);
            if (yf * ya < 0.0)
                bb = dlambda;
            else
            {
                a = dlambda;
                ya = yf;
            }
        }
    }
    v.dlam[i] = dlambda;
    v.alpha0[i] = al0 + dlambda;
    for (int c = 0; c < 6; c++)
    {
        dpl[c] = dlambda * 1.5 * sl[c] / seq;
        v.beta0[(int64_t)c * n + i] += C * dpl[c];
    }
    for (int j = 0; j < nb0; j++)
    {
        const int64_t b = BI(j, i);
        const double cx = v.csx[b], cy = v.csy[b], cz = v.csz[b];
        const double d = v.L0[b] * (dpl[0] * cx * cx + dpl[1] * cy * cy + dpl[2] * cz * cz + 2 * dpl[3] * cy * cz +
                                    2 * dpl[4] * cx * cz + 2 * dpl[5] * cx * cy);
        v.ddLp[b] = d;
        v.dLp0[b] += d;
    }
}

int lpmx_law3(int t)
{
    const dim3 G(nblocks(g.n)), B(LPMX_BLOCK);
    hipLaunchKernelGGL(k_j2e_particle, G, B, 0, g_stream, g, t);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_avg_force, G, B, 0, g_stream, g, 1);
    LAUNCHCHK();
    return LPMX_OK;
}

int lpmx_law5(void)
{
    const dim3 G(nblocks(g.n)), B(LPMX_BLOCK);
    hipLaunchKernelGGL(k_stretch_slot0, G, B, 0, g_stream, g);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_j2nl_update, G, B, 0, g_stream, g, g.J2C);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_stretch_slot0, G, B, 0, g_stream, g);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_avg_force, G, B, 0, g_stream, g, 0);
    LAUNCHCHK();
    return LPMX_OK;
}

// ----------------------------------------------------------- damage variants -----
// Local damage rate (1 + A triaxiality) dlambda of a particle, 0 if the factor is
// not positive (constitutive.c:1548-1551).
__device__ __forceinline__ double local_rate(const DV &v, int i)
{
    const double f = (1.0 + v.dA * v.triax[i]);
    return f > 0.0 ? v.dlam[i] * (1.0 + v.dA * v.triax[i]) : 0.0;
}

// updateDuctileDamagePwiseLocal (constitutive.c:1698-1755): local damage from the
// multiplier; a particle passing the threshold (and not yet at 1) is set to 1 and
// its bonds and their reverse entries break (every writer writes 0, so concurrent
// writes agree).  Each decision reads only the particle's own state.
__global__ void k_dmg_pl_particle(DV v, int *failed)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    double d = v.dloc0[i];
    const double f = (1.0 + v.dA * v.triax[i]);
    if (f > 0.0 && d <= v.dthr)
        d += f * v.dlam[i];
    if (d > v.dthr && fabs(d - 1.0) > LPMX_EPS)
    {
        d = 1.0;
        for (int j = 0; j < v.nb0[i]; j++)
        {
            const int64_t b = BI(j, i);
            v.db[b] = 0.0;
            const int r = v.rev[b];
            if (r >= 0)
                v.db[BI(r, v.nbr[b])] = 0.0;
        }
        atomicAdd(failed, 1);
    }
    v.dloc0[i] = d;
}

// bond damage the larger of the two ends' local damage, weight 1 - D (:1745-1753)
__global__ void k_dmg_pl_bond(DV v)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.nb)
        return;
    const int i = (int)(t % v.n), j = (int)(t / v.n);
    if (j >= v.nb0[i])
        return;
    const double a = v.dloc0[i], c = v.dloc0[v.nbr[t]];
    v.D0[t] = a < c ? c : a;
    v.dw[t] = 1.0 - v.D0[t];
}

// updateDuctileDamageBwiseNonlocal (constitutive.c:1529-1605): every particle below
// the threshold grows by the Gaussian-weighted mean (DAM_PHI of the initial bond
// length, include/lpm.h:51) of the local rates of itself and its list.
__global__ void k_dmg_bn_particle(DV v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const double dn = v.dnl0[i];
    if (dn > v.dthr)
    {
        if (dn > 1.0)
            v.dnl0[i] = 1.0;
        return;
    }
    const double L = v.dmgL, V = v.volume;
    double Ddot = local_rate(v, i) * V;
    double A = V;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const double x = v.L0[b];
        const double phi = 1.0 / L / sqrt(2 * 3.14159265358979323846) * exp(-0.5 * x * x / L / L);
        Ddot += local_rate(v, v.nbr[b]) * phi * V;
        A += phi * V;
    }
    if (Ddot > 0.0)
        v.dnl0[i] = dn + 1.0 / A * Ddot;
}

// a directed bond whose two ends average above the threshold breaks (:1583-1596)
__global__ void k_dmg_bn_break(DV v, int *count)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.nb)
        return;
    const int i = (int)(t % v.n), j = (int)(t / v.n);
    if (j >= v.nb0[i])
        return;
    const int q = v.nbr[t];
    if (0.5 * (v.dnl0[i] + v.dnl0[q]) > v.dthr && fabs(v.db[t]) > LPMX_EPS)
    {
        v.db[t] = 0.0;
        v.D0[t] = 1.0;
        atomicAdd(count, 1);
    }
}

__global__ void k_dmg_bn_weight(DV v)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.nb)
        return;
    const int i = (int)(t % v.n), j = (int)(t / v.n);
    if (j >= v.nb0[i])
        return;
    if (fabs(v.db[t]) > LPMX_EPS)
        v.D0[t] = 0.5 * (v.dnl0[i] + v.dnl0[v.nbr[t]]);
    v.dw[t] = 1.0 - v.D0[t];
}

// updateDuctileDamageBwiseLocal (constitutive.c:1607-1695).  Local damage from the
// multiplier; the bond damage is the mean of the two ends; a bond passing the
// threshold breaks with its reverse entry.  Either end breaking a pair breaks both
// entries, so an entry ends broken when it was, or when the mean passes the
// threshold and either entry was intact; the reference counts each pair once, at
// the first end that finds an intact entry, here at the lower index.
__global__ void k_dmg_bl_local(DV v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const double f = (1.0 + v.dA * v.triax[i]);
    double d = v.dloc0[i];
    if (f > 0.0 && d <= v.dthr)
        d += f * v.dlam[i];
    else if (d > 1.0)
        d = 1.0;
    v.dloc0[i] = d;
}

__global__ void k_dmg_bl_break(DV v, double *nbk, int *count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    int nbl = v.nb0[i], pairs = 0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b], r = v.rev[b];
        const double mean = 0.5 * (v.dloc0[i] + v.dloc0[q]);
        const double bi_ = v.db[b], bq = r >= 0 ? v.db[BI(r, q)] : 0.0;
        const bool intact = bi_ > LPMX_EPS || bq > LPMX_EPS;
        double D = mean;
        double nb_ = bi_;
        if (mean > v.dthr && intact)
        {
            D = 1.0, nb_ = 0.0;
            pairs += i < q;
        }
        nbk[b] = nb_; // the entry's new damage_broken, applied by k_dmg_bl_apply
        v.D0[b] = D;
        if (nb_ <= LPMX_EPS)
            nbl--;
    }
    v.nbcur[i] = nbl;
    if (pairs)
        atomicAdd(count, pairs);
}

__global__ void k_dmg_bl_apply(DV v, const double *nbk)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.nb)
        return;
    const int i = (int)(t % v.n), j = (int)(t / v.n);
    if (j < v.nb0[i])
        v.db[t] = nbk[t];
}

__global__ void k_dmg_bl_weight(DV v)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.nb)
        return;
    const int i = (int)(t % v.n), j = (int)(t / v.n);
    if (j >= v.nb0[i])
        return;
    const int q = v.nbr[t];
    if (fabs(v.db[t]) < LPMX_EPS || v.nbcur[i] == 0 || v.nbcur[q] == 0)
        v.D0[t] = 1.0;
    v.dw[t] = 1.0 - v.D0[t];
}

// Counter of the damage updates: cleared by a fill kernel, read back once.
__global__ void k_zero_int(int *p) { *p = 0; }

static int *l_cnt = nullptr;
static double *l_nbk = nullptr; // new damage_broken of every entry (bond-wise local damage)
static int64_t l_nbk_cap = 0;

void lpmx_laws_free(void)
{
    if (l_cnt)
        hipFree(l_cnt);
    if (l_nbk)
        hipFree(l_nbk);
    l_cnt = nullptr, l_nbk = nullptr, l_nbk_cap = 0;
}

// law: LPMX_DMG_PWISE_LOCAL, LPMX_DMG_BWISE_NONLOCAL or LPMX_DMG_BWISE_LOCAL.
int lpmx_damage_variant(int law, int *count)
{
    if (!l_cnt)
        HIPCHK(hipMalloc(&l_cnt, sizeof(int)));
    const int64_t nbn = (int64_t)g.n * g.nb;
    const dim3 Gp(nblocks(g.n)), Gb(nblocks(nbn)), B(LPMX_BLOCK);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, g_stream, l_cnt);
    LAUNCHCHK();
    if (law == LPMX_DMG_PWISE_LOCAL)
    {
        hipLaunchKernelGGL(k_dmg_pl_particle, Gp, B, 0, g_stream, g, l_cnt);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_dmg_pl_bond, Gb, B, 0, g_stream, g);
    }
    else if (law == LPMX_DMG_BWISE_NONLOCAL)
    {
        hipLaunchKernelGGL(k_dmg_bn_particle, Gp, B, 0, g_stream, g);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_dmg_bn_break, Gb, B, 0, g_stream, g, l_cnt);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_dmg_bn_weight, Gb, B, 0, g_stream, g);
    }
    else if (law == LPMX_DMG_BWISE_LOCAL)
    {
        if (l_nbk_cap < nbn)
        {
            if (l_nbk)
                hipFree(l_nbk);
            l_nbk = nullptr, l_nbk_cap = 0;
            HIPCHK(hipMalloc(&l_nbk, sizeof(double) * nbn));
            l_nbk_cap = nbn;
        }
        hipLaunchKernelGGL(k_dmg_bl_local, Gp, B, 0, g_stream, g);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_dmg_bl_break, Gp, B, 0, g_stream, g, l_nbk, l_cnt);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_dmg_bl_apply, Gb, B, 0, g_stream, g, l_nbk);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_dmg_bl_weight, Gb, B, 0, g_stream, g);
    }
    else
        return lpmx_fail(LPMX_ERR_ARG, "damage law %d", law);
    LAUNCHCHK();
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, l_cnt, sizeof(int), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    if (count)
        *count = h;
    return LPMX_OK;
}
