// Device helpers shared by the per-bond kernels: the slot-plane index, distances,
// the opposite-bond factor of the stress sums and one stretch pass of a particle.
#pragma once

#include "common.h"

#define BI(j, i) ((int64_t)(j) * v.n + (i))

__device__ __forceinline__ double dist_xyz(const double *x, int i, int j)
{
    double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// Opposite-bond factor of the stress sums (constitutive.c:540-559,
// lpm_basic.c:71-90); the first opposite slot is precomputed in opp.
__device__ __forceinline__ double opp_flag(const DV &v, int i, int j)
{
    if (v.nbcur[i] == v.nb)
        return 0.5;
    const int o = v.opp[BI(j, i)];
    if (o < 0)
        return 1.0;
    return v.db[BI(o, i)] <= LPMX_EPS ? 1.0 : 0.5;
}

// One stretch pass of particle i (constitutive.c:495-515): dL, cs and the shell
// sums, with plastic stretch from `xp` (slot plane) or, if plastic == nullptr, none.
__device__ __forceinline__ void stretch_pass(const DV &v, int i, const double *xp, double &a0, double &a1, double &b0,
                                             double &b1)
{
    a0 = 0, b0 = 0;
    a1 = 0, b1 = 0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b];
        const double dis = dist_xyz(v.x, i, q);
        double dl = dis - v.L0[b];
        dl -= xp[b];
        dl *= v.db[b];
        v.dL[b] = dl;
        if (v.nsign[b] == 0)
            a0 += dl, b0 += v.Tv[b] * dl;
        else
            a1 += dl, b1 += v.Tv[b] * dl;
        v.csx[b] = (v.x[3 * i] - v.x[3 * q]) / dis;
        v.csy[b] = (v.x[3 * i + 1] - v.x[3 * q + 1]) / dis;
        v.csz[b] = (v.x[3 * i + 2] - v.x[3 * q + 2]) / dis;
    }
}
