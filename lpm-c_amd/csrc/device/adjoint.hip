// Reverse mode of the J2 load path on the device: vector-Jacobian products of one
// force evaluation (the J2 law of src/constitutive.c:466-686 as evaluated by
// k_j2_particle / k_j2_force), and the discrete adjoint of a load history built on
// them (lpmx_adjoint_run).
//
// A force evaluation maps positions x, the slot-0 state s = (dLp, beta, alpha) and
// the parameters (per-particle modulus factor e, yield stress sigmay, H) to the
// internal force Pin and the slot-2 state.  Its reverse sweep, for seeds on Pin and
// the slot-2 state, runs in three kernels, all deterministic (gathers, no atomics):
//   k_vjp_force     per bond of particle i: adjoint of F_ij (= Pbar_i . cs_ij), the
//                   geometric term of Pin = sum cs F, the bond-constant adjoints of F;
//   k_vjp_particle  per particle: gathers the adjoints of its second-pass stretches
//                   and sums from its own and its partners' bonds (reverse index),
//                   then reverses its second stretch pass, radial return, trial
//                   stress and first stretch pass, recomputed in the forward's order
//                   (so every yield decision is the forward's);
//   k_vjp_gather    positions from the bond-vector adjoints (x_i - x_q), modulus
//                   factor from the bond constants (SC averages both ends).
#include <math.h>
#include <string.h>

#include <vector>

#include "common.h"

#define BI(j, i) ((int64_t)(j) * v.n + (i))


struct AdjWork
{
    double *Fb = nullptr, *gb = nullptr, *Knb = nullptr, *Tvb = nullptr, *c44b = nullptr;
    int64_t cap = 0;
};
static AdjWork AW;

void lpmx_adjoint_free_work(void)
{
    void *ps[] = {AW.Fb, AW.gb, AW.Knb, AW.Tvb, AW.c44b};
    for (void *p : ps)
        if (p)
            hipFree(p);
    AW = AdjWork();
}

static int adj_work(void)
{
    const int64_t nbn = (int64_t)g.n * g.nb;
    if (AW.cap == nbn && AW.Fb)
        return LPMX_OK;
    lpmx_adjoint_free_work();
    HIPCHK(hipMalloc(&AW.Fb, sizeof(double) * nbn));
    HIPCHK(hipMalloc(&AW.gb, sizeof(double) * 3 * nbn));
    HIPCHK(hipMalloc(&AW.Knb, sizeof(double) * nbn));
    HIPCHK(hipMalloc(&AW.Tvb, sizeof(double) * nbn));
    HIPCHK(hipMalloc(&AW.c44b, sizeof(double) * g.n));
    AW.cap = nbn;
    return LPMX_OK;
}

__device__ __forceinline__ double vjp_opp(const DV &v, int i, int j)
{
    if (v.nbcur[i] == v.nb)
        return 0.5;
    const int o = v.opp[BI(j, i)];
    if (o < 0)
        return 1.0;
    return v.db[BI(o, i)] <= LPMX_EPS ? 1.0 : 0.5;
}

__device__ __forceinline__ void cc6(double cx, double cy, double cz, double *o)
{
    o[0] = cx * cx, o[1] = cy * cy, o[2] = cz * cz, o[3] = cy * cz, o[4] = cx * cz, o[5] = cx * cy;
}

__device__ __forceinline__ void cc6_bar(double cx, double cy, double cz, const double *ob, double *cb)
{
    cb[0] += 2 * cx * ob[0] + cz * ob[4] + cy * ob[5];
    cb[1] += 2 * cy * ob[1] + cz * ob[3] + cx * ob[5];
    cb[2] += 2 * cz * ob[2] + cy * ob[3] + cx * ob[4];
}

// Force pass: F_ij of the last evaluation (v.F, v.dLave, v.A, v.Bt hold it).
__global__ void k_vjp_force(DV v, VjpIO io, AdjWork a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const double p0 = io.Pb[3 * i], p1 = io.Pb[3 * i + 1], p2 = io.Pb[3 * i + 2];
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b], s = v.nsign[b];
        const double dx = v.x[3 * i] - v.x[3 * q], dy = v.x[3 * i + 1] - v.x[3 * q + 1],
                     dz = v.x[3 * i + 2] - v.x[3 * q + 2];
        const double dis = sqrt(dx * dx + dy * dy + dz * dz);
        const double cx = v.csx[b], cy = v.csy[b], cz = v.csz[b], F = v.F[b];
        const double fb = p0 * cx + p1 * cy + p2 * cz;
        const double bx = p0 * F, by = p1 * F, bz = p2 * F;
        const double dot = cx * bx + cy * by + cz * bz;
        a.Fb[b] = fb;
        a.gb[b] = (bx - dot * cx) / dis;
        a.gb[(int64_t)v.n * v.nb + b] = (by - dot * cy) / dis;
        a.gb[2 * (int64_t)v.n * v.nb + b] = (bz - dot * cz) / dis;
        const double t = fb * v.dw[b];
        a.Knb[b] = t * 2.0 * v.dLave[b];
        a.Tvb[b] = t * 0.5 * (v.A[(int64_t)s * v.n + i] + v.A[(int64_t)s * v.n + q]);
    }
}

// Bond (i, j) of the forward recomputation: distance, unit vector (x_i - x_q)/d and
// first-pass stretch, in k_j2_particle's expression order.
struct VjpBond
{
    double dis, c[3], dl;
};

__device__ __forceinline__ VjpBond vjp_bond(const DV &v, int i, int64_t b)
{
    VjpBond o;
    const int q = v.nbr[b];
    const double dx = v.x[3 * i] - v.x[3 * q], dy = v.x[3 * i + 1] - v.x[3 * q + 1],
                 dz = v.x[3 * i + 2] - v.x[3 * q + 2];
    o.dis = sqrt(dx * dx + dy * dy + dz * dz);
    double dl = o.dis - v.L0[b];
    dl -= v.dLp0[b];
    dl *= v.db[b];
    o.dl = dl;
    o.c[0] = (v.x[3 * i] - v.x[3 * q]) / o.dis;
    o.c[1] = (v.x[3 * i + 1] - v.x[3 * q + 1]) / o.dis;
    o.c[2] = (v.x[3 * i + 2] - v.x[3 * q + 2]) / o.dis;
    return o;
}

// Position adjoint of bond b from the adjoints of its distance (disb) and unit
// vector (cb): d(dis)/dx_i = c, dc/dx_i = (I - c c^T)/dis.  Linear in (disb, cb), so
// each reverse pass adds its own part.
__device__ __forceinline__ void vjp_geom(const DV &v, const AdjWork &a, int64_t b, int64_t NB, const VjpBond &g,
                                         double disb, const double *cb)
{
    const double dot = g.c[0] * cb[0] + g.c[1] * cb[1] + g.c[2] * cb[2];
    a.gb[b] += disb * g.c[0] + (cb[0] - dot * g.c[0]) / g.dis;
    a.gb[NB + b] += disb * g.c[1] + (cb[1] - dot * g.c[1]) / g.dis;
    a.gb[2 * NB + b] += disb * g.c[2] + (cb[2] - dot * g.c[2]) / g.dis;
}

// Per particle: second pass, return map and trial pass in reverse.  Nothing per bond
// is held across passes: the geometry, stretches and trial forces are recomputed
// from the cached positions in the forward's expressions (so the same values and
// the same yield decisions), and each pass adds its part of the position adjoint to
// a.gb, so no per-bond array lives in scratch memory.  (Keeping the geometry in
// LDS or in bond planes of HBM instead measured slower: 893 and 639 against 547 us
// at 216k particles.)
__global__ void k_vjp_particle(DV v, VjpIO io, AdjWork a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int n = v.n, nbi = v.nb0[i];
    const int64_t NB = (int64_t)n * v.nb;
    // sums of the adjoints of the second-pass stretches, from both ends' forces
    double A2b[2] = {0, 0}, B2b[2] = {0, 0};
    for (int j = 0; j < nbi; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b], s = v.nsign[b], r = v.rev[b];
        const double t = a.Fb[b] * v.dw[b];
        A2b[s] += t * 0.5 * v.Tv[b];
        B2b[s] += t * 0.5;
        if (r >= 0)
        {
            const int64_t br = BI(r, q);
            const double tr = a.Fb[br] * v.dw[br];
            A2b[s] += tr * 0.5 * v.Tv[br];
            B2b[s] += tr * 0.5;
        }
    }
    // forward recomputation (k_j2_particle's order)
    double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    for (int j = 0; j < nbi; j++)
    {
        const int64_t b = BI(j, i);
        const VjpBond g = vjp_bond(v, i, b);
        if (v.nsign[b] == 0)
            a0 += g.dl, b0 += v.Tv[b] * g.dl;
        else
            a1 += g.dl, b1 += v.Tv[b] * g.dl;
    }
    // trial force of bond b from its first-pass stretch
    auto ftrial = [&](int64_t b, double dl) {
        const int s = v.nsign[b];
        double Fij = 2.0 * v.Kn[b] * dl + (s == 0 ? b0 : b1) + v.Tv[b] * (s == 0 ? a0 : a1);
        return Fij * v.dw[b];
    };
    double sl[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nbi; j++)
    {
        const int64_t b = BI(j, i);
        const VjpBond g = vjp_bond(v, i, b);
        const double Fij = ftrial(b, g.dl);
        const double of = vjp_opp(v, i, j);
        const double cx = g.c[0], cy = g.c[1], cz = g.c[2], L0 = v.L0[b];
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
    for (int c = 0; c < 6; c++)
        sl[c] -= v.beta0[(int64_t)c * n + i];
    double seq = 0.0;
    for (int c = 0; c < 6; c++)
        seq += c < 3 ? sl[c] * sl[c] : 2.0 * sl[c] * sl[c];
    seq = sqrt(3.0 / 2.0 * seq);
    const double alpha0 = v.alpha0[i], H = v.H, xi = v.xi, c44 = v.c44[i];
    const double yf = seq - (v.sigmay[i] + (1.0 - xi) * H * alpha0);
    const bool plastic = yf > 0.0;
    const double dlam = plastic ? yf / (3 * c44 + H) : 0.0;
    double dep[6] = {0, 0, 0, 0, 0, 0};
    if (fabs(seq) > LPMX_EPS)
        for (int c = 0; c < 6; c++)
            dep[c] = dlam * 1.5 * sl[c] / seq;
    // second pass in reverse (dL2 of the evaluation is in v.dL)
    double depb[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nbi; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b], s = v.nsign[b], r = v.rev[b];
        const double t = a.Fb[b] * v.dw[b];
        double dL2b = r >= 0 ? t * v.Kn[b] : 0.0;
        if (r >= 0)
        {
            const int64_t br = BI(r, q);
            const double tr = a.Fb[br] * v.dw[br];
            dL2b += tr * v.Kn[br];
        }
        const VjpBond g = vjp_bond(v, i, b);
        const double db = v.db[b];
        const double tot = dL2b + A2b[s] + B2b[s] * v.Tv[b];
        a.Tvb[b] += B2b[s] * v.dL[b];
        const double xpb = -tot * db + io.dLp2b[b] * db;
        io.dLp0b[b] = xpb;
        double cc[6], ccb[6];
        cc6(g.c[0], g.c[1], g.c[2], cc);
        for (int c = 0; c < 6; c++)
        {
            const double wc = c < 3 ? 1.0 : 2.0;
            depb[c] += xpb * v.L0[b] * db * wc * cc[c];
            ccb[c] = xpb * v.L0[b] * db * wc * dep[c];
        }
        double cb[3] = {0.0, 0.0, 0.0};
        cc6_bar(g.c[0], g.c[1], g.c[2], ccb, cb);
        vjp_geom(v, a, b, NB, g, tot * db, cb);
    }
    for (int j = nbi; j < v.nb; j++)
        io.dLp0b[BI(j, i)] = io.dLp2b[BI(j, i)] * v.db[BI(j, i)];
    // slot-2 state
    double dlamb = io.alpha2b[i], Hb = 0.0;
    double alpha0b = io.alpha2b[i];
    double beta0b[6];
    for (int c = 0; c < 6; c++)
    {
        const double bb = io.beta2b[(int64_t)c * n + i];
        beta0b[c] = bb;
        depb[c] += 2. / 3. * xi * H * bb;
        Hb += 2. / 3. * xi * dep[c] * bb;
    }
    double sdevb[6] = {0, 0, 0, 0, 0, 0}, seqb = 0.0;
    if (fabs(seq) > LPMX_EPS)
        for (int c = 0; c < 6; c++)
        {
            dlamb += depb[c] * 1.5 * sl[c] / seq;
            sdevb[c] += depb[c] * 1.5 * dlam / seq;
            seqb -= depb[c] * 1.5 * dlam * sl[c] / (seq * seq);
        }
    double c44b = 0.0, syb = 0.0;
    if (plastic)
    {
        const double den = 3 * c44 + H;
        const double yfb = dlamb / den;
        c44b -= dlamb * yf * 3.0 / (den * den);
        Hb -= dlamb * yf / (den * den);
        seqb += yfb;
        syb -= yfb;
        Hb -= yfb * (1.0 - xi) * alpha0;
        alpha0b -= yfb * (1.0 - xi) * H;
    }
    if (seq > 0.0)
        for (int c = 0; c < 6; c++)
            sdevb[c] += seqb * 1.5 * (c < 3 ? 1.0 : 2.0) * sl[c] / seq;
    double slb[6];
    const double mb = (sdevb[0] + sdevb[1] + sdevb[2]) / 3.0;
    for (int c = 0; c < 6; c++)
    {
        beta0b[c] -= sdevb[c];
        slb[c] = c < 3 ? sdevb[c] - mb : sdevb[c];
    }
    // trial stress, trial force: adjoint of the trial force of bond b, and the part of
    // its unit-vector adjoint that comes through the trial stress
    auto ftrial_bar = [&](int j, int64_t b, const VjpBond &g, double *cb) {
        double cc[6], ccb[6];
        cc6(g.c[0], g.c[1], g.c[2], cc);
        const double w = vjp_opp(v, i, j) / v.volume * v.L0[b];
        const double Ftr = ftrial(b, g.dl);
        double Ftrb = 0.0;
        for (int c = 0; c < 6; c++)
            Ftrb += slb[c] * w * cc[c], ccb[c] = slb[c] * w * Ftr;
        cb[0] = cb[1] = cb[2] = 0.0;
        cc6_bar(g.c[0], g.c[1], g.c[2], ccb, cb);
        return Ftrb * v.dw[b];
    };
    double A1b[2] = {0, 0}, B1b[2] = {0, 0};
    for (int j = 0; j < nbi; j++)
    {
        const int64_t b = BI(j, i);
        const int s = v.nsign[b];
        const VjpBond g = vjp_bond(v, i, b);
        double cb[3];
        const double t = ftrial_bar(j, b, g, cb);
        B1b[s] += t;
        A1b[s] += t * v.Tv[b];
        a.Knb[b] += t * 2.0 * g.dl;
        a.Tvb[b] += t * (s == 0 ? a0 : a1);
        vjp_geom(v, a, b, NB, g, 0.0, cb);
    }
    // first pass and geometry
    for (int j = 0; j < nbi; j++)
    {
        const int64_t b = BI(j, i);
        const int s = v.nsign[b];
        const VjpBond g = vjp_bond(v, i, b);
        double cb[3];
        const double dL1b = ftrial_bar(j, b, g, cb) * 2.0 * v.Kn[b];
        const double tot = dL1b + A1b[s] + B1b[s] * v.Tv[b];
        a.Tvb[b] += B1b[s] * g.dl;
        io.dLp0b[b] -= tot * v.db[b];
        const double zero[3] = {0.0, 0.0, 0.0};
        vjp_geom(v, a, b, NB, g, tot * v.db[b], zero);
    }
    for (int c = 0; c < 6; c++)
        io.beta0b[(int64_t)c * n + i] = beta0b[c];
    io.alpha0b[i] = alpha0b;
    io.syb[i] = syb;
    io.Hb[i] = Hb;
    a.c44b[i] = c44b;
}

// Positions from the bond vectors; the modulus factor from the bond constants.
__global__ void k_vjp_gather(DV v, VjpIO io, AdjWork a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const int64_t NB = (int64_t)v.n * v.nb;
    const int ti = v.type[i];
    double x0 = 0, x1 = 0, x2 = 0, e = 0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int64_t b = BI(j, i);
        const int q = v.nbr[b], s = v.nsign[b], r = v.rev[b];
        x0 += a.gb[b], x1 += a.gb[NB + b], x2 += a.gb[2 * NB + b];
        const double kn = v.lattice == 1 ? v.ktype_t[3 * ti] : v.ktype_t[3 * ti + s];
        const double tv = v.lattice == 1 ? v.ktype_t[3 * ti + 1] : v.ktype_t[3 * ti + 2];
        if (r >= 0)
        {
            const int64_t br = BI(r, q);
            x0 -= a.gb[br], x1 -= a.gb[NB + br], x2 -= a.gb[2 * NB + br];
            if (v.lattice == 2)
                e += 0.5 * (a.Knb[b] * kn + a.Tvb[b] * tv) + 0.5 * (a.Knb[br] * kn + a.Tvb[br] * tv);
        }
        else if (v.lattice == 2)
            e += 0.5 * (a.Knb[b] * kn + a.Tvb[b] * tv);
        if (v.lattice != 2)
            e += a.Knb[b] * kn + a.Tvb[b] * tv;
    }
    io.xb[3 * i] = x0, io.xb[3 * i + 1] = x1, io.xb[3 * i + 2] = x2;
    io.eb[i] = e + a.c44b[i] * v.ce_t[3 * ti + 2];
}

// The reverse sweep on device buffers, at the forward values left by the last
// lpmx_eval_j2 (positions and slot-0 state unchanged since).
int lpmx_vjp_device(const VjpIO &io)
{
    RETCHK(adj_work());
    const dim3 G(nblocks(g.n)), B(LPMX_BLOCK);
    hipLaunchKernelGGL(k_vjp_force, G, B, 0, g_stream, g, io, AW);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_vjp_particle, G, B, 0, g_stream, g, io, AW);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_vjp_gather, G, B, 0, g_stream, g, io, AW);
    LAUNCHCHK();
    return LPMX_OK;
}

// rows [n][w] <-> planes [w][n]
__global__ void k_adj_transpose(const double *src, double *dst, int n, int w, int to_planes)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * w)
        return;
    const int64_t i = t / w, c = t - i * w;
    if (to_planes)
        dst[c * n + i] = src[t];
    else
        dst[t] = src[c * n + i];
}

// One J2 evaluation at the current state, then its reverse sweep, host arrays in
// the reference layouts (Pbar [3n], dLp2bar [n][nb], beta2bar [n][6], alpha2bar [n];
// outputs likewise, Hbar a scalar).
extern "C" int lpmx_vjp(const double *Pb, const double *dLp2b, const double *beta2b, const double *alpha2b,
                        double *xb, double *dLp0b, double *beta0b, double *alpha0b, double *eb, double *syb,
                        double *Hb)
{
    RETCHK(lpmx_require_model());
    if (!Pb || !dLp2b || !beta2b || !alpha2b || !xb || !dLp0b || !beta0b || !alpha0b || !eb || !syb || !Hb)
        return lpmx_fail(LPMX_ERR_ARG, "lpmx_vjp: null argument");
    const int n = g.n, nb = g.nb;
    const int64_t NB = (int64_t)n * nb;
    // device scratch: seeds and outputs
    std::vector<double *> bufs;
    auto alloc = [&](int64_t c) -> double * {
        double *p = nullptr;
        if (hipMalloc(&p, sizeof(double) * c) != hipSuccess)
            return nullptr;
        bufs.push_back(p);
        return p;
    };
    double *dPb = alloc(3 * (int64_t)n), *tmp = alloc(NB > 6 * (int64_t)n ? NB : 6 * (int64_t)n);
    double *ddLp2b = alloc(NB), *dbeta2b = alloc(6 * (int64_t)n), *dalpha2b = alloc(n);
    double *dxb = alloc(3 * (int64_t)n), *ddLp0b = alloc(NB), *dbeta0b = alloc(6 * (int64_t)n);
    double *dalpha0b = alloc(n), *deb = alloc(n), *dsyb = alloc(n), *dHb = alloc(n);
    int rc = LPMX_OK;
    for (double *p : bufs)
        if (!p)
            rc = lpmx_fail(LPMX_ERR_HIP, "lpmx_vjp: allocation");
    if (rc == LPMX_OK)
        do
        {
            const dim3 B(LPMX_BLOCK);
            if ((rc = lpmx_eval_j2()) != LPMX_OK)
                break;
            hipMemcpyAsync(dPb, Pb, sizeof(double) * 3 * n, hipMemcpyHostToDevice, g_stream);
            hipMemcpyAsync(tmp, dLp2b, sizeof(double) * NB, hipMemcpyHostToDevice, g_stream);
            hipLaunchKernelGGL(k_adj_transpose, dim3(nblocks(NB)), B, 0, g_stream, tmp, ddLp2b, n, nb, 1);
            hipMemcpyAsync(tmp, beta2b, sizeof(double) * 6 * n, hipMemcpyHostToDevice, g_stream);
            hipLaunchKernelGGL(k_adj_transpose, dim3(nblocks(6 * (int64_t)n)), B, 0, g_stream, tmp, dbeta2b, n, 6, 1);
            hipMemcpyAsync(dalpha2b, alpha2b, sizeof(double) * n, hipMemcpyHostToDevice, g_stream);
            VjpIO io = {dPb, ddLp2b, dbeta2b, dalpha2b, dxb, ddLp0b, dbeta0b, dalpha0b, deb, dsyb, dHb};
            if ((rc = lpmx_vjp_device(io)) != LPMX_OK)
                break;
            hipMemcpyAsync(xb, dxb, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, g_stream);
            hipLaunchKernelGGL(k_adj_transpose, dim3(nblocks(NB)), B, 0, g_stream, ddLp0b, tmp, n, nb, 0);
            hipMemcpyAsync(dLp0b, tmp, sizeof(double) * NB, hipMemcpyDeviceToHost, g_stream);
            hipStreamSynchronize(g_stream);
            hipLaunchKernelGGL(k_adj_transpose, dim3(nblocks(6 * (int64_t)n)), B, 0, g_stream, dbeta0b, tmp, n, 6, 0);
            hipMemcpyAsync(beta0b, tmp, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, g_stream);
            hipMemcpyAsync(alpha0b, dalpha0b, sizeof(double) * n, hipMemcpyDeviceToHost, g_stream);
            hipMemcpyAsync(eb, deb, sizeof(double) * n, hipMemcpyDeviceToHost, g_stream);
            hipMemcpyAsync(syb, dsyb, sizeof(double) * n, hipMemcpyDeviceToHost, g_stream);
            double hsum = 0.0;
            if ((rc = lpmx_dev_sum(dHb, n, &hsum)) != LPMX_OK)
                break;
            *Hb = hsum;
            if (hipStreamSynchronize(g_stream) != hipSuccess || hipGetLastError() != hipSuccess)
                rc = lpmx_fail(LPMX_ERR_HIP, "lpmx_vjp: device error");
        } while (0);
    for (double *p : bufs)
        if (p)
            hipFree(p);
    return rc;
}

void lpmx_adjoint_free(void)
{
    lpmx_adjoint_free_work();
    lpmx_history_free();
}
