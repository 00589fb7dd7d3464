// Discrete adjoint of a J2 load history (no counterpart in the reference).
//
// Forward: the load steps of the reference driver (src/lpmc_project.c:382-546) with
// the damage update off, so every step n is an equilibrium
//     P_f(x_n, s_{n-1}; theta) = Pex_f,   s_n = S(x_n, s_{n-1}; theta)
// on the free DOFs f (x_n on the prescribed DOFs is fixed by the BC tables), with
// s the converged slot state (plastic bond stretch, back stress, hardening variable)
// and theta the per-particle modulus factor and yield stress and H.  The objective
// is J(R_1..R_N) with R_n the writeReaction value -sum_{type t1} Pin_axis.
//
// Reverse, step n = N..1, with Rbar_n = dJ/dR_n and the state adjoint sbar_n (zero
// after the last step):
//     b      = [VJP_x(P = c Rbar_n, s = sbar_n)]_f                c: the reaction seed
//     J_ff^T lambda = b                                           the adjoint solve
//     (sbar_{n-1}, theta_bar_n) = VJP(P = c Rbar_n - lambda, s = sbar_n)
// and theta_bar = sum_n theta_bar_n.  J is the consistent Jacobian of the return
// map, which the forward's modified Newton never forms; it enters only through VJPs
// (adjoint.hip).  The adjoint solve is the fixed-point iteration
//     lambda <- lambda + K^-1 (b - J^T lambda)
// preconditioned by the step's elastic tangent K (constraints eliminated, the same
// block-Jacobi PCG as the forward) and Anderson-accelerated like the forward's
// Newton loop; lambda of step n+1 starts step n.
//
// Memory: a checkpoint before every ckpt_every-th step (everything a step reads
// before writing it: positions, bond forces and direction cosines, Pin, Pex, the
// converged slot state, constraint masks) and records of the steps of one segment; the reverse sweep re-runs a segment's steps from its checkpoint
// (bitwise the same trajectory: every reduction has a fixed order).
#include <chrono>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "common.h"

static double now_s(void)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A line per step on stderr when the environment variable LPMX_PROGRESS is set
// (long histories otherwise run silently inside one call).
static bool progress(void)
{
    static int on = -1;
    if (on < 0)
        on = getenv("LPMX_PROGRESS") != nullptr;
    return on == 1;
}

// State before a step (a checkpoint) or after the last one.
struct Snap
{
    double *x = nullptr, *F = nullptr, *Pex = nullptr, *Pin = nullptr, *dLave = nullptr;
    double *cs[3] = {nullptr, nullptr, nullptr}; // the predictor projects on the last direction cosines
    double *dLp = nullptr, *beta = nullptr, *alpha = nullptr;
    int *dispBC = nullptr, *fix = nullptr;
};

// What the reverse sweep needs of one step: converged positions x_n, incoming
// state s_{n-1}, the step's constraint masks.
struct Rec
{
    double *x = nullptr, *dLp = nullptr, *beta = nullptr, *alpha = nullptr;
    int *dispBC = nullptr, *fix = nullptr;
};

static struct Hist
{
    bool valid = false;
    int n = 0, nb = 0, dim = 0, nsteps = 0, k = 0, maxbc = 0, t1 = 0;
    char axis = 'z';
    lpmx_solver s{};
    std::vector<int> nbd, nbf;
    std::vector<lpmx_dispbc> dbc;
    std::vector<lpmx_forcebc> fbc;
    std::vector<Snap> ck;  // ck[j]: before step j*k
    Snap fin;              // after the last step
    std::vector<Rec> rec;  // rec[st - rec_first] for the steps of one segment
    int rec_first = 0, rec_count = 0;
    std::vector<double> reac;
    double t_forward = 0.0, bytes = 0.0;
    std::vector<void *> mem;
} H;

static AaHist *g_adj_aa = nullptr; // Anderson history of the adjoint iteration

static int halloc(void **p, size_t bytes)
{
    HIPCHK(hipMalloc(p, bytes));
    H.mem.push_back(*p);
    H.bytes += (double)bytes;
    return LPMX_OK;
}
template <typename T>
static int halloc_n(T **p, int64_t count)
{
    return halloc((void **)p, sizeof(T) * (size_t)count);
}

extern "C" void lpmx_history_free(void)
{
    for (void *p : H.mem)
        hipFree(p);
    H = Hist();
    lpmx_aa_destroy(g_adj_aa);
    g_adj_aa = nullptr;
}

static int alloc_snap(Snap &c)
{
    const int64_t n = H.n, nbn = (int64_t)H.n * H.nb, m = (int64_t)H.n * H.dim;
    RETCHK(halloc_n(&c.x, 3 * n));
    RETCHK(halloc_n(&c.F, nbn));
    RETCHK(halloc_n(&c.Pex, m));
    RETCHK(halloc_n(&c.Pin, 3 * n));
    RETCHK(halloc_n(&c.dLave, nbn));
    for (int q = 0; q < 3; q++)
        RETCHK(halloc_n(&c.cs[q], nbn));
    RETCHK(halloc_n(&c.dLp, nbn));
    RETCHK(halloc_n(&c.beta, 6 * n));
    RETCHK(halloc_n(&c.alpha, n));
    RETCHK(halloc_n(&c.dispBC, m));
    RETCHK(halloc_n(&c.fix, m));
    return LPMX_OK;
}

static int alloc_rec(Rec &r)
{
    const int64_t n = H.n, nbn = (int64_t)H.n * H.nb, m = (int64_t)H.n * H.dim;
    RETCHK(halloc_n(&r.x, 3 * n));
    RETCHK(halloc_n(&r.dLp, nbn));
    RETCHK(halloc_n(&r.beta, 6 * n));
    RETCHK(halloc_n(&r.alpha, n));
    RETCHK(halloc_n(&r.dispBC, m));
    RETCHK(halloc_n(&r.fix, m));
    return LPMX_OK;
}

template <typename T>
static int dcopy(T *dst, const T *src, int64_t count)
{
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * (size_t)count, hipMemcpyDeviceToDevice, g_stream));
    return LPMX_OK;
}

// The converged state lives in slot 1 between steps (switchStateV(1) at the end of
// a step); the next step's first switchStateV(0) copies it into slot 0.
static int save_snap(Snap &c)
{
    const int64_t n = H.n, nbn = (int64_t)H.n * H.nb, m = (int64_t)H.n * H.dim;
    RETCHK(dcopy(c.x, g.x, 3 * n));
    RETCHK(dcopy(c.F, g.F, nbn));
    RETCHK(dcopy(c.Pex, g.Pex, m));
    RETCHK(dcopy(c.Pin, g.Pin, 3 * n));
    RETCHK(dcopy(c.dLave, g.dLave, nbn));
    RETCHK(dcopy(c.cs[0], g.csx, nbn));
    RETCHK(dcopy(c.cs[1], g.csy, nbn));
    RETCHK(dcopy(c.cs[2], g.csz, nbn));
    RETCHK(dcopy(c.dLp, g.dLp1, nbn));
    RETCHK(dcopy(c.beta, g.beta1, 6 * n));
    RETCHK(dcopy(c.alpha, g.alpha1, n));
    RETCHK(dcopy(c.dispBC, g.dispBC, m));
    RETCHK(dcopy(c.fix, g.fix, m));
    return LPMX_OK;
}

static int load_snap(const Snap &c)
{
    const int64_t n = H.n, nbn = (int64_t)H.n * H.nb, m = (int64_t)H.n * H.dim;
    RETCHK(dcopy(g.x, c.x, 3 * n));
    RETCHK(dcopy(g.F, c.F, nbn));
    RETCHK(dcopy(g.Pex, c.Pex, m));
    RETCHK(dcopy(g.Pin, c.Pin, 3 * n));
    RETCHK(dcopy(g.dLave, c.dLave, nbn));
    RETCHK(dcopy(g.csx, c.cs[0], nbn));
    RETCHK(dcopy(g.csy, c.cs[1], nbn));
    RETCHK(dcopy(g.csz, c.cs[2], nbn));
    RETCHK(dcopy(g.dLp0, c.dLp, nbn));
    RETCHK(dcopy(g.dLp1, c.dLp, nbn));
    RETCHK(dcopy(g.beta0, c.beta, 6 * n));
    RETCHK(dcopy(g.beta1, c.beta, 6 * n));
    RETCHK(dcopy(g.alpha0, c.alpha, n));
    RETCHK(dcopy(g.alpha1, c.alpha, n));
    RETCHK(dcopy(g.dispBC, c.dispBC, m));
    RETCHK(dcopy(g.fix, c.fix, m));
    return LPMX_OK;
}

// One forward step st with its record in r (r may be null: no record).
static int forward_step(int st, Rec *r, lpmx_step_stats *o)
{
    const int64_t n = H.n, nbn = (int64_t)H.n * H.nb, m = (int64_t)H.n * H.dim;
    if (r)
    {
        RETCHK(dcopy(r->dLp, g.dLp1, nbn));
        RETCHK(dcopy(r->beta, g.beta1, 6 * n));
        RETCHK(dcopy(r->alpha, g.alpha1, n));
    }
    lpmx_step_stats so;
    RETCHK(lpmx_run_steps_impl(&H.s, 1, &H.nbd[st], &H.dbc[(size_t)st * H.maxbc], &H.nbf[st],
                               &H.fbc[(size_t)st * H.maxbc], H.maxbc, H.axis, H.t1, &so, 1, 1));
    if (r)
    {
        RETCHK(dcopy(r->x, g.x, 3 * n));
        RETCHK(dcopy(r->dispBC, g.dispBC, m));
        RETCHK(dcopy(r->fix, g.fix, m));
    }
    if (o)
        *o = so;
    return LPMX_OK;
}

extern "C" int lpmx_history_forward(const lpmx_solver *s, int ckpt_every, int nsteps, const int *nbd,
                                    const lpmx_dispbc *dbc, const int *nbf, const lpmx_forcebc *fbc, int maxbc,
                                    char out_axis, int out_type, double *reactions, lpmx_step_stats *out)
{
    RETCHK(lpmx_require_pattern());
    if (!s || nsteps <= 0 || !nbd || !nbf || maxbc <= 0 || ckpt_every < 0 || !reactions)
        return lpmx_fail(LPMX_ERR_ARG, "invalid arguments to lpmx_history_forward");
    if (g.plmode != 0 || g.dim != 3)
        return lpmx_fail(LPMX_ERR_ARG, "the history adjoint covers the 3D J2 law (plmode 0), not plmode %d in %dD",
                         g.plmode, g.dim);
    lpmx_history_free();
    const double t0 = now_s();
    H.n = g.n, H.nb = g.nb, H.dim = g.dim;
    H.nsteps = nsteps, H.maxbc = maxbc, H.axis = out_axis, H.t1 = out_type, H.s = *s;
    H.k = ckpt_every == 0 || ckpt_every > nsteps ? nsteps : ckpt_every;
    H.nbd.assign(nbd, nbd + nsteps);
    H.nbf.assign(nbf, nbf + nsteps);
    H.dbc.assign((size_t)nsteps * maxbc, lpmx_dispbc{});
    H.fbc.assign((size_t)nsteps * maxbc, lpmx_forcebc{});
    if (dbc)
        memcpy(H.dbc.data(), dbc, sizeof(lpmx_dispbc) * (size_t)nsteps * maxbc);
    if (fbc)
        memcpy(H.fbc.data(), fbc, sizeof(lpmx_forcebc) * (size_t)nsteps * maxbc);
    const int nseg = (nsteps + H.k - 1) / H.k;
    H.ck.resize(nseg > 1 ? nseg : 0);
    for (auto &c : H.ck)
        RETCHK(alloc_snap(c));
    RETCHK(alloc_snap(H.fin));
    H.rec.resize(H.k);
    for (auto &r : H.rec)
        RETCHK(alloc_rec(r));
    H.reac.assign(nsteps, 0.0);
    for (int st = 0; st < nsteps; st++)
    {
        const int seg = st / H.k;
        if (st % H.k == 0)
        {
            if (!H.ck.empty())
                RETCHK(save_snap(H.ck[seg]));
            H.rec_first = st;
            H.rec_count = 0;
        }
        lpmx_step_stats o;
        RETCHK(forward_step(st, &H.rec[st - H.rec_first], &o));
        if (progress())
            fprintf(stderr, "[lpmx] history step %d/%d: Newton %d (restarts %d), CG %d, %.2f s, status %d, "
                    "reaction %.6e\n", st + 1, nsteps, o.newton_iters, o.accel_restarts, o.cg_iters, o.t_total,
                    o.status, o.reaction);
        H.rec_count++;
        H.reac[st] = o.reaction;
        reactions[st] = o.reaction;
        if (out)
            out[st] = o;
    }
    RETCHK(save_snap(H.fin));
    HIPCHK(hipStreamSynchronize(g_stream));
    H.t_forward = now_s() - t0;
    H.valid = true;
    return LPMX_OK;
}

// ------------------------------------------------------------------ kernels ------
// Pb = val on component c of the particles of type t1, 0 elsewhere ([3n]).
__global__ void k_reaction_seed(int n, const int *type, int t1, int c, double val, double *Pb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const bool on = type[i] == t1;
    for (int k = 0; k < 3; k++)
        Pb[3 * i + k] = on && k == c ? val : 0.0;
}

__device__ __forceinline__ bool dof_free(const int *dispBC, const int *fix, int64_t t)
{
    return dispBC[t] != 0 && fix[t] != 0;
}

// out = free ? (b ? b - a : a) : 0 over n*3 DOFs
__global__ void k_free_diff(int64_t m, const int *dispBC, const int *fix, const double *b, const double *a,
                            double *out)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= m)
        return;
    out[t] = dof_free(dispBC, fix, t) ? (b ? b[t] - a[t] : a[t]) : 0.0;
}

// lam = free ? lam : 0 (a warm start under the step's own constraints)
__global__ void k_free_only(int64_t m, const int *dispBC, const int *fix, double *lam)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < m && !dof_free(dispBC, fix, t))
        lam[t] = 0.0;
}

// y += a x
__global__ void k_axpy(int64_t m, double a, const double *x, double *y)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < m)
        y[t] += a * x[t];
}

// Pb = seed - lam
__global__ void k_seed_minus(int64_t m, const double *seed, const double *lam, double *Pb)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < m)
        Pb[t] = seed[t] - lam[t];
}

// ------------------------------------------------------------ reverse sweep ------
struct Rev
{
    double *seed = nullptr, *Pb = nullptr, *xb = nullptr, *b = nullptr, *r = nullptr, *z = nullptr, *lam = nullptr;
    double *sdLp[2] = {nullptr, nullptr}, *sbeta[2] = {nullptr, nullptr}, *salpha[2] = {nullptr, nullptr};
    double *zdLp = nullptr, *zbeta = nullptr, *zalpha = nullptr; // zero seeds
    double *tdLp = nullptr, *tbeta = nullptr, *talpha = nullptr; // discarded outputs
    double *eb = nullptr, *syb = nullptr, *Hb = nullptr;
    double *ge = nullptr, *gy = nullptr, *gH = nullptr;
    std::vector<void *> mem;
};

static int ralloc(Rev &R, double **p, int64_t count, bool zero)
{
    HIPCHK(hipMalloc(p, sizeof(double) * (size_t)count));
    R.mem.push_back(*p);
    if (zero)
        HIPCHK(hipMemsetAsync(*p, 0, sizeof(double) * (size_t)count, g_stream));
    return LPMX_OK;
}

static void rfree(Rev &R)
{
    for (void *p : R.mem)
        hipFree(p);
    R.mem.clear();
}

static int restore_rec(const Rec &r)
{
    const int64_t n = H.n, nbn = (int64_t)H.n * H.nb, m = (int64_t)H.n * H.dim;
    RETCHK(dcopy(g.x, r.x, 3 * n));
    RETCHK(dcopy(g.dLp0, r.dLp, nbn));
    RETCHK(dcopy(g.beta0, r.beta, 6 * n));
    RETCHK(dcopy(g.alpha0, r.alpha, n));
    RETCHK(dcopy(g.dispBC, r.dispBC, m));
    RETCHK(dcopy(g.fix, r.fix, m));
    return LPMX_OK;
}

static int reverse_step(Rev &R, int cur, const lpmx_adjoint_opts *o, double Rbar, lpmx_adjoint_stats *st)
{
    const int n = H.n;
    const int64_t m = 3 * (int64_t)n;
    const dim3 B(LPMX_BLOCK), Gm(nblocks(m));
    const int c = H.axis == 'x' ? 0 : (H.axis == 'y' ? 1 : 2);
    RETCHK(lpmx_eval_j2());
    // preconditioner: the elastic tangent at x_n with the step's constraints
    RETCHK(lpmx_tangent());
    RETCHK(lpmx_bc_eliminate_ex(1));
    // R = -sum_{type t1} Pin_c, so the seed on Pin is -Rbar there
    hipLaunchKernelGGL(k_reaction_seed, dim3(nblocks(n)), B, 0, g_stream, n, g.type, H.t1, c, -Rbar, R.seed);
    LAUNCHCHK();
    const int nxt = 1 - cur;
    VjpIO io0 = {R.seed, R.sdLp[cur], R.sbeta[cur], R.salpha[cur], R.xb, R.tdLp, R.tbeta, R.talpha,
                 R.eb, R.syb, R.Hb};
    RETCHK(lpmx_vjp_device(io0));
    hipLaunchKernelGGL(k_free_diff, Gm, B, 0, g_stream, m, g.dispBC, g.fix, nullptr, R.xb, R.b);
    LAUNCHCHK();
    double bb = 0.0;
    RETCHK(lpmx_dev_sumsq(R.b, m, &bb));
    hipLaunchKernelGGL(k_free_only, Gm, B, 0, g_stream, m, g.dispBC, g.fix, R.lam);
    LAUNCHCHK();
    // J^T lambda = b by the preconditioned, accelerated fixed point
    if (!g_adj_aa)
        g_adj_aa = lpmx_aa_create();
    AaHist *aa = g_adj_aa;
    RETCHK(lpmx_aa_begin(aa, o->depth));
    VjpIO io1 = {R.lam, R.zdLp, R.zbeta, R.zalpha, R.xb, R.tdLp, R.tbeta, R.talpha, R.eb, R.syb, R.Hb};
    int it = 0;
    double rr = 0.0;
    for (;;)
    {
        if (!(bb > 0.0))
        {
            HIPCHK(hipMemsetAsync(R.lam, 0, sizeof(double) * m, g_stream));
            break;
        }
        RETCHK(lpmx_vjp_device(io1));
        hipLaunchKernelGGL(k_free_diff, Gm, B, 0, g_stream, m, g.dispBC, g.fix, R.b, R.xb, R.r);
        LAUNCHCHK();
        RETCHK(lpmx_dev_sumsq(R.r, m, &rr));
        if (!isfinite(rr))
            return lpmx_fail(LPMX_ERR_NOCONV, "adjoint iteration diverged");
        if (rr <= o->tol * o->tol * bb)
            break;
        if (it >= o->maxit)
            return lpmx_fail(LPMX_ERR_NOCONV, "adjoint solve: %d iterations left |r|/|b| = %.3e", it, sqrt(rr / bb));
        int cg = 0, stt = 0;
        double r2 = 0.0;
        RETCHK(lpmx_solve_into(o->precision, o->rtol2, 0.0, o->cg_maxit, H.s.inner_rtol2, H.s.ir_maxit, R.r, R.z,
                               &cg, &r2, &stt));
        if (stt == 3)
            return lpmx_fail(LPMX_ERR_NOCONV, "adjoint preconditioner solve broke down");
        st->adj_cg_iters += cg;
        hipLaunchKernelGGL(k_axpy, Gm, B, 0, g_stream, m, 1.0, R.z, R.lam);
        LAUNCHCHK();
        int mixed = 0;
        RETCHK(lpmx_aa_step(aa, R.lam, R.z, &mixed));
        it++;
    }
    st->adj_iters += it;
    if (it > st->adj_iters_max)
        st->adj_iters_max = it;
    if (bb > 0.0 && sqrt(rr / bb) > st->max_rel_res)
        st->max_rel_res = sqrt(rr / bb);
    // state and parameter adjoints
    hipLaunchKernelGGL(k_seed_minus, Gm, B, 0, g_stream, m, R.seed, R.lam, R.Pb);
    LAUNCHCHK();
    VjpIO io2 = {R.Pb, R.sdLp[cur], R.sbeta[cur], R.salpha[cur], R.xb, R.sdLp[nxt], R.sbeta[nxt], R.salpha[nxt],
                 R.eb, R.syb, R.Hb};
    RETCHK(lpmx_vjp_device(io2));
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(n)), B, 0, g_stream, (int64_t)n, 1.0, R.eb, R.ge);
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(n)), B, 0, g_stream, (int64_t)n, 1.0, R.syb, R.gy);
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(n)), B, 0, g_stream, (int64_t)n, 1.0, R.Hb, R.gH);
    LAUNCHCHK();
    if (bb > 0.0)
    {
        hipLaunchKernelGGL(k_free_diff, Gm, B, 0, g_stream, m, g.dispBC, g.fix, nullptr, R.xb, R.r);
        LAUNCHCHK();
        double ff = 0.0;
        RETCHK(lpmx_dev_sumsq(R.r, m, &ff));
        if (sqrt(ff / bb) > st->max_rel_free)
            st->max_rel_free = sqrt(ff / bb);
    }
    return LPMX_OK;
}

extern "C" int lpmx_history_backward(const lpmx_adjoint_opts *o, const double *Rbar, double *g_emod,
                                     double *g_sigmay, double *g_H, lpmx_adjoint_stats *st)
{
    RETCHK(lpmx_require_pattern());
    if (!H.valid || H.n != g.n || H.nb != g.nb)
        return lpmx_fail(LPMX_ERR_STATE, "no recorded history for this model (lpmx_history_forward first)");
    if (!o || !Rbar || !g_emod || !g_sigmay || !g_H || o->depth < 0 || o->depth > 32 || o->maxit < 0 ||
        !(o->tol > 0.0))
        return lpmx_fail(LPMX_ERR_ARG, "invalid arguments to lpmx_history_backward");
    // the K^-1 solves use the forward's preconditioner and mixed-form segments
    RETCHK(lpmx_set_preconditioner(H.s.precond));
    lpmx_adjoint_stats S;
    memset(&S, 0, sizeof(S));
    S.t_forward = H.t_forward;
    S.checkpoints = (int)H.ck.size();
    S.ckpt_bytes = H.bytes;
    const int n = H.n;
    const int64_t nbn = (int64_t)n * H.nb, m = 3 * (int64_t)n;
    Rev R;
    int rc = LPMX_OK;
    auto body = [&]() -> int {
        RETCHK(ralloc(R, &R.seed, m, false));
        RETCHK(ralloc(R, &R.Pb, m, false));
        RETCHK(ralloc(R, &R.xb, m, false));
        RETCHK(ralloc(R, &R.b, m, false));
        RETCHK(ralloc(R, &R.r, m, false));
        RETCHK(ralloc(R, &R.z, m, false));
        RETCHK(ralloc(R, &R.lam, m, true));
        for (int q = 0; q < 2; q++)
        {
            RETCHK(ralloc(R, &R.sdLp[q], nbn, true));
            RETCHK(ralloc(R, &R.sbeta[q], 6 * (int64_t)n, true));
            RETCHK(ralloc(R, &R.salpha[q], n, true));
        }
        RETCHK(ralloc(R, &R.zdLp, nbn, true));
        RETCHK(ralloc(R, &R.zbeta, 6 * (int64_t)n, true));
        RETCHK(ralloc(R, &R.zalpha, n, true));
        RETCHK(ralloc(R, &R.tdLp, nbn, false));
        RETCHK(ralloc(R, &R.tbeta, 6 * (int64_t)n, false));
        RETCHK(ralloc(R, &R.talpha, n, false));
        RETCHK(ralloc(R, &R.eb, n, false));
        RETCHK(ralloc(R, &R.syb, n, false));
        RETCHK(ralloc(R, &R.Hb, n, false));
        RETCHK(ralloc(R, &R.ge, n, true));
        RETCHK(ralloc(R, &R.gy, n, true));
        RETCHK(ralloc(R, &R.gH, n, true));
        const double t0 = now_s();
        double t_re = 0.0;
        int cur = 0;
        const int nseg = (H.nsteps + H.k - 1) / H.k;
        for (int seg = nseg - 1; seg >= 0; seg--)
        {
            const int a = seg * H.k, e = a + H.k < H.nsteps ? a + H.k : H.nsteps;
            if (H.rec_first != a)
            {
                // re-run the segment from its checkpoint, recording its steps
                const double tr = now_s();
                RETCHK(load_snap(H.ck[seg]));
                for (int s2 = a; s2 < e; s2++)
                {
                    RETCHK(forward_step(s2, &H.rec[s2 - a], nullptr));
                    if (progress())
                        fprintf(stderr, "[lpmx] recomputed step %d\n", s2 + 1);
                }
                H.rec_first = a;
                H.rec_count = e - a;
                S.recomputed_steps += e - a;
                HIPCHK(hipStreamSynchronize(g_stream));
                t_re += now_s() - tr;
            }
            for (int s2 = e - 1; s2 >= a; s2--)
            {
                const double ts = now_s();
                const int i0 = S.adj_iters, c0 = S.adj_cg_iters;
                RETCHK(restore_rec(H.rec[s2 - a]));
                RETCHK(reverse_step(R, cur, o, Rbar[s2], &S));
                cur = 1 - cur;
                if (progress())
                    fprintf(stderr, "[lpmx] adjoint step %d: %d iterations, CG %d, %.2f s\n", s2 + 1,
                            S.adj_iters - i0, S.adj_cg_iters - c0, now_s() - ts);
            }
        }
        HIPCHK(hipMemcpyAsync(g_emod, R.ge, sizeof(double) * n, hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipMemcpyAsync(g_sigmay, R.gy, sizeof(double) * n, hipMemcpyDeviceToHost, g_stream));
        RETCHK(lpmx_dev_sum(R.gH, n, g_H));
        // the model as the forward left it
        RETCHK(load_snap(H.fin));
        HIPCHK(hipStreamSynchronize(g_stream));
        S.t_recompute = t_re;
        S.t_backward = now_s() - t0 - t_re;
        return LPMX_OK;
    };
    rc = body();
    rfree(R);
    if (st)
        *st = S;
    return rc;
}
