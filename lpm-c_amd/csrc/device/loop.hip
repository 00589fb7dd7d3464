// The implicit load-step loop of the reference driver (src/lpmc_project.c:382-546)
// with every array resident on the device.  The host only takes the convergence
// decisions: two FP64 norms per Newton iteration come back, nothing else.
// With accel_depth > 0 the modified-Newton iterates are Anderson-accelerated
// (accel.hip); an accelerated iterate whose residual norm is not finite or grew
// more than 100-fold is replaced by the plain iterate and the history restarts, and
// the history also restarts when the residual has not halved over the last `depth`
// iterations (differences taken before particles switched between elastic and
// plastic mislead the mixing; on the hard steps of long plastic histories this
// halves the iteration count).
#include <chrono>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "common.h"

int lpmx_last_ir_passes(void);

static double now_s(void)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define RING 64 // residual norms of the last iterations (> the largest Anderson depth)

static int sync(void)
{
    HIPCHK(hipStreamSynchronize(g_stream));
    return LPMX_OK;
}

// LPMX_PROGRESS set: a line to stderr every 50 Newton iterations of a step
static bool loop_progress(void)
{
    static int on = -1;
    if (on < 0)
        on = getenv("LPMX_PROGRESS") != nullptr;
    return on;
}

// A re-solve after bond breakage keeps the Anderson differences of the step's earlier
// solves (lpmx_accel_rebase): the broken bonds change the map little, and on the 1M
// brittle shear the re-solves of a cascade take half the iterations of a new history,
// with the same bonds broken.  LPMX_AA_WARM=0 starts a new history instead.
static bool aa_warm(void)
{
    static int on = -1;
    if (on < 0)
    {
        const char *e = getenv("LPMX_AA_WARM");
        on = !(e && atoi(e) == 0);
    }
    return on;
}

int lpmx_run_steps_impl(const lpmx_solver *s, int nsteps, const int *nbd, const lpmx_dispbc *dbc, const int *nbf,
                        const lpmx_forcebc *fbc, int maxbc, char out_axis, int out_type, lpmx_step_stats *out,
                        int skip_damage, int newton_only)
{
    RETCHK(lpmx_require_pattern());
    if (!s || nsteps < 0 || (nsteps > 0 && (!nbd || !nbf)) || maxbc <= 0)
        return lpmx_fail(LPMX_ERR_ARG, "invalid arguments to lpmx_run_steps");
    const int n = g.n, d = g.dim;
    const int64_t nbn = (int64_t)n * g.nb;
    int rc_all = LPMX_OK;
    RETCHK(lpmx_set_preconditioner(s->precond));
    for (int st = 0; st < nsteps; st++)
    {
        lpmx_step_stats o;
        memset(&o, 0, sizeof(o));
        const double t0 = now_s();
        // xyz_temp, F_temp, Pex_temp (lpmc_project.c:387-389)
        HIPCHK(hipMemcpyAsync(g.xt, g.x, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, g_stream));
        HIPCHK(hipMemcpyAsync(g.Ft, g.F, sizeof(double) * nbn, hipMemcpyDeviceToDevice, g_stream));
        HIPCHK(hipMemcpyAsync(g.Pext, g.Pex, sizeof(double) * (int64_t)d * n, hipMemcpyDeviceToDevice, g_stream));
        // elastic tangent of the step (lpmc_project.c:391-396)
        RETCHK(lpmx_tangent());
        RETCHK(sync());
        const double t1 = now_s();
        o.t_tangent += t1 - t0;
        RETCHK(lpmx_set_disp_bc(nbd[st], dbc + (int64_t)st * maxbc));
        RETCHK(lpmx_set_force_bc(nbf[st], fbc + (int64_t)st * maxbc));
        RETCHK(lpmx_bond_force(4)); // predictor (lpmc_project.c:405)
        bool resolve = false;
        for (;;)
        {
            double nres = 0, nrea = 0;
            RETCHK(lpmx_update_rr(&nres, &nrea, nullptr));
            const double tolm = nres > nrea ? nres : nrea;
            int ni = 0;
            if (resolve && aa_warm())
                lpmx_accel_rebase();
            else
                RETCHK(lpmx_accel_begin(s->accel_depth, s->precision));
            double nprev = nres;
            double ring[RING];
            int since = 0;
            ring[0] = nres;
            // Within the loop every force evaluation reads slot 0 (= slot 1 from the first
            // iteration on) and writes slot 2 and the forces; the stress and the slot
            // 2 -> 0 copy that follow each evaluation in the reference
            // (constitutive.c:144-145) are needed only after the last one, below.
            while (nres > s->newton_tol * tolm && ni < s->newton_maxit)
            {
                if (ni == 0)
                    RETCHK(lpmx_switch_state(0));
                // K and the constraints stay within the loop: the first pass eliminates,
                // the later ones only zero the constrained residual
                RETCHK(ni == 0 ? lpmx_bc_eliminate_ex(1) : lpmx_bc_eliminate_repeat());
                double ts = now_s();
                int it = 0, stt = 0;
                double rr = 0;
                RETCHK(lpmx_solve_ex(s->precision, s->rtol2, s->atol2, s->maxit, s->inner_rtol2, s->ir_maxit, &it,
                                     &rr, &stt));
                int mixed = 0;
                RETCHK(lpmx_accel_step(&mixed));
                // no synchronisation for the timing split: the mixing step has read its
                // dot products back (and a step without mixing is timed to the launch)
                double tf = now_s();
                o.t_solve += tf - ts;
                o.cg_iters += it;
                o.ir_passes += lpmx_last_ir_passes();
                if (stt != 1)
                    o.status |= stt == 3 ? 4 : 2;
                // plmode 5 advances slot 0 in place: every evaluation starts from the
                // converged state, as the reference's switchStateV(0) at each iteration gives
                if (g.plmode == 5)
                    RETCHK(lpmx_switch_state(0));
                RETCHK(lpmx_bond_force_range(g.plmode, 0, n));
                RETCHK(lpmx_update_rr(&nres, nullptr, nullptr));
                if (mixed)
                {
                    o.accel_mixed++;
                    if (!(nres <= 1e2 * nprev))
                    {
                        o.accel_reverts++;
                        RETCHK(lpmx_accel_revert());
                        if (g.plmode == 5)
                            RETCHK(lpmx_switch_state(0));
                        RETCHK(lpmx_bond_force_range(g.plmode, 0, n));
                        RETCHK(lpmx_update_rr(&nres, nullptr, nullptr));
                    }
                }
                o.t_force += now_s() - tf;
                nprev = nres;
                ni++;
                since++;
                if (s->accel_depth > 0 && since > s->accel_depth &&
                    nres > 0.5 * ring[(since - s->accel_depth) % RING])
                {
                    lpmx_accel_forget();
                    o.accel_restarts++;
                    since = 0;
                }
                ring[since % RING] = nres;
                if (ni % 50 == 0 && loop_progress())
                    fprintf(stderr, "[lpmx] step %d: Newton %d, |R| %.3e (target %.3e), CG %d, %.2f s\n", st, ni,
                            nres, s->newton_tol * tolm, o.cg_iters, now_s() - t0);
            }
            if (ni > 0)
            {
                RETCHK(lpmx_compute_stress());
                RETCHK(lpmx_switch_state(2));
            }
            if (ni >= s->newton_maxit)
                o.status |= 1;
            o.newton_iters += ni;
            o.res_norm = nres;
            o.tol_mult = tolm;
            int broken = 0;
            if (!skip_damage)
                RETCHK(lpmx_damage(g.plmode, &broken));
            RETCHK(lpmx_update_crack());
            RETCHK(lpmx_switch_state(1));
            RETCHK(lpmx_reaction(out_axis, out_type, &o.reaction, &o.disp));
            o.broken += broken;
            if (broken > 0)
            {
                if (++o.resolves > 100000)
                    return lpmx_fail(LPMX_ERR_NOCONV, "bond breakage did not settle in step %d", st);
                double tb = now_s();
                RETCHK(lpmx_tangent());
                RETCHK(sync());
                o.t_tangent += now_s() - tb;
                resolve = true;
                continue;
            }
            break;
        }
        if (g.plmode == 1)
        {
            int fails = 0;
            RETCHK(lpmx_cp_failures(&fails));
            if (fails > 0) // the reference exits here (constitutive.c:1235-1241)
                return lpmx_fail(LPMX_ERR_NOCONV, "singular slip Jacobian in %d return maps of step %d", fails, st);
        }
        RETCHK(sync());
        o.t_total = now_s() - t0;
        o.t_other = o.t_total - o.t_tangent - o.t_solve - o.t_force;
        if (newton_only ? (o.status & 1) : o.status)
            rc_all = LPMX_ERR_NOCONV;
        if (out)
            out[st] = o;
    }
    if (rc_all != LPMX_OK)
        return lpmx_fail(rc_all, "a solve or Newton loop hit its cap (see lpmx_step_stats.status)");
    return LPMX_OK;
}

extern "C" int lpmx_run_steps(const lpmx_solver *s, int nsteps, const int *nbd, const lpmx_dispbc *dbc,
                              const int *nbf, const lpmx_forcebc *fbc, int maxbc, char out_axis, int out_type,
                              lpmx_step_stats *out)
{
    return lpmx_run_steps_impl(s, nsteps, nbd, dbc, nbf, fbc, maxbc, out_axis, out_type, out, 0, 0);
}
