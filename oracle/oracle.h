/*
 * oracle.h -- CPU restatement of the LPM-C implicit load-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Nothing in the product (liblpm.so) links or calls it.
 *
 * Every function restates one reference function on flat arrays (row-major
 * [particle][slot] instead of the reference's double** rows) and cites the
 * reference file:line it follows.  Arithmetic is written in the reference's
 * evaluation order (left-to-right C semantics, x*x for pow(x,2) as gcc -O2 folds
 * it), and the library is built with -ffp-contract=off so that no FMA is formed.
 * Execution is serial, which is the race-free specification of the reference's
 * OpenMP loops.
 *
 * Pins (what each function is checked against; see tests/test_oracle.py):
 *   orc_create_lattice      closed-form particle counts and spacings
 *   orc_neighbor_search     closed-form lattice shells; brute force vs cell list
 *   orc_afem_connectivity   closed form |conn| = 61 (SC/FCC interior), 31 (hex);
 *                           K_pointer rule 9k-3 / 4k-1; brute-force set definition
 *   orc_kntv                closed forms of Kn/Tv per lattice
 *   orc_tangent_fd          symmetry, translation invariance, analytic Jacobian
 *   orc_tangent_analytic    central finite differences of orc_internal_force
 *   orc_bond_force_*        zero force at the reference configuration; affine patch
 *                           test (interior stress = C:eps); radial-return closed form;
 *                           the predictor continues the elastic law between two states
 *   orc_bond_force_j2energy serial order = Jacobi form bit for bit; the elastic branch
 *                           is the elastic law; a plastic return whose root lies in
 *                           the bisection bracket lands on the hardened yield surface
 *                           when its energy is recomputed from the updated stretches
 *   orc_bond_force_j2nonlinear the elastic branch is the elastic law; the Jacobi
 *                           multiplier is the root of the hardening law to TOLITER
 *   orc_damage_bwise_local  reverse particle order gives the same state; pairwise
 *                           brute force of the rule, breaks, counts and weights
 *   orc_damage_pwise_local, orc_damage_bwise_nonlocal  the same (order, brute force)
 *   orc_solve_cg            dense direct solve (numpy) on small systems
 *   orc_damage_*            uniform-field closed form; brute force shell sort
 *   orc_slip_systems_fcc    the 24 Schmid tensors: traceless, unit norm, [001] Schmid
 *                           factors 1/sqrt(6) on the eight tied systems
 *   orc_lu_solve            dense direct solve (numpy)
 *   orc_compute_cab         C_ab = -dRSS_a/dgamma_b of the return map's stress path
 *   orc_bond_force_cp       single-slip return map against its scalar closed form
 *   orc_run_steps           Anderson-accelerated and plain loops reach one equilibrium
 *   orc_vjp_j2              central differences of orc_eval_j2 along random directions
 *   orc_adjoint_run         central differences of the whole forward history
 *   orc_compute_strain      exact recovery of a strain from stretches linear in it;
 *                           weighted least squares by numpy lstsq; the stored
 *                           right-hand side of a collinear chain's singular system
 * MKL's internal iterate sequence (dcg) is closed source: parity of solverCG is
 * checked on converged solutions only ("parity unpinned" for iterates).
 */
#ifndef LPM_ORACLE_H
#define LPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model state.  All arrays are owned by the caller (numpy in the tests).
 * Bond arrays are [n][nbmax]; slot arrays are [n][nbmax][nslot]. */
typedef struct orc_model
{
    /* sizes and lattice constants */
    int dim, lattice, n, nbmax, nb1max, nb2max, connmax, ntype;
    double radius, cut1, cut2, volume;

    /* material and damage parameters */
    int plmode;
    double J2_H, J2_xi;
    double damage_L, damage_threshold, damagec_A, critical_bstrain;
    int nbreak;

    /* per-type and per-particle inputs */
    int *type;           /* n */
    double *Ce;          /* ntype*3: C11, C12, C44 */
    double *sigmay;      /* n */
    double *emod;        /* n or NULL: per-particle modulus factor (extension) */
    double *crit;        /* n*nbmax or NULL: per-bond critical stretch (extension) */

    /* geometry */
    double *x, *x0, *xt; /* n*3: xyz, xyz_initial, xyz_temp */
    int *nbr, *nsign;    /* n*nbmax, -1 padded */
    int *nbr1, *nbr2;    /* n*nb1max, n*nb2max */
    int *nb, *nb_initial;
    double *L0;          /* n*nbmax: distance_initial */
    double *cs0;         /* n*nbmax*3: csx/csy/csz_initial */
    int *conn, *nb_conn; /* n*connmax, n */
    int *kptr;           /* (n+1)*2: K_pointer */

    /* bond fields */
    double *Kn, *Tv, *dL, *dL_ave, *ddL, *ddLp, *F, *Ftemp, *bond_stress;
    double *cs;          /* n*nbmax*3: csx/csy/csz */
    double *dLp;         /* n*nbmax*3 slots */
    double *dmgD;        /* n*nbmax*2 slots */
    double *dmg_broken, *dmg_w;

    /* particle fields */
    double *dLtot, *TdLtot, *ddLtot, *TddLtot; /* n*2 */
    double *beta;        /* n*6*3 slots */
    double *alpha, *beta_eq; /* n*3 slots */
    double *dlambda, *stress, *seq, *sm, *triax; /* n, n*6, n, n, n */
    int *pl_flag;
    double *dnl, *dloc;  /* n*2 slots: damage_nonlocal, damage_local */
    double *damage_visual;
    double *Pin;         /* n*3 */
    double *Pex, *Pex_temp, *residual, *disp; /* n*dim */
    int *dispBC, *fixidx; /* n*dim */
    double *reaction;    /* n*dim capacity */
    int nreaction;

    /* tangent in the reference's 1-based upper CSR */
    int64_t nK;
    double *K;
    int *IK, *JK;

    /* crystal plasticity (plmode 1); NULL pointers when unused */
    double *dist;        /* n*nbmax: distance, set by orc_computedL */
    int nslip;           /* slip systems (24 for FCC: 12 {111}<110> x 2 senses) */
    double *schmid;      /* nslip*6: s11 s22 s33 s23 s13 s12 (engineering shear) */
    double cp_tau0, cp_taus, cp_h0, cp_q, cp_eta, cp_p, dtime;
    int cp_maxloop;
    double *cp_gy, *cp_A_single; /* n*nslip*3 slots */
    double *cp_A;        /* n*3 slots */
    double *cp_dA;       /* n */
    double *cp_dgy, *cp_dA_single, *cp_RSS; /* n*nslip */
    int *cp_Jact;        /* n*nslip */
    double *cp_Cab;      /* n*nslip*nslip */
    int *state_v;        /* n: return map done in the current sweep */
    int cp_fail;         /* singular slip Jacobians met */

    /* plmode 5: kinematic modulus J2_C; 1 for the one-update-per-particle form in
     * orc_bond_force_general (0: the reference's serial order) */
    double J2_C;
    int j2nl_jacobi;
} orc_model;

typedef struct orc_dispbc
{
    int type;
    char flag;
    double step;
} orc_dispbc;

typedef struct orc_forcebc
{
    int type;
    double step[3];
} orc_forcebc;

/* Lattice description filled by orc_lattice_params (reference src/initialization.c). */
typedef struct orc_lattice
{
    int dim, lattice, nbmax, nb1max, nb2max, nafem;
    double radius, cut1, cut2, volume;
} orc_lattice;

/* Solver / loop settings; reference values are rtol2=1e-8, atol2=1e-12, maxit=dim*n,
 * newton_tol=TOLITER, newton_maxit=MAXITER. */
typedef struct orc_solver
{
    double rtol2, atol2;
    int maxit;
    double newton_tol;
    int newton_maxit;
    int tangent;         /* 0: forward FD (reference), 1: analytic elastic Jacobian */
    int fixed_diag;      /* 1: constrained-DOF diagonal fixed per tangent assembly */
    int nthreads;        /* >1 only for timing (CPU baseline); 1 is the specification */
    int accel;           /* >0: Anderson acceleration with that many differences, to reach
                          * the same equilibrium where modified Newton crawls (off in the
                          * specification runs) */
    int skip_damage;     /* 1: no damage update after a step (the differentiated history) */
} orc_solver;

/* Per-step outputs of orc_run_steps. */
typedef struct orc_stepout
{
    double reaction;     /* writeReaction value after the step */
    double disp;         /* writeDisp value after the step */
    int newton_iters;    /* Newton iterations over all re-solves of the step */
    int cg_iters;        /* CG iterations summed over the step */
    int resolves;        /* re-entries after bond breakage */
    int broken;          /* bonds broken during the step */
    double res_norm;     /* final residual norm */
    double tol_mult;     /* tolerance multiplier of the last solve */
} orc_stepout;

/* ---- geometry ------------------------------------------------------------- */
void orc_lattice_params(int lattice, double radius, double box_z, orc_lattice *out);
int orc_create_lattice(int lattice, double radius, const double *box, int eulerflag,
                       double a1, double a2, double a3, double *x, int cap);

/* ---- neighbours ----------------------------------------------------------- */
int orc_neighbor_search(orc_model *m);
int orc_neighbor_search_cells(orc_model *m);
int64_t orc_afem_connectivity(orc_model *m);

/* ---- bond constants, kinematics, constitutive ----------------------------- */
void orc_kntv(orc_model *m);
void orc_computedL(orc_model *m);
void orc_bond_force_elastic(orc_model *m, int ii);
void orc_bond_force_j2(orc_model *m, int ii);
void orc_bond_force_incremental(orc_model *m, int ii);
/* crystal plasticity: slip systems, slip-stiffness matrices, return map */
int orc_slip_systems_fcc(const double *R, double *schmid);
void orc_compute_cab(orc_model *m);
void orc_bond_force_cp(orc_model *m, int ii);
int orc_lu_solve(int n, double *a, double *b);
void orc_cp_rss(orc_model *m, int i, const double *gamma, double *rss);

/* ---- adjoint of the J2 load path ------------------------------------------------- */
/* One J2 evaluation (orc_bond_force_j2 for every particle, no stress, no switch). */
void orc_eval_j2(orc_model *m);
/* Reverse mode of orc_eval_j2 at the current positions and slot-0 state: seeds on
 * Pin [3n], dLp slot 2 [n*nbmax], beta slot 2 [n*6], alpha slot 2 [n]; adjoints of
 * the positions [3n], dLp slot 0, beta slot 0, alpha slot 0, the per-particle
 * modulus factor [n], the yield stress [n] and H. */
void orc_vjp_j2(orc_model *m, const double *Pb, const double *dLp2b, const double *beta2b, const double *alpha2b,
                double *xb, double *dLp0b, double *beta0b, double *alpha0b, double *eb, double *syb, double *Hb);
/* Discrete adjoint of a J2 load history (damage off): J = sum_n (R_n - target_n)^2 with
 * R_n the reaction along `axis` of type t1 after step n; gradients with respect to the
 * per-particle modulus factor [n], the yield stress [n] and H.  Each step's adjoint
 * system is formed densely and solved directly (small models only).  Returns 0, or
 * the first step whose Newton loop failed, +1. */
int orc_adjoint_run(orc_model *m, const orc_solver *s, int nsteps, const int *nbd, const orc_dispbc *dbc,
                    const int *nbf, const orc_forcebc *fbc, int maxbc, char axis, int t1, const double *target,
                    double *J, double *reactions, double *g_emod, double *g_sigmay, double *g_H);
/* computeBondForceJ2energyReturnMap (plmode 3, reference src/constitutive.c:286-463)
 * for every particle with load indicator t (+1 tension, -1 compression), then the
 * stress and switchStateV(2).  jacobi = 0: the reference's serial order, each
 * particle recomputing its list from slot 0; jacobi = 1: every particle's update
 * once, all second stretch passes, then every force (the form a parallel version
 * takes; without damage it gives the same values). */
void orc_bond_force_j2energy(orc_model *m, int t, int jacobi);
/* computeBondForceJ2nonlinearIso (plmode 5, reference src/constitutive.c:689-863)
 * with kinematic modulus C (the reference's J2_C), then the stress and
 * switchStateV(2).  jacobi = 0: the reference's serial order, in which every member
 * of a particle's list advances its slot-0 state in place, so a particle is advanced
 * once per list it is in; jacobi = 1: every particle advanced once from the state at
 * the start.  Neither writes slot 2, so the switch that follows restores slot 0 from
 * it, as in the reference. */
void orc_bond_force_j2nonlinear(orc_model *m, double C, int jacobi);
void orc_compute_stress(orc_model *m);
void orc_compute_strain(const orc_model *m, double *strain);
void orc_switch_state(orc_model *m, int flag);
void orc_bond_force_general(orc_model *m, int mode);
void orc_internal_force(orc_model *m, const double *x, double *pin);

/* ---- tangent, residual, boundary conditions -------------------------------- */
void orc_tangent_fd(orc_model *m);
void orc_tangent_analytic(orc_model *m);
void orc_update_rr(orc_model *m);
void orc_set_disp_bc(orc_model *m, int nbd, const orc_dispbc *bc);
void orc_set_force_bc(orc_model *m, int nbf, const orc_forcebc *bc);
void orc_bc_stiffness_update(orc_model *m);
void orc_bc_stiffness_update_nd(orc_model *m, double nd);
double orc_nrm2(int n, const double *x);

/* ---- linear solve ----------------------------------------------------------- */
void orc_spmv_upper(const orc_model *m, const double *x, double *y);
int orc_solve_cg(orc_model *m, const orc_solver *s, double *final_rr);
/* Timing only (CPU baseline): seconds of `iters` threaded CG iterations. */
double orc_cg_baseline(const orc_model *m, int nthreads, int iters, double *rr_out);

/* ---- damage and cracks ------------------------------------------------------ */
int orc_damage_pwise_nonlocal(orc_model *m);
/* updateDuctileDamageBwiseNonlocal and updateDuctileDamagePwiseLocal, the other
 * ductile damage variants of the reference (selectable in updateDamageGeneral).
 * Return the directed bond entries broken and the particles failed. */
int orc_damage_bwise_nonlocal(orc_model *m);
int orc_damage_pwise_local(orc_model *m, int order);
/* updateDuctileDamageBwiseLocal (the damage of plmode 5): order 0 is the reference's
 * particle order, 1 the reverse (same result).  Returns the bonds broken. */
int orc_damage_bwise_local(orc_model *m, int order);
int orc_damage_brittle(orc_model *m, int *broken_i, int *broken_j, int cap);
void orc_update_crack(orc_model *m);
int orc_shell_sort_select(int k, const double *bstrain, int nbreak, int *order);

/* ---- outputs ------------------------------------------------------------------- */
double orc_reaction(const orc_model *m, char axis, int t1);
double orc_mean_disp(const orc_model *m, char axis, int t1);

/* ---- load-step loop ------------------------------------------------------------- */
/* Reference-format files that orc_run_steps appends to (NULL: none): the broken-bond
 * file of the damage update, and writeReaction / writeDisp of (axis, t1) after every
 * Newton loop of a step; step numbers are the step index + 1 + step_offset. */
void orc_set_output(const char *bond_file, const char *reaction_file, const char *disp_file, char axis, int t1,
                    int step_offset);
int orc_run_steps(orc_model *m, const orc_solver *s, int nsteps,
                  const int *nbd, const orc_dispbc *dbc, const int *nbf, const orc_forcebc *fbc,
                  int maxbc, char out_axis, int out_type, orc_stepout *out);

#ifdef __cplusplus
}
#endif
#endif
