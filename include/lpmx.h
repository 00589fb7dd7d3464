/*
 * lpmx.h -- C-ABI of the MI355X device layer behind the drop-in lpm.h API.
 *
 * Plain pointers and sizes only.  One process drives one GPU (device 0 unless
 * LPMX_DEVICE is set); the library keeps a single device-resident model, like the
 * reference keeps its state in globals, and is not reentrant.  Every entry point
 * returns LPMX_OK (0) or a negative status; lpmx_last_error() describes the last
 * failure.  No entry point falls back to host computation: without a usable
 * gfx950 device the calls fail with LPMX_ERR_NOGPU.
 *
 * Array layouts at this boundary are the reference's, flattened row-major:
 * positions [n][3], bond arrays [n][nneighbors], DOF vectors [n*dim],
 * per-particle state [n].  Inside the device the bond arrays are stored as
 * slot planes ([slot][particle]) and the tangent as 3x3 (2x2) blocks in a
 * sliced-ELL layout of 64 rows per slice; the getters transpose back.
 */
#ifndef LPMX_H
#define LPMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum
{
    LPMX_OK = 0,
    LPMX_ERR_HIP = -1,      /* a HIP runtime call failed */
    LPMX_ERR_ARG = -2,      /* invalid argument or size */
    LPMX_ERR_STATE = -3,    /* call out of order (no model, no pattern, ...) */
    LPMX_ERR_CAPACITY = -4, /* a neighbour or conn list exceeded its capacity */
    LPMX_ERR_NOGPU = -5,    /* no gfx950 device available */
    LPMX_ERR_NOCONV = -6    /* a solve or Newton loop hit its iteration cap */
};

const char *lpmx_last_error(void);

/* ---- model --------------------------------------------------------------------- */
typedef struct lpmx_model_desc
{
    int dim;                  /* 2 or 3 */
    int lattice;              /* 0 square, 1 hexagon, 2 SC, 3 FCC, 4 BCC */
    int n;                    /* particles */
    int nneighbors, nneighbors1, nneighbors2, nneighbors_AFEM;
    double radius, neighbor1_cutoff, neighbor2_cutoff, particle_volume;
} lpmx_model_desc;

typedef struct lpmx_params
{
    int plmode;               /* 0 J2 mixed-linear, 1 crystal plasticity, 3 energy-based J2,
                               * 5 nonlinear-isotropic J2, 6 elastic-brittle */
    double J2_H, J2_xi;
    double damage_L, damage_threshold, damagec_A, critical_bstrain;
    int nbreak;
    double J2_C;              /* kinematic hardening modulus of plmode 5 */
} lpmx_params;

/* Create the device model from positions (xyz = xyz_initial = xyz_in, [n][3]). */
int lpmx_model_create(const lpmx_model_desc *desc, const double *xyz_in);
int lpmx_model_destroy(void);
int lpmx_set_params(const lpmx_params *p);

/* ---- setup on the device ------------------------------------------------------- */
/* searchNormalNeighbor: cell-list build of the bond lists (reference neighbor.c:9-46). */
int lpmx_neighbor_build(void);
/* searchAFEMNeighbor: conn sets, tangent pattern; *len_K = K_pointer[n][1]. */
int lpmx_connectivity_build(int64_t *len_K);
/* Particle types, C11/C12/C44 per type, yield stress, optional per-particle modulus
 * factor (NULL: none), then calcKnTv on the device. */
int lpmx_set_materials(int ntype, const int *type, const double *Ce, const double *sigmay, const double *emod);
/* Optional per-bond critical stretch [n][nneighbors] of the brittle law (NULL: global). */
int lpmx_set_bond_critical(const double *crit);

/* createCuboid's candidate grid (reference src/initialization.c:6-560): rows of points
 * along x, row r = (layer k - 1) * rows + (j - 1), each candidate rotated by R
 * (row-major, p = R t) and kept when inside the closed box (x, y; z too in 3D). */
typedef struct lpmx_lattice_spec
{
    int lattice, dim;          /* 0 square, 1 hexagon (2D); 2 SC, 3 FCC, 4 BCC (3D) */
    int pfr, rows, layers;     /* particles_first_row, rows, layers (1 in 2D) */
    double hx, hy, hz, radius; /* spacings and the particle radius */
    double bt[6];              /* candidate range: -a, a per axis, a the box diagonal */
    double box[6];             /* the closed box that keeps a point */
    double R[9];               /* R_matrix */
} lpmx_lattice_spec;
/* The kept points of createCuboid's lattice, generated on the device (needs no model):
 * *count = their number; with xyz != NULL and cap >= *count also their positions
 * [count][3] in the reference's generation order (z = 0 in 2D). */
int lpmx_lattice_points(const lpmx_lattice_spec *spec, double *xyz, int cap, int *count);
/* computeStrain (reference src/lpm_basic.c:127-249) on the device from the model's dL:
 * strain [n][6] (xx yy zz yz xz xy), in and out -- components the reference leaves
 * alone keep their values. */
int lpmx_compute_strain(double *strain);

/* ---- field transfer (reference layouts) ------------------------------------------ */
enum lpmx_field
{
    LPMX_F_XYZ = 0, LPMX_F_XYZ_INITIAL, LPMX_F_XYZ_TEMP,                 /* double [n][3] */
    LPMX_F_NEIGHBORS, LPMX_F_NSIGN,                                      /* int [n][nb] */
    LPMX_F_NB, LPMX_F_NB_INITIAL, LPMX_F_NB_CONN, LPMX_F_TYPE, LPMX_F_PL_FLAG, /* int [n] */
    LPMX_F_CONN,                                                         /* int [n][nafem+1] */
    LPMX_F_K_POINTER,                                                    /* int [n+1][2] */
    LPMX_F_DISTANCE_INITIAL, LPMX_F_CSX_INITIAL, LPMX_F_CSY_INITIAL, LPMX_F_CSZ_INITIAL,
    LPMX_F_KN, LPMX_F_TV, LPMX_F_DL, LPMX_F_DL_AVE, LPMX_F_DDL, LPMX_F_DDLP, LPMX_F_F, LPMX_F_F_TEMP,
    LPMX_F_CSX, LPMX_F_CSY, LPMX_F_CSZ, LPMX_F_BOND_STRESS, LPMX_F_DISTANCE,
    LPMX_F_DAMAGE_BROKEN, LPMX_F_DAMAGE_W,                               /* double [n][nb] */
    LPMX_F_DLP0, LPMX_F_DLP1, LPMX_F_DLP2, LPMX_F_DAMAGE_D0, LPMX_F_DAMAGE_D1, /* double [n][nb] */
    LPMX_F_DL_TOTAL, LPMX_F_TDL_TOTAL, LPMX_F_DDL_TOTAL, LPMX_F_TDDL_TOTAL,   /* double [n][2] */
    LPMX_F_J2_BETA0, LPMX_F_J2_BETA1, LPMX_F_J2_BETA2, LPMX_F_STRESS,    /* double [n][6] */
    LPMX_F_J2_ALPHA0, LPMX_F_J2_ALPHA1, LPMX_F_J2_ALPHA2,                 /* double [n] */
    LPMX_F_J2_DLAMBDA, LPMX_F_J2_STRESSEQ, LPMX_F_J2_STRESSM, LPMX_F_J2_TRIAXIALITY,
    LPMX_F_DAMAGE_NONLOCAL0, LPMX_F_DAMAGE_NONLOCAL1, LPMX_F_DAMAGE_VISUAL, LPMX_F_SIGMAY,
    LPMX_F_PIN,                                                          /* double [n*3] */
    LPMX_F_PEX, LPMX_F_PEX_TEMP, LPMX_F_RESIDUAL, LPMX_F_DISP,           /* double [n*dim] */
    LPMX_F_DISPBC_INDEX, LPMX_F_FIX_INDEX,                               /* int [n*dim] */
    LPMX_F_CP_GY0, LPMX_F_CP_GY1, LPMX_F_CP_GY2,                         /* double [n][nslip] */
    LPMX_F_CP_A_SINGLE0, LPMX_F_CP_A_SINGLE1, LPMX_F_CP_A_SINGLE2,       /* double [n][nslip] */
    LPMX_F_CP_A0, LPMX_F_CP_A1, LPMX_F_CP_A2, LPMX_F_CP_DA,              /* double [n] */
    LPMX_F_CP_DGY, LPMX_F_CP_DA_SINGLE, LPMX_F_CP_RSS,                   /* double [n][nslip] */
    LPMX_F_CP_JACT,                                                      /* int [n][nslip] */
    LPMX_F_CP_CAB,                                                       /* double [n][nslip*nslip] */
    LPMX_F_J2_BETA_EQ0, LPMX_F_J2_BETA_EQ1, LPMX_F_J2_BETA_EQ2,          /* double [n] */
    LPMX_F_DAMAGE_LOCAL0, LPMX_F_DAMAGE_LOCAL1,                          /* double [n] */
    LPMX_F_COUNT
};

int lpmx_get(int field, void *host, int64_t count);
int lpmx_put(int field, const void *host, int64_t count);

/* ---- tangent stiffness ------------------------------------------------------------ */
/* Analytic elastic tangent (the Jacobian the reference differences), symmetrised as
 * the reference's 0.5(A + A^T), into the device block matrix. */
int lpmx_tangent(void);
/* The device matrix in the reference's 1-based upper CSR (IK [n*dim+1], JK, K [len_K]). */
int lpmx_export_upper_csr(int *IK, int *JK, double *K);
/* Replace the device matrix values by a host upper CSR of the same pattern. */
int lpmx_import_upper_csr(const double *K);
/* setDispBC_stiffnessUpdate: zero constrained rows/columns, diagonal = |diag|_2 of
 * the current matrix (the reference's rule, src/boundary.c:159-281). */
int lpmx_bc_eliminate(void);
/* fixed_diag = 1 keeps the diagonal value of the first elimination after each
 * tangent assembly instead of re-deriving it from the already modified matrix
 * (the reference's value grows by ~sqrt(#constrained DOFs) per call and overflows
 * after ~150 Newton iterations of one step; it never changes the solution). */
int lpmx_bc_eliminate_ex(int fixed_diag);

/* ---- linear solve ------------------------------------------------------------------- */
enum { LPMX_FP64 = 0, LPMX_MIXED = 1 };
/* K disp = residual by block-Jacobi PCG, stopping when |r|^2 <= rtol2 |r0|^2 + atol2
 * (the reference's dcg test) or after maxit iterations; then xyz += disp.
 * LPMX_MIXED runs FP32 values, vectors and compensated dots inside FP64 iterative
 * refinement.  *iters gets the inner CG iterations, *rr the final |r|^2. */
int lpmx_solve(int precision, double rtol2, double atol2, int maxit, int *iters, double *rr);
/* The same with the mixed form's refinement settings: one FP32 PCG whose residual is
 * replaced by the FP64 residual b - K x (x accumulated in FP64) each time the FP32
 * recurrence has dropped by the squared factor inner_rtol2 (or below a quarter of the
 * outer threshold), the search direction kept; at most ir_maxit replacements.
 * *status: 1 converged, 2 iteration cap, 3 breakdown (p.Kp <= 0).  Returns LPMX_OK
 * whatever the status. */
int lpmx_solve_ex(int precision, double rtol2, double atol2, int maxit, double inner_rtol2, int ir_maxit,
                  int *iters, double *rr, int *status);

/* ---- crystal plasticity (plmode 1) ---------------------------------------------------- */
/* Parameters of the reference's rate-dependent slip law (cp_tau0[0], cp_taus[0], cp_h0,
 * cp_q, cp_eta, cp_p, dtime, cp_maxloop in include/lpm.h). */
typedef struct lpmx_cp_params
{
    double tau0, taus, h0, q, eta, cp_p, dtime;
    int maxloop;
} lpmx_cp_params;
/* slipSysDefine3D's output for the device: nslip Schmid tensors [nslip][6]
 * (s11 s22 s33 s23 s13 s12, the reference's schmid_tensor), parameters; the slip
 * state is set to the initial yield stress tau0 and zero slip. */
int lpmx_set_slip_systems(int nslip, const double *schmid, const lpmx_cp_params *p);
/* computeCab (reference constitutive.c:1864-1917) at the current geometry. */
int lpmx_compute_cab(void);
/* Number of return maps that met a singular slip Jacobian (the reference exits). */
int lpmx_cp_failures(int *count);

/* ---- adjoint of the J2 load path (no counterpart in the reference) --------------------- */
/* One J2 force evaluation at the current positions and slot-0 state (no stress, no
 * state switch), then its vector-Jacobian product: seeds on Pin [3n] and on the slot-2
 * state dLp [n][nneighbors], J2_beta [n][6], J2_alpha [n]; out the adjoints of the
 * positions [3n], slot-0 dLp, J2_beta, J2_alpha, the per-particle modulus factor [n],
 * the yield stress [n] and the hardening modulus H. */
int lpmx_vjp(const double *Pbar, const double *dLp2bar, const double *beta2bar, const double *alpha2bar,
             double *xbar, double *dLp0bar, double *beta0bar, double *alpha0bar, double *emodbar,
             double *sigmaybar, double *Hbar);

/* ---- constitutive, residual, boundary conditions, damage ------------------------------ */
int lpmx_computedL(void);
/* computeBondForceGeneral(mode, t) (reference src/constitutive.c:88-146): 0 J2 mixed-linear,
 * 1 crystal plasticity, 3 energy-based J2 with load indicator t (+1 tension, -1
 * compression), 4 predictor, 5 nonlinear-isotropic J2, 6 elastic; then stress and
 * switchStateV(2).  lpmx_bond_force(mode) is t = 1.  The _range form runs the
 * per-particle law for [first, first+count) only (no stress, no state switch);
 * modes 3 and 5 take the whole range only.  Mode 5 advances each particle's slot-0
 * state once per evaluation, where the reference advances it once per list it is
 * in, in its particle order. */
int lpmx_bond_force(int mode);
int lpmx_bond_force_ex(int mode, int t);
int lpmx_bond_force_range(int mode, int first, int count);
int lpmx_bond_force_range_ex(int mode, int t, int first, int count);
int lpmx_compute_stress(void);
int lpmx_switch_state(int flag);
/* updateRR: residual and reaction on the device; norms returned (NULL allowed). */
int lpmx_update_rr(double *norm_residual, double *norm_reaction, int *nreaction);
/* lpmx_update_rr without any allocation on any call (no counterpart in the reference:
 * for callers that cannot allocate device memory, such as code under HIP graph
 * capture or a tight serving loop).  Same kernels and results, bit for bit; its
 * intermediates live in one static device array of the library (16 MiB per GPU,
 * held until the process ends), so models of at most LPMX_NOALLOC_MAX_PARTICLES
 * particles (larger ones fail with LPMX_ERR_ARG), and calls that overlap on one GPU
 * race. */
#define LPMX_NOALLOC_MAX_PARTICLES 1048576
int lpmx_update_rr_noalloc(double *norm_residual, double *norm_reaction, int *nreaction);
int lpmx_get_reaction_force(double *host, int count);

typedef struct lpmx_dispbc { int type; char flag; double step; } lpmx_dispbc;
typedef struct lpmx_forcebc { int type; double step[3]; } lpmx_forcebc;
int lpmx_set_disp_bc(int nbd, const lpmx_dispbc *bc);
int lpmx_set_force_bc(int nbf, const lpmx_forcebc *bc);

/* updateDamageGeneral for plmode 0 (nonlocal particle-wise ductile), 5 (local bond-wise
 * ductile) and 6 (brittle); plmode 3 breaks nothing;
 * *broken gets the reference's return value.  Broken pairs of the brittle law can be
 * read back with lpmx_get_broken (up to cap pairs). */
int lpmx_damage(int plmode, int *broken);
/* One damage update of the reference by name: updateDuctileDamagePwiseNonlocal (0),
 * updateDuctileDamagePwiseLocal (1), updateDuctileDamageBwiseNonlocal (2),
 * updateDuctileDamageBwiseLocal (3), updateBrittleDamage (6) (reference
 * src/constitutive.c:1437-1862).  *count gets the reference's return value: directed
 * entries broken (0, 2), particles failed (1), pairs broken (3), candidates (6). */
enum
{
    LPMX_DMG_PWISE_NONLOCAL = 0,
    LPMX_DMG_PWISE_LOCAL = 1,
    LPMX_DMG_BWISE_NONLOCAL = 2,
    LPMX_DMG_BWISE_LOCAL = 3,
    LPMX_DMG_BRITTLE = 6
};
int lpmx_damage_law(int law, int *count);
int lpmx_get_broken(int *bi, int *bj, int cap, int *count);
int lpmx_update_crack(void);
/* writeReaction / writeDisp values for particles of type t1 along axis 'x','y','z'. */
int lpmx_reaction(char axis, int t1, double *force, double *mean_disp);

/* ---- device-resident load-step loop (reference src/lpmc_project.c:382-546) -------- */
typedef struct lpmx_solver
{
    int precision;          /* LPMX_FP64 or LPMX_MIXED */
    double rtol2, atol2;    /* CG test on squared norms (reference: 1e-8, 1e-12) */
    int maxit;              /* CG cap (reference: n*dim) */
    double newton_tol;      /* reference TOLITER */
    int newton_maxit;       /* reference MAXITER */
    double inner_rtol2;     /* mixed form: squared reduction of an FP32 segment between
                             * FP64 residual replacements (1e-2) */
    int ir_maxit;           /* mixed form: cap on residual replacements per solve */
    int accel_depth;        /* 0: the reference's modified Newton; 1..32: Anderson
                             * acceleration of it with that many differences (same
                             * fixed points, far fewer iterations in plastic flow) */
    int precond;            /* CG preconditioner: 0 block-Jacobi, 1 smoothed-aggregation
                             * AMG (V-cycle, block-Jacobi Chebyshev smoothing; hierarchy
                             * built on the first solve, kept while the pattern and the
                             * constraint set stay); 2 (fused loop only) one AMG V-cycle
                             * per Newton iteration in place of the CG solve, the
                             * fixed-point map u + B R(u) accelerated by Anderson mixing
                             * (solverCG treats 2 as 1) */
} lpmx_solver;

typedef struct lpmx_step_stats
{
    double reaction, disp;  /* writeReaction / writeDisp values after the step */
    int newton_iters, cg_iters, resolves, broken, ir_passes, status;
    double res_norm, tol_mult;
    double t_tangent, t_solve, t_force, t_other, t_total; /* seconds */
    int accel_mixed;        /* Newton iterations that took the accelerated iterate */
    int accel_reverts;      /* accelerated iterates rejected for the plain one */
    int accel_restarts;     /* histories restarted because the residual stagnated */
} lpmx_step_stats;

/* Preconditioner of the following solves (lpmx_solve*, solverCG): 0 block-Jacobi
 * (default), 1 smoothed-aggregation AMG, 2 one AMG V-cycle instead of a solve.
 * lpmx_run_steps sets it from its solver. */
int lpmx_set_preconditioner(int kind);

int lpmx_run_steps(const lpmx_solver *s, int nsteps, const int *nbd, const lpmx_dispbc *dbc,
                   const int *nbf, const lpmx_forcebc *fbc, int maxbc, char out_axis, int out_type,
                   lpmx_step_stats *out);

/* ---- adjoint of a J2 load history (no counterpart in the reference) ------------------- */
/* The forward half: lpmx_run_steps with the damage update switched off (J2, plmode 0,
 * 3D), recording what the reverse sweep needs.  Every ckpt_every steps the state before
 * the step is kept (positions, bond forces, Pin, Pex, the converged slot state,
 * constraint masks); the steps of the last segment are kept whole.  ckpt_every = 0
 * keeps every step (no recomputation).  reactions[nsteps] gets writeReaction of
 * (out_axis, out_type) after each step, which is what the reverse sweep differentiates.
 * Fails (LPMX_ERR_NOCONV) when a Newton loop hits its cap; a linear solve that hits
 * its own cap inside a Newton loop that then converges is only reported in out[].status. */
int lpmx_history_forward(const lpmx_solver *s, int ckpt_every, int nsteps, const int *nbd,
                         const lpmx_dispbc *dbc, const int *nbf, const lpmx_forcebc *fbc, int maxbc,
                         char out_axis, int out_type, double *reactions, lpmx_step_stats *out);

typedef struct lpmx_adjoint_opts
{
    double tol;             /* relative residual |b - J^T lambda| / |b| of each step's
                             * adjoint solve (J: consistent Jacobian of the free DOFs) */
    int maxit;              /* cap on the adjoint iterations of one step */
    int depth;              /* Anderson depth of the K-preconditioned fixed-point
                             * iteration lambda <- lambda + K^-1 (b - J^T lambda), 0..32 */
    int precision;          /* LPMX_FP64 or LPMX_MIXED for the K^-1 solves (which use the
                             * forward solver's preconditioner and mixed-form segments) */
    double rtol2;           /* CG test of the K^-1 solves (squared, relative) */
    int cg_maxit;           /* their cap */
} lpmx_adjoint_opts;

typedef struct lpmx_adjoint_stats
{
    double t_forward;       /* lpmx_history_forward, seconds */
    double t_recompute;     /* forward steps re-run from checkpoints */
    double t_backward;      /* reverse sweep without the recomputation */
    int checkpoints, recomputed_steps;
    int adj_iters, adj_cg_iters; /* totals over the steps */
    int adj_iters_max;      /* largest per step */
    double max_rel_res;     /* largest final |b - J^T lambda| / |b| */
    double max_rel_free;    /* largest |xbar on free DOFs| / |b| after the last VJP of a step */
    double ckpt_bytes;      /* device memory held by the history */
} lpmx_adjoint_stats;

/* The reverse half: given Rbar[nsteps] = dJ/dR_n for the reactions of the recorded
 * history, the gradient of J through every step's equilibrium and return map:
 * g_emod [n] (per-particle modulus factor), g_sigmay [n] (per-particle yield stress),
 * *g_H (hardening modulus).  dJ/dE of a global modulus is sum_i emod_i g_emod_i / E.
 * Leaves the model in the state after the last forward step. */
int lpmx_history_backward(const lpmx_adjoint_opts *o, const double *Rbar, double *g_emod, double *g_sigmay,
                          double *g_H, lpmx_adjoint_stats *st);
/* Release the recorded history. */
void lpmx_history_free(void);

/* ---- measurement ---------------------------------------------------------------------- */
/* Time `reps` launches of the CG's block SpMV kernel (w = K z over the gathered columns,
 * p = z + beta p, q = w + beta q, partial p.q; FP64 or FP32) with HIP events on the
 * library's stream;
 * *avg_ms gets the mean duration, *bytes the algorithmic bytes per launch as returned
 * by lpmx_spmv_bytes. */
int lpmx_bench_spmv(int precision, int reps, double *avg_ms, double *bytes);
/* Algorithmic HBM bytes of one CG SpMV launch on the current pattern:
 * nnzb*(dim^2*s_val + 4) + 4*(nslices+1) + 4*n*dim*s_vec (z and p_old read, p and q
 * written once; s = 8 for FP64, 4 for the mixed form). */
int lpmx_spmv_bytes(int precision, double *bytes);
/* Counters of the device matrix: block rows, stored blocks (nnzb), slots incl. padding. */
int lpmx_matrix_info(int64_t *nnzb, int64_t *padded_slots, int *nslices);
/* y = K x with the device matrix (FP64 or FP32 values), vectors of n*dim doubles. */
int lpmx_spmv(int precision, const double *x, double *y);
/* HIP-event timing of the solver's SpMV launches (up to 64 per solve) while on;
 * switching on resets the totals. */
void lpmx_profile_spmv(int on);
void lpmx_profile_spmv_read(double *total_ms, int64_t *launches);
/* With the AMG preconditioner: HIP-event totals of the finest level's smoother SpMV
 * launches (y = b - K x, y -= K x) in the CG iterations that ran, while
 * lpmx_profile_spmv is on. */
void lpmx_profile_amg_read(double *total_ms, int64_t *launches);
/* Algorithmic HBM bytes of one such launch: nnzb*(dim^2*s_val + 4) + 4*(nslices+1)
 * + 3*n*dim*s_vec (x read once, b or y read, y written; s = 8 FP64, 4 mixed). */
int lpmx_amg_smoother_bytes(int precision, double *bytes);
/* The current AMG hierarchy: levels (0 when none), wall time of its last build,
 * device memory it holds. */
int lpmx_amg_info(int *levels, double *setup_seconds, int64_t *device_bytes);
/* No counterpart in the reference; for tests of the AMG set-up.  The stored blocks of
 * level `level`'s operator (which 0, levels >= 1), prolongator (1) or restriction (2)
 * as block COO: rows[k], cols[k], blocks vals[k*br*bc ...] row-major (a sliced-ELL
 * level includes its zero padding blocks).  dims receives nr, nc, br, bc and the
 * block count; with rows null only dims is filled. */
int lpmx_amg_export(int level, int which, int64_t cap, int *rows, int *cols, double *vals, int *dims);
/* HIP-event timing of the J2 return-map kernel (k_j2_particle, one launch per force
 * evaluation) while on; switching on resets the totals. */
void lpmx_profile_force(int on);
int lpmx_profile_force_read(double *total_ms, int64_t *launches);
/* Algorithmic HBM bytes of one such launch: n*(4+4+24+24+48) read and n*(32+48+16+4)
 * written per particle, nbond*(4+4+48) read and nbond*40 written per bond, and the
 * slot-2 plastic stretch of all n*nneighbors list slots (8 B each). */
int lpmx_j2_particle_bytes(double *bytes);

#ifdef __cplusplus
}
#endif
#endif
