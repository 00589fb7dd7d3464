// Internal header of the MI355X device layer: the device-resident model, launch
// helpers and deterministic reductions.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lpmx.h"

#define LPMX_SLICE 64           // rows per slice of the block matrix = one wavefront
#define LPMX_BLOCK 256          // default work-group size
#define LPMX_EPS 1e-6           // the reference's EPS (include/lpm.h:42)
#define LPMX_MAXNB 32           // largest neighbour capacity handled (SC/FCC: 18)
#define LPMX_MAXCONN 64         // largest conn row handled (SC/FCC: 61, hex: 31)
#define LPMX_MAXBC 20           // the reference's MAXSMALL BC entries per step
#define LPMX_MAXSLIP 24         // slip systems handled (FCC: 24 unidirectional)

// All device pointers and sizes of the model, passed by value to every kernel.
// Bond arrays are slot planes: element (particle i, slot j) at [j * n + i].
// Per-particle vectors with components are planes too ([c * n + i]), except the
// AoS position and force vectors xyz/Pin ([i * 3 + c]) and the DOF vectors
// residual/disp/Pex ([i * dim + c]) that keep the reference's interleaving.
struct DV
{
    int n, dim, nb, nb1, nb2, cm, lattice, d2;
    double radius, cut1, cut2, volume;
    // parameters
    int plmode, nbreak;
    double H, xi, dmgL, dthr, dA, crit0, J2C; // J2C: the kinematic modulus of plmode 5
    // positions
    double *x, *x0, *xt;
    // per particle
    int *type, *nbcur, *nb0, *pl_flag, *nb_conn, *kcnt;
    double *sigmay, *c44, *kt; // kt: per-particle bond constants [3][n]
    // bond planes
    int *nbr, *nsign, *rev, *opp;
    double *L0, *cs0x, *cs0y, *cs0z, *Kn, *Tv, *db, *dw, *D0, *D1;
    double *dLp0, *dLp1, *dLp2, *dL, *dLave, *ddLp, *ddL, *F, *Ft, *csx, *csy, *csz, *bstress, *crit, *dist;
    // per-particle planes
    double *A, *Bt, *dA_, *dB_;                 // [2][n]: dL_total, TdL_total, ddL_total, TddL_total
    double *beta0, *beta1, *beta2;              // [6][n]
    double *alpha0, *alpha1, *alpha2, *beq0, *beq1, *beq2;
    double *dlam, *stress, *seq, *sm, *triax;   // stress: [6][n]
    double *dnl0, *dnl1, *dloc0, *dloc1, *dvis;
    double *Pin, *Pex, *Pext, *res, *disp;
    int *dispBC, *fix;
    // block matrix: sliced ELL, 64 rows per slice, slot k of slice s at sptr[s]+k
    int nslices;
    int64_t nslots, nnzb;
    int *sptr, *col, *tpos, *dslot, *conn;      // conn: [cm][n] planes
    int64_t *rowptr;                            // row-major block offsets (conn order)
    double *K, *Minv;
    float *K32, *Minv32;
    // static nonlocal-damage neighbourhoods (index order) and weights
    int *dptr, *didx;
    double *dwt;
    // crystal plasticity (plmode 1): slip systems, parameters, state planes
    int nslip, cp_maxloop;
    double cp_tau0, cp_taus, cp_h0, cp_q, cp_eta, cp_p, dtime;
    double *schmid;                                          // [nslip][6]
    double *cp_gy0, *cp_gy1, *cp_gy2, *cp_As0, *cp_As1, *cp_As2; // [nslip][n]
    double *cp_A0, *cp_A1, *cp_A2, *cp_dA;                    // [n]
    double *cp_dgy, *cp_dAs, *cp_RSS;                         // [nslip][n]
    int *cp_Jact;                                            // [nslip][n]
    double *cp_Cab, *cp_D;                                   // [nslip*nslip][n]
    int *cp_fail;                                            // singular slip Jacobians met
    // per-type constants kept for the adjoint: bond constants per unit modulus
    // factor [ntype][3], elastic constants [ntype][3], the factor per particle [n]
    int ntype;
    double *ktype_t, *ce_t, *emod;
};

extern DV g;
extern hipStream_t g_stream;

// ------------------------------------------------------------------ errors -------
int lpmx_fail(int code, const char *fmt, ...);
int lpmx_hip_fail(hipError_t e, const char *what, const char *file, int line);
#define HIPCHK(call)                                                        \
    do                                                                      \
    {                                                                       \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess)                                               \
            return lpmx_hip_fail(e_, #call, __FILE__, __LINE__);            \
    } while (0)
#define LAUNCHCHK() HIPCHK(hipGetLastError())
#define RETCHK(x)                                                           \
    do                                                                      \
    {                                                                       \
        int r_ = (x);                                                       \
        if (r_ != LPMX_OK)                                                  \
            return r_;                                                      \
    } while (0)

int lpmx_require_model(void);
// selects and checks the gfx950 device and creates the library stream (idempotent)
int lpmx_device_init(void);
// updateRR's kernels (bondforce.hip), shared by lpmx_update_rr and its allocation-free form
__global__ void k_update_rr(DV v, double *sq_res, double *sq_rea);
__global__ void k_count_bc(const int *dispBC, int m, int *cnt);
// laws.hip: plmode 3 (load indicator t) and plmode 5 force evaluations of every
// particle, and the damage variants (codes of include/lpmx.h)
int lpmx_law3(int t);
int lpmx_law5(void);
int lpmx_damage_variant(int law, int *count);
void lpmx_laws_free(void);
int lpmx_require_pattern(void);
// changes with every neighbour-list or pattern build (for caches of pattern data)
int lpmx_pattern_version(void);
static inline int nblocks(int64_t n, int bs = LPMX_BLOCK) { return (int)((n + bs - 1) / bs); }

// -------------------------------------------------------- deterministic sums -----
// Wavefront sum in a fixed butterfly order (64 lanes).
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Work-groups are dealt to the 8 XCDs round robin; remap so that each XCD walks a
// contiguous range of rows and its L2 keeps the gathered neighbour columns.
__device__ __forceinline__ int xcd_block(void)
{
    const int G = gridDim.x, b = blockIdx.x;
    const int per = G / 8, rem = G % 8;
    const int x = b % 8, k = b / 8;
    // XCDs 0..rem-1 own per+1 groups, the others per
    const int base = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    return base + k;
}

// Work-group sum (blockDim.x multiple of 64, at most 1024): wave sums, then the
// wave totals in wave order.  All threads get the result.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0)
        sh[w] = v;
    __syncthreads();
    T t = 0;
    for (int k = 0; k < nw; k++)
        t += sh[k];
    return t;
}

// ------------------------------------------------------ host-side reductions -----
// Deterministic device sums of up to 4 per-element values produced by a functor;
// used for norms and small outputs outside the CG inner loop.
#define LPMX_RED_PARTS 1024       // partials of the deterministic device sums
int lpmx_dev_sum_into(const double *src, int64_t count, int sq, double *parts, double *out_host);
int lpmx_dev_sum(const double *src, int64_t count, double *out_host);
int lpmx_dev_sumsq(const double *src, int64_t count, double *out_host);
int lpmx_dev_scan_int(const int *in, int *out_excl, int count, int *total_host);

// scratch: a device buffer reused by host-side helpers (grown on demand)
double *lpmx_scratch(int64_t bytes);

// solver workspace release (called from lpmx_model_destroy)
void lpmx_solver_free(void);
void lpmx_damage_free(void);
void lpmx_accel_free(void);
void lpmx_cp_free(void);
void lpmx_adjoint_free(void);
int lpmx_eval_j2(void); // one J2 evaluation of every particle, no stress, no slot switch

// Seeds and outputs of the reverse sweep of one J2 evaluation (adjoint.hip); the
// outputs are overwritten, not accumulated.
struct VjpIO
{
    const double *Pb;     // [3n], Pin layout
    const double *dLp2b;  // [nb][n] planes
    const double *beta2b; // [6][n]
    const double *alpha2b;
    double *xb;           // [3n], xyz layout
    double *dLp0b, *beta0b, *alpha0b, *eb, *syb, *Hb; // planes / [n]
};
int lpmx_vjp_device(const VjpIO &io);
int lpmx_cp_particles(void);

// Anderson acceleration of the Newton loop (accel.hip)
int lpmx_accel_begin(int depth, int precision);
void lpmx_accel_rebase(void);
int lpmx_accel_step(int *mixed);
int lpmx_accel_revert(void);
void lpmx_accel_forget(void);
// Further Anderson histories over DOF vectors (x and d both [n*dim])
struct AaHist;
AaHist *lpmx_aa_create(void);
void lpmx_aa_destroy(AaHist *a);
int lpmx_aa_begin(AaHist *a, int depth);
int lpmx_aa_step(AaHist *a, double *x, const double *d, int *mixed);

// smoothed-aggregation AMG (amg.hip): hierarchy from the current eliminated tangent,
// V-cycle z = B r on fine DOF vectors (kernels return at once while *done is set)
int lpmx_amg_build(void);
bool lpmx_amg_current(void);
int lpmx_amg_apply_f64(const double *r, double *z, const int *done);
int lpmx_amg_apply_f32(const float *r, float *z, const int *done);
int lpmx_amg_levels(void);
double lpmx_amg_setup_seconds(void);
void lpmx_amg_free(void);
int lpmx_tangent_epoch(void); // increments with every tangent assembly (tangent.hip)
int lpmx_bc_eliminate_repeat(void); // residual-only repeat of lpmx_bc_eliminate_ex(1) (tangent.hip)
// HIP-event timing of the finest level's smoother SpMV launches: the CG tags each
// V-cycle with its iteration in the batch and, after the batch, keeps the launches
// of the iterations that ran
void lpmx_amg_prof_enable(bool on);
void lpmx_amg_prof_tag(int k);
int lpmx_amg_prof_flush(int nreal);

// K x = b on the current constraint-eliminated tangent (solver.hip)
int lpmx_solve_into(int precision, double rtol2, double atol2, int maxit, double inner_rtol2, int ir_maxit,
                    const double *b, double *x, int *iters, double *rr, int *status);
// the load-step loop with or without the damage update after each step; newton_only:
// only a Newton loop that hit its cap fails the call, not a linear solve that did
// (loop.hip)
int lpmx_run_steps_impl(const lpmx_solver *s, int nsteps, const int *nbd, const lpmx_dispbc *dbc, const int *nbf,
                        const lpmx_forcebc *fbc, int maxbc, char out_axis, int out_type, lpmx_step_stats *out,
                        int skip_damage, int newton_only);
