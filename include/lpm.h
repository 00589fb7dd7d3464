/*
 * lpm.h -- drop-in C API of the nonlocal lattice particle method (LPM) load-step path,
 * served by hand-written HIP kernels on one MI355X (gfx950).
 *
 * Same names, C types, macros and structs as the reference API header
 * (include/lpm.h:34-99 of ymlasu/LPM-C), so that an unmodified driver such as
 * src/lpmc_project.c compiles against this header and links against liblpm.so.
 * Differences from the reference header:
 *   - no Intel MKL headers: the two MKL routines that drivers call directly
 *     (cblas_dnrm2, mkl_free_buffers) are declared here and implemented by liblpm.so;
 *   - the extension entry points (device-resident load-step loop, precision mode,
 *     per-particle modulus, per-bond critical stretch) live in lpmx.h.
 *
 * Ownership and threading follow the reference: the driver defines the globals, the
 * library allocates most arrays, every call is made from one host thread, and the
 * library is not reentrant.
 */
#pragma once
#ifndef LPM3D_H
#define LPM3D_H

#include <stdio.h>
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <memory.h>
#include <time.h>
#include <sys/types.h>
#include <sys/stat.h>
#include <unistd.h>
#include <omp.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Constants (reference include/lpm.h:34-42). */
#define PI 3.14159265358979323846
#define NDIM 3          /* storage dimension of coordinates and internal forces, also in 2D */
#define TOL 1e-3        /* equality tolerance used when matching bond directions */
#define MAXLINE 500     /* maximum number of load steps in a boundary-condition table */
#define MAXSMALL 20     /* maximum BC entries per step, small inner iteration caps */
#define MAXITER 100     /* maximum Newton iterations per load step */
#define MAXSLIPSYS 50   /* maximum number of slip systems */
#define TOLITER 1e-4    /* Newton tolerance relative to max(|residual0|, |reaction0|) */
#define EPS 1e-6        /* finite-difference step factor and broken-bond threshold */

/* Helper macros (reference include/lpm.h:45-51). */
#define LEN(arr) (sizeof(arr) / sizeof(arr[0]))
#define MAX(x, y) ((x) < (y) ? (y) : (x))
#define MIN(x, y) ((x) > (y) ? (y) : (x))
#define SIGN(x) ((x) < (0) ? (-1.0) : (1))
#define SY(x) (620.0 + 3300.0 * (1.0 - exp(-0.4 * x)))
#define DAM_PHI(x) (1.0 / damage_L / sqrt(2 * PI) * exp(-0.5 * x * x / damage_L / damage_L))

/* Global model state; defined by the driver (or by liblpm.so when no driver defines it). */
extern int ntype, particles_first_row, nparticle, rows, layers, nneighbors, nneighbors1, nneighbors2, dim;
extern int lattice, nneighbors_AFEM, nneighbors_AFEM1, nneighbors_AFEM2, plmode, eulerflag, nslip_face;
extern int max_nslip_vector, nslipSys, nbreak;

extern int *IK, *JK, *type, *dispBC_index, *fix_index, *nslip_vector, *lacknblist, pbc[NDIM], *pl_flag;
extern int *nb, *nb_initial, *nb_conn, *state_v;
extern int **neighbors, **neighbors1, **neighbors2, **neighbors_AFEM;
extern int **K_pointer, **conn, **nsign, **cp_Jact;

extern double radius, hx, hy, hz, neighbor1_cutoff, neighbor2_cutoff, angle1, angle2, angle3, particle_volume;
extern double R_matrix[NDIM * NDIM], box[2 * NDIM], cp_tau0[3], cp_taus[3], cp_eta, cp_p, cp_h0, cp_q, cp_maxloop;
extern double box_x, box_y, box_z, cp_q, dtime, cp_theta, J2_H, J2_xi, J2_C, damage_L;
extern double damage_threshold, damageb_A, damagec_A, critical_bstrain;

extern double *K_global, *plastic_K_global, *residual, *Pin, *Pex, *Pex_temp, *disp, *sigmay, *cp_dA;
extern double *reaction_force, *damage_visual;
extern double *J2_dlambda, *J2_stresseq, *J2_stressm, *J2_triaxiality;

extern double **xyz, **xyz_initial, **xyz_temp, **distance, **distance_initial, **KnTve, **F, **csx, **csy, **csz;
extern double **dL_total, **TdL_total, **csx_initial, **csy_initial, **csz_initial, **Ce;
extern double **slip_normal, **schmid_tensor, **schmid_tensor_local, **cp_RSS, **stress_tensor;
extern double **cp_Cab, **strain_tensor, **dL_ave, **ddL_total, **TddL_total, **F_temp, **ddLp;
extern double **dL, **ddL, **bond_stress, **damage_broken, **damage_w, **bond_stretch, **bond_vector, **bond_force;

extern double **Kn, **Tv, **J2_alpha, **damage_local, **damage_nonlocal, **cp_A, **cp_dgy, **cp_dA_single, **J2_beta_eq;
extern double ***slip_vector, ***dLp, ***J2_beta, ***damage_D, ***cp_gy, ***cp_A_single;

/* Boundary-condition tables, one row per load step (reference include/lpm.h:83-99). */
struct dispBCPara
{
    int type;    /* particle type the entry applies to */
    char flag;   /* 'x', 'y' or 'z' */
    double step; /* displacement increment per load step */
};

struct forceBCPara
{
    int type;
    char flag1;
    double step1; /* x force increment, shared by all particles of the type */
    char flag2;
    double step2; /* y force increment */
    char flag3;
    double step3; /* z force increment (3D) */
};

/* The two MKL routines that drivers call directly, with their public signatures. */
double cblas_dnrm2(const int n, const double *x, const int incx);
void mkl_free_buffers(void);

#ifdef __cplusplus
}
#endif

#include "boundary.h"
#include "data_handler.h"
#include "initialization.h"
#include "neighbor.h"
#include "solver.h"
#include "stiffness.h"
#include "lpm_basic.h"
#include "constitutive.h"

#endif
