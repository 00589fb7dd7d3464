/*
 * Definitions of the drop-in API's global model state (the names and C types of
 * the reference's include/lpm.h:55-81).  A driver that defines these globals itself
 * (as src/lpmc_project.c:17-45 does) takes precedence at load time: the library
 * refers to them through the dynamic symbol table, so every reference below
 * resolves to the driver's copies.  Without such a driver (ctypes, tests) these
 * definitions are used.
 */
#include "lpm.h"

int ntype, particles_first_row, nparticle, rows, layers, nneighbors, nneighbors1, nneighbors2, dim;
int lattice, nneighbors_AFEM, nneighbors_AFEM1, nneighbors_AFEM2, plmode, eulerflag, nslip_face;
int max_nslip_vector, nslipSys, nbreak;

int *IK, *JK, *type, *dispBC_index, *fix_index, *nslip_vector, *lacknblist, pbc[NDIM], *pl_flag;
int *nb, *nb_initial, *nb_conn, *state_v;
int **neighbors, **neighbors1, **neighbors2, **neighbors_AFEM;
int **K_pointer, **conn, **nsign, **cp_Jact;

double radius, hx, hy, hz, neighbor1_cutoff, neighbor2_cutoff, angle1, angle2, angle3, particle_volume;
double R_matrix[NDIM * NDIM], box[2 * NDIM], cp_tau0[3], cp_taus[3], cp_eta, cp_p, cp_h0, cp_q, cp_maxloop;
double box_x, box_y, box_z, dtime, cp_theta, J2_H, J2_xi, J2_C, damage_L;
double damage_threshold, damageb_A, damagec_A, critical_bstrain;

double *K_global, *plastic_K_global, *residual, *Pin, *Pex, *Pex_temp, *disp, *sigmay, *cp_dA;
double *reaction_force, *damage_visual;
double *J2_dlambda, *J2_stresseq, *J2_stressm, *J2_triaxiality;

double **xyz, **xyz_initial, **xyz_temp, **distance, **distance_initial, **KnTve, **F, **csx, **csy, **csz;
double **dL_total, **TdL_total, **csx_initial, **csy_initial, **csz_initial, **Ce;
double **slip_normal, **schmid_tensor, **schmid_tensor_local, **cp_RSS, **stress_tensor;
double **cp_Cab, **strain_tensor, **dL_ave, **ddL_total, **TddL_total, **F_temp, **ddLp;
double **dL, **ddL, **bond_stress, **damage_broken, **damage_w, **bond_stretch, **bond_vector, **bond_force;

double **Kn, **Tv, **J2_alpha, **damage_local, **damage_nonlocal, **cp_A, **cp_dgy, **cp_dA_single, **J2_beta_eq;
double ***slip_vector, ***dLp, ***J2_beta, ***damage_D, ***cp_gy, ***cp_A_single;
