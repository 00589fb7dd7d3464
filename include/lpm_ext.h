/*
 * lpm_ext.h -- extensions of the drop-in API that the reference does not have.
 * They keep the reference's global data model: call them from a driver that uses
 * lpm.h.  None of them changes the behaviour of a reference entry point unless the
 * driver calls it.
 */
#ifndef LPM_EXT_H
#define LPM_EXT_H

#include "lpm.h"
#include "lpmx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-particle factor on the elastic constants (E_i = factor_i * E) of the current
 * lattice, applied by every following calcKnTv(); NULL switches it off.  The nparticle
 * values are copied; a new lattice (searchNormalNeighbor) drops them.  Needs the
 * lattice (call after searchNormalNeighbor); LPMX_ERR_STATE otherwise. */
int lpm_ext_set_modulus(const double *factor);

/* Per-bond critical stretch [nparticle][nneighbors] of the brittle law (plmode 6) of
 * the current lattice, symmetric in (i, j); NULL restores the global
 * critical_bstrain.  The values are copied; a new lattice drops them.  Needs the
 * lattice (call after searchNormalNeighbor); LPMX_ERR_STATE otherwise. */
int lpm_ext_set_bond_critical(const double *crit);

/* Precision and tolerances of solverCG(); NULL restores the reference settings
 * (FP64, dpar[0] = 1e-8, dpar[1] = 1e-12 on squared norms). */
int lpm_ext_set_solver(const lpmx_solver *s);

/* Load steps first .. first+nsteps-1 of the driver's BC tables, with every field
 * resident on the device (the loop of the reference's src/lpmc_project.c:382-546).
 * The host globals are pushed before and pulled after; per-step statistics go to
 * out[nsteps] when out is not NULL. */
int lpm_ext_run_steps(const lpmx_solver *s, int first, int nsteps, const int *nbd,
                      struct dispBCPara (*dBP)[MAXSMALL], const int *nbf, struct forceBCPara (*fBP)[MAXSMALL],
                      char axis, int out_type, lpmx_step_stats *out);

#ifdef __cplusplus
}
#endif
#endif
