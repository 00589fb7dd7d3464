/* Constitutive laws, damage and crack bookkeeping (reference include/constitutive.h,
 * src/constitutive.c). */
#pragma once
#ifndef PLASTICITY_H
#define PLASTICITY_H

#ifdef __cplusplus
extern "C" {
#endif

/* conv_flag 0: slot[1] -> slot[0]; 1: slot[0] -> slot[1]; 2: slot[2] -> slot[0]. */
void switchStateV(int conv_flag);

void computeCab();

/* Dispatch on plmode: 0 J2 mixed-linear, 1 crystal plasticity, 4 incremental
 * predictor, 6 elastic; then stress and switchStateV(2). */
void computeBondForceGeneral(int plmode, int temp);
void computeBondForceElastic(int i);
void computeBondForceJ2mixedLinear3D(int ii);
void computeBondForceJ2nonlinearIso(int ii);
void computeBondForceCPMiehe(int ii);
void computeBondForceIncrementalUpdating(int ii);
void computeBondForceJ2energyReturnMap(int ii, int load_indicator);

/* Each returns the number of bonds (or particles) newly broken in this call. */
int updateDamageGeneral(const char *dataName, int tstep, int plmode);
int updateBrittleDamage(const char *dataName, int tstep, int nbreak);
int updateDuctileDamageBwiseLocal(const char *dataName, int tstep);
int updateDuctileDamagePwiseLocal(const char *dataName, int tstep);
int updateDuctileDamageBwiseNonlocal(const char *dataName, int tstep);
int updateDuctileDamagePwiseNonlocal(const char *dataName, int tstep);

void updateCrack();

struct bondStrain
{
    int i_index;
    int j_index;
    double bstrain;
};

#ifdef __cplusplus
}
#endif
#endif
