/* Bond constants, tangent stiffness and residual (reference include/stiffness.h,
 * src/stiffness.c). */
#pragma once
#ifndef STIFFNESS_H
#define STIFFNESS_H

#ifdef __cplusplus
extern "C" {
#endif

void calcKnTv();
void updateRR(); /* unprototyped on purpose: drivers call updateRR(ni++) */
void calcStiffness2DFiniteDifference(int plmode);
void calcStiffness3DFiniteDifference(int plmode);

#ifdef __cplusplus
}
#endif
#endif
