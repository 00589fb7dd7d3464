/* Boundary conditions (reference include/boundary.h, src/boundary.c:12-281). */
#pragma once
#ifndef BOUNDARY_H
#define BOUNDARY_H

#ifdef __cplusplus
extern "C" {
#endif

struct dispBCPara;
struct forceBCPara;

/* Zero the rows and columns of displacement-constrained DOFs in K_global and put
 * the 2-norm of diag(K_global) on their diagonal; residual of those DOFs is zeroed. */
void setDispBC_stiffnessUpdate2D();
void setDispBC_stiffnessUpdate3D();

/* Move the particles of each entry's type by step and mark the DOF constrained. */
void setDispBC(int nboundDisp, struct dispBCPara *dBP);
/* Add step/count to the external force of every particle of each entry's type. */
void setForceBC(int nboundForce, struct forceBCPara *fBP);

#ifdef __cplusplus
}
#endif
#endif
