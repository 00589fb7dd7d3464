/* Linear solves of the Newton step (reference include/solver.h, src/solver.c). */
#pragma once
#ifndef SOLVER_H
#define SOLVER_H

#ifdef __cplusplus
extern "C" {
#endif

void solverPARDISO();
/* K_global * disp = residual, then xyz += disp (reference src/solver.c:188-270). */
void solverCG();

#ifdef __cplusplus
}
#endif
#endif
