/* Neighbour lists (reference include/neighbor.h, src/neighbor.c:9-141). */
#pragma once
#ifndef NEIGHBOR_H
#define NEIGHBOR_H

#ifdef __cplusplus
extern "C" {
#endif

/* First/second-shell bond lists in increasing particle index (cell-list build on the GPU). */
void searchNormalNeighbor();
/* conn sets, K_pointer, and allocation of IK/JK/K_global; returns len(K_global). */
int searchAFEMNeighbor();

#ifdef __cplusplus
}
#endif
#endif
