// Internal header: device-side construction of the smoothed-aggregation hierarchy
// (amg_setup.hip), used by the AMG build in amg.hip.  Not part of the C-ABI.
#pragma once

#include "common.h"

// Block CSR on the device: blocks row-major br x bc, columns ascending per row.
struct DBsr
{
    int nr = 0, nc = 0, br = 0, bc = 0;
    int nnzb = 0;
    int maxrow = 0;          // longest row (set by the products)
    int *ptr = nullptr;      // nr + 1 offsets
    int *idx = nullptr;      // nnzb columns
    double *val = nullptr;   // nnzb * br * bc
};
void amgd_free(DBsr &M);

// Left operand of a product: a DBsr or the model's sliced ELL tangent (g.sptr, g.col,
// g.K, row lengths g.nb_conn).
struct AmgLeft
{
    bool ell = false;
    int nr = 0, br = 0, bk = 0;
    const int *ptr = nullptr, *len = nullptr, *idx = nullptr;
    const double *val = nullptr;
};
AmgLeft amgd_left(const DBsr &A);
AmgLeft amgd_left_tangent(void);

// C = A B (columns of every row of C ascending, sums in a fixed order).
int amgd_spgemm(const AmgLeft &A, const DBsr &B, DBsr *C);
// T = A^T (blocks transposed).
int amgd_transpose(const DBsr &A, DBsr *T);
// Ascending distinct bond partners of every particle (self excluded): aptr[n + 1], adj.
int amgd_bond_graph(int **aptr, int **adj, int *nadj);
// Tentative prolongator of the aggregates (members mem[mptr[a] .. mptr[a + 1]),
// ascending): P0 (one d x nm block per row, at column agg[i]) and the coarse null space
// Bc [na*nm][nm] from the fine one B [nr*d][nm].
int amgd_tentative(int nr, int na, int d, int nm, const int *agg, const int *mptr, const int *mem, const double *B,
                   DBsr *P0, double *Bc);
// Fine null space: rigid-body modes about the aggregate centroids, zero except on the
// DOFs both displacement-constrained and fixed.
int amgd_null_space(int n, int d, int nm, int na, const int *agg, const int *mptr, const int *mem, double *B);
// In place: AP0 <- P0 - w D^-1 (A P0) with the planar inverses Minv [e*nr + i].
int amgd_smooth(DBsr &AP0, const double *Minv, const int *agg, const DBsr &P0, double w);
// Planar inverses of the diagonal blocks (FP64 and FP32), [e*nr + i].
int amgd_diag_inverses(const DBsr &A, double *Minv, float *Minv32);
// Layouts of a level operator: block-wise WROW values, sliced-ELL widths and fill.
int amgd_rptr64(const DBsr &A, int64_t *rptr);
int amgd_wrow_vals(const DBsr &A, double *val);
int amgd_slice_widths(const DBsr &A, int *sptr, int *nslots, const int *perm = nullptr);
int amgd_slice_fill(const DBsr &A, const int *sptr, int64_t nslots, int *col, double *val,
                    const int *perm = nullptr);
int amgd_to_f32(const double *src, float *dst, int64_t count);
// y = Minv x with planar block inverses (d x d), the power iteration's step.
int amgd_block_diag_mv(int n, int d, const double *Minv, const double *x, double *y);
int amgd_scale(double *x, int64_t count, double div);
