// Device-side construction of the smoothed-aggregation hierarchy (no counterpart in
// the reference): block sparse products, transposes, the tentative prolongator, the
// prolongator smoothing, diagonal inverses and the level layouts, all on the library
// stream.  Only the aggregation itself (a greedy pass in index order) stays on the
// host, in amg.hip.
//
// Products C = A B run in two passes of one wavefront per row of C:
//   symbolic: the row's candidate columns (the columns of the rows of B that A's row
//             names) go into a hash set in LDS; the distinct columns are counted,
//             and, after a scan of the counts, written in ascending order (each
//             column's rank among the set);
//   numeric:  each lane owns one column of the row; for every entry of A's row, in
//             order, the lanes look their column up in the named row of B (its
//             columns cached in LDS, binary search) and accumulate the block product.
// Every sum runs in a fixed order, so the hierarchy is the same from run to run.
#include <math.h>
#include <stdio.h>

#include <vector>

#include "amg_setup.h"

#define AMGD_HC 1024   // hash slots of a row's distinct columns
#define AMGD_UMAX 768  // distinct columns per row handled (the table stays 3/4 full)
#define AMGD_BCAP 1024 // right-operand row cached in LDS (its rows have <= AMGD_UMAX)

template <typename T>
static int dalloc(T **p, int64_t count)
{
    HIPCHK(hipMalloc((void **)p, sizeof(T) * (size_t)(count > 0 ? count : 1)));
    return LPMX_OK;
}

void amgd_free(DBsr &M)
{
    hipFree(M.ptr);
    hipFree(M.idx);
    hipFree(M.val);
    M = DBsr();
}

AmgLeft amgd_left(const DBsr &A)
{
    AmgLeft L;
    L.ell = false, L.nr = A.nr, L.br = A.br, L.bk = A.bc;
    L.ptr = A.ptr, L.idx = A.idx, L.val = A.val;
    return L;
}

AmgLeft amgd_left_tangent(void)
{
    AmgLeft L;
    L.ell = true, L.nr = g.n, L.br = L.bk = g.dim;
    L.ptr = g.sptr, L.len = g.nb_conn, L.idx = g.col, L.val = g.K;
    return L;
}

// ------------------------------------------------------------ left operand -------
template <bool ELL>
__device__ __forceinline__ int l_len(const AmgLeft &L, int i)
{
    return ELL ? L.len[i] : L.ptr[i + 1] - L.ptr[i];
}

template <bool ELL>
__device__ __forceinline__ int l_col(const AmgLeft &L, int i, int k)
{
    if (ELL)
        return L.idx[(int64_t)(L.ptr[i >> 6] + k) * 64 + (i & 63)];
    return L.idx[L.ptr[i] + k];
}

template <bool ELL, int BS>
__device__ __forceinline__ double l_val(const AmgLeft &L, int i, int k, int e)
{
    if (ELL)
        return L.val[((int64_t)(L.ptr[i >> 6] + k) * BS + e) * 64 + (i & 63)];
    return L.val[(int64_t)(L.ptr[i] + k) * BS + e];
}

__device__ __forceinline__ unsigned col_hash(int c)
{
    return ((unsigned)c * 0x9E3779B1u) >> 22; // 10 bits: AMGD_HC slots
}

// Symbolic pass, one wavefront (work-group) per row.  Pass 1 (cidx null): cnt[i] =
// distinct columns, flag[0] set on overflow, flag[1] = longest row.  Pass 2: the
// ascending columns at cidx[cptr[i]...].
template <bool ELL>
__global__ __launch_bounds__(64) void k_sym(AmgLeft L, const int *__restrict__ bptr, const int *__restrict__ bidx,
                                            int *__restrict__ cnt, const int *__restrict__ cptr,
                                            int *__restrict__ cidx, int *flag)
{
    __shared__ int tab[AMGD_HC];
    __shared__ int list[AMGD_UMAX];
    __shared__ int nu, nl;
    const int i = blockIdx.x, lane = threadIdx.x;
    for (int t = lane; t < AMGD_HC; t += 64)
        tab[t] = -1;
    if (lane == 0)
        nu = 0, nl = 0;
    __syncthreads();
    const int len = l_len<ELL>(L, i);
    for (int k = lane; k < len; k += 64)
    {
        const int j = l_col<ELL>(L, i, k);
        const int q1 = bptr[j + 1];
        for (int q = bptr[j]; q < q1; q++)
        {
            const int c = bidx[q];
            unsigned h = col_hash(c);
            for (int probe = 0; probe < AMGD_HC; probe++) // bounded: ends on a full table
            {
                const int prev = atomicCAS(&tab[h], -1, c);
                if (prev == -1)
                {
                    atomicAdd(&nu, 1);
                    break;
                }
                if (prev == c)
                    break;
                h = (h + 1) & (AMGD_HC - 1);
            }
        }
    }
    __syncthreads();
    const int u = nu;
    if (u > AMGD_UMAX) // pass 1 reports it; pass 2 does not run then
    {
        if (lane == 0 && flag)
            atomicOr(flag, 1);
        return;
    }
    if (!cidx)
    {
        if (lane == 0)
        {
            cnt[i] = u;
            atomicMax(flag + 1, u);
        }
        return;
    }
    for (int t = lane; t < AMGD_HC; t += 64)
        if (tab[t] != -1)
            list[atomicAdd(&nl, 1)] = tab[t];
    __syncthreads();
    int *out = cidx + cptr[i];
    for (int e = lane; e < u; e += 64)
    {
        const int v = list[e];
        int r = 0;
        for (int x = 0; x < u; x++)
            r += list[x] < v;
        out[r] = v;
    }
}

// Numeric pass, one wavefront (work-group) per row; lanes own the row's columns, 64 at
// a time.
template <bool ELL, int BR, int BK, int BC>
__global__ __launch_bounds__(64) void k_num(AmgLeft L, const int *__restrict__ bptr, const int *__restrict__ bidx,
                                            const double *__restrict__ bval, const int *__restrict__ cptr,
                                            const int *__restrict__ cidx, double *__restrict__ cval)
{
    __shared__ int bcol[AMGD_BCAP];
    const int i = blockIdx.x, lane = threadIdx.x;
    const int c0 = cptr[i], u = cptr[i + 1] - c0;
    const int len = l_len<ELL>(L, i);
    for (int s0 = 0; s0 < u; s0 += 64)
    {
        const int s = s0 + lane;
        const int mine = s < u ? cidx[c0 + s] : -1;
        double acc[BR * BC];
#pragma unroll
        for (int e = 0; e < BR * BC; e++)
            acc[e] = 0.0;
        for (int k = 0; k < len; k++)
        {
            const int j = l_col<ELL>(L, i, k);
            const int b0 = bptr[j], bl = bptr[j + 1] - b0;
            __syncthreads(); // the previous row of B is no longer read
            for (int t = lane; t < bl; t += 64)
                bcol[t] = bidx[b0 + t];
            __syncthreads();
            int lo = 0, hi = bl;
            while (lo < hi)
            {
                const int mid = (lo + hi) >> 1;
                if (bcol[mid] < mine)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (s < u && lo < bl && bcol[lo] == mine)
            {
                const double *bb = bval + (int64_t)(b0 + lo) * (BK * BC);
#pragma unroll
                for (int m = 0; m < BK; m++)
                {
                    double bm[BC];
#pragma unroll
                    for (int c = 0; c < BC; c++)
                        bm[c] = bb[m * BC + c];
#pragma unroll
                    for (int r = 0; r < BR; r++)
                    {
                        const double a = l_val<ELL, BR * BK>(L, i, k, r * BK + m);
#pragma unroll
                        for (int c = 0; c < BC; c++)
                            acc[r * BC + c] += a * bm[c];
                    }
                }
            }
        }
        if (s < u)
#pragma unroll
            for (int e = 0; e < BR * BC; e++)
                cval[(int64_t)(c0 + s) * (BR * BC) + e] = acc[e];
    }
}

__global__ void k_zero_int(int *p, int64_t n)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n)
        p[t] = 0;
}

template <bool ELL, int BR, int BK, int BC>
static void launch_num(const AmgLeft &A, const DBsr &B, const DBsr &C)
{
    hipLaunchKernelGGL((k_num<ELL, BR, BK, BC>), dim3(A.nr), dim3(64), 0, g_stream, A, B.ptr, B.idx, B.val, C.ptr,
                       C.idx, C.val);
}

int amgd_spgemm(const AmgLeft &A, const DBsr &B, DBsr *C)
{
    if (A.bk != B.br)
        return lpmx_fail(LPMX_ERR_STATE, "AMG product: inner block sizes %d and %d differ", A.bk, B.br);
    if (B.maxrow > AMGD_BCAP)
        return lpmx_fail(LPMX_ERR_CAPACITY, "AMG product: right operand rows of %d blocks (at most %d)", B.maxrow,
                         AMGD_BCAP);
    DBsr R;
    R.nr = A.nr, R.nc = B.nc, R.br = A.br, R.bc = B.bc;
    int *cnt = nullptr, *flag = nullptr;
    RETCHK(dalloc(&cnt, (int64_t)A.nr + 1));
    RETCHK(dalloc(&flag, 2));
    RETCHK(dalloc(&R.ptr, (int64_t)A.nr + 1));
    hipLaunchKernelGGL(k_zero_int, dim3(nblocks((int64_t)A.nr + 1)), dim3(LPMX_BLOCK), 0, g_stream, cnt,
                       (int64_t)A.nr + 1);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(LPMX_BLOCK), 0, g_stream, flag, (int64_t)2);
    if (A.ell)
        hipLaunchKernelGGL(k_sym<true>, dim3(A.nr), dim3(64), 0, g_stream, A, B.ptr, B.idx, cnt, nullptr, nullptr,
                           flag);
    else
        hipLaunchKernelGGL(k_sym<false>, dim3(A.nr), dim3(64), 0, g_stream, A, B.ptr, B.idx, cnt, nullptr, nullptr,
                           flag);
    LAUNCHCHK();
    int fl[2];
    HIPCHK(hipMemcpyAsync(fl, flag, sizeof(fl), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    int rc = LPMX_OK;
    if (fl[0])
        rc = lpmx_fail(LPMX_ERR_CAPACITY, "AMG product: a row has more than %d distinct columns", AMGD_UMAX);
    if (rc == LPMX_OK)
        rc = lpmx_dev_scan_int(cnt, R.ptr, A.nr + 1, &R.nnzb);
    hipFree(cnt);
    hipFree(flag);
    if (rc != LPMX_OK)
    {
        amgd_free(R);
        return rc;
    }
    R.maxrow = fl[1];
    const int bs = R.br * R.bc;
    if (dalloc(&R.idx, R.nnzb) != LPMX_OK || dalloc(&R.val, (int64_t)R.nnzb * bs) != LPMX_OK)
    {
        amgd_free(R);
        return lpmx_fail(LPMX_ERR_HIP, "AMG product: allocation of %d blocks failed", R.nnzb);
    }
    if (A.ell)
        hipLaunchKernelGGL(k_sym<true>, dim3(A.nr), dim3(64), 0, g_stream, A, B.ptr, B.idx, nullptr, R.ptr, R.idx,
                           nullptr);
    else
        hipLaunchKernelGGL(k_sym<false>, dim3(A.nr), dim3(64), 0, g_stream, A, B.ptr, B.idx, nullptr, R.ptr, R.idx,
                           nullptr);
    LAUNCHCHK();
    const int key = (A.ell ? 1000 : 0) + A.br * 100 + A.bk * 10 + B.bc;
    switch (key)
    {
    case 1336: launch_num<true, 3, 3, 6>(A, B, R); break;
    case 1223: launch_num<true, 2, 2, 3>(A, B, R); break;
    case 636: launch_num<false, 6, 3, 6>(A, B, R); break;
    case 323: launch_num<false, 3, 2, 3>(A, B, R); break;
    case 666: launch_num<false, 6, 6, 6>(A, B, R); break;
    case 333: launch_num<false, 3, 3, 3>(A, B, R); break;
    default:
        amgd_free(R);
        return lpmx_fail(LPMX_ERR_STATE, "AMG product: block sizes %dx%d by %dx%d not built", A.br, A.bk, B.br,
                         B.bc);
    }
    LAUNCHCHK();
    *C = R;
    return LPMX_OK;
}

// ------------------------------------------------------------------ transpose ----
__global__ void k_tr_count(int64_t nnz, const int *__restrict__ idx, int *__restrict__ cnt)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < nnz)
        atomicAdd(&cnt[idx[t]], 1);
}

// Entries of T's rows in any order (atomic cursors); sorted by k_tr_sort.
__global__ void k_tr_fill(int nr, const int *__restrict__ ptr, const int *__restrict__ idx,
                          const int *__restrict__ tptr, int *__restrict__ cur, int *__restrict__ trow,
                          int *__restrict__ tsrc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr)
        return;
    for (int k = ptr[i]; k < ptr[i + 1]; k++)
    {
        const int c = idx[k];
        const int p = tptr[c] + atomicAdd(&cur[c], 1);
        trow[p] = i;
        tsrc[p] = k;
    }
}

// One wavefront per row of T: each entry's place is its rank by source row (the
// source rows of a column are distinct); the block is copied transposed.
template <int BR, int BC>
__global__ __launch_bounds__(64) void k_tr_sort(int ntr, const int *__restrict__ tptr, const int *__restrict__ trow,
                                                const int *__restrict__ tsrc, const double *__restrict__ val,
                                                int *__restrict__ tidx, double *__restrict__ tval)
{
    __shared__ int rows[AMGD_BCAP];
    const int c = blockIdx.x, lane = threadIdx.x;
    const int a = tptr[c], L = tptr[c + 1] - a;
    const bool lds = L <= AMGD_BCAP;
    if (lds)
        for (int e = lane; e < L; e += 64)
            rows[e] = trow[a + e];
    __syncthreads();
    for (int e = lane; e < L; e += 64)
    {
        const int r = trow[a + e];
        int rank = 0;
        if (lds)
            for (int x = 0; x < L; x++)
                rank += rows[x] < r;
        else
            for (int x = 0; x < L; x++)
                rank += trow[a + x] < r;
        const int64_t d = a + rank, s = tsrc[a + e];
        tidx[d] = r;
#pragma unroll
        for (int p = 0; p < BR; p++)
#pragma unroll
            for (int q = 0; q < BC; q++)
                tval[d * (BR * BC) + q * BR + p] = val[s * (BR * BC) + p * BC + q];
    }
}

int amgd_transpose(const DBsr &A, DBsr *T)
{
    DBsr R;
    R.nr = A.nc, R.nc = A.nr, R.br = A.bc, R.bc = A.br, R.nnzb = A.nnzb;
    int *cnt = nullptr, *cur = nullptr, *trow = nullptr, *tsrc = nullptr;
    RETCHK(dalloc(&cnt, (int64_t)R.nr + 1));
    RETCHK(dalloc(&cur, (int64_t)R.nr));
    RETCHK(dalloc(&trow, A.nnzb));
    RETCHK(dalloc(&tsrc, A.nnzb));
    RETCHK(dalloc(&R.ptr, (int64_t)R.nr + 1));
    RETCHK(dalloc(&R.idx, A.nnzb));
    RETCHK(dalloc(&R.val, (int64_t)A.nnzb * A.br * A.bc));
    hipLaunchKernelGGL(k_zero_int, dim3(nblocks((int64_t)R.nr + 1)), dim3(LPMX_BLOCK), 0, g_stream, cnt,
                       (int64_t)R.nr + 1);
    hipLaunchKernelGGL(k_zero_int, dim3(nblocks(R.nr)), dim3(LPMX_BLOCK), 0, g_stream, cur, (int64_t)R.nr);
    hipLaunchKernelGGL(k_tr_count, dim3(nblocks(A.nnzb)), dim3(LPMX_BLOCK), 0, g_stream, (int64_t)A.nnzb, A.idx, cnt);
    LAUNCHCHK();
    int total = 0;
    RETCHK(lpmx_dev_scan_int(cnt, R.ptr, R.nr + 1, &total));
    hipLaunchKernelGGL(k_tr_fill, dim3(nblocks(A.nr)), dim3(LPMX_BLOCK), 0, g_stream, A.nr, A.ptr, A.idx, R.ptr, cur,
                       trow, tsrc);
    const int key = A.br * 10 + A.bc;
    switch (key)
    {
#define LPMX_TR(BR_, BC_)                                                                                           \
    case BR_ * 10 + BC_:                                                                                            \
        hipLaunchKernelGGL((k_tr_sort<BR_, BC_>), dim3(R.nr), dim3(64), 0, g_stream, R.nr, R.ptr, trow, tsrc, A.val, \
                           R.idx, R.val);                                                                           \
        break;
        LPMX_TR(3, 6)
        LPMX_TR(2, 3)
        LPMX_TR(6, 6)
        LPMX_TR(3, 3)
#undef LPMX_TR
    default:
        return lpmx_fail(LPMX_ERR_STATE, "AMG transpose: %dx%d blocks not built", A.br, A.bc);
    }
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(cnt);
    hipFree(cur);
    hipFree(trow);
    hipFree(tsrc);
    // rows of T: at most the rows of A's column; the longest is not needed as a right
    // operand (T is only ever the left one)
    R.maxrow = 0;
    *T = R;
    return LPMX_OK;
}

// ----------------------------------------------------------------- bond graph ----
// Ascending distinct partners of particle i (self and empty slots excluded) into
// planes tmp[k*n + i]; cnt[i] their number.
__global__ void k_bond_rows(DV v, int *__restrict__ tmp, int *__restrict__ cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    int o[LPMX_MAXNB];
    int m = 0;
    for (int j = 0; j < v.nb0[i]; j++)
    {
        const int q = v.nbr[(int64_t)j * v.n + i];
        if (q < 0 || q == i)
            continue;
        int p = m;
        while (p > 0 && o[p - 1] > q)
            p--;
        if (p > 0 && o[p - 1] == q)
            continue;
        for (int t = m; t > p; t--)
            o[t] = o[t - 1];
        o[p] = q;
        m++;
    }
    for (int k = 0; k < m; k++)
        tmp[(int64_t)k * v.n + i] = o[k];
    cnt[i] = m;
}

__global__ void k_bond_fill(int n, const int *__restrict__ tmp, const int *__restrict__ aptr, int *__restrict__ adj)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int a = aptr[i], m = aptr[i + 1] - a;
    for (int k = 0; k < m; k++)
        adj[a + k] = tmp[(int64_t)k * n + i];
}

int amgd_bond_graph(int **aptr, int **adj, int *nadj)
{
    const int n = g.n;
    if (g.nb > LPMX_MAXNB)
        return lpmx_fail(LPMX_ERR_CAPACITY, "bond graph: %d bonds per particle (at most %d)", g.nb, LPMX_MAXNB);
    int *tmp = nullptr, *cnt = nullptr, *ptr = nullptr, *a = nullptr;
    RETCHK(dalloc(&tmp, (int64_t)n * g.nb));
    RETCHK(dalloc(&cnt, (int64_t)n + 1));
    RETCHK(dalloc(&ptr, (int64_t)n + 1));
    hipLaunchKernelGGL(k_zero_int, dim3(nblocks((int64_t)n + 1)), dim3(LPMX_BLOCK), 0, g_stream, cnt, (int64_t)n + 1);
    hipLaunchKernelGGL(k_bond_rows, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, tmp, cnt);
    LAUNCHCHK();
    int total = 0;
    RETCHK(lpmx_dev_scan_int(cnt, ptr, n + 1, &total));
    RETCHK(dalloc(&a, total));
    hipLaunchKernelGGL(k_bond_fill, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, n, tmp, ptr, a);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(tmp);
    hipFree(cnt);
    *aptr = ptr, *adj = a, *nadj = total;
    return LPMX_OK;
}

// ------------------------------------------------------ null space, tentative ----
__global__ void k_centroid(int na, const int *__restrict__ mptr, const int *__restrict__ mem,
                           const double *__restrict__ x, double *__restrict__ cx)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= na)
        return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int q = mptr[a]; q < mptr[a + 1]; q++)
        for (int c = 0; c < 3; c++)
            s[c] += x[(int64_t)3 * mem[q] + c];
    const int m = mptr[a + 1] - mptr[a];
    for (int c = 0; c < 3; c++)
        cx[(int64_t)a * 3 + c] = s[c] / m;
}

template <int D, int NM>
__global__ void k_null0(int n, const double *__restrict__ x, const int *__restrict__ agg,
                        const double *__restrict__ cx, const int *__restrict__ db, const int *__restrict__ fx,
                        double *__restrict__ B)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t a = agg[i];
    const double px = x[(int64_t)3 * i] - cx[a * 3], py = x[(int64_t)3 * i + 1] - cx[a * 3 + 1],
                 pz = x[(int64_t)3 * i + 2] - cx[a * 3 + 2];
    double b[D][NM];
    if constexpr (D == 3)
    {
        const double m6[3][6] = {{1, 0, 0, 0, pz, -py}, {0, 1, 0, -pz, 0, px}, {0, 0, 1, py, -px, 0}};
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 6; c++)
                b[r][c] = m6[r][c];
    }
    else
    {
        const double m3[2][3] = {{1, 0, -py}, {0, 1, px}};
        for (int r = 0; r < 2; r++)
            for (int c = 0; c < 3; c++)
                b[r][c] = m3[r][c];
    }
    for (int r = 0; r < D; r++)
    {
        const bool keep = db[(int64_t)i * D + r] && fx[(int64_t)i * D + r];
        for (int c = 0; c < NM; c++)
            B[((int64_t)i * D + r) * NM + c] = keep ? b[r][c] : 0.0;
    }
}

int amgd_null_space(int n, int d, int nm, int na, const int *agg, const int *mptr, const int *mem, double *B)
{
    double *cx = nullptr;
    RETCHK(dalloc(&cx, (int64_t)na * 3));
    hipLaunchKernelGGL(k_centroid, dim3(nblocks(na)), dim3(LPMX_BLOCK), 0, g_stream, na, mptr, mem, g.x, cx);
    if (d == 3 && nm == 6)
        hipLaunchKernelGGL((k_null0<3, 6>), dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, n, g.x, agg, cx,
                           g.dispBC, g.fix, B);
    else if (d == 2 && nm == 3)
        hipLaunchKernelGGL((k_null0<2, 3>), dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, n, g.x, agg, cx,
                           g.dispBC, g.fix, B);
    else
        return lpmx_fail(LPMX_ERR_STATE, "AMG null space: %d DOFs and %d modes not built", d, nm);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(cx);
    return LPMX_OK;
}

// One thread per aggregate: the aggregate's rows of B orthonormalised in place in P0
// (modified Gram-Schmidt; a column below 1e-10 of the largest column norm becomes
// zero), the triangular factor into Bc.
template <int D, int NM>
__global__ void k_tentative(int na, const int *__restrict__ mptr, const int *__restrict__ mem,
                            const double *__restrict__ B, double *__restrict__ Q, double *__restrict__ Bc)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= na)
        return;
    const int q0 = mptr[a], q1 = mptr[a + 1];
#define QV(q, r, c) Q[((int64_t)mem[q] * D + (r)) * NM + (c)]
    for (int q = q0; q < q1; q++)
        for (int r = 0; r < D; r++)
            for (int c = 0; c < NM; c++)
                QV(q, r, c) = B[((int64_t)mem[q] * D + r) * NM + c];
    double *R = Bc + (int64_t)a * NM * NM;
    for (int e = 0; e < NM * NM; e++)
        R[e] = 0.0;
    double ref = 0.0;
    for (int c = 0; c < NM; c++)
    {
        double s = 0;
        for (int q = q0; q < q1; q++)
            for (int r = 0; r < D; r++)
                s += QV(q, r, c) * QV(q, r, c);
        ref = fmax(ref, sqrt(s));
    }
    for (int c = 0; c < NM; c++)
    {
        for (int p = 0; p < c; p++)
        {
            double s = 0;
            for (int q = q0; q < q1; q++)
                for (int r = 0; r < D; r++)
                    s += QV(q, r, p) * QV(q, r, c);
            R[p * NM + c] = s;
            for (int q = q0; q < q1; q++)
                for (int r = 0; r < D; r++)
                    QV(q, r, c) -= s * QV(q, r, p);
        }
        double s = 0;
        for (int q = q0; q < q1; q++)
            for (int r = 0; r < D; r++)
                s += QV(q, r, c) * QV(q, r, c);
        s = sqrt(s);
        const bool keep = s > 1e-10 * ref && ref > 0;
        R[c * NM + c] = keep ? s : 0.0;
        for (int q = q0; q < q1; q++)
            for (int r = 0; r < D; r++)
                QV(q, r, c) = keep ? QV(q, r, c) / s : 0.0;
    }
#undef QV
}

__global__ void k_iota(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n)
        p[i] = i;
}

int amgd_tentative(int nr, int na, int d, int nm, const int *agg, const int *mptr, const int *mem, const double *B,
                   DBsr *P0, double *Bc)
{
    DBsr P;
    P.nr = nr, P.nc = na, P.br = d, P.bc = nm, P.nnzb = nr, P.maxrow = 1;
    RETCHK(dalloc(&P.ptr, (int64_t)nr + 1));
    RETCHK(dalloc(&P.idx, nr));
    RETCHK(dalloc(&P.val, (int64_t)nr * d * nm));
    hipLaunchKernelGGL(k_iota, dim3(nblocks((int64_t)nr + 1)), dim3(LPMX_BLOCK), 0, g_stream, P.ptr, nr);
    HIPCHK(hipMemcpyAsync(P.idx, agg, sizeof(int) * nr, hipMemcpyDeviceToDevice, g_stream));
    const int key = d * 10 + nm;
    switch (key)
    {
#define LPMX_TT(D_, NM_)                                                                                            \
    case D_ * 10 + NM_:                                                                                             \
        hipLaunchKernelGGL((k_tentative<D_, NM_>), dim3(nblocks(na, 64)), dim3(64), 0, g_stream, na, mptr, mem, B,   \
                           P.val, Bc);                                                                              \
        break;
        LPMX_TT(3, 6)
        LPMX_TT(2, 3)
        LPMX_TT(6, 6)
        LPMX_TT(3, 3)
#undef LPMX_TT
    default:
        amgd_free(P);
        return lpmx_fail(LPMX_ERR_STATE, "AMG tentative prolongator: %d DOFs and %d modes not built", d, nm);
    }
    LAUNCHCHK();
    *P0 = P;
    return LPMX_OK;
}

// -------------------------------------------------------------- smoothing --------
template <int D, int NM>
__global__ void k_smooth(int nr, const int *__restrict__ ptr, const int *__restrict__ idx, double *__restrict__ val,
                         const double *__restrict__ Minv, const int *__restrict__ agg, const double *__restrict__ P0,
                         double w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr)
        return;
    double Di[D * D];
#pragma unroll
    for (int e = 0; e < D * D; e++)
        Di[e] = Minv[(int64_t)e * nr + i];
    const int ai = agg[i];
    for (int k = ptr[i]; k < ptr[i + 1]; k++)
    {
        double *blk = val + (int64_t)k * D * NM;
        double t[D * NM];
#pragma unroll
        for (int r = 0; r < D; r++)
#pragma unroll
            for (int c = 0; c < NM; c++)
            {
                double s = 0;
#pragma unroll
                for (int q = 0; q < D; q++)
                    s += Di[r * D + q] * blk[q * NM + c];
                t[r * NM + c] = -w * s;
            }
        if (idx[k] == ai)
#pragma unroll
            for (int e = 0; e < D * NM; e++)
                t[e] += P0[(int64_t)i * D * NM + e];
#pragma unroll
        for (int e = 0; e < D * NM; e++)
            blk[e] = t[e];
    }
}

int amgd_smooth(DBsr &AP0, const double *Minv, const int *agg, const DBsr &P0, double w)
{
    const int key = AP0.br * 10 + AP0.bc;
    const dim3 G(nblocks(AP0.nr)), T(LPMX_BLOCK);
    switch (key)
    {
#define LPMX_SM(D_, NM_)                                                                                            \
    case D_ * 10 + NM_:                                                                                             \
        hipLaunchKernelGGL((k_smooth<D_, NM_>), G, T, 0, g_stream, AP0.nr, AP0.ptr, AP0.idx, AP0.val, Minv, agg,      \
                           P0.val, w);                                                                              \
        break;
        LPMX_SM(3, 6)
        LPMX_SM(2, 3)
        LPMX_SM(6, 6)
        LPMX_SM(3, 3)
#undef LPMX_SM
    default:
        return lpmx_fail(LPMX_ERR_STATE, "AMG smoothing: %dx%d blocks not built", AP0.br, AP0.bc);
    }
    LAUNCHCHK();
    return LPMX_OK;
}

// --------------------------------------------------------- diagonal inverses -----
// Gauss-Jordan with partial pivoting; a pivot below 1e-14 of the largest entry marks
// a decoupled zero row and column (an unused coarse mode), which gets the identity.
template <int D>
__device__ void dev_small_inverse(double *m, double *inv)
{
    for (int e = 0; e < D * D; e++)
        inv[e] = 0.0;
    for (int r = 0; r < D; r++)
        inv[r * D + r] = 1.0;
    double scale = 0.0;
    for (int e = 0; e < D * D; e++)
        scale = fmax(scale, fabs(m[e]));
    for (int c = 0; c < D; c++)
    {
        int p = c;
        for (int r = c + 1; r < D; r++)
            if (fabs(m[r * D + c]) > fabs(m[p * D + c]))
                p = r;
        if (!(fabs(m[p * D + c]) > 1e-14 * scale))
        {
            for (int k = 0; k < D; k++)
                m[c * D + k] = m[k * D + c] = 0.0, inv[c * D + k] = inv[k * D + c] = 0.0;
            m[c * D + c] = 1.0, inv[c * D + c] = 1.0;
            continue;
        }
        if (p != c)
            for (int k = 0; k < D; k++)
            {
                const double t = m[p * D + k], u = inv[p * D + k];
                m[p * D + k] = m[c * D + k], inv[p * D + k] = inv[c * D + k];
                m[c * D + k] = t, inv[c * D + k] = u;
            }
        const double piv = m[c * D + c];
        for (int k = 0; k < D; k++)
            m[c * D + k] /= piv, inv[c * D + k] /= piv;
        for (int r = 0; r < D; r++)
            if (r != c && m[r * D + c] != 0.0)
            {
                const double f = m[r * D + c];
                for (int k = 0; k < D; k++)
                    m[r * D + k] -= f * m[c * D + k], inv[r * D + k] -= f * inv[c * D + k];
            }
    }
}

template <int D>
__global__ void k_diag_inv(int nr, const int *__restrict__ ptr, const int *__restrict__ idx,
                           const double *__restrict__ val, double *__restrict__ Minv, float *__restrict__ Minv32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr)
        return;
    double m[D * D], inv[D * D];
    for (int e = 0; e < D * D; e++)
        m[e] = 0.0;
    int lo = ptr[i], hi = ptr[i + 1];
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (idx[mid] < i)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < ptr[i + 1] && idx[lo] == i)
        for (int e = 0; e < D * D; e++)
            m[e] = val[(int64_t)lo * D * D + e];
    dev_small_inverse<D>(m, inv);
    for (int e = 0; e < D * D; e++)
    {
        Minv[(int64_t)e * nr + i] = inv[e];
        Minv32[(int64_t)e * nr + i] = (float)inv[e];
    }
}

int amgd_diag_inverses(const DBsr &A, double *Minv, float *Minv32)
{
    const dim3 G(nblocks(A.nr)), T(LPMX_BLOCK);
    if (A.br == 6 && A.bc == 6)
        hipLaunchKernelGGL(k_diag_inv<6>, G, T, 0, g_stream, A.nr, A.ptr, A.idx, A.val, Minv, Minv32);
    else if (A.br == 3 && A.bc == 3)
        hipLaunchKernelGGL(k_diag_inv<3>, G, T, 0, g_stream, A.nr, A.ptr, A.idx, A.val, Minv, Minv32);
    else
        return lpmx_fail(LPMX_ERR_STATE, "AMG diagonal inverses: %dx%d blocks not built", A.br, A.bc);
    LAUNCHCHK();
    return LPMX_OK;
}

// ------------------------------------------------------------------ layouts ------
__global__ void k_rptr64(int n, const int *__restrict__ p, int64_t *__restrict__ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n)
        q[i] = p[i];
}

int amgd_rptr64(const DBsr &A, int64_t *rptr)
{
    hipLaunchKernelGGL(k_rptr64, dim3(nblocks((int64_t)A.nr + 1)), dim3(LPMX_BLOCK), 0, g_stream, A.nr, A.ptr, rptr);
    LAUNCHCHK();
    return LPMX_OK;
}

// WROW values: entry e of the row's j-th block at val[k0*bs + e*len + j]; one
// wavefront per row, 4 rows per work-group.
__global__ __launch_bounds__(256) void k_wrow_vals(int nr, int bs, const int *__restrict__ ptr,
                                                   const double *__restrict__ src, double *__restrict__ dst)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= nr)
        return;
    const int64_t k0 = ptr[i];
    const int len = ptr[i + 1] - (int)k0;
    for (int j = lane; j < len; j += 64)
        for (int e = 0; e < bs; e++)
            dst[k0 * bs + (int64_t)e * len + j] = src[(k0 + j) * bs + e];
}

int amgd_wrow_vals(const DBsr &A, double *val)
{
    hipLaunchKernelGGL(k_wrow_vals, dim3(nblocks(A.nr, 4)), dim3(256), 0, g_stream, A.nr, A.br * A.bc, A.ptr, A.val,
                       val);
    LAUNCHCHK();
    return LPMX_OK;
}

__global__ void k_slice_width(int nr, int nslices, const int *__restrict__ ptr, const int *__restrict__ perm,
                              int *__restrict__ w)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nslices)
        return;
    int m = 0;
    if (s < nslices)
        for (int l = 0; l < 64 && s * 64 + l < nr; l++)
        {
            const int i = perm ? perm[s * 64 + l] : s * 64 + l;
            m = max(m, ptr[i + 1] - ptr[i]);
        }
    w[s] = m;
}

int amgd_slice_widths(const DBsr &A, int *sptr, int *nslots, const int *perm)
{
    const int ns = (A.nr + 63) / 64;
    int *w = nullptr;
    RETCHK(dalloc(&w, (int64_t)ns + 1));
    hipLaunchKernelGGL(k_slice_width, dim3(nblocks((int64_t)ns + 1)), dim3(LPMX_BLOCK), 0, g_stream, A.nr, ns, A.ptr,
                       perm, w);
    LAUNCHCHK();
    const int rc = lpmx_dev_scan_int(w, sptr, ns + 1, nslots);
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(w);
    return rc;
}

// Sliced ELL: the row at position p (row perm[p], or p without a map) has its k-th
// block in slot sptr[p/64] + k, lane p%64; padding slots keep a zero block at a valid
// column (the row itself for a square matrix).
__global__ void k_slice_fill(int nr, int nc, int bs, const int *__restrict__ ptr, const int *__restrict__ idx,
                             const double *__restrict__ src, const int *__restrict__ perm, const int *__restrict__ sptr,
                             int *__restrict__ col, double *__restrict__ val)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nr)
        return;
    const int i = perm ? perm[p] : p;
    const int s = p >> 6, l = p & 63;
    const int p0 = ptr[i], len = ptr[i + 1] - p0, w = sptr[s + 1] - sptr[s];
    for (int k = 0; k < w; k++)
    {
        const int64_t slot = sptr[s] + k;
        if (k < len)
        {
            col[slot * 64 + l] = idx[p0 + k];
            for (int e = 0; e < bs; e++)
                val[(slot * bs + e) * 64 + l] = src[(int64_t)(p0 + k) * bs + e];
        }
        else
            col[slot * 64 + l] = nr == nc ? i : 0;
    }
}

int amgd_slice_fill(const DBsr &A, const int *sptr, int64_t nslots, int *col, double *val, const int *perm)
{
    const int bs = A.br * A.bc;
    HIPCHK(hipMemsetAsync(col, 0, sizeof(int) * nslots * 64, g_stream));
    HIPCHK(hipMemsetAsync(val, 0, sizeof(double) * nslots * bs * 64, g_stream));
    hipLaunchKernelGGL(k_slice_fill, dim3(nblocks(A.nr)), dim3(LPMX_BLOCK), 0, g_stream, A.nr, A.nc, bs, A.ptr, A.idx,
                       A.val, perm, sptr, col, val);
    LAUNCHCHK();
    return LPMX_OK;
}

__global__ void k_f32(const double *__restrict__ s, float *__restrict__ d, int64_t n)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n)
        d[t] = (float)s[t];
}

int amgd_to_f32(const double *src, float *dst, int64_t count)
{
    hipLaunchKernelGGL(k_f32, dim3(nblocks(count)), dim3(LPMX_BLOCK), 0, g_stream, src, dst, count);
    LAUNCHCHK();
    return LPMX_OK;
}

// ------------------------------------------------------------ power iteration ----
template <int D>
__global__ void k_bdiag(int n, const double *__restrict__ Minv, const double *__restrict__ x, double *__restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    for (int a = 0; a < D; a++)
    {
        double t = 0;
        for (int q = 0; q < D; q++)
            t += Minv[(int64_t)(a * D + q) * n + i] * x[(int64_t)i * D + q];
        y[(int64_t)i * D + a] = t;
    }
}

int amgd_block_diag_mv(int n, int d, const double *Minv, const double *x, double *y)
{
    const dim3 G(nblocks(n)), T(LPMX_BLOCK);
    if (d == 3)
        hipLaunchKernelGGL(k_bdiag<3>, G, T, 0, g_stream, n, Minv, x, y);
    else if (d == 6)
        hipLaunchKernelGGL(k_bdiag<6>, G, T, 0, g_stream, n, Minv, x, y);
    else if (d == 2)
        hipLaunchKernelGGL(k_bdiag<2>, G, T, 0, g_stream, n, Minv, x, y);
    else
        return lpmx_fail(LPMX_ERR_STATE, "block diagonal product: %d DOFs not built", d);
    LAUNCHCHK();
    return LPMX_OK;
}

__global__ void k_div(double *x, int64_t n, double d)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n)
        x[t] /= d;
}

int amgd_scale(double *x, int64_t count, double div)
{
    hipLaunchKernelGGL(k_div, dim3(nblocks(count)), dim3(LPMX_BLOCK), 0, g_stream, x, count, div);
    LAUNCHCHK();
    return LPMX_OK;
}
