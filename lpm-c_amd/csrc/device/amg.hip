// Smoothed-aggregation algebraic multigrid as the CG preconditioner (no counterpart
// in the reference, whose solverCG is unpreconditioned).
//
// Setup, from the current constraint-eliminated tangent:
//   - aggregates (host): greedy in index order over a graph (a root and its free
//     neighbours), then leftovers join a neighbouring aggregate; the graph is the
//     bond graph on the finest level and, on each coarser level, the adjacency of the
//     aggregates of the level above (a and b adjacent when a member of a has a
//     neighbour in b), so every level coarsens by about the size of a neighbourhood
//     and the Galerkin stencils stay compact;
//   - tentative prolongator (device): the rigid-body modes of each aggregate (3
//     translations and 3 rotations in 3D, 2 + 1 in 2D) about its centroid, zero on
//     constrained DOFs, orthonormalised per aggregate (modified Gram-Schmidt; a
//     dependent mode gets a zero column); the triangular factor carries the modes to
//     the next level;
//   - smoothed prolongator P = (I - w D^-1 A) P0, w = 4 / (3 lambda_max(D^-1 A))
//     (device: power iteration, block sparse product);
//   - Galerkin coarse matrix P^T A P (device), down to at most ~1200 coarse DOFs,
//     which are inverted densely (host).
// The device kernels of the set-up are in amg_setup.hip.
// The hierarchy is built on the first preconditioned solve and kept while the
// pattern and the constraint set stay the same: the smoother of the finest level
// always uses the current tangent and its block-Jacobi inverse, the coarse levels
// the tangent of the build (the elastic tangent moves with the geometry only).
// Solve (device): one V-cycle per CG iteration with a Chebyshev polynomial smoother
// (block-Jacobi scaled, interval [1.15 lambda_max / ratio, 1.15 lambda_max]) of the
// same degree before and after the coarse correction, so the preconditioner is
// symmetric positive definite.  Levels with many rows are stored like the fine
// tangent (sliced ELL blocks, 64 rows per slice), levels with few, long rows as block
// CSR read by one wavefront per row, in FP64 and FP32.  Everything is per row, in a
// fixed order: results are deterministic.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "amg_setup.h"
#include "common.h"

// ------------------------------------------------------------------ host side ----
namespace
{
// In-place inverse of a small dense matrix (Gauss-Jordan, partial pivoting); a
// singular matrix gets the pseudo-inverse of its nonzero diagonal part.
void small_inverse(int d, double *a)
{
    std::vector<double> m(a, a + d * d), inv(d * d, 0.0);
    for (int i = 0; i < d; i++)
        inv[i * d + i] = 1.0;
    double scale = 0.0;
    for (int i = 0; i < d * d; i++)
        scale = std::max(scale, fabs(m[i]));
    for (int c = 0; c < d; c++)
    {
        int p = c;
        for (int r = c + 1; r < d; r++)
            if (fabs(m[r * d + c]) > fabs(m[p * d + c]))
                p = r;
        if (!(fabs(m[p * d + c]) > 1e-14 * scale))
        {
            // a decoupled zero row and column (an unused coarse mode): identity there
            for (int k = 0; k < d; k++)
                m[c * d + k] = m[k * d + c] = 0.0, inv[c * d + k] = inv[k * d + c] = 0.0;
            m[c * d + c] = 1.0, inv[c * d + c] = 1.0;
            continue;
        }
        if (p != c)
            for (int k = 0; k < d; k++)
                std::swap(m[p * d + k], m[c * d + k]), std::swap(inv[p * d + k], inv[c * d + k]);
        const double piv = m[c * d + c];
        for (int k = 0; k < d; k++)
            m[c * d + k] /= piv, inv[c * d + k] /= piv;
        for (int r = 0; r < d; r++)
            if (r != c && m[r * d + c] != 0.0)
            {
                const double f = m[r * d + c];
                for (int k = 0; k < d; k++)
                    m[r * d + k] -= f * m[c * d + k], inv[r * d + k] -= f * inv[c * d + k];
            }
    }
    memcpy(a, inv.data(), sizeof(double) * d * d);
}

// Greedy aggregation over an adjacency list (ascending neighbours, self excluded).
std::vector<int> aggregate(int n, const std::vector<int64_t> &aptr, const std::vector<int> &adj, int *nagg)
{
    std::vector<int> agg(n, -1);
    int na = 0;
    for (int i = 0; i < n; i++)
    {
        if (agg[i] >= 0)
            continue;
        bool free_ = true;
        for (int64_t k = aptr[i]; k < aptr[i + 1] && free_; k++)
            free_ = agg[adj[k]] < 0;
        if (!free_)
            continue;
        agg[i] = na;
        for (int64_t k = aptr[i]; k < aptr[i + 1]; k++)
            agg[adj[k]] = na;
        na++;
    }
    // leftovers join the first aggregated neighbour's aggregate (of the first pass)
    std::vector<int> a1 = agg;
    for (int i = 0; i < n; i++)
        if (a1[i] < 0)
            for (int64_t k = aptr[i]; k < aptr[i + 1]; k++)
                if (a1[adj[k]] >= 0)
                {
                    agg[i] = a1[adj[k]];
                    break;
                }
    // isolated remainder: new aggregates with their unaggregated neighbours
    for (int i = 0; i < n; i++)
        if (agg[i] < 0)
        {
            agg[i] = na;
            for (int64_t k = aptr[i]; k < aptr[i + 1]; k++)
                if (agg[adj[k]] < 0)
                    agg[adj[k]] = na;
            na++;
        }
    *nagg = na;
    return agg;
}

// Graph of the aggregates: a and b are adjacent when a member of a has a neighbour in
// b (ascending, self excluded).  It is the aggregation graph of the next level.
void quotient_graph(const std::vector<int> &agg, int na, std::vector<int64_t> &aptr, std::vector<int> &adj)
{
    const int n = (int)agg.size();
    std::vector<std::vector<int>> members(na);
    for (int i = 0; i < n; i++)
        members[agg[i]].push_back(i);
    std::vector<int64_t> qptr(na + 1, 0);
    std::vector<int> qadj, mark(na, -1), nb;
    for (int a = 0; a < na; a++)
    {
        nb.clear();
        for (int i : members[a])
            for (int64_t k = aptr[i]; k < aptr[i + 1]; k++)
            {
                const int b = agg[adj[k]];
                if (b != a && mark[b] != a)
                {
                    mark[b] = a;
                    nb.push_back(b);
                }
            }
        std::sort(nb.begin(), nb.end());
        qadj.insert(qadj.end(), nb.begin(), nb.end());
        qptr[a + 1] = (int64_t)qadj.size();
    }
    aptr.swap(qptr);
    adj.swap(qadj);
}
} // namespace

// ---------------------------------------------------------------- device side ----
// A level operator on the device, in one of two layouts:
//   SLICE: sliced ELL, 64 rows per slice, slot-major within a slice (the layout of
//          the model's tangent): entry e of lane l's block in slot k at
//          val[(k*BR*BC + e)*64 + l]; one thread per row.  For matrices with many
//          rows (enough threads to fill the device).
//   WROW:  block CSR, one wavefront per row: entry e of the row's j-th block at
//          val[rptr[i]*BR*BC + e*len + j] (len = row length), so the row's lanes read
//          consecutive blocks of one entry in one access.  For matrices with few,
//          long rows (the coarse Galerkin operators and the restrictions).
enum
{
    LAYOUT_SLICE = 0,
    LAYOUT_WROW = 1
};
struct Mat
{
    int layout = LAYOUT_SLICE;
    int nr = 0, nc = 0, br = 0, bc = 0, nslices = 0;
    int64_t nslots = 0, nnzb = 0;
    int *sptr = nullptr, *col = nullptr;
    int *perm = nullptr; // SLICE: the row at each position (null: position = row)
    int64_t *rptr = nullptr;
    double *val = nullptr;
    float *val32 = nullptr;
};

struct Level
{
    int n = 0, d = 0;   // block rows and block size
    Mat A;              // level matrix (level 0: the model's tangent, not owned)
    double *Minv = nullptr;
    float *Minv32 = nullptr;
    Mat P, R;           // to and from the next level
    double *b = nullptr, *x = nullptr, *r = nullptr, *dd = nullptr;
    float *b32 = nullptr, *x32 = nullptr, *r32 = nullptr, *d32 = nullptr;
    double lmax = 2.0;
};

static struct Amg
{
    bool ready = false;
    int nlev = 0;           // levels with a smoother; the last one is solved densely
    std::vector<Level> L;
    int nc_dense = 0;       // DOFs of the dense coarsest level
    double *Ainv = nullptr; // dense inverse
    float *Ainv32 = nullptr;
    std::vector<void *> mem;
    int64_t n_model = -1;
    uint64_t bc_hash = 0;
    double t_setup = 0.0;
    int degree = 1;         // Chebyshev degree on the finest level
    int cdegree = 1;        // ... and on the coarser levels
    double ratio = 10.0;
    int post = 1;           // post-smoothing (0: V(1,0) cycles, not symmetric)
    int64_t bytes = 0;      // device memory of the hierarchy
} AM;

// rows from which one thread per row fills the device (256 CUs x 1024 threads)
static const int SLICE_MIN_ROWS = 262144;

// LPMX_AMG_SLICE_MIN=<rows> (a test hook): the sliced layout from that many rows on,
// with the rows always sorted, so that small models exercise it
static int slice_min_override(void)
{
    const char *e = getenv("LPMX_AMG_SLICE_MIN");
    return e ? atoi(e) : 0;
}

// Copies between the host and the model's data, ordered on the library's stream (a
// non-blocking stream: a copy on the null stream would not wait for the kernels
// queued on it, e.g. the constraint elimination right before a solve).
static int d2h(void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return LPMX_OK;
}

static int h2d(void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return LPMX_OK;
}

template <typename T>
static int amg_alloc(T **p, int64_t count)
{
    const size_t by = sizeof(T) * (size_t)std::max<int64_t>(count, 1);
    HIPCHK(hipMalloc((void **)p, by));
    AM.mem.push_back(*p);
    AM.bytes += (int64_t)by;
    return LPMX_OK;
}

void lpmx_amg_free(void)
{
    for (void *p : AM.mem)
        hipFree(p);
    const int deg = AM.degree, cdeg = AM.cdegree, post = AM.post;
    const double ratio = AM.ratio;
    AM = Amg();
    AM.degree = deg, AM.cdegree = cdeg, AM.ratio = ratio, AM.post = post;
}


static int alloc_vectors(Level &v)
{
    const int64_t m = (int64_t)v.n * v.d;
    RETCHK(amg_alloc(&v.b, m));
    RETCHK(amg_alloc(&v.x, m));
    RETCHK(amg_alloc(&v.r, m));
    RETCHK(amg_alloc(&v.dd, m));
    RETCHK(amg_alloc(&v.b32, m));
    RETCHK(amg_alloc(&v.x32, m));
    RETCHK(amg_alloc(&v.r32, m));
    RETCHK(amg_alloc(&v.d32, m));
    return LPMX_OK;
}

// A fingerprint of the constraint set: the sum, mod 2^64, of the splitmix64 finaliser
// (Steele, Lea and Flood, OOPSLA 2014) of the index of every DOF that is both
// displacement-constrained and fixed.  Integer sums do not depend on their order, so
// per-work-group partials summed on the host give the same value every time, and
// only the partials leave the device.
#define SIG_GRID 256
__device__ __forceinline__ uint64_t sig_mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_mask_sig(int64_t m, const int *__restrict__ db, const int *__restrict__ fx,
                                                   uint64_t *__restrict__ parts)
{
    __shared__ uint64_t sh[256];
    uint64_t s = 0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m; t += (int64_t)gridDim.x * blockDim.x)
        if (db[t] != 0 && fx[t] != 0)
            s += sig_mix((uint64_t)t);
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint64_t b = 0;
        for (int k = 0; k < (int)blockDim.x; k++)
            b += sh[k];
        parts[blockIdx.x] = b;
    }
}

static uint64_t *g_sig_parts = nullptr; // SIG_GRID partials, kept until the process ends

static int mask_signature(uint64_t *sig)
{
    if (!g_sig_parts)
        HIPCHK(hipMalloc((void **)&g_sig_parts, sizeof(uint64_t) * SIG_GRID));
    const int64_t m = (int64_t)g.n * g.dim;
    hipLaunchKernelGGL(k_mask_sig, dim3(SIG_GRID), dim3(256), 0, g_stream, m, g.dispBC, g.fix, g_sig_parts);
    LAUNCHCHK();
    uint64_t parts[SIG_GRID];
    RETCHK(d2h(parts, g_sig_parts, sizeof(parts)));
    uint64_t h = 0;
    for (int b = 0; b < SIG_GRID; b++)
        h += parts[b];
    *sig = h ^ (uint64_t)m;
    return LPMX_OK;
}


// Whether the hierarchy fits the current model and constraint set.
bool lpmx_amg_current(void)
{
    if (!AM.ready || AM.n_model != g.n)
        return false;
    uint64_t sig = 0;
    if (mask_signature(&sig) != LPMX_OK)
        return false;
    return sig == AM.bc_hash;
}

double lpmx_amg_setup_seconds(void) { return AM.t_setup; }
int lpmx_amg_levels(void) { return AM.ready ? AM.nlev + 1 : 0; }


// ------------------------------------------------------------------ kernels ------
// y = A x (MODE 0), y = b - A x (1), y += A x (2), y -= A x (3) for the lane's row.
template <int BR, typename TX, int MODE>
__device__ __forceinline__ void mv_store(int64_t i, const TX *acc, TX *__restrict__ y, const TX *__restrict__ b)
{
#pragma unroll
    for (int r = 0; r < BR; r++)
    {
        const int64_t o = i * BR + r;
        if (MODE == 0)
            y[o] = acc[r];
        else if (MODE == 1)
            y[o] = b[o] - acc[r];
        else if (MODE == 2)
            y[o] += acc[r];
        else
            y[o] -= acc[r];
    }
}

template <int BR, int BC, typename TV, typename TX>
__device__ __forceinline__ void slice_mac(const TV *__restrict__ kb, const TX *__restrict__ x, int m, TX *acc)
{
    TX xm[BC];
#pragma unroll
    for (int c = 0; c < BC; c++)
        xm[c] = x[(int64_t)m * BC + c];
#pragma unroll
    for (int r = 0; r < BR; r++)
#pragma unroll
        for (int c = 0; c < BC; c++)
            acc[r] += (TX)__builtin_nontemporal_load(kb + (r * BC + c) * 64) * xm[c];
}

// SLICE layout, one thread per row; U slots loaded ahead per step; work-groups
// remapped so that each XCD walks a contiguous range of rows.
template <int BR, int BC, typename TV, typename TX, int MODE, int U>
__global__ __launch_bounds__(256) void k_slice_mv(int nr, const int *__restrict__ sptr, const int *__restrict__ col,
                                                  const TV *__restrict__ val, const TX *__restrict__ x,
                                                  TX *__restrict__ y, const TX *__restrict__ b, const int *done,
                                                  const int *__restrict__ perm)
{
    if (done && *done)
        return;
    const int i = xcd_block() * 256 + threadIdx.x; // position in the sliced layout
    if (i >= nr)
        return;
    const int s = i >> 6, l = i & 63;
    const int k1 = sptr[s + 1];
    TX acc[BR];
#pragma unroll
    for (int r = 0; r < BR; r++)
        acc[r] = 0;
    int k = sptr[s];
    for (; k + U <= k1; k += U)
    {
        int m[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            m[u] = __builtin_nontemporal_load(col + (int64_t)(k + u) * 64 + l);
#pragma unroll
        for (int u = 0; u < U; u++)
            slice_mac<BR, BC, TV, TX>(val + (int64_t)(k + u) * BR * BC * 64 + l, x, m[u], acc);
    }
    for (; k < k1; k++)
        slice_mac<BR, BC, TV, TX>(val + (int64_t)k * BR * BC * 64 + l, x,
                                  __builtin_nontemporal_load(col + (int64_t)k * 64 + l), acc);
    mv_store<BR, TX, MODE>(perm ? perm[i] : i, acc, y, b);
}

// WROW layout, one wavefront per row (4 rows per work-group); the lanes' partial
// sums are combined in a fixed butterfly order.  RND = 2 takes two 64-block rounds
// per loop trip with their loads in flight together (the per-lane sums keep the
// order of one round after the other, so the bits are those of RND = 1): faster on
// long rows (level-1 operator, 182 blocks per row at 1M: 6-8%; restriction, 298:
// 11%), slower on short ones, whose extra registers halve the waves per SIMD.
template <int BR, int BC, typename TV, typename TX, int MODE, int RND>
__global__ __launch_bounds__(256) void k_wrow_mv(int nr, const int64_t *__restrict__ rptr, const int *__restrict__ col,
                                                 const TV *__restrict__ val, const TX *__restrict__ x,
                                                 TX *__restrict__ y, const TX *__restrict__ b, const int *done)
{
    if (done && *done)
        return;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= nr) // uniform over the wavefront
        return;
    const int64_t k0 = rptr[i];
    const int len = (int)(rptr[i + 1] - k0);
    const TV *v = val + k0 * (BR * BC);
    TX acc[BR];
#pragma unroll
    for (int r = 0; r < BR; r++)
        acc[r] = 0;
    if (RND == 1)
        for (int j = lane; j < len; j += 64)
        {
            const int m = col[k0 + j];
            TX xm[BC];
#pragma unroll
            for (int c = 0; c < BC; c++)
                xm[c] = x[(int64_t)m * BC + c];
#pragma unroll
            for (int r = 0; r < BR; r++)
#pragma unroll
                for (int c = 0; c < BC; c++)
                    acc[r] += (TX)__builtin_nontemporal_load(v + (int64_t)(r * BC + c) * len + j) * xm[c];
        }
    else
        for (int j = lane; j < len; j += 128)
        {
            const bool two = j + 64 < len;
            const int m0 = col[k0 + j], m1 = two ? col[k0 + j + 64] : m0;
            TX x0[BC], x1[BC], v0[BR * BC], v1[BR * BC];
#pragma unroll
            for (int c = 0; c < BC; c++)
                x0[c] = x[(int64_t)m0 * BC + c], x1[c] = x[(int64_t)m1 * BC + c];
#pragma unroll
            for (int e = 0; e < BR * BC; e++)
            {
                v0[e] = (TX)__builtin_nontemporal_load(v + (int64_t)e * len + j);
                v1[e] = two ? (TX)__builtin_nontemporal_load(v + (int64_t)e * len + j + 64) : (TX)0;
            }
#pragma unroll
            for (int r = 0; r < BR; r++)
#pragma unroll
                for (int c = 0; c < BC; c++)
                    acc[r] += v0[r * BC + c] * x0[c];
            if (two)
#pragma unroll
                for (int r = 0; r < BR; r++)
#pragma unroll
                    for (int c = 0; c < BC; c++)
                        acc[r] += v1[r * BC + c] * x1[c];
        }
#pragma unroll
    for (int r = 0; r < BR; r++)
        acc[r] = wave_sum(acc[r]);
    if (lane == 0)
        mv_store<BR, TX, MODE>(i, acc, y, b);
}

// Chebyshev start: d = c M f; x = d (SET) or x += d; r = f (SET only).
template <int D, typename TM, typename TX, bool SET>
__global__ __launch_bounds__(256) void k_cheb_start(int n, const TM *__restrict__ Minv, const TX *__restrict__ f,
                                                    TX *__restrict__ d, TX *__restrict__ x, TX *__restrict__ r,
                                                    double c, const int *done)
{
    if (done && *done)
        return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    TX fv[D];
#pragma unroll
    for (int q = 0; q < D; q++)
        fv[q] = f[(int64_t)i * D + q];
#pragma unroll
    for (int a = 0; a < D; a++)
    {
        TX z = 0;
#pragma unroll
        for (int q = 0; q < D; q++)
            z += (TX)Minv[(int64_t)(a * D + q) * n + i] * fv[q];
        const int64_t o = (int64_t)i * D + a;
        const TX dv = (TX)c * z;
        d[o] = dv;
        if (SET)
        {
            x[o] = dv;
            r[o] = fv[a];
        }
        else
            x[o] += dv;
    }
}

// Chebyshev step: d = c1 d + c2 M r; x += d.
template <int D, typename TM, typename TX>
__global__ __launch_bounds__(256) void k_cheb_step(int n, const TM *__restrict__ Minv, const TX *__restrict__ r,
                                                   TX *__restrict__ d, TX *__restrict__ x, double c1, double c2,
                                                   const int *done)
{
    if (done && *done)
        return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    TX rv[D];
#pragma unroll
    for (int q = 0; q < D; q++)
        rv[q] = r[(int64_t)i * D + q];
#pragma unroll
    for (int a = 0; a < D; a++)
    {
        TX z = 0;
#pragma unroll
        for (int q = 0; q < D; q++)
            z += (TX)Minv[(int64_t)(a * D + q) * n + i] * rv[q];
        const int64_t o = (int64_t)i * D + a;
        const TX dv = (TX)c1 * d[o] + (TX)c2 * z;
        d[o] = dv;
        x[o] += dv;
    }
}

// x = Ainv b (dense, row-major m x m), one wavefront per row.
template <typename TX>
__global__ __launch_bounds__(256) void k_dense_mv(int m, const TX *__restrict__ A, const TX *__restrict__ b,
                                                  TX *__restrict__ x, const int *done)
{
    if (done && *done)
        return;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= m)
        return;
    TX s = 0;
    for (int k = lane; k < m; k += 64)
        s += A[(int64_t)i * m + k] * b[k];
    s = wave_sum(s);
    if (lane == 0)
        x[i] = s;
}

template <int BR, int BC, typename TX, int MODE>
static void launch_mv_t(const Mat &B, const TX *x, TX *y, const TX *b, const int *done)
{
    if (B.layout == LAYOUT_WROW)
    {
        const dim3 G(nblocks(B.nr, 4)), T(256);
        const bool lng = B.nnzb >= 96 * (int64_t)B.nr; // rows of 96 blocks and more on average
        if (sizeof(TX) == 8)
            hipLaunchKernelGGL((k_wrow_mv<BR, BC, double, TX, MODE, 1>), G, T, 0, g_stream, B.nr, B.rptr, B.col,
                               (const double *)B.val, x, y, b, done);
        else if (lng)
            hipLaunchKernelGGL((k_wrow_mv<BR, BC, float, TX, MODE, 2>), G, T, 0, g_stream, B.nr, B.rptr, B.col,
                               (const float *)B.val32, x, y, b, done);
        else
            hipLaunchKernelGGL((k_wrow_mv<BR, BC, float, TX, MODE, 1>), G, T, 0, g_stream, B.nr, B.rptr, B.col,
                               (const float *)B.val32, x, y, b, done);
        return;
    }
    const dim3 G(nblocks(B.nr, 256)), T(256);
    if (sizeof(TX) == 8)
        hipLaunchKernelGGL((k_slice_mv<BR, BC, double, TX, MODE, 1>), G, T, 0, g_stream, B.nr, B.sptr, B.col,
                           (const double *)B.val, x, y, b, done, B.perm);
    else
        hipLaunchKernelGGL((k_slice_mv<BR, BC, float, TX, MODE, 4>), G, T, 0, g_stream, B.nr, B.sptr, B.col,
                           (const float *)B.val32, x, y, b, done, B.perm);
}

template <typename TX, int MODE>
static void launch_mv(const Mat &B, const TX *x, TX *y, const TX *b, const int *done)
{
#define LPMX_MV(BR_, BC_)                                                                                           \
    if (B.br == BR_ && B.bc == BC_)                                                                                 \
        return launch_mv_t<BR_, BC_, TX, MODE>(B, x, y, b, done);
    LPMX_MV(3, 3)
    LPMX_MV(6, 6)
    LPMX_MV(3, 6)
    LPMX_MV(6, 3)
    LPMX_MV(2, 2)
    LPMX_MV(2, 3)
    LPMX_MV(3, 2)
#undef LPMX_MV
}

template <typename TX>
static TX *lv(double *a, float *b)
{
    return sizeof(TX) == 8 ? (TX *)a : (TX *)b;
}

// Level matrix of level l (the model's tangent, SLICE layout, on level 0).
static Mat level_matrix(int l)
{
    if (l > 0)
        return AM.L[l].A;
    Mat B;
    B.layout = LAYOUT_SLICE;
    B.nr = B.nc = g.n, B.br = B.bc = g.dim, B.nslices = g.nslices, B.nslots = g.nslots;
    B.sptr = g.sptr, B.col = g.col, B.val = g.K, B.val32 = g.K32;
    return B;
}

template <typename TX>
static const void *level_minv(int l)
{
    if (l > 0)
        return sizeof(TX) == 8 ? (const void *)AM.L[l].Minv : (const void *)AM.L[l].Minv32;
    return sizeof(TX) == 8 ? (const void *)g.Minv : (const void *)g.Minv32;
}

template <typename TX, bool SET>
static void cheb_start(int d, int n, const void *Minv, const TX *f, TX *dd, TX *x, TX *r, double c, const int *done)
{
    const dim3 G(nblocks(n)), T(LPMX_BLOCK);
    typedef typename std::conditional<sizeof(TX) == 8, double, float>::type TM;
    if (d == 3)
        hipLaunchKernelGGL((k_cheb_start<3, TM, TX, SET>), G, T, 0, g_stream, n, (const TM *)Minv, f, dd, x, r, c, done);
    else if (d == 6)
        hipLaunchKernelGGL((k_cheb_start<6, TM, TX, SET>), G, T, 0, g_stream, n, (const TM *)Minv, f, dd, x, r, c, done);
    else
        hipLaunchKernelGGL((k_cheb_start<2, TM, TX, SET>), G, T, 0, g_stream, n, (const TM *)Minv, f, dd, x, r, c, done);
}

template <typename TX>
static void cheb_step(int d, int n, const void *Minv, const TX *r, TX *dd, TX *x, double c1, double c2, const int *done)
{
    const dim3 G(nblocks(n)), T(LPMX_BLOCK);
    typedef typename std::conditional<sizeof(TX) == 8, double, float>::type TM;
    if (d == 3)
        hipLaunchKernelGGL((k_cheb_step<3, TM, TX>), G, T, 0, g_stream, n, (const TM *)Minv, r, dd, x, c1, c2, done);
    else if (d == 6)
        hipLaunchKernelGGL((k_cheb_step<6, TM, TX>), G, T, 0, g_stream, n, (const TM *)Minv, r, dd, x, c1, c2, done);
    else
        hipLaunchKernelGGL((k_cheb_step<2, TM, TX>), G, T, 0, g_stream, n, (const TM *)Minv, r, dd, x, c1, c2, done);
}

// ---------------------------------------------------------------- measuring ----
static struct AmgProf
{
    bool on = false, have_ev = false;
    enum { MAXEV = 128 };
    hipEvent_t ev[2 * MAXEV];
    int tag[MAXEV];
    int used = 0, cur = -1;
    double ms = 0.0;
    int64_t n = 0;
} AP;

void lpmx_amg_prof_enable(bool on)
{
    AP.on = on, AP.used = 0, AP.cur = -1, AP.ms = 0.0, AP.n = 0;
}

void lpmx_amg_prof_tag(int k) { AP.cur = k; }

int lpmx_amg_prof_flush(int nreal)
{
    if (AP.on)
        for (int u = 0; u < AP.used; u++)
            if (AP.tag[u] >= 0 && AP.tag[u] < nreal)
            {
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, AP.ev[2 * u], AP.ev[2 * u + 1]));
                AP.ms += ms, AP.n++;
            }
    AP.used = 0, AP.cur = -1;
    return LPMX_OK;
}

// A finest-level SpMV of the smoother, timed when profiling is on.
template <typename TX, int MODE>
static void fine_mv(const Mat &A, const TX *x, TX *y, const TX *b, const int *done)
{
    const bool timed = AP.on && AP.cur >= 0 && AP.used < AmgProf::MAXEV;
    if (timed)
    {
        if (!AP.have_ev)
        {
            for (int e = 0; e < 2 * AmgProf::MAXEV; e++)
                hipEventCreate(&AP.ev[e]);
            AP.have_ev = true;
        }
        hipEventRecord(AP.ev[2 * AP.used], g_stream);
    }
    launch_mv<TX, MODE>(A, x, y, b, done);
    if (timed)
    {
        hipEventRecord(AP.ev[2 * AP.used + 1], g_stream);
        AP.tag[AP.used++] = AP.cur;
    }
}

extern "C" void lpmx_profile_amg_read(double *total_ms, int64_t *launches)
{
    *total_ms = AP.ms;
    *launches = AP.n;
}

// Algorithmic HBM bytes of one finest-level smoother SpMV (y = b - K x or y -= K x):
// every stored block and its column index once, the slice pointers, x read once (its
// gathers are re-reads from cache), and two vector streams (b or y read, y written).
extern "C" int lpmx_amg_smoother_bytes(int precision, double *bytes)
{
    RETCHK(lpmx_require_pattern());
    const double sv = precision == LPMX_MIXED ? 4.0 : 8.0;
    const int D = g.dim;
    *bytes = (double)g.nnzb * (D * D * sv + 4) + 4.0 * (g.nslices + 1) + 3.0 * g.n * D * sv;
    return LPMX_OK;
}

extern "C" int lpmx_amg_info(int *levels, double *setup_seconds, int64_t *device_bytes)
{
    *levels = AM.ready ? AM.nlev + 1 : 0;
    *setup_seconds = AM.t_setup;
    *device_bytes = AM.bytes;
    return LPMX_OK;
}

// Chebyshev smoothing of A x = f on level l (Saad, Iterative Methods for Sparse
// Linear Systems, alg. 12.1): `degree` corrections, from x = 0 (SET) or from the
// current x.  With SET (the pre-smoother) r holds f - A x on return, for the
// restriction; the post-smoother leaves r stale (nothing reads it).
template <typename TX, bool SET>
static void smooth(int l, const TX *f, TX *x, TX *r, TX *dd, const int *done)
{
    const Level &v = AM.L[l];
    const Mat A = level_matrix(l);
    const void *Minv = level_minv<TX>(l);
    const double hi = 1.15 * v.lmax, lo = hi / AM.ratio;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    auto mv = [&](auto mode, const TX *xin, TX *y, const TX *b) {
        if (l == 0)
            fine_mv<TX, decltype(mode)::value>(A, xin, y, b, done);
        else
            launch_mv<TX, decltype(mode)::value>(A, xin, y, b, done);
    };
    if (SET)
        cheb_start<TX, true>(v.d, v.n, Minv, f, dd, x, r, 1.0 / theta, done);
    else
    {
        mv(std::integral_constant<int, 1>(), x, r, f); // r = f - A x
        cheb_start<TX, false>(v.d, v.n, Minv, r, dd, x, nullptr, 1.0 / theta, done);
    }
    const int degree = l == 0 ? AM.degree : AM.cdegree;
    for (int k = 1; k < degree; k++)
    {
        mv(std::integral_constant<int, 3>(), dd, r, nullptr); // r -= A d
        const double rn = 1.0 / (2.0 * sigma - rho);
        cheb_step<TX>(v.d, v.n, Minv, r, dd, x, rn * rho, 2.0 * rn / delta, done);
        rho = rn;
    }
    if (SET)
        mv(std::integral_constant<int, 3>(), dd, r, nullptr); // r = f - A x
}

template <typename TX>
static void vcycle(int l, const TX *f, TX *x, const int *done)
{
    Level &v = AM.L[l];
    if (l == AM.nlev)
    {
        const TX *Ai = sizeof(TX) == 8 ? (const TX *)AM.Ainv : (const TX *)AM.Ainv32;
        hipLaunchKernelGGL(k_dense_mv<TX>, dim3(nblocks(AM.nc_dense, 4)), dim3(256), 0, g_stream, AM.nc_dense, Ai, f,
                           x, done);
        return;
    }
    TX *r = lv<TX>(v.r, v.r32), *dd = lv<TX>(v.dd, v.d32);
    smooth<TX, true>(l, f, x, r, dd, done);
    Level &c = AM.L[l + 1];
    TX *cb = lv<TX>(c.b, c.b32), *cx = lv<TX>(c.x, c.x32);
    launch_mv<TX, 0>(v.R, r, cb, nullptr, done);
    vcycle<TX>(l + 1, cb, cx, done);
    launch_mv<TX, 2>(v.P, cx, x, nullptr, done);
    if (AM.post)
        smooth<TX, false>(l, f, x, r, dd, done);
}

// z = B r with the V-cycle (r, z: fine DOF vectors); kernels return at once while
// *done is set.
int lpmx_amg_apply_f64(const double *r, double *z, const int *done)
{
    if (!AM.ready)
        return lpmx_fail(LPMX_ERR_STATE, "AMG hierarchy not built");
    vcycle<double>(0, r, z, done);
    LAUNCHCHK();
    return LPMX_OK;
}

int lpmx_amg_apply_f32(const float *r, float *z, const int *done)
{
    if (!AM.ready)
        return lpmx_fail(LPMX_ERR_STATE, "AMG hierarchy not built");
    vcycle<float>(0, r, z, done);
    LAUNCHCHK();
    return LPMX_OK;
}

// ------------------------------------------------------------------- set-up -----
int lpmx_bj_inverse(void);

namespace
{
// Device temporaries of the set-up, released on every return.
struct Scratch
{
    std::vector<void *> p;
    template <typename T>
    int alloc(T **q, int64_t count)
    {
        HIPCHK(hipMalloc((void **)q, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
        p.push_back(*q);
        return LPMX_OK;
    }
    void release(void *q)
    {
        for (void *&x : p)
            if (q && x == q)
            {
                hipFree(x);
                x = nullptr;
            }
    }
    ~Scratch()
    {
        for (void *q : p)
            if (q)
                hipFree(q);
    }
};

struct OwnedBsr
{
    DBsr m;
    ~OwnedBsr() { amgd_free(m); }
};
} // namespace

// A level operator from block CSR into its layout, FP64 and FP32 values.
// Row order of a sliced layout: within windows of SLICE_WINDOW rows, rows by
// descending length (ties in row order), so that a slice pads to rows of about its
// own length; kept only when it saves at least 5% of the slots.  The SpMV writes row
// perm[p] from position p and sums each row's blocks in their own order, so the
// results are those of the unsorted layout.  Prolongator at 1M: 8-19 blocks per row,
// 23% padding unsorted.
#define SLICE_WINDOW 4096
static int slice_order(const DBsr &H, int **perm)
{
    *perm = nullptr;
    std::vector<int> ptr((size_t)H.nr + 1);
    RETCHK(d2h(ptr.data(), H.ptr, sizeof(int) * ptr.size()));
    std::vector<int> ord(H.nr);
    int64_t plain = 0, sorted = 0;
    for (int w0 = 0; w0 < H.nr; w0 += SLICE_WINDOW)
    {
        const int w1 = std::min(H.nr, w0 + SLICE_WINDOW);
        for (int i = w0; i < w1; i++)
            ord[i] = i;
        std::stable_sort(ord.begin() + w0, ord.begin() + w1,
                         [&](int a, int b) { return ptr[a + 1] - ptr[a] > ptr[b + 1] - ptr[b]; });
    }
    for (int s0 = 0; s0 < H.nr; s0 += 64)
    {
        int mp = 0, ms = 0;
        for (int p = s0; p < std::min(H.nr, s0 + 64); p++)
            mp = std::max(mp, ptr[p + 1] - ptr[p]), ms = std::max(ms, ptr[ord[p] + 1] - ptr[ord[p]]);
        plain += mp, sorted += ms;
    }
    if (sorted > plain - plain / 20 && slice_min_override() <= 0)
        return LPMX_OK;
    RETCHK(amg_alloc(perm, (int64_t)H.nr));
    HIPCHK(hipMemcpyAsync(*perm, ord.data(), sizeof(int) * (size_t)H.nr, hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return LPMX_OK;
}

static int to_mat(const DBsr &H, Mat &B)
{
    B.nr = H.nr, B.nc = H.nc, B.br = H.br, B.bc = H.bc, B.nnzb = H.nnzb;
    const int64_t bs = (int64_t)H.br * H.bc;
    int64_t nval = 0;
    // one thread per row (SLICE) for many rows, or for a fair number of short rows (a
    // wavefront on a row of a few blocks leaves most lanes idle); a wavefront per row
    // (WROW) for long rows
    const int smin = slice_min_override();
    const bool slice = (smin > 0 && H.nr >= smin) || H.nr >= SLICE_MIN_ROWS ||
                       (H.nr >= 65536 && (int64_t)H.nnzb <= 48LL * H.nr);
    if (!slice)
    {
        B.layout = LAYOUT_WROW;
        nval = (int64_t)H.nnzb * bs;
        RETCHK(amg_alloc(&B.rptr, (int64_t)H.nr + 1));
        RETCHK(amg_alloc(&B.col, (int64_t)H.nnzb));
        RETCHK(amg_alloc(&B.val, nval));
        RETCHK(amgd_rptr64(H, B.rptr));
        HIPCHK(hipMemcpyAsync(B.col, H.idx, sizeof(int) * (size_t)H.nnzb, hipMemcpyDeviceToDevice, g_stream));
        RETCHK(amgd_wrow_vals(H, B.val));
    }
    else
    {
        B.layout = LAYOUT_SLICE;
        B.nslices = (H.nr + 63) / 64;
        RETCHK(amg_alloc(&B.sptr, (int64_t)B.nslices + 1));
        RETCHK(slice_order(H, &B.perm));
        int ns = 0;
        RETCHK(amgd_slice_widths(H, B.sptr, &ns, B.perm));
        B.nslots = ns;
        nval = B.nslots * bs * 64;
        RETCHK(amg_alloc(&B.col, B.nslots * 64));
        RETCHK(amg_alloc(&B.val, nval));
        RETCHK(amgd_slice_fill(H, B.sptr, B.nslots, B.col, B.val, B.perm));
    }
    RETCHK(amg_alloc(&B.val32, nval));
    return amgd_to_f32(B.val, B.val32, nval);
}

// lambda_max(D^-1 A) of level l: 20 power-iteration steps from a fixed pseudo-random
// start (xorshift64); D^-1 A is similar to a symmetric matrix.
static int level_lambda_max(int l, const double *Minv, double *lam)
{
    const Level &v = AM.L[l];
    const int64_t m = (int64_t)v.n * v.d;
    std::vector<double> x0(m);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (int64_t i = 0; i < m; i++)
    {
        s ^= s << 13, s ^= s >> 7, s ^= s << 17;
        x0[i] = (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    Scratch S;
    double *x = nullptr, *y = nullptr, *z = nullptr;
    RETCHK(S.alloc(&x, m));
    RETCHK(S.alloc(&y, m));
    RETCHK(S.alloc(&z, m));
    RETCHK(h2d(x, x0.data(), sizeof(double) * m));
    const Mat A = level_matrix(l);
    double val = 0.0;
    for (int it = 0; it < 20; it++)
    {
        double nx2 = 0.0;
        RETCHK(lpmx_dev_sumsq(x, m, &nx2));
        const double nx = sqrt(nx2);
        if (!(nx > 0))
        {
            val = 1.0;
            break;
        }
        RETCHK(amgd_scale(x, m, nx));
        launch_mv<double, 0>(A, x, y, nullptr, nullptr);
        LAUNCHCHK();
        RETCHK(amgd_block_diag_mv(v.n, v.d, Minv, y, z));
        double nz2 = 0.0;
        RETCHK(lpmx_dev_sumsq(z, m, &nz2));
        val = sqrt(nz2);
        std::swap(x, z);
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    *lam = val;
    return LPMX_OK;
}

// The coarsest level as a dense row-major m x m matrix: from its block CSR, or (a
// one-level hierarchy) from the model's tangent.
static int dense_level(const DBsr *H, int m, std::vector<double> &M)
{
    M.assign((size_t)m * m, 0.0);
    if (H)
    {
        const int bs = H->br * H->bc;
        std::vector<int> ptr(H->nr + 1), idx(H->nnzb);
        std::vector<double> val((size_t)H->nnzb * bs);
        RETCHK(d2h(ptr.data(), H->ptr, sizeof(int) * ptr.size()));
        RETCHK(d2h(idx.data(), H->idx, sizeof(int) * idx.size()));
        RETCHK(d2h(val.data(), H->val, sizeof(double) * val.size()));
        for (int i = 0; i < H->nr; i++)
            for (int k = ptr[i]; k < ptr[i + 1]; k++)
                for (int r = 0; r < H->br; r++)
                    for (int c = 0; c < H->bc; c++)
                        M[(size_t)(i * H->br + r) * m + idx[k] * H->bc + c] = val[(size_t)k * bs + r * H->bc + c];
        return LPMX_OK;
    }
    const int n = g.n, D = g.dim, D2 = D * D;
    std::vector<int> sptr(g.nslices + 1), col((size_t)g.nslots * 64), len(n);
    std::vector<double> K((size_t)g.nslots * D2 * 64);
    RETCHK(d2h(sptr.data(), g.sptr, sizeof(int) * sptr.size()));
    RETCHK(d2h(col.data(), g.col, sizeof(int) * col.size()));
    RETCHK(d2h(K.data(), g.K, sizeof(double) * K.size()));
    RETCHK(d2h(len.data(), g.nb_conn, sizeof(int) * n));
    for (int i = 0; i < n; i++)
    {
        const int s = i / 64, l = i % 64;
        for (int k = 0; k < len[i]; k++)
        {
            const int64_t slot = sptr[s] + k;
            const int c = col[slot * 64 + l];
            for (int r = 0; r < D; r++)
                for (int q = 0; q < D; q++)
                    M[(size_t)(i * D + r) * m + c * D + q] = K[(slot * D2 + r * D + q) * 64 + l];
        }
    }
    return LPMX_OK;
}

// Build the hierarchy from the current (constraint-eliminated) tangent.
int lpmx_amg_build(void)
{
    RETCHK(lpmx_require_pattern());
    RETCHK(lpmx_bj_inverse());
    const auto t0 = std::chrono::steady_clock::now();
    lpmx_amg_free();
    if (const char *e = getenv("LPMX_AMG_DEGREE"))
        AM.degree = std::max(1, std::min(6, atoi(e)));
    if (const char *e = getenv("LPMX_AMG_CDEGREE"))
        AM.cdegree = std::max(1, std::min(6, atoi(e)));
    if (const char *e = getenv("LPMX_AMG_POST"))
        AM.post = atoi(e) != 0;
    if (const char *e = getenv("LPMX_AMG_RATIO"))
        AM.ratio = std::max(2.0, atof(e));
    const int n = g.n, D = g.dim, nm = D == 3 ? 6 : 3;
    const bool progress = getenv("LPMX_PROGRESS") != nullptr;
    auto phase = [&](const char *what) {
        if (progress)
            fprintf(stderr, "[lpmx] AMG setup: %s at %.3f s\n", what,
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    };
    // the bond graph: the aggregation graph of the finest level
    std::vector<int64_t> aptr;
    std::vector<int> adj;
    {
        int *dp = nullptr, *da = nullptr, nadj = 0;
        RETCHK(amgd_bond_graph(&dp, &da, &nadj));
        std::vector<int> p32(n + 1);
        adj.resize(nadj);
        int rc = d2h(p32.data(), dp, sizeof(int) * p32.size());
        if (rc == LPMX_OK)
            rc = d2h(adj.data(), da, sizeof(int) * adj.size());
        hipFree(dp);
        hipFree(da);
        RETCHK(rc);
        aptr.assign(p32.begin(), p32.end());
    }
    phase("bond graph");
    const int max_levels = 8, dense_max = 1200;
    std::vector<int64_t> nnzb;
    OwnedBsr Acur;            // the operator of the current level (levels > 0)
    Scratch S;
    double *Bnull = nullptr;  // null space of the current level, [rows*d][nm]
    AM.L.reserve(max_levels);
    int lev = 0;
    for (;; lev++)
    {
        AM.L.emplace_back();
        Level &v = AM.L[lev];
        v.n = lev ? Acur.m.nr : n;
        v.d = lev ? Acur.m.br : D;
        nnzb.push_back(lev ? (int64_t)Acur.m.nnzb : g.nnzb);
        if ((int64_t)v.n * v.d <= dense_max || lev == max_levels - 1)
            break;
        const double *Minv = g.Minv; // the finest level smooths with the model's inverses
        if (lev > 0)
        {
            RETCHK(to_mat(Acur.m, v.A));
            const int64_t bs = (int64_t)v.d * v.d;
            RETCHK(amg_alloc(&v.Minv, bs * v.n));
            RETCHK(amg_alloc(&v.Minv32, bs * v.n));
            RETCHK(amgd_diag_inverses(Acur.m, v.Minv, v.Minv32));
            Minv = v.Minv;
        }
        RETCHK(level_lambda_max(lev, Minv, &v.lmax));
        phase("diagonal inverses and spectral bound");
        int na = 0;
        const std::vector<int> agg = aggregate(v.n, aptr, adj, &na);
        if (na >= v.n) // no coarsening possible: this level is solved densely
            break;
        // members of every aggregate, ascending
        std::vector<int> mptr(na + 1, 0), mem(v.n);
        for (int i = 0; i < v.n; i++)
            mptr[agg[i] + 1]++;
        for (int a = 0; a < na; a++)
            mptr[a + 1] += mptr[a];
        {
            std::vector<int> pos(mptr.begin(), mptr.end() - 1);
            for (int i = 0; i < v.n; i++)
                mem[pos[agg[i]]++] = i;
        }
        Scratch T;
        int *dagg = nullptr, *dmptr = nullptr, *dmem = nullptr;
        RETCHK(T.alloc(&dagg, v.n));
        RETCHK(T.alloc(&dmptr, (int64_t)na + 1));
        RETCHK(T.alloc(&dmem, v.n));
        RETCHK(h2d(dagg, agg.data(), sizeof(int) * agg.size()));
        RETCHK(h2d(dmptr, mptr.data(), sizeof(int) * mptr.size()));
        RETCHK(h2d(dmem, mem.data(), sizeof(int) * mem.size()));
        if (lev == 0)
        {
            RETCHK(S.alloc(&Bnull, (int64_t)n * D * nm));
            RETCHK(amgd_null_space(n, D, nm, na, dagg, dmptr, dmem, Bnull));
        }
        double *Bc = nullptr;
        RETCHK(S.alloc(&Bc, (int64_t)na * nm * nm));
        OwnedBsr P0, P, R, AP, Ac;
        RETCHK(amgd_tentative(v.n, na, v.d, nm, dagg, dmptr, dmem, Bnull, &P0.m, Bc));
        HIPCHK(hipStreamSynchronize(g_stream));
        S.release(Bnull);
        Bnull = Bc;
        phase("aggregates and tentative prolongator");
        const AmgLeft Al = lev ? amgd_left(Acur.m) : amgd_left_tangent();
        RETCHK(amgd_spgemm(Al, P0.m, &P.m));
        RETCHK(amgd_smooth(P.m, Minv, dagg, P0.m, 4.0 / (3.0 * v.lmax)));
        HIPCHK(hipStreamSynchronize(g_stream));
        amgd_free(P0.m);
        phase("smoothed prolongator");
        RETCHK(amgd_transpose(P.m, &R.m));
        RETCHK(amgd_spgemm(Al, P.m, &AP.m));
        HIPCHK(hipStreamSynchronize(g_stream));
        phase("A P");
        RETCHK(amgd_spgemm(amgd_left(R.m), AP.m, &Ac.m));
        RETCHK(to_mat(P.m, v.P));
        RETCHK(to_mat(R.m, v.R));
        RETCHK(alloc_vectors(v));
        HIPCHK(hipStreamSynchronize(g_stream));
        amgd_free(Acur.m);
        Acur.m = Ac.m;
        Ac.m = DBsr();
        quotient_graph(agg, na, aptr, adj);
        if (progress)
            fprintf(stderr, "[lpmx] AMG setup: level %d -> %d block rows of %d\n", lev, Acur.m.nr, Acur.m.br);
        phase("Galerkin product done");
    }
    AM.nlev = lev;
    phase("levels built");
    // dense coarsest level
    {
        Level &v = AM.L[lev];
        const int m = v.n * v.d;
        if (m > 6000)
        {
            lpmx_amg_free();
            return lpmx_fail(LPMX_ERR_STATE, "AMG coarsening stalled at %d DOFs", m);
        }
        std::vector<double> M;
        RETCHK(dense_level(lev ? &Acur.m : nullptr, m, M));
        small_inverse(m, M.data());
        std::vector<float> M32(M.size());
        for (size_t q = 0; q < M.size(); q++)
            M32[q] = (float)M[q];
        AM.nc_dense = m;
        RETCHK(amg_alloc(&AM.Ainv, (int64_t)M.size()));
        RETCHK(amg_alloc(&AM.Ainv32, (int64_t)M.size()));
        RETCHK(h2d(AM.Ainv, M.data(), sizeof(double) * M.size()));
        RETCHK(h2d(AM.Ainv32, M32.data(), sizeof(float) * M32.size()));
        RETCHK(alloc_vectors(v));
    }
    AM.n_model = n;
    RETCHK(mask_signature(&AM.bc_hash));
    AM.ready = true;
    AM.t_setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (progress)
    {
        fprintf(stderr, "[lpmx] AMG setup %.3f s:", AM.t_setup);
        for (int l = 0; l <= AM.nlev; l++)
            fprintf(stderr, " L%d %d x %d (%lld blocks, lmax %.3f)", l, AM.L[l].n, AM.L[l].d, (long long)nnzb[l],
                    AM.L[l].lmax);
        fprintf(stderr, "\n");
    }
    return LPMX_OK;
}

// Test access to the stored blocks of a level (block COO, rows ascending).
extern "C" int lpmx_amg_export(int level, int which, int64_t cap, int *rows, int *cols, double *vals, int *dims)
{
    if (!AM.ready)
        return lpmx_fail(LPMX_ERR_STATE, "AMG hierarchy not built");
    if (level < 0 || level >= AM.nlev || which < 0 || which > 2 || (which == 0 && level == 0) || !dims)
        return lpmx_fail(LPMX_ERR_ARG, "no AMG matrix %d on level %d (levels 0..%d)", which, level, AM.nlev - 1);
    const Mat &B = which == 0 ? AM.L[level].A : which == 1 ? AM.L[level].P : AM.L[level].R;
    const int bs = B.br * B.bc;
    std::vector<int> ri, ci;
    std::vector<double> vv;
    if (B.layout == LAYOUT_WROW)
    {
        std::vector<int64_t> rptr(B.nr + 1);
        std::vector<int> col(B.nnzb);
        std::vector<double> val((size_t)B.nnzb * bs);
        RETCHK(d2h(rptr.data(), B.rptr, sizeof(int64_t) * rptr.size()));
        RETCHK(d2h(col.data(), B.col, sizeof(int) * col.size()));
        RETCHK(d2h(val.data(), B.val, sizeof(double) * val.size()));
        for (int i = 0; i < B.nr; i++)
        {
            const int64_t k0 = rptr[i], len = rptr[i + 1] - k0;
            for (int64_t j = 0; j < len; j++)
            {
                ri.push_back(i), ci.push_back(col[k0 + j]);
                for (int e = 0; e < bs; e++)
                    vv.push_back(val[(size_t)(k0 * bs + e * len + j)]);
            }
        }
    }
    else
    {
        std::vector<int> sptr(B.nslices + 1), col((size_t)B.nslots * 64), perm(B.perm ? B.nr : 0);
        std::vector<double> val((size_t)B.nslots * bs * 64);
        RETCHK(d2h(sptr.data(), B.sptr, sizeof(int) * sptr.size()));
        RETCHK(d2h(col.data(), B.col, sizeof(int) * col.size()));
        RETCHK(d2h(val.data(), B.val, sizeof(double) * val.size()));
        if (B.perm)
            RETCHK(d2h(perm.data(), B.perm, sizeof(int) * perm.size()));
        for (int i = 0; i < B.nr; i++) // position i holds row perm[i]
        {
            const int s = i / 64, l = i % 64;
            for (int64_t slot = sptr[s]; slot < sptr[s + 1]; slot++)
            {
                ri.push_back(B.perm ? perm[i] : i), ci.push_back(col[slot * 64 + l]);
                for (int e = 0; e < bs; e++)
                    vv.push_back(val[(slot * bs + e) * 64 + l]);
            }
        }
    }
    const int64_t count = (int64_t)ri.size();
    dims[0] = B.nr, dims[1] = B.nc, dims[2] = B.br, dims[3] = B.bc, dims[4] = (int)count;
    if (!rows)
        return LPMX_OK;
    if (cap < count || !cols || !vals)
        return lpmx_fail(LPMX_ERR_ARG, "export buffers hold %lld blocks, %lld needed", (long long)cap,
                         (long long)count);
    memcpy(rows, ri.data(), sizeof(int) * count);
    memcpy(cols, ci.data(), sizeof(int) * count);
    memcpy(vals, vv.data(), sizeof(double) * vv.size());
    return LPMX_OK;
}
