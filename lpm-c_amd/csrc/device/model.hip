// Device model: lifetime, transfers, neighbour lists (cell list), conn sets and the
// sliced-ELL block pattern of the tangent.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "common.h"

DV g;
hipStream_t g_stream = nullptr;

static char g_err[512] = "";
static bool g_have_model = false, g_have_pattern = false, g_have_nbr = false;
static int g_pattern_version = 0; // bumped by every neighbour or pattern build
int lpmx_pattern_version(void) { return g_pattern_version; }
static double g_lo[3], g_hi[3];
static double *g_scratch = nullptr;
static int64_t g_scratch_bytes = 0;

extern "C" const char *lpmx_last_error(void) { return g_err; }

int lpmx_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int lpmx_hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    return lpmx_fail(e == hipErrorNoDevice || e == hipErrorInvalidDevice ? LPMX_ERR_NOGPU : LPMX_ERR_HIP,
                     "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

int lpmx_require_model(void)
{
    return g_have_model ? LPMX_OK : lpmx_fail(LPMX_ERR_STATE, "no device model (lpmx_model_create first)");
}

int lpmx_require_pattern(void)
{
    RETCHK(lpmx_require_model());
    return g_have_pattern ? LPMX_OK : lpmx_fail(LPMX_ERR_STATE, "no tangent pattern (lpmx_connectivity_build first)");
}

static int init_device(void)
{
    if (g_stream)
        return LPMX_OK;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return lpmx_fail(LPMX_ERR_NOGPU, "no HIP device visible (%s)", hipGetErrorString(e));
    int dev = 0;
    const char *env = getenv("LPMX_DEVICE");
    if (env)
        dev = atoi(env);
    HIPCHK(hipSetDevice(dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return lpmx_fail(LPMX_ERR_NOGPU, "device %d is %s, this library is built for gfx950", dev, prop.gcnArchName);
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    return LPMX_OK;
}

int lpmx_device_init(void) { return init_device(); }

double *lpmx_scratch(int64_t bytes)
{
    if (bytes > g_scratch_bytes)
    {
        if (g_scratch)
            hipFree(g_scratch);
        g_scratch = nullptr;
        if (hipMalloc(&g_scratch, bytes) != hipSuccess)
        {
            g_scratch_bytes = 0;
            return nullptr;
        }
        g_scratch_bytes = bytes;
    }
    return g_scratch;
}

// ---------------------------------------------------------------- allocation -----
static std::vector<void *> g_allocs;

template <typename T>
static int dalloc(T **p, int64_t count, int fill_byte = 0)
{
    *p = nullptr;
    if (count <= 0)
        count = 1;
    HIPCHK(hipMalloc((void **)p, sizeof(T) * count));
    HIPCHK(hipMemsetAsync(*p, fill_byte, sizeof(T) * count, g_stream));
    g_allocs.push_back(*p);
    return LPMX_OK;
}

__global__ void k_fill_d(double *p, int64_t n, double v)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ void k_fill_i(int *p, int64_t n, int v)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

static int fill_d(double *p, int64_t n, double v)
{
    hipLaunchKernelGGL(k_fill_d, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, p, n, v);
    LAUNCHCHK();
    return LPMX_OK;
}

static int fill_i(int *p, int64_t n, int v)
{
    hipLaunchKernelGGL(k_fill_i, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, p, n, v);
    LAUNCHCHK();
    return LPMX_OK;
}

void lpmx_tangent_free(void);

static void free_all(void)
{
    for (void *p : g_allocs)
        hipFree(p);
    g_allocs.clear();
    void *pat[] = {g.sptr, g.col, g.tpos, g.K, g.K32, g.Minv, g.Minv32, g.crit, g.ktype_t, g.ce_t, g.emod};
    for (void *p : pat)
        if (p)
            hipFree(p);
    lpmx_tangent_free();
    lpmx_solver_free();
    lpmx_damage_free();
    lpmx_laws_free();
    lpmx_accel_free();
    lpmx_amg_free();
    lpmx_cp_free();
    lpmx_adjoint_free();
    memset(&g, 0, sizeof(g));
    g_have_model = g_have_pattern = g_have_nbr = false;
}

extern "C" int lpmx_model_destroy(void)
{
    if (g_stream)
        hipStreamSynchronize(g_stream);
    free_all();
    return LPMX_OK;
}

extern "C" int lpmx_model_create(const lpmx_model_desc *d, const double *xyz)
{
    RETCHK(init_device());
    if (!d || !xyz || d->n <= 0 || (d->dim != 2 && d->dim != 3) || d->nneighbors <= 0 ||
        d->nneighbors > LPMX_MAXNB || d->nneighbors_AFEM + 1 > LPMX_MAXCONN)
        return lpmx_fail(LPMX_ERR_ARG, "invalid model description");
    free_all();
    const int n = d->n, nb = d->nneighbors, dim = d->dim;
    g.n = n, g.dim = dim, g.nb = nb, g.nb1 = d->nneighbors1, g.nb2 = d->nneighbors2;
    g.cm = d->nneighbors_AFEM + 1, g.lattice = d->lattice, g.d2 = dim * dim;
    g.radius = d->radius, g.cut1 = d->neighbor1_cutoff, g.cut2 = d->neighbor2_cutoff, g.volume = d->particle_volume;
    g.plmode = 0, g.nbreak = 20, g.H = 0, g.xi = 0, g.dmgL = 0.5, g.dthr = 0.9, g.dA = 0, g.crit0 = 1e-2;
    const int64_t nbn = (int64_t)n * nb;
    RETCHK(dalloc(&g.x, 3 * (int64_t)n));
    RETCHK(dalloc(&g.x0, 3 * (int64_t)n));
    RETCHK(dalloc(&g.xt, 3 * (int64_t)n));
    RETCHK(dalloc(&g.type, n));
    RETCHK(dalloc(&g.nbcur, n));
    RETCHK(dalloc(&g.nb0, n));
    RETCHK(dalloc(&g.pl_flag, n));
    RETCHK(dalloc(&g.nb_conn, n));
    RETCHK(dalloc(&g.kcnt, n));
    RETCHK(dalloc(&g.sigmay, n));
    RETCHK(dalloc(&g.c44, n));
    RETCHK(dalloc(&g.kt, 3 * (int64_t)n));
    int **ib[] = {&g.nbr, &g.nsign, &g.rev, &g.opp};
    for (int **p : ib)
        RETCHK(dalloc(p, nbn, 0xff)); // -1
    double **bd[] = {&g.L0, &g.cs0x, &g.cs0y, &g.cs0z, &g.Kn, &g.Tv, &g.db, &g.dw, &g.D0, &g.D1, &g.dLp0, &g.dLp1,
                     &g.dLp2, &g.dL, &g.dLave, &g.ddLp, &g.ddL, &g.F, &g.Ft, &g.csx, &g.csy, &g.csz, &g.bstress,
                     &g.dist};
    for (double **p : bd)
        RETCHK(dalloc(p, nbn));
    RETCHK(fill_d(g.db, nbn, 1.0));
    RETCHK(fill_d(g.dw, nbn, 1.0));
    double **p2[] = {&g.A, &g.Bt, &g.dA_, &g.dB_};
    for (double **p : p2)
        RETCHK(dalloc(p, 2 * (int64_t)n));
    double **p6[] = {&g.beta0, &g.beta1, &g.beta2, &g.stress};
    for (double **p : p6)
        RETCHK(dalloc(p, 6 * (int64_t)n));
    double **p1[] = {&g.alpha0, &g.alpha1, &g.alpha2, &g.beq0, &g.beq1, &g.beq2, &g.dlam, &g.seq, &g.sm,
                     &g.triax, &g.dnl0, &g.dnl1, &g.dloc0, &g.dloc1, &g.dvis};
    for (double **p : p1)
        RETCHK(dalloc(p, n));
    RETCHK(dalloc(&g.Pin, 3 * (int64_t)n));
    RETCHK(dalloc(&g.Pex, (int64_t)dim * n));
    RETCHK(dalloc(&g.Pext, (int64_t)dim * n));
    RETCHK(dalloc(&g.res, (int64_t)dim * n));
    RETCHK(dalloc(&g.disp, (int64_t)dim * n));
    RETCHK(dalloc(&g.dispBC, (int64_t)dim * n));
    RETCHK(dalloc(&g.fix, (int64_t)dim * n));
    RETCHK(fill_i(g.dispBC, (int64_t)dim * n, 1));
    RETCHK(fill_i(g.fix, (int64_t)dim * n, 1));
    RETCHK(fill_i(g.nbcur, n, -1));
    RETCHK(fill_i(g.nb0, n, -1));
    RETCHK(fill_d(g.sigmay, n, 200.0));
    HIPCHK(hipMemcpyAsync(g.x, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipMemcpyAsync(g.x0, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, g_stream));
    for (int c = 0; c < 3; c++)
        g_lo[c] = 1e300, g_hi[c] = -1e300;
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++)
        {
            g_lo[c] = fmin(g_lo[c], xyz[3 * i + c]);
            g_hi[c] = fmax(g_hi[c], xyz[3 * i + c]);
        }
    HIPCHK(hipStreamSynchronize(g_stream));
    g_have_model = true;
    return LPMX_OK;
}

extern "C" int lpmx_set_params(const lpmx_params *p)
{
    RETCHK(lpmx_require_model());
    if (!p)
        return lpmx_fail(LPMX_ERR_ARG, "null params");
    g.plmode = p->plmode, g.H = p->J2_H, g.xi = p->J2_xi;
    g.dmgL = p->damage_L, g.dthr = p->damage_threshold, g.dA = p->damagec_A, g.crit0 = p->critical_bstrain;
    g.nbreak = p->nbreak;
    g.J2C = p->J2_C;
    return LPMX_OK;
}

// -------------------------------------------------------------------- scans ------
// Exclusive scan of int counts, deterministic, three passes of 1024-wide groups.
__global__ void k_scan_block(const int *in, int *out, int *bsum, int n)
{
    __shared__ int sh[1024];
    int i = blockIdx.x * 1024 + threadIdx.x;
    int v = i < n ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1)
    {
        int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n)
        out[i] = sh[threadIdx.x] - v;
    if (threadIdx.x == 1023)
        bsum[blockIdx.x] = sh[1023];
}

__global__ void k_scan_add(int *out, const int *boff, int n)
{
    int i = blockIdx.x * 1024 + threadIdx.x;
    if (i < n)
        out[i] += boff[blockIdx.x];
}

static int *g_scan_buf = nullptr;
static int64_t g_scan_cap = 0;
static double *g_red_buf = nullptr;

int lpmx_dev_scan_int(const int *in, int *out, int count, int *total)
{
    if (count <= 0)
    {
        if (total)
            *total = 0;
        return LPMX_OK;
    }
    int nb = (count + 1023) / 1024;
    if (2 * (int64_t)(nb + 1) > g_scan_cap)
    {
        if (g_scan_buf)
            hipFree(g_scan_buf);
        g_scan_cap = 2 * (int64_t)(nb + 1);
        HIPCHK(hipMalloc(&g_scan_buf, sizeof(int) * g_scan_cap));
    }
    int *bs = g_scan_buf;
    int *boff = bs + nb + 1;
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(1024), 0, g_stream, in, out, bs, count);
    LAUNCHCHK();
    std::vector<int> h(nb), o(nb + 1);
    HIPCHK(hipMemcpyAsync(h.data(), bs, sizeof(int) * nb, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    int64_t acc = 0;
    for (int k = 0; k < nb; k++)
    {
        o[k] = (int)acc;
        acc += h[k];
    }
    if (acc > 0x7fffffff)
        return lpmx_fail(LPMX_ERR_CAPACITY, "scan total overflows int");
    HIPCHK(hipMemcpyAsync(boff, o.data(), sizeof(int) * nb, hipMemcpyHostToDevice, g_stream));
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, g_stream, out, boff, count);
    LAUNCHCHK();
    if (total)
        *total = (int)acc;
    return LPMX_OK;
}

// ------------------------------------------------------------- reductions --------
__global__ void k_partial_sum(const double *src, int64_t n, int sq, double *parts)
{
    __shared__ double sh[16];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += sq ? src[i] * src[i] : src[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0)
        parts[blockIdx.x] = s;
}

__global__ void k_final_sum(const double *parts, int np, double *out)
{
    __shared__ double sh[16];
    double s = 0.0;
    for (int k = threadIdx.x; k < np; k += blockDim.x)
        s += parts[k];
    s = block_sum(s, sh);
    if (threadIdx.x == 0)
        *out = s;
}

// Sum (sq = 0) or sum of squares (sq = 1) of src[0, n) in a fixed order: LPMX_RED_PARTS
// grid-stride partials, then one work-group over them.  parts: LPMX_RED_PARTS + 8
// device doubles of the caller's.
int lpmx_dev_sum_into(const double *src, int64_t n, int sq, double *parts, double *out_host)
{
    const int np = LPMX_RED_PARTS;
    hipLaunchKernelGGL(k_partial_sum, dim3(np), dim3(LPMX_BLOCK), 0, g_stream, src, n, sq, parts);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(1024), 0, g_stream, parts, np, parts + np);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(out_host, parts + np, sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return LPMX_OK;
}

static int dev_sum_impl(const double *src, int64_t n, int sq, double *out_host)
{
    if (!g_red_buf)
        HIPCHK(hipMalloc(&g_red_buf, sizeof(double) * (LPMX_RED_PARTS + 8)));
    return lpmx_dev_sum_into(src, n, sq, g_red_buf, out_host);
}

int lpmx_dev_sum(const double *src, int64_t n, double *out) { return dev_sum_impl(src, n, 0, out); }
int lpmx_dev_sumsq(const double *src, int64_t n, double *out) { return dev_sum_impl(src, n, 1, out); }

// ---------------------------------------------------------------- transfers ------
// row-major [n][w] (host layout) <-> planes [w][n] (device layout)
template <typename T>
__global__ void k_rows_to_planes(const T *src, T *dst, int n, int w)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * w)
        return;
    int i = (int)(t / w), j = (int)(t % w);
    dst[(int64_t)j * n + i] = src[t];
}

template <typename T>
__global__ void k_planes_to_rows(const T *src, T *dst, int n, int w)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * w)
        return;
    int i = (int)(t / w), j = (int)(t % w);
    dst[t] = src[(int64_t)j * n + i];
}

// conn planes with -1 beyond nb_conn, as the reference's conn rows
__global__ void k_conn_rows(const int *conn, const int *ncn, int *dst, int n, int cm)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * cm)
        return;
    int i = (int)(t / cm), k = (int)(t % cm);
    dst[t] = k < ncn[i] ? conn[(int64_t)k * n + i] : -1;
}

struct FieldDesc
{
    void *ptr;
    int width;  // elements per particle
    int planes; // 1: stored as planes [w][n]; 0: same layout as host
    int isint;
};

static int field_desc(int f, FieldDesc *fd)
{
    const int nb = g.nb, dim = g.dim;
    void *p = nullptr;
    int w = 1, pl = 0, ii = 0;
    switch (f)
    {
    case LPMX_F_XYZ: p = g.x, w = 3; break;
    case LPMX_F_XYZ_INITIAL: p = g.x0, w = 3; break;
    case LPMX_F_XYZ_TEMP: p = g.xt, w = 3; break;
    case LPMX_F_NEIGHBORS: p = g.nbr, w = nb, pl = 1, ii = 1; break;
    case LPMX_F_NSIGN: p = g.nsign, w = nb, pl = 1, ii = 1; break;
    case LPMX_F_NB: p = g.nbcur, ii = 1; break;
    case LPMX_F_NB_INITIAL: p = g.nb0, ii = 1; break;
    case LPMX_F_NB_CONN: p = g.nb_conn, ii = 1; break;
    case LPMX_F_TYPE: p = g.type, ii = 1; break;
    case LPMX_F_PL_FLAG: p = g.pl_flag, ii = 1; break;
    case LPMX_F_DISTANCE_INITIAL: p = g.L0, w = nb, pl = 1; break;
    case LPMX_F_CSX_INITIAL: p = g.cs0x, w = nb, pl = 1; break;
    case LPMX_F_CSY_INITIAL: p = g.cs0y, w = nb, pl = 1; break;
    case LPMX_F_CSZ_INITIAL: p = g.cs0z, w = nb, pl = 1; break;
    case LPMX_F_KN: p = g.Kn, w = nb, pl = 1; break;
    case LPMX_F_TV: p = g.Tv, w = nb, pl = 1; break;
    case LPMX_F_DL: p = g.dL, w = nb, pl = 1; break;
    case LPMX_F_DL_AVE: p = g.dLave, w = nb, pl = 1; break;
    case LPMX_F_DDL: p = g.ddL, w = nb, pl = 1; break;
    case LPMX_F_DDLP: p = g.ddLp, w = nb, pl = 1; break;
    case LPMX_F_F: p = g.F, w = nb, pl = 1; break;
    case LPMX_F_F_TEMP: p = g.Ft, w = nb, pl = 1; break;
    case LPMX_F_CSX: p = g.csx, w = nb, pl = 1; break;
    case LPMX_F_CSY: p = g.csy, w = nb, pl = 1; break;
    case LPMX_F_CSZ: p = g.csz, w = nb, pl = 1; break;
    case LPMX_F_BOND_STRESS: p = g.bstress, w = nb, pl = 1; break;
    case LPMX_F_DISTANCE: p = g.dist, w = nb, pl = 1; break;
    case LPMX_F_DAMAGE_BROKEN: p = g.db, w = nb, pl = 1; break;
    case LPMX_F_DAMAGE_W: p = g.dw, w = nb, pl = 1; break;
    case LPMX_F_DLP0: p = g.dLp0, w = nb, pl = 1; break;
    case LPMX_F_DLP1: p = g.dLp1, w = nb, pl = 1; break;
    case LPMX_F_DLP2: p = g.dLp2, w = nb, pl = 1; break;
    case LPMX_F_DAMAGE_D0: p = g.D0, w = nb, pl = 1; break;
    case LPMX_F_DAMAGE_D1: p = g.D1, w = nb, pl = 1; break;
    case LPMX_F_DL_TOTAL: p = g.A, w = 2, pl = 1; break;
    case LPMX_F_TDL_TOTAL: p = g.Bt, w = 2, pl = 1; break;
    case LPMX_F_DDL_TOTAL: p = g.dA_, w = 2, pl = 1; break;
    case LPMX_F_TDDL_TOTAL: p = g.dB_, w = 2, pl = 1; break;
    case LPMX_F_J2_BETA0: p = g.beta0, w = 6, pl = 1; break;
    case LPMX_F_J2_BETA1: p = g.beta1, w = 6, pl = 1; break;
    case LPMX_F_J2_BETA2: p = g.beta2, w = 6, pl = 1; break;
    case LPMX_F_STRESS: p = g.stress, w = 6, pl = 1; break;
    case LPMX_F_J2_ALPHA0: p = g.alpha0; break;
    case LPMX_F_J2_ALPHA1: p = g.alpha1; break;
    case LPMX_F_J2_ALPHA2: p = g.alpha2; break;
    case LPMX_F_J2_DLAMBDA: p = g.dlam; break;
    case LPMX_F_J2_STRESSEQ: p = g.seq; break;
    case LPMX_F_J2_STRESSM: p = g.sm; break;
    case LPMX_F_J2_TRIAXIALITY: p = g.triax; break;
    case LPMX_F_DAMAGE_NONLOCAL0: p = g.dnl0; break;
    case LPMX_F_DAMAGE_NONLOCAL1: p = g.dnl1; break;
    case LPMX_F_DAMAGE_VISUAL: p = g.dvis; break;
    case LPMX_F_J2_BETA_EQ0: p = g.beq0; break;
    case LPMX_F_J2_BETA_EQ1: p = g.beq1; break;
    case LPMX_F_J2_BETA_EQ2: p = g.beq2; break;
    case LPMX_F_DAMAGE_LOCAL0: p = g.dloc0; break;
    case LPMX_F_DAMAGE_LOCAL1: p = g.dloc1; break;
    case LPMX_F_SIGMAY: p = g.sigmay; break;
    case LPMX_F_PIN: p = g.Pin, w = 3; break;
    case LPMX_F_PEX: p = g.Pex, w = dim; break;
    case LPMX_F_PEX_TEMP: p = g.Pext, w = dim; break;
    case LPMX_F_RESIDUAL: p = g.res, w = dim; break;
    case LPMX_F_DISP: p = g.disp, w = dim; break;
    case LPMX_F_DISPBC_INDEX: p = g.dispBC, w = dim, ii = 1; break;
    case LPMX_F_FIX_INDEX: p = g.fix, w = dim, ii = 1; break;
    case LPMX_F_CP_GY0: p = g.cp_gy0, w = g.nslip, pl = 1; break;
    case LPMX_F_CP_GY1: p = g.cp_gy1, w = g.nslip, pl = 1; break;
    case LPMX_F_CP_GY2: p = g.cp_gy2, w = g.nslip, pl = 1; break;
    case LPMX_F_CP_A_SINGLE0: p = g.cp_As0, w = g.nslip, pl = 1; break;
    case LPMX_F_CP_A_SINGLE1: p = g.cp_As1, w = g.nslip, pl = 1; break;
    case LPMX_F_CP_A_SINGLE2: p = g.cp_As2, w = g.nslip, pl = 1; break;
    case LPMX_F_CP_A0: p = g.cp_A0; break;
    case LPMX_F_CP_A1: p = g.cp_A1; break;
    case LPMX_F_CP_A2: p = g.cp_A2; break;
    case LPMX_F_CP_DA: p = g.cp_dA; break;
    case LPMX_F_CP_DGY: p = g.cp_dgy, w = g.nslip, pl = 1; break;
    case LPMX_F_CP_DA_SINGLE: p = g.cp_dAs, w = g.nslip, pl = 1; break;
    case LPMX_F_CP_RSS: p = g.cp_RSS, w = g.nslip, pl = 1; break;
    case LPMX_F_CP_JACT: p = g.cp_Jact, w = g.nslip, pl = 1, ii = 1; break;
    case LPMX_F_CP_CAB: p = g.cp_Cab, w = g.nslip * g.nslip, pl = 1; break;
    default: return lpmx_fail(LPMX_ERR_ARG, "unknown field %d", f);
    }
    fd->ptr = p, fd->width = w, fd->planes = pl, fd->isint = ii;
    return LPMX_OK;
}

__global__ void k_kptr(const int *kcnt, const int *start, int *dst, int n, int d2c)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n)
        return;
    dst[2 * i] = i < n ? kcnt[i] : 0;
    dst[2 * i + 1] = start[i];
    (void)d2c;
}

__global__ void k_row_len(const int *kcnt, int *len, int n, int d2, int sub)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        len[i] = d2 * kcnt[i] - sub;
    else if (i == n)
        len[i] = 0;
}

// K_pointer[i][1] on the device (n+1 entries), from the conn counts
static int kptr_start(int **out, int *total)
{
    const int n = g.n;
    int *start = nullptr, *len2 = nullptr;
    HIPCHK(hipMalloc(&start, sizeof(int) * ((int64_t)n + 1)));
    HIPCHK(hipMalloc(&len2, sizeof(int) * ((int64_t)n + 1)));
    hipLaunchKernelGGL(k_row_len, dim3(nblocks(n + 1)), dim3(LPMX_BLOCK), 0, g_stream, g.kcnt, len2, n, g.d2,
                       g.dim == 2 ? 1 : 3);
    LAUNCHCHK();
    int rc = lpmx_dev_scan_int(len2, start, n + 1, total);
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(len2);
    if (rc != LPMX_OK)
    {
        hipFree(start);
        return rc;
    }
    *out = start;
    return LPMX_OK;
}

extern "C" int lpmx_get(int f, void *host, int64_t count)
{
    RETCHK(lpmx_require_model());
    if (!host)
        return lpmx_fail(LPMX_ERR_ARG, "null host buffer");
    const int n = g.n;
    if (f == LPMX_F_CONN)
    {
        RETCHK(lpmx_require_pattern());
        if (count < (int64_t)n * g.cm)
            return lpmx_fail(LPMX_ERR_ARG, "conn buffer too small");
        int *tmp = (int *)lpmx_scratch(sizeof(int) * (int64_t)n * g.cm);
        if (!tmp)
            return lpmx_fail(LPMX_ERR_HIP, "scratch allocation failed");
        hipLaunchKernelGGL(k_conn_rows, dim3(nblocks((int64_t)n * g.cm)), dim3(LPMX_BLOCK), 0, g_stream, g.conn,
                           g.nb_conn, tmp, n, g.cm);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(host, tmp, sizeof(int) * (int64_t)n * g.cm, hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        return LPMX_OK;
    }
    if (f == LPMX_F_K_POINTER)
    {
        RETCHK(lpmx_require_pattern());
        if (count < 2 * ((int64_t)n + 1))
            return lpmx_fail(LPMX_ERR_ARG, "K_pointer buffer too small");
        int *start = nullptr, total = 0;
        RETCHK(kptr_start(&start, &total));
        int *tmp = nullptr;
        HIPCHK(hipMalloc(&tmp, sizeof(int) * 2 * ((int64_t)n + 1)));
        hipLaunchKernelGGL(k_kptr, dim3(nblocks(n + 1)), dim3(LPMX_BLOCK), 0, g_stream, g.kcnt, start, tmp, n, 0);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(host, tmp, sizeof(int) * 2 * ((int64_t)n + 1), hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        hipFree(tmp);
        hipFree(start);
        return LPMX_OK;
    }
    FieldDesc fd;
    RETCHK(field_desc(f, &fd));
    const int64_t tot = (int64_t)n * fd.width;
    if (count < tot)
        return lpmx_fail(LPMX_ERR_ARG, "buffer of %lld elements < %lld", (long long)count, (long long)tot);
    const size_t es = fd.isint ? sizeof(int) : sizeof(double);
    if (!fd.planes)
    {
        HIPCHK(hipMemcpyAsync(host, fd.ptr, es * tot, hipMemcpyDeviceToHost, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        return LPMX_OK;
    }
    void *tmp = lpmx_scratch(es * tot);
    if (!tmp)
        return lpmx_fail(LPMX_ERR_HIP, "scratch allocation failed");
    if (fd.isint)
        hipLaunchKernelGGL(k_planes_to_rows<int>, dim3(nblocks(tot)), dim3(LPMX_BLOCK), 0, g_stream,
                           (const int *)fd.ptr, (int *)tmp, n, fd.width);
    else
        hipLaunchKernelGGL(k_planes_to_rows<double>, dim3(nblocks(tot)), dim3(LPMX_BLOCK), 0, g_stream,
                           (const double *)fd.ptr, (double *)tmp, n, fd.width);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(host, tmp, es * tot, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return LPMX_OK;
}

extern "C" int lpmx_put(int f, const void *host, int64_t count)
{
    RETCHK(lpmx_require_model());
    if (!host)
        return lpmx_fail(LPMX_ERR_ARG, "null host buffer");
    const int n = g.n;
    FieldDesc fd;
    RETCHK(field_desc(f, &fd));
    const int64_t tot = (int64_t)n * fd.width;
    if (count < tot)
        return lpmx_fail(LPMX_ERR_ARG, "buffer of %lld elements < %lld", (long long)count, (long long)tot);
    const size_t es = fd.isint ? sizeof(int) : sizeof(double);
    if (!fd.planes)
    {
        HIPCHK(hipMemcpyAsync(fd.ptr, host, es * tot, hipMemcpyHostToDevice, g_stream));
        HIPCHK(hipStreamSynchronize(g_stream));
        return LPMX_OK;
    }
    void *tmp = lpmx_scratch(es * tot);
    if (!tmp)
        return lpmx_fail(LPMX_ERR_HIP, "scratch allocation failed");
    HIPCHK(hipMemcpyAsync(tmp, host, es * tot, hipMemcpyHostToDevice, g_stream));
    if (fd.isint)
        hipLaunchKernelGGL(k_rows_to_planes<int>, dim3(nblocks(tot)), dim3(LPMX_BLOCK), 0, g_stream,
                           (const int *)tmp, (int *)fd.ptr, n, fd.width);
    else
        hipLaunchKernelGGL(k_rows_to_planes<double>, dim3(nblocks(tot)), dim3(LPMX_BLOCK), 0, g_stream,
                           (const double *)tmp, (double *)fd.ptr, n, fd.width);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    return LPMX_OK;
}

// ------------------------------------------------------------ neighbour build ----
struct CellGrid
{
    double lo[3], h;
    int nc[3];
};

__device__ __forceinline__ int cell_coord(double v, double lo, double h, int nc)
{
    int c = (int)floor((v - lo) / h);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void k_cell_count(const double *x, int n, CellGrid cg, int *cid, int *cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int a = cell_coord(x[3 * i], cg.lo[0], cg.h, cg.nc[0]);
    int b = cell_coord(x[3 * i + 1], cg.lo[1], cg.h, cg.nc[1]);
    int c = cell_coord(x[3 * i + 2], cg.lo[2], cg.h, cg.nc[2]);
    int id = (c * cg.nc[1] + b) * cg.nc[0] + a;
    cid[i] = id;
    atomicAdd(&cnt[id], 1);
}

__global__ void k_cell_fill(const int *cid, int n, const int *start, int *fillc, int *list)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int id = cid[i];
    list[start[id] + atomicAdd(&fillc[id], 1)] = i;
}

// searchNormalNeighbor (reference src/neighbor.c:9-46) from a cell list.  Accepted
// partners are kept sorted by index, so the lists equal the reference's
// increasing-j scan; the predicate and the arithmetic (sqrt of the sum of
// squared differences, no contraction) are the reference's.
__global__ void k_neighbors(DV v, CellGrid cg, const int *cstart, const int *list, int *overflow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    const double xi = v.x[3 * i], yi = v.x[3 * i + 1], zi = v.x[3 * i + 2];
    const double c1 = 1.01 * v.cut1, c2 = 1.01 * v.cut2;
    int js[LPMX_MAXNB], sh[LPMX_MAXNB];
    int cnt = 0, bad = 0;
    int ca = cell_coord(xi, cg.lo[0], cg.h, cg.nc[0]);
    int cb = cell_coord(yi, cg.lo[1], cg.h, cg.nc[1]);
    int cc = cell_coord(zi, cg.lo[2], cg.h, cg.nc[2]);
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++)
            {
                int qa = ca + dx, qb = cb + dy, qc = cc + dz;
                if (qa < 0 || qb < 0 || qc < 0 || qa >= cg.nc[0] || qb >= cg.nc[1] || qc >= cg.nc[2])
                    continue;
                int id = (qc * cg.nc[1] + qb) * cg.nc[0] + qa;
                for (int s = cstart[id]; s < cstart[id + 1]; s++)
                {
                    int j = list[s];
                    double ddx = v.x[3 * j] - xi, ddy = v.x[3 * j + 1] - yi, ddz = v.x[3 * j + 2] - zi;
                    double dis = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                    int shell;
                    if ((dis < c1) && (j != i))
                        shell = 0;
                    else if ((dis > c1) && (dis < c2))
                        shell = 1;
                    else
                        continue;
                    if (cnt >= v.nb)
                    {
                        bad = 1;
                        continue;
                    }
                    int p = cnt++;
                    while (p > 0 && js[p - 1] > j)
                    {
                        js[p] = js[p - 1];
                        sh[p] = sh[p - 1];
                        p--;
                    }
                    js[p] = j;
                    sh[p] = shell;
                }
            }
    int c1n = 0, c2n = 0;
    for (int a = 0; a < cnt; a++)
    {
        int j = js[a];
        double ddx = v.x[3 * j] - xi, ddy = v.x[3 * j + 1] - yi, ddz = v.x[3 * j + 2] - zi;
        double dis = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        int64_t b = (int64_t)a * v.n + i;
        v.nbr[b] = j;
        v.nsign[b] = sh[a];
        v.L0[b] = dis;
        v.cs0x[b] = (xi - v.x[3 * j]) / dis;
        v.cs0y[b] = (yi - v.x[3 * j + 1]) / dis;
        v.cs0z[b] = (zi - v.x[3 * j + 2]) / dis;
        if (sh[a] == 0)
            c1n++;
        else
            c2n++;
    }
    if (c1n > v.nb1 || c2n > v.nb2)
        bad = 1;
    v.nbcur[i] = cnt;
    v.nb0[i] = cnt;
    if (bad)
        atomicOr(overflow, 1);
}

// Reverse slot of every bond and the first opposite-direction slot
// (constitutive.c:540-559, 654-658).
__global__ void k_rev_opp(DV v)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.nb)
        return;
    int i = (int)(t % v.n), a = (int)(t / v.n);
    if (a >= v.nb0[i])
        return;
    int j = v.nbr[t];
    int r = -1;
    for (int e = 0; e < v.nb; e++)
        if (v.nbr[(int64_t)e * v.n + j] == i)
            r = e;
    v.rev[t] = r;
    int o = -1;
    const double cx = v.cs0x[t], cy = v.cs0y[t], cz = v.cs0z[t];
    for (int k = 0; k < v.nb0[i]; k++)
    {
        int64_t bk = (int64_t)k * v.n + i;
        if (fabs(v.cs0x[bk] + cx) < LPMX_EPS && fabs(v.cs0y[bk] + cy) < LPMX_EPS && fabs(v.cs0z[bk] + cz) < LPMX_EPS)
        {
            o = k;
            break;
        }
    }
    v.opp[t] = o;
}

extern "C" int lpmx_neighbor_build(void)
{
    RETCHK(lpmx_require_model());
    const int n = g.n;
    CellGrid cg;
    cg.h = 1.01 * g.cut2 * 1.0001;
    int64_t ncell = 1;
    for (int c = 0; c < 3; c++)
    {
        cg.lo[c] = g_lo[c];
        cg.nc[c] = 1 + (int)floor((g_hi[c] - g_lo[c]) / cg.h);
        ncell *= cg.nc[c];
    }
    if (ncell > 0x3fffffff)
        return lpmx_fail(LPMX_ERR_CAPACITY, "cell grid too large");
    int *cid, *cnt, *start, *fillc, *list, *ovf;
    HIPCHK(hipMalloc(&cid, sizeof(int) * n));
    HIPCHK(hipMalloc(&cnt, sizeof(int) * (ncell + 1)));
    HIPCHK(hipMalloc(&start, sizeof(int) * (ncell + 1)));
    HIPCHK(hipMalloc(&fillc, sizeof(int) * (ncell + 1)));
    HIPCHK(hipMalloc(&list, sizeof(int) * n));
    HIPCHK(hipMalloc(&ovf, sizeof(int)));
    HIPCHK(hipMemsetAsync(cnt, 0, sizeof(int) * (ncell + 1), g_stream));
    HIPCHK(hipMemsetAsync(fillc, 0, sizeof(int) * (ncell + 1), g_stream));
    HIPCHK(hipMemsetAsync(ovf, 0, sizeof(int), g_stream));
    hipLaunchKernelGGL(k_cell_count, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g.x, n, cg, cid, cnt);
    LAUNCHCHK();
    int total = 0;
    RETCHK(lpmx_dev_scan_int(cnt, start, (int)ncell + 1, &total));
    hipLaunchKernelGGL(k_cell_fill, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, cid, n, start, fillc, list);
    LAUNCHCHK();
    RETCHK(fill_i(g.nbr, (int64_t)n * g.nb, -1));
    RETCHK(fill_i(g.nsign, (int64_t)n * g.nb, -1));
    hipLaunchKernelGGL(k_neighbors, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, cg, start, list, ovf);
    LAUNCHCHK();
    RETCHK(fill_i(g.rev, (int64_t)n * g.nb, -1));
    RETCHK(fill_i(g.opp, (int64_t)n * g.nb, -1));
    hipLaunchKernelGGL(k_rev_opp, dim3(nblocks((int64_t)n * g.nb)), dim3(LPMX_BLOCK), 0, g_stream, g);
    LAUNCHCHK();
    int hovf = 0;
    HIPCHK(hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(cid), hipFree(cnt), hipFree(start), hipFree(fillc), hipFree(list), hipFree(ovf);
    if (hovf)
        return lpmx_fail(LPMX_ERR_CAPACITY, "a neighbour list exceeds nneighbors/nneighbors1/nneighbors2");
    g_have_nbr = true;
    g_have_pattern = false;
    g_pattern_version++;
    return LPMX_OK;
}

// ------------------------------------------------------------ connectivity -------
// conn[i] (reference src/neighbor.c:49-112): the sorted set of i's bond partners and
// the same-shell partners of each partner.  One thread per particle keeps its
// sorted set in LDS.
__global__ __launch_bounds__(64) void k_conn(DV v, int *overflow)
{
    __shared__ int set[64][LPMX_MAXCONN + 1];
    int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= v.n)
        return;
    int *s = set[threadIdx.x];
    int cnt = 0, bad = 0;
    for (int a = 0; a < v.nb0[i]; a++)
    {
        int64_t b = (int64_t)a * v.n + i;
        int j = v.nbr[b], sh = v.nsign[b];
        for (int e = -1; e < v.nb0[j]; e++)
        {
            int q;
            if (e < 0)
                q = j;
            else
            {
                int64_t be = (int64_t)e * v.n + j;
                if (v.nsign[be] != sh)
                    continue;
                q = v.nbr[be];
            }
            // insert q into the sorted set if absent
            int lo = 0, hi = cnt;
            while (lo < hi)
            {
                int mid = (lo + hi) >> 1;
                if (s[mid] < q)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (lo < cnt && s[lo] == q)
                continue;
            if (cnt >= v.cm)
            {
                bad = 1;
                continue;
            }
            for (int k = cnt; k > lo; k--)
                s[k] = s[k - 1];
            s[lo] = q;
            cnt++;
        }
    }
    int kc = 0;
    for (int k = 0; k < v.cm; k++)
    {
        int q = k < cnt ? s[k] : -1;
        v.conn[(int64_t)k * v.n + i] = q;
        if (k < cnt && q >= i)
            kc++;
    }
    v.nb_conn[i] = cnt;
    v.kcnt[i] = kc;
    if (bad)
        atomicOr(overflow, 1);
}

__global__ void k_slice_width(const int *nbc, int n, int nslices, int *w)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices)
        return;
    int m = 0;
    for (int l = 0; l < LPMX_SLICE; l++)
    {
        int i = s * LPMX_SLICE + l;
        if (i < n && nbc[i] > m)
            m = nbc[i];
    }
    w[s] = m;
}

// Column of every (row, slot) of the sliced ELL; padding slots point at the row
// itself with zero values.  dslot: slot of the diagonal block.
__global__ void k_sell_cols(DV v)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int i = (int)(t % v.n);
    int k = (int)(t / v.n);
    if (t >= (int64_t)v.n * v.cm)
        return;
    int s = i / LPMX_SLICE, l = i % LPMX_SLICE;
    int w = v.sptr[s + 1] - v.sptr[s];
    if (k >= w)
        return;
    int64_t slot = (int64_t)(v.sptr[s] + k) * LPMX_SLICE + l;
    int q = k < v.nb_conn[i] ? v.conn[(int64_t)k * v.n + i] : i;
    v.col[slot] = q;
    if (k < v.nb_conn[i] && q == i)
        v.dslot[i] = k;
    // transposed block: slot of i in row q (conn is symmetric)
    int tp = -1;
    if (k < v.nb_conn[i])
    {
        int lo = 0, hi = v.nb_conn[q] - 1;
        while (lo <= hi)
        {
            int mid = (lo + hi) >> 1;
            int c = v.conn[(int64_t)mid * v.n + q];
            if (c == i)
            {
                tp = (int)((v.sptr[q / LPMX_SLICE] + mid) * (int64_t)LPMX_SLICE + q % LPMX_SLICE);
                break;
            }
            if (c < i)
                lo = mid + 1;
            else
                hi = mid - 1;
        }
    }
    v.tpos[slot] = tp;
}

__global__ void k_rowptr_blocks(const int *nbc, int64_t *rowptr, int n, const int64_t *boff, int per)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t acc = boff[b];
    for (int i = b * per; i < n && i < (b + 1) * per; i++)
    {
        rowptr[i] = acc;
        acc += nbc[i];
    }
    if (b * per <= n && n < (b + 1) * per)
        rowptr[n] = acc;
}

__global__ void k_rowlen_sum(const int *nbc, int n, int per, int64_t *bsum, int nb)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb)
        return;
    int64_t s = 0;
    for (int i = b * per; i < n && i < (b + 1) * per; i++)
        s += nbc[i];
    bsum[b] = s;
}

extern "C" int lpmx_connectivity_build(int64_t *len_K)
{
    RETCHK(lpmx_require_model());
    if (!g_have_nbr)
        return lpmx_fail(LPMX_ERR_STATE, "no neighbour lists (lpmx_neighbor_build first)");
    const int n = g.n;
    if (!g.conn)
    {
        RETCHK(dalloc(&g.conn, (int64_t)n * g.cm));
        RETCHK(dalloc(&g.dslot, n));
        RETCHK(dalloc(&g.rowptr, (int64_t)n + 1));
    }
    int *ovf;
    HIPCHK(hipMalloc(&ovf, sizeof(int)));
    HIPCHK(hipMemsetAsync(ovf, 0, sizeof(int), g_stream));
    hipLaunchKernelGGL(k_conn, dim3(nblocks(n, 64)), dim3(64), 0, g_stream, g, ovf);
    LAUNCHCHK();
    int hovf = 0;
    HIPCHK(hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(ovf);
    if (hovf)
        return lpmx_fail(LPMX_ERR_CAPACITY, "a conn row exceeds nneighbors_AFEM+1");
    // slices
    const int ns = (n + LPMX_SLICE - 1) / LPMX_SLICE;
    int *w = nullptr;
    if (g.sptr)
        hipFree(g.sptr);
    HIPCHK(hipMalloc(&g.sptr, sizeof(int) * (ns + 1)));
    HIPCHK(hipMalloc(&w, sizeof(int) * (ns + 1)));
    HIPCHK(hipMemsetAsync(w, 0, sizeof(int) * (ns + 1), g_stream));
    hipLaunchKernelGGL(k_slice_width, dim3(nblocks(ns)), dim3(LPMX_BLOCK), 0, g_stream, g.nb_conn, n, ns, w);
    LAUNCHCHK();
    int tot = 0;
    RETCHK(lpmx_dev_scan_int(w, g.sptr, ns + 1, &tot));
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(w);
    g.nslices = ns;
    g.nslots = tot;
    // (re)allocate the block storage
    const int64_t nval = (int64_t)tot * LPMX_SLICE;
    if (g.col)
    {
        hipFree(g.col), hipFree(g.tpos), hipFree(g.K), hipFree(g.K32);
        g.col = g.tpos = nullptr, g.K = nullptr, g.K32 = nullptr;
    }
    HIPCHK(hipMalloc(&g.col, sizeof(int) * nval));
    HIPCHK(hipMalloc(&g.tpos, sizeof(int) * nval));
    HIPCHK(hipMalloc(&g.K, sizeof(double) * nval * g.d2));
    HIPCHK(hipMalloc(&g.K32, sizeof(float) * nval * g.d2));
    HIPCHK(hipMemsetAsync(g.K, 0, sizeof(double) * nval * g.d2, g_stream));
    HIPCHK(hipMemsetAsync(g.K32, 0, sizeof(float) * nval * g.d2, g_stream));
    HIPCHK(hipMemsetAsync(g.tpos, 0xff, sizeof(int) * nval, g_stream));
    // padding of the last slice beyond n: column 0 (any valid index), zero values
    HIPCHK(hipMemsetAsync(g.col, 0, sizeof(int) * nval, g_stream));
    hipLaunchKernelGGL(k_sell_cols, dim3(nblocks((int64_t)n * g.cm)), dim3(LPMX_BLOCK), 0, g_stream, g);
    LAUNCHCHK();
    // row-major block offsets (the first block of each row in conn order) and nnzb
    const int per = 1024;
    const int nbk = (n + 1 + per - 1) / per;
    int64_t *bsum = nullptr;
    HIPCHK(hipMalloc(&bsum, sizeof(int64_t) * (nbk + 1)));
    hipLaunchKernelGGL(k_rowlen_sum, dim3(nblocks(nbk)), dim3(LPMX_BLOCK), 0, g_stream, g.nb_conn, n, per, bsum, nbk);
    LAUNCHCHK();
    std::vector<int64_t> hb(nbk), ho(nbk);
    HIPCHK(hipMemcpyAsync(hb.data(), bsum, sizeof(int64_t) * nbk, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    int64_t acc = 0;
    for (int b = 0; b < nbk; b++)
        ho[b] = acc, acc += hb[b];
    HIPCHK(hipMemcpyAsync(bsum, ho.data(), sizeof(int64_t) * nbk, hipMemcpyHostToDevice, g_stream));
    hipLaunchKernelGGL(k_rowptr_blocks, dim3(nblocks(nbk)), dim3(LPMX_BLOCK), 0, g_stream, g.nb_conn, g.rowptr, n,
                       bsum, per);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    hipFree(bsum);
    g.nnzb = acc;
    // K_pointer[n][1]
    int *start = nullptr, total = 0;
    RETCHK(kptr_start(&start, &total));
    hipFree(start);
    if (len_K)
        *len_K = total;
    g_have_pattern = true;
    g_pattern_version++;
    return LPMX_OK;
}

extern "C" int lpmx_matrix_info(int64_t *nnzb, int64_t *padded, int *nslices)
{
    RETCHK(lpmx_require_pattern());
    if (nnzb)
        *nnzb = g.nnzb;
    if (padded)
        *padded = g.nslots * LPMX_SLICE;
    if (nslices)
        *nslices = g.nslices;
    return LPMX_OK;
}

// --------------------------------------------------------------- materials -------
// calcKnTv (reference src/stiffness.c:11-268) per bond from the per-particle
// constants kt[c][i] = emod_i * KnTve[type_i][c]; SC averages both ends.
__global__ void k_kntv(DV v)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)v.n * v.nb)
        return;
    int i = (int)(t % v.n), a = (int)(t / v.n);
    if (a >= v.nb0[i])
        return;
    int j = v.nbr[t], s = v.nsign[t];
    if (v.lattice == 1)
    {
        v.Kn[t] = v.kt[i];
        v.Tv[t] = v.kt[(int64_t)v.n + i];
    }
    else if (v.lattice == 2)
    {
        v.Kn[t] = 0.5 * (v.kt[(int64_t)s * v.n + i] + v.kt[(int64_t)s * v.n + j]);
        v.Tv[t] = 0.5 * (v.kt[2 * (int64_t)v.n + i] + v.kt[2 * (int64_t)v.n + j]);
    }
    else
    {
        v.Kn[t] = v.kt[(int64_t)s * v.n + i];
        v.Tv[t] = v.kt[2 * (int64_t)v.n + i];
    }
}

__global__ void k_particle_constants(DV v, const double *ktype, const double *ce, const double *emod, int ntype)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n)
        return;
    int t = v.type[i];
    if (t < 0 || t >= ntype)
        t = 0;
    double e = emod ? emod[i] : 1.0;
    for (int c = 0; c < 3; c++)
        v.kt[(int64_t)c * v.n + i] = e * ktype[3 * t + c];
    v.c44[i] = ce[3 * t + 2] * e;
}

// KnTve per type: the reference's mapping matrices times (C11, C12, C44),
// scaled by the radius for the 3D lattices (stiffness.c:17-23, 51-56, 179-185,
// 210-216, 242-248).
static void kntve_host(int lattice, double r, const double *ce, double *kt)
{
    const double c11 = ce[0], c12 = ce[1], c44 = ce[2];
    if (lattice == 0)
    {
        kt[0] = 1 / 2.0 * c11 + -1 / 2.0 * c12 + 0.0 * c44;
        kt[1] = 0.0 * c11 + 0.0 * c12 + 1 / 2.0 * c44;
        kt[2] = 0.0 * c11 + 1.0 / 12.0 * c12 + -1.0 / 12.0 * c44;
    }
    else if (lattice == 1)
    {
        kt[0] = sqrt(3.0) / 12.0 * c11 + -sqrt(3.0) / 12.0 * c12;
        kt[1] = -sqrt(3.0) / 144.0 * c11 + sqrt(3.0) / 48.0 * c12;
        kt[2] = 0.0;
    }
    else if (lattice == 2)
    {
        kt[0] = r * (1 * c11 + -1 * c12 + -1 * c44);
        kt[1] = r * (0 * c11 + 0 * c12 + 1 * c44);
        kt[2] = r * (0 * c11 + 1.0 / 18.0 * c12 + -1.0 / 18.0 * c44);
    }
    else if (lattice == 3)
    {
        kt[0] = r * (0 * c11 + 0 * c12 + sqrt(2.0) * c44);
        kt[1] = r * (sqrt(2.0) / 4.0 * c11 + -sqrt(2.0) / 4.0 * c12 + -sqrt(2.0) / 4.0 * c44);
        kt[2] = r * (0 * c11 + sqrt(2.0) / 24.0 * c12 + -sqrt(2.0) / 24.0 * c44);
    }
    else
    {
        kt[0] = r * (0. * c11 + 0. * c12 + sqrt(3.0) * c44);
        kt[1] = r * (1. / sqrt(3.0) * c11 + -1. / sqrt(3.0) * c12 + 0. * c44);
        kt[2] = r * (0. * c11 + sqrt(3.0) / 14.0 * c12 + sqrt(3.0) / 14.0 * c44);
    }
}

extern "C" int lpmx_set_materials(int ntype, const int *type, const double *Ce, const double *sigmay,
                                  const double *emod)
{
    RETCHK(lpmx_require_model());
    if (!g_have_nbr)
        return lpmx_fail(LPMX_ERR_STATE, "no neighbour lists (lpmx_neighbor_build first)");
    if (ntype <= 0 || !type || !Ce)
        return lpmx_fail(LPMX_ERR_ARG, "materials need ntype > 0, type[] and Ce[]");
    const int n = g.n;
    for (int i = 0; i < n; i++)
        if (type[i] < 0 || type[i] >= ntype)
            return lpmx_fail(LPMX_ERR_ARG, "type[%d] = %d outside [0, %d)", i, type[i], ntype);
    std::vector<double> kt(3 * ntype);
    for (int t = 0; t < ntype; t++)
        kntve_host(g.lattice, g.radius, Ce + 3 * t, kt.data() + 3 * t);
    double *dk, *dce, *dem = nullptr;
    HIPCHK(hipMalloc(&dk, sizeof(double) * 3 * ntype));
    HIPCHK(hipMalloc(&dce, sizeof(double) * 3 * ntype));
    HIPCHK(hipMemcpyAsync(dk, kt.data(), sizeof(double) * 3 * ntype, hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipMemcpyAsync(dce, Ce, sizeof(double) * 3 * ntype, hipMemcpyHostToDevice, g_stream));
    if (emod)
    {
        HIPCHK(hipMalloc(&dem, sizeof(double) * n));
        HIPCHK(hipMemcpyAsync(dem, emod, sizeof(double) * n, hipMemcpyHostToDevice, g_stream));
    }
    HIPCHK(hipMemcpyAsync(g.type, type, sizeof(int) * n, hipMemcpyHostToDevice, g_stream));
    if (sigmay)
        HIPCHK(hipMemcpyAsync(g.sigmay, sigmay, sizeof(double) * n, hipMemcpyHostToDevice, g_stream));
    hipLaunchKernelGGL(k_particle_constants, dim3(nblocks(n)), dim3(LPMX_BLOCK), 0, g_stream, g, dk, dce, dem, ntype);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_kntv, dim3(nblocks((int64_t)n * g.nb)), dim3(LPMX_BLOCK), 0, g_stream, g);
    LAUNCHCHK();
    // kept for the adjoint (d Kn / d factor, d c44 / d factor)
    if (g.ktype_t)
        hipFree(g.ktype_t), hipFree(g.ce_t);
    if (!g.emod)
        HIPCHK(hipMalloc(&g.emod, sizeof(double) * n));
    g.ktype_t = dk, g.ce_t = dce, g.ntype = ntype;
    if (dem)
        HIPCHK(hipMemcpyAsync(g.emod, dem, sizeof(double) * n, hipMemcpyDeviceToDevice, g_stream));
    else
    {
        std::vector<double> ones(n, 1.0);
        HIPCHK(hipMemcpyAsync(g.emod, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice, g_stream));
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    if (dem)
        hipFree(dem);
    return LPMX_OK;
}

extern "C" int lpmx_set_bond_critical(const double *crit)
{
    RETCHK(lpmx_require_model());
    const int64_t nbn = (int64_t)g.n * g.nb;
    if (!crit)
    {
        if (g.crit)
            hipFree(g.crit);
        g.crit = nullptr;
        return LPMX_OK;
    }
    if (!g.crit)
        HIPCHK(hipMalloc(&g.crit, sizeof(double) * nbn));
    double *tmp = lpmx_scratch(sizeof(double) * nbn);
    if (!tmp)
        return lpmx_fail(LPMX_ERR_HIP, "scratch allocation failed");
    HIPCHK(hipMemcpyAsync(tmp, crit, sizeof(double) * nbn, hipMemcpyHostToDevice, g_stream));
    hipLaunchKernelGGL(k_rows_to_planes<double>, dim3(nblocks(nbn)), dim3(LPMX_BLOCK), 0, g_stream, tmp, g.crit, g.n,
                       g.nb);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(g_stream));
    return LPMX_OK;
}
