/*
 * The drop-in hot path: the reference's load-step entry points (src/neighbor.c,
 * src/stiffness.c, src/boundary.c, src/solver.c, src/constitutive.c,
 * src/lpm_basic.c) on the MI355X device layer.
 *
 * The reference keeps its state in host globals (double** rows) that drivers read
 * and write directly.  Every entry point here therefore pushes the globals it
 * reads to the device model, runs the device kernels, and pulls back the globals
 * it writes, so between calls the host globals hold exactly what the reference's
 * would.  The transfers cost PCIe time; lpm_ext_run_steps() runs whole load steps
 * with every field resident on the device instead.
 *
 * Fatal conditions print and exit(1), as the reference does; a CG that reaches its
 * cap prints and carries on with the displacement it has, as the reference does.
 *
 * Resident mode (LPMX_DROPIN_RESIDENT=1 in the environment when the device model is
 * made): the device copy of every field is the current one.  A call no longer pulls
 * what it writes but marks the host copy stale, a push skips a field the device
 * already holds, and library code that reads a host array (the writers, copyDouble*,
 * countNEqual, cblas_dnrm2, the damage file lines) first brings back the stale fields
 * it reads.  The tangent K_global stays on the device the same way.  It is for
 * drivers that touch the model's arrays only through library calls, as the
 * reference's drivers do; the device computes the same values in both modes.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dropin.h"
#include "lpm_ext.h"

static int d_have_model = 0, d_n = -1;
/* extension data of the current lattice, copied when set and dropped when a new
 * lattice is created (searchNormalNeighbor) */
static double *ext_modulus = NULL; /* [nparticle] or NULL */
static double *ext_crit = NULL;    /* [nparticle * nneighbors] or NULL */
static lpmx_solver ext_solver;
static int ext_solver_set = 0;
/* resident mode: per field, the host copy is behind the device's (h_stale), the
 * device copy is current (d_fresh), and the host array the device copy came from */
static int resident = 0;
static unsigned char h_stale[LPMX_F_COUNT], d_fresh[LPMX_F_COUNT];
static const void *d_from[LPMX_F_COUNT];
static int k_h_stale = 0, k_d_fresh = 0, k_have_pattern = 0; /* the same for K_global */

void lpm_dropin_fatal(const char *where, int rc)
{
    fprintf(stderr, "%s: device call failed (%d): %s\n", where, rc, lpmx_last_error());
    exit(1);
}

#define CHK(where, x)                                                          \
    do                                                                         \
    {                                                                          \
        int rc_ = (x);                                                         \
        if (rc_ != LPMX_OK)                                                    \
            lpm_dropin_fatal(where, rc_);                                      \
    } while (0)

/* ---------------------------------------------------------- field bindings --- */
enum kind
{
    K_D2,  /* double **a, a[i][j], width w */
    K_D3S, /* double ***a, a[i][j][slot], width w */
    K_D2S, /* double **a, a[i][slot] */
    K_D1,  /* double *a, n*w contiguous */
    K_I2,  /* int **a, a[i][j] */
    K_I1   /* int *a, n*w contiguous */
};

struct hb
{
    int kind, width, slot;
    void *p;
};

static int bind(int f, struct hb *b)
{
    const int nb_ = nneighbors;
    b->slot = 0;
    b->width = 1;
    switch (f)
    {
    case LPMX_F_XYZ: b->kind = K_D2, b->p = xyz, b->width = 3; break;
    case LPMX_F_XYZ_INITIAL: b->kind = K_D2, b->p = xyz_initial, b->width = 3; break;
    case LPMX_F_XYZ_TEMP: b->kind = K_D2, b->p = xyz_temp, b->width = 3; break;
    case LPMX_F_NEIGHBORS: b->kind = K_I2, b->p = neighbors, b->width = nb_; break;
    case LPMX_F_NSIGN: b->kind = K_I2, b->p = nsign, b->width = nb_; break;
    case LPMX_F_NB: b->kind = K_I1, b->p = nb; break;
    case LPMX_F_NB_INITIAL: b->kind = K_I1, b->p = nb_initial; break;
    case LPMX_F_NB_CONN: b->kind = K_I1, b->p = nb_conn; break;
    case LPMX_F_TYPE: b->kind = K_I1, b->p = type; break;
    case LPMX_F_PL_FLAG: b->kind = K_I1, b->p = pl_flag; break;
    case LPMX_F_DISTANCE_INITIAL: b->kind = K_D2, b->p = distance_initial, b->width = nb_; break;
    case LPMX_F_CSX_INITIAL: b->kind = K_D2, b->p = csx_initial, b->width = nb_; break;
    case LPMX_F_CSY_INITIAL: b->kind = K_D2, b->p = csy_initial, b->width = nb_; break;
    case LPMX_F_CSZ_INITIAL: b->kind = K_D2, b->p = csz_initial, b->width = nb_; break;
    case LPMX_F_KN: b->kind = K_D2, b->p = Kn, b->width = nb_; break;
    case LPMX_F_TV: b->kind = K_D2, b->p = Tv, b->width = nb_; break;
    case LPMX_F_DL: b->kind = K_D2, b->p = dL, b->width = nb_; break;
    case LPMX_F_DL_AVE: b->kind = K_D2, b->p = dL_ave, b->width = nb_; break;
    case LPMX_F_DDL: b->kind = K_D2, b->p = ddL, b->width = nb_; break;
    case LPMX_F_DDLP: b->kind = K_D2, b->p = ddLp, b->width = nb_; break;
    case LPMX_F_F: b->kind = K_D2, b->p = F, b->width = nb_; break;
    case LPMX_F_F_TEMP: b->kind = K_D2, b->p = F_temp, b->width = nb_; break;
    case LPMX_F_CSX: b->kind = K_D2, b->p = csx, b->width = nb_; break;
    case LPMX_F_CSY: b->kind = K_D2, b->p = csy, b->width = nb_; break;
    case LPMX_F_CSZ: b->kind = K_D2, b->p = csz, b->width = nb_; break;
    case LPMX_F_BOND_STRESS: b->kind = K_D2, b->p = bond_stress, b->width = nb_; break;
    case LPMX_F_DISTANCE: b->kind = K_D2, b->p = distance, b->width = nb_; break;
    case LPMX_F_DAMAGE_BROKEN: b->kind = K_D2, b->p = damage_broken, b->width = nb_; break;
    case LPMX_F_DAMAGE_W: b->kind = K_D2, b->p = damage_w, b->width = nb_; break;
    case LPMX_F_DLP0: case LPMX_F_DLP1: case LPMX_F_DLP2:
        b->kind = K_D3S, b->p = dLp, b->width = nb_, b->slot = f - LPMX_F_DLP0; break;
    case LPMX_F_DAMAGE_D0: case LPMX_F_DAMAGE_D1:
        b->kind = K_D3S, b->p = damage_D, b->width = nb_, b->slot = f - LPMX_F_DAMAGE_D0; break;
    case LPMX_F_DL_TOTAL: b->kind = K_D2, b->p = dL_total, b->width = 2; break;
    case LPMX_F_TDL_TOTAL: b->kind = K_D2, b->p = TdL_total, b->width = 2; break;
    case LPMX_F_DDL_TOTAL: b->kind = K_D2, b->p = ddL_total, b->width = 2; break;
    case LPMX_F_TDDL_TOTAL: b->kind = K_D2, b->p = TddL_total, b->width = 2; break;
    case LPMX_F_J2_BETA0: case LPMX_F_J2_BETA1: case LPMX_F_J2_BETA2:
        b->kind = K_D3S, b->p = J2_beta, b->width = 6, b->slot = f - LPMX_F_J2_BETA0; break;
    case LPMX_F_STRESS: b->kind = K_D2, b->p = stress_tensor, b->width = 6; break;
    case LPMX_F_J2_ALPHA0: case LPMX_F_J2_ALPHA1: case LPMX_F_J2_ALPHA2:
        b->kind = K_D2S, b->p = J2_alpha, b->slot = f - LPMX_F_J2_ALPHA0; break;
    case LPMX_F_J2_DLAMBDA: b->kind = K_D1, b->p = J2_dlambda; break;
    case LPMX_F_J2_STRESSEQ: b->kind = K_D1, b->p = J2_stresseq; break;
    case LPMX_F_J2_STRESSM: b->kind = K_D1, b->p = J2_stressm; break;
    case LPMX_F_J2_TRIAXIALITY: b->kind = K_D1, b->p = J2_triaxiality; break;
    case LPMX_F_DAMAGE_NONLOCAL0: case LPMX_F_DAMAGE_NONLOCAL1:
        b->kind = K_D2S, b->p = damage_nonlocal, b->slot = f - LPMX_F_DAMAGE_NONLOCAL0; break;
    case LPMX_F_DAMAGE_VISUAL: b->kind = K_D1, b->p = damage_visual; break;
    case LPMX_F_SIGMAY: b->kind = K_D1, b->p = sigmay; break;
    case LPMX_F_PIN: b->kind = K_D1, b->p = Pin, b->width = 3; break;
    case LPMX_F_PEX: b->kind = K_D1, b->p = Pex, b->width = dim; break;
    case LPMX_F_PEX_TEMP: b->kind = K_D1, b->p = Pex_temp, b->width = dim; break;
    case LPMX_F_RESIDUAL: b->kind = K_D1, b->p = residual, b->width = dim; break;
    case LPMX_F_DISP: b->kind = K_D1, b->p = disp, b->width = dim; break;
    case LPMX_F_DISPBC_INDEX: b->kind = K_I1, b->p = dispBC_index, b->width = dim; break;
    case LPMX_F_FIX_INDEX: b->kind = K_I1, b->p = fix_index, b->width = dim; break;
    case LPMX_F_CP_GY0: case LPMX_F_CP_GY1: case LPMX_F_CP_GY2:
        b->kind = K_D3S, b->p = cp_gy, b->width = nslipSys, b->slot = f - LPMX_F_CP_GY0; break;
    case LPMX_F_CP_A_SINGLE0: case LPMX_F_CP_A_SINGLE1: case LPMX_F_CP_A_SINGLE2:
        b->kind = K_D3S, b->p = cp_A_single, b->width = nslipSys, b->slot = f - LPMX_F_CP_A_SINGLE0; break;
    case LPMX_F_CP_A0: case LPMX_F_CP_A1: case LPMX_F_CP_A2:
        b->kind = K_D2S, b->p = cp_A, b->slot = f - LPMX_F_CP_A0; break;
    case LPMX_F_CP_DA: b->kind = K_D1, b->p = cp_dA; break;
    case LPMX_F_CP_DGY: b->kind = K_D2, b->p = cp_dgy, b->width = nslipSys; break;
    case LPMX_F_CP_DA_SINGLE: b->kind = K_D2, b->p = cp_dA_single, b->width = nslipSys; break;
    case LPMX_F_CP_RSS: b->kind = K_D2, b->p = cp_RSS, b->width = nslipSys; break;
    case LPMX_F_CP_JACT: b->kind = K_I2, b->p = cp_Jact, b->width = nslipSys; break;
    case LPMX_F_CP_CAB: b->kind = K_D2, b->p = cp_Cab, b->width = nslipSys * nslipSys; break;
    case LPMX_F_J2_BETA_EQ0: case LPMX_F_J2_BETA_EQ1: case LPMX_F_J2_BETA_EQ2:
        b->kind = K_D2S, b->p = J2_beta_eq, b->slot = f - LPMX_F_J2_BETA_EQ0; break;
    case LPMX_F_DAMAGE_LOCAL0: case LPMX_F_DAMAGE_LOCAL1:
        b->kind = K_D2S, b->p = damage_local, b->slot = f - LPMX_F_DAMAGE_LOCAL0; break;
    default: return 0;
    }
    if (f >= LPMX_F_CP_GY0 && f <= LPMX_F_CP_CAB && nslipSys <= 0)
        return 0;
    return b->p != NULL;
}

static double *dbuf = NULL;
static int *ibuf = NULL;
static long dcap = 0, icap = 0;

static double *dtmp(long n)
{
    if (n > dcap)
    {
        free(dbuf);
        dbuf = malloc(sizeof(double) * n);
        dcap = n;
    }
    return dbuf;
}

static int *itmp(long n)
{
    if (n > icap)
    {
        free(ibuf);
        ibuf = malloc(sizeof(int) * n);
        icap = n;
    }
    return ibuf;
}

void lpm_dropin_push(int f)
{
    struct hb b;
    if (!bind(f, &b))
        return;
    /* resident: nothing to send while the device holds this array's current values */
    if (resident && d_from[f] == b.p && (d_fresh[f] || h_stale[f]))
        return;
    const long n = nparticle, w = b.width, tot = n * w;
    int rc;
    if (b.kind == K_I1 || b.kind == K_I2)
    {
        int *t = b.kind == K_I1 ? (int *)b.p : itmp(tot);
        if (b.kind == K_I2)
            for (long i = 0; i < n; i++)
                memcpy(t + i * w, ((int **)b.p)[i], sizeof(int) * w);
        rc = lpmx_put(f, t, tot);
    }
    else
    {
        double *t = b.kind == K_D1 ? (double *)b.p : dtmp(tot);
        if (b.kind == K_D2)
            for (long i = 0; i < n; i++)
                memcpy(t + i * w, ((double **)b.p)[i], sizeof(double) * w);
        else if (b.kind == K_D3S)
            for (long i = 0; i < n; i++)
                for (long j = 0; j < w; j++)
                    t[i * w + j] = ((double ***)b.p)[i][j][b.slot];
        else if (b.kind == K_D2S)
            for (long i = 0; i < n; i++)
                t[i] = ((double **)b.p)[i][b.slot];
        rc = lpmx_put(f, t, tot);
    }
    CHK("push", rc);
    d_fresh[f] = 1, h_stale[f] = 0, d_from[f] = b.p;
}

static void pull_now(int f);

void lpm_dropin_pull(int f)
{
    struct hb b;
    if (!bind(f, &b))
        return;
    if (resident)
    {
        h_stale[f] = 1, d_fresh[f] = 1, d_from[f] = b.p;
        return;
    }
    pull_now(f);
}

/* the device copy of field f into the host array */
static void pull_now(int f)
{
    struct hb b;
    if (!bind(f, &b))
        return;
    const long n = nparticle, w = b.width, tot = n * w;
    if (b.kind == K_I1 || b.kind == K_I2)
    {
        int *t = b.kind == K_I1 ? (int *)b.p : itmp(tot);
        CHK("pull", lpmx_get(f, t, tot));
        if (b.kind == K_I2)
            for (long i = 0; i < n; i++)
                memcpy(((int **)b.p)[i], t + i * w, sizeof(int) * w);
    }
    else
    {
        double *t = b.kind == K_D1 ? (double *)b.p : dtmp(tot);
        CHK("pull", lpmx_get(f, t, tot));
        if (b.kind == K_D2)
            for (long i = 0; i < n; i++)
                memcpy(((double **)b.p)[i], t + i * w, sizeof(double) * w);
        else if (b.kind == K_D3S)
            for (long i = 0; i < n; i++)
                for (long j = 0; j < w; j++)
                    ((double ***)b.p)[i][j][b.slot] = t[i * w + j];
        else if (b.kind == K_D2S)
            for (long i = 0; i < n; i++)
                ((double **)b.p)[i][b.slot] = t[i];
    }
    h_stale[f] = 0, d_fresh[f] = 1, d_from[f] = b.p;
}

/* Host code is about to read (write = 0) or change (write = 1) the array at p: bring
 * back the fields stored there whose host copy is stale; after a change the device
 * copy is behind. */
void lpm_dropin_host_access(const void *p, int write)
{
    if (!resident || !p)
        return;
    if (p == (const void *)K_global)
    {
        if (k_h_stale)
            CHK("K export", lpmx_export_upper_csr(NULL, NULL, K_global));
        k_h_stale = 0;
        if (write)
            k_d_fresh = 0;
        return;
    }
    for (int f = 0; f < LPMX_F_COUNT; f++)
    {
        struct hb b;
        if (!bind(f, &b) || b.p != p)
            continue;
        if (h_stale[f])
            pull_now(f);
        if (write)
            d_fresh[f] = 0;
    }
}

/* Every stale host copy brought back (before host code that reads many arrays). */
void lpm_dropin_sync_host(void)
{
    if (!resident)
        return;
    for (int f = 0; f < LPMX_F_COUNT; f++)
        if (h_stale[f])
            pull_now(f);
    if (k_h_stale)
        CHK("K export", lpmx_export_upper_csr(NULL, NULL, K_global));
    k_h_stale = 0;
}

static void host_read(int f)
{
    if (resident && h_stale[f])
        pull_now(f);
}

/* K_global to the device unless it holds the current values, and back unless the
 * host copy may wait (resident) */
static void k_import(void)
{
    if (resident && k_d_fresh)
        return;
    CHK("K import", lpmx_import_upper_csr(K_global));
    k_d_fresh = 1;
}

static void k_export(void)
{
    if (resident)
    {
        k_h_stale = 1, k_d_fresh = 1;
        return;
    }
    CHK("K export", lpmx_export_upper_csr(NULL, NULL, K_global));
}

static void push_list(const int *f, int k)
{
    for (int i = 0; i < k; i++)
        lpm_dropin_push(f[i]);
}

static void pull_list(const int *f, int k)
{
    for (int i = 0; i < k; i++)
        lpm_dropin_pull(f[i]);
}

#define PUSH(...)                                                              \
    do                                                                         \
    {                                                                          \
        const int fl_[] = {__VA_ARGS__};                                       \
        push_list(fl_, (int)(sizeof(fl_) / sizeof(int)));                      \
    } while (0)
#define PULL(...)                                                              \
    do                                                                         \
    {                                                                          \
        const int fl_[] = {__VA_ARGS__};                                       \
        pull_list(fl_, (int)(sizeof(fl_) / sizeof(int)));                      \
    } while (0)
#define PULL_NOW(...)                                                          \
    do                                                                         \
    {                                                                          \
        const int fl_[] = {__VA_ARGS__};                                       \
        for (int k_ = 0; k_ < (int)(sizeof(fl_) / sizeof(int)); k_++)           \
            pull_now(fl_[k_]);                                                 \
    } while (0)

/* ------------------------------------------------------------ device model --- */
static void push_params(void)
{
    lpmx_params p = {plmode, J2_H, J2_xi, damage_L, damage_threshold, damagec_A, critical_bstrain, nbreak, J2_C};
    CHK("params", lpmx_set_params(&p));
}

/* (Re)create the device model from the host positions. */
static void create_model(void)
{
    lpmx_model_desc d = {dim, lattice, nparticle, nneighbors, nneighbors1, nneighbors2, nneighbors_AFEM,
                         radius, neighbor1_cutoff, neighbor2_cutoff, particle_volume};
    double *x = dtmp(3L * nparticle);
    for (long i = 0; i < nparticle; i++)
        memcpy(x + 3 * i, xyz[i], sizeof(double) * 3);
    CHK("lpmx_model_create", lpmx_model_create(&d, x));
    const char *rm = getenv("LPMX_DROPIN_RESIDENT");
    resident = rm && atoi(rm) != 0;
    memset(h_stale, 0, sizeof(h_stale));
    memset(d_fresh, 0, sizeof(d_fresh));
    memset(d_from, 0, sizeof(d_from));
    k_h_stale = k_d_fresh = k_have_pattern = 0;
    if (xyz_initial)
        lpm_dropin_push(LPMX_F_XYZ_INITIAL);
    d_have_model = 1;
    d_n = nparticle;
    /* per-particle and per-bond extension data belonged to the previous lattice */
    free(ext_modulus), free(ext_crit);
    ext_modulus = ext_crit = NULL;
    push_params();
}

static void need_model(const char *where)
{
    if (!d_have_model || d_n != nparticle)
    {
        fprintf(stderr, "%s: no device model; call searchNormalNeighbor() first\n", where);
        exit(1);
    }
    push_params();
}

/* ---------------------------------------------------------- neighbour lists --- */
/* searchNormalNeighbor (reference src/neighbor.c:9-46) by the device cell list. */
void searchNormalNeighbor()
{
    create_model();
    CHK("searchNormalNeighbor", lpmx_neighbor_build());
    PULL_NOW(LPMX_F_NEIGHBORS, LPMX_F_NSIGN, LPMX_F_NB, LPMX_F_NB_INITIAL, LPMX_F_DISTANCE_INITIAL,
             LPMX_F_CSX_INITIAL, LPMX_F_CSY_INITIAL, LPMX_F_CSZ_INITIAL);
    /* the per-shell lists are the shell-ordered subsequences of neighbors[i] */
    for (int i = 0; i < nparticle; i++)
    {
        int a = 0, b = 0;
        for (int k = 0; k < nneighbors1; k++)
            neighbors1[i][k] = -1;
        for (int k = 0; k < nneighbors2; k++)
            neighbors2[i][k] = -1;
        for (int j = 0; j < nb_initial[i]; j++)
        {
            if (nsign[i][j] == 0 && a < nneighbors1)
                neighbors1[i][a++] = neighbors[i][j];
            else if (nsign[i][j] == 1 && b < nneighbors2)
                neighbors2[i][b++] = neighbors[i][j];
        }
    }
}

/* searchAFEMNeighbor (reference src/neighbor.c:49-141). */
int searchAFEMNeighbor()
{
    need_model("searchAFEMNeighbor");
    int64_t len = 0;
    CHK("searchAFEMNeighbor", lpmx_connectivity_build(&len));
    pull_now(LPMX_F_NB_CONN);
    const long cm = nneighbors_AFEM + 1;
    int *t = itmp(nparticle * cm);
    CHK("conn", lpmx_get(LPMX_F_CONN, t, nparticle * cm));
    for (long i = 0; i < nparticle; i++)
        memcpy(conn[i], t + i * cm, sizeof(int) * cm);
    t = itmp(2L * (nparticle + 1));
    CHK("K_pointer", lpmx_get(LPMX_F_K_POINTER, t, 2L * (nparticle + 1)));
    for (long i = 0; i <= nparticle; i++)
        K_pointer[i][0] = t[2 * i], K_pointer[i][1] = t[2 * i + 1];
    JK = allocInt1D((int)len, -1);
    IK = allocInt1D(dim * nparticle + 1, -1);
    K_global = allocDouble1D((int)len, -1);
    return (int)len;
}

/* ------------------------------------------------------------- bond constants --- */
static void kntve_table(void)
{
    const int w = lattice == 1 ? 2 : 3;
    KnTve = allocDouble2D(ntype, w, 0);
    for (int t = 0; t < ntype; t++)
    {
        const double c11 = Ce[t][0], c12 = Ce[t][1], c44 = Ce[t][2], r = radius;
        double *k = KnTve[t];
        if (lattice == 0)
        {
            k[0] = 1 / 2.0 * c11 + -1 / 2.0 * c12 + 0.0 * c44;
            k[1] = 0.0 * c11 + 0.0 * c12 + 1 / 2.0 * c44;
            k[2] = 0.0 * c11 + 1.0 / 12.0 * c12 + -1.0 / 12.0 * c44;
        }
        else if (lattice == 1)
        {
            k[0] = sqrt(3.0) / 12.0 * c11 + -sqrt(3.0) / 12.0 * c12;
            k[1] = -sqrt(3.0) / 144.0 * c11 + sqrt(3.0) / 48.0 * c12;
        }
        else if (lattice == 2)
        {
            k[0] = r * (1 * c11 + -1 * c12 + -1 * c44);
            k[1] = r * (0 * c11 + 0 * c12 + 1 * c44);
            k[2] = r * (0 * c11 + 1.0 / 18.0 * c12 + -1.0 / 18.0 * c44);
        }
        else if (lattice == 3)
        {
            k[0] = r * (0 * c11 + 0 * c12 + sqrt(2.0) * c44);
            k[1] = r * (sqrt(2.0) / 4.0 * c11 + -sqrt(2.0) / 4.0 * c12 + -sqrt(2.0) / 4.0 * c44);
            k[2] = r * (0 * c11 + sqrt(2.0) / 24.0 * c12 + -sqrt(2.0) / 24.0 * c44);
        }
        else
        {
            k[0] = r * (0. * c11 + 0. * c12 + sqrt(3.0) * c44);
            k[1] = r * (1. / sqrt(3.0) * c11 + -1. / sqrt(3.0) * c12 + 0. * c44);
            k[2] = r * (0. * c11 + sqrt(3.0) / 14.0 * c12 + sqrt(3.0) / 14.0 * c44);
        }
    }
}

/* calcKnTv (reference src/stiffness.c:11-268): per-type table on the host, per-bond
 * constants on the device. */
void calcKnTv()
{
    need_model("calcKnTv");
    kntve_table();
    double *ce = malloc(sizeof(double) * 3 * (ntype > 0 ? ntype : 1));
    for (int t = 0; t < ntype; t++)
        for (int c = 0; c < 3; c++)
            ce[3 * t + c] = Ce[t][c];
    CHK("calcKnTv", lpmx_set_materials(ntype, type, ce, sigmay, ext_modulus));
    free(ce);
    PULL_NOW(LPMX_F_KN, LPMX_F_TV);
}

/* --------------------------------------------------------------------- tangent --- */
static void stiffness(int pm)
{
    need_model("calcStiffnessFiniteDifference");
    PUSH(LPMX_F_XYZ, LPMX_F_DLP0, LPMX_F_DAMAGE_BROKEN, LPMX_F_KN, LPMX_F_TV);
    CHK("tangent", lpmx_tangent());
    if (!resident || !k_have_pattern)
    {
        CHK("tangent export", lpmx_export_upper_csr(IK, JK, K_global));
        k_have_pattern = 1, k_h_stale = 0, k_d_fresh = 1;
    }
    else
        k_export();
    if (pm != 6) /* the reference only evaluates the elastic law (plmode 6) */
    {
        memset(K_global, 0, sizeof(double) * K_pointer[nparticle][1]);
        k_h_stale = 0, k_d_fresh = 0;
    }
}

/* calcStiffness{3,2}DFiniteDifference (reference src/stiffness.c:271-516): the
 * analytic Jacobian of the same elastic law replaces the forward differences. */
void calcStiffness3DFiniteDifference(int pm) { stiffness(pm); }
void calcStiffness2DFiniteDifference(int pm) { stiffness(pm); }

/* ---------------------------------------------------------- boundary conditions --- */
void setDispBC(int nboundDisp, struct dispBCPara *dBP)
{
    need_model("setDispBC");
    PUSH(LPMX_F_XYZ, LPMX_F_TYPE, LPMX_F_DISPBC_INDEX);
    lpmx_dispbc e[MAXSMALL];
    int k = nboundDisp < MAXSMALL ? nboundDisp : MAXSMALL;
    for (int i = 0; i < k; i++)
        e[i].type = dBP[i].type, e[i].flag = dBP[i].flag, e[i].step = dBP[i].step;
    CHK("setDispBC", lpmx_set_disp_bc(k, e));
    PULL(LPMX_F_XYZ, LPMX_F_DISPBC_INDEX);
    free(reaction_force);
    reaction_force = allocDouble1D(countNEqual(dispBC_index, nparticle * dim, 1), 0.);
}

void setForceBC(int nboundForce, struct forceBCPara *fBP)
{
    need_model("setForceBC");
    PUSH(LPMX_F_TYPE, LPMX_F_PEX);
    lpmx_forcebc e[MAXSMALL];
    int k = nboundForce < MAXSMALL ? nboundForce : MAXSMALL;
    for (int i = 0; i < k; i++)
    {
        e[i].type = fBP[i].type;
        e[i].step[0] = fBP[i].step1, e[i].step[1] = fBP[i].step2, e[i].step[2] = fBP[i].step3;
    }
    CHK("setForceBC", lpmx_set_force_bc(k, e));
    lpm_dropin_pull(LPMX_F_PEX);
}

static void bc_stiffness(void)
{
    need_model("setDispBC_stiffnessUpdate");
    k_import();
    PUSH(LPMX_F_DISPBC_INDEX, LPMX_F_FIX_INDEX, LPMX_F_RESIDUAL);
    CHK("setDispBC_stiffnessUpdate", lpmx_bc_eliminate());
    k_export();
    lpm_dropin_pull(LPMX_F_RESIDUAL);
}

void setDispBC_stiffnessUpdate3D() { bc_stiffness(); }
void setDispBC_stiffnessUpdate2D() { bc_stiffness(); }

/* ---------------------------------------------------------------------- solvers --- */
static void solve(double rtol2, double atol2, int announce)
{
    need_model("solverCG");
    k_import();
    PUSH(LPMX_F_RESIDUAL, LPMX_F_XYZ);
    int prec = ext_solver_set ? ext_solver.precision : LPMX_FP64;
    double inner = ext_solver_set ? ext_solver.inner_rtol2 : 1e-2;
    int irmax = ext_solver_set ? ext_solver.ir_maxit : 100;
    int it = 0, st = 0;
    double rr = 0;
    /* a solve to tolerance: the single-V-cycle mode (2) of the fused loop means AMG here */
    CHK("preconditioner", lpmx_set_preconditioner(ext_solver_set ? (ext_solver.precond ? 1 : 0) : 0));
    CHK("solverCG", lpmx_solve_ex(prec, rtol2, atol2, dim * nparticle, inner, irmax, &it, &rr, &st));
    if (announce)
    {
        if (st == 1)
            printf("The system has been solved after %d iterations\n", it);
        else
            printf("The computation FAILED as the solver has returned the ERROR code %d\n", -st);
    }
    PULL(LPMX_F_DISP, LPMX_F_XYZ);
}

/* solverCG (reference src/solver.c:188-270): the dcg residual test with the
 * reference's dpar[0] = 1e-8, dpar[1] = 1e-12 on squared norms, cap dim*nparticle;
 * block-Jacobi PCG on the device instead of unpreconditioned CG. */
void solverCG()
{
    double rtol2 = ext_solver_set ? ext_solver.rtol2 : 1e-8;
    double atol2 = ext_solver_set ? ext_solver.atol2 : 1e-12;
    solve(rtol2, atol2, 1);
}

/* solverPARDISO (reference src/solver.c:3-92) asks for a direct solve; the same
 * system is solved here by PCG to a relative residual of 1e-14. */
void solverPARDISO()
{
    solve(1e-28, 0.0, 0);
    printf("\nSolve completed at iteration: %d\n", 1);
}

/* ------------------------------------------------------------- state and laws --- */

/* switchStateV (reference src/constitutive.c:10-85). */
void switchStateV(int conv_flag)
{
    need_model("switchStateV");
    PUSH(LPMX_F_DLP0, LPMX_F_DLP1, LPMX_F_DLP2, LPMX_F_DAMAGE_D0, LPMX_F_DAMAGE_D1, LPMX_F_J2_BETA0, LPMX_F_J2_BETA1,
         LPMX_F_J2_BETA2, LPMX_F_J2_ALPHA0, LPMX_F_J2_ALPHA1, LPMX_F_J2_ALPHA2, LPMX_F_DAMAGE_NONLOCAL0,
         LPMX_F_DAMAGE_NONLOCAL1, LPMX_F_CP_GY0, LPMX_F_CP_GY1, LPMX_F_CP_GY2, LPMX_F_CP_A_SINGLE0,
         LPMX_F_CP_A_SINGLE1, LPMX_F_CP_A_SINGLE2, LPMX_F_CP_A0, LPMX_F_CP_A1, LPMX_F_CP_A2, LPMX_F_J2_BETA_EQ0,
         LPMX_F_J2_BETA_EQ1, LPMX_F_J2_BETA_EQ2, LPMX_F_DAMAGE_LOCAL0, LPMX_F_DAMAGE_LOCAL1);
    CHK("switchStateV", lpmx_switch_state(conv_flag));
    if (conv_flag == 0)
    {
        PULL(LPMX_F_J2_BETA_EQ0, LPMX_F_DAMAGE_LOCAL0);
        PULL(LPMX_F_DLP0, LPMX_F_DAMAGE_D0, LPMX_F_J2_BETA0, LPMX_F_J2_ALPHA0, LPMX_F_DAMAGE_NONLOCAL0, LPMX_F_CP_GY0,
             LPMX_F_CP_A_SINGLE0, LPMX_F_CP_A0);
    }
    else if (conv_flag == 1)
    {
        PULL(LPMX_F_J2_BETA_EQ1, LPMX_F_DAMAGE_LOCAL1);
        PULL(LPMX_F_DLP1, LPMX_F_DAMAGE_D1, LPMX_F_J2_BETA1, LPMX_F_J2_ALPHA1, LPMX_F_DAMAGE_NONLOCAL1, LPMX_F_CP_GY1,
             LPMX_F_CP_A_SINGLE1, LPMX_F_CP_A1);
    }
    else
        PULL(LPMX_F_DLP0, LPMX_F_J2_BETA0, LPMX_F_J2_ALPHA0, LPMX_F_CP_GY0, LPMX_F_CP_A_SINGLE0, LPMX_F_CP_A0,
             LPMX_F_J2_BETA_EQ0);
}

static void law_inputs(void)
{
    PUSH(LPMX_F_XYZ, LPMX_F_XYZ_TEMP, LPMX_F_DLP0, LPMX_F_DLP2, LPMX_F_J2_BETA0, LPMX_F_J2_BETA2, LPMX_F_J2_ALPHA0,
         LPMX_F_J2_ALPHA2, LPMX_F_DAMAGE_BROKEN, LPMX_F_DAMAGE_W, LPMX_F_KN, LPMX_F_TV, LPMX_F_CSX, LPMX_F_CSY,
         LPMX_F_CSZ, LPMX_F_F_TEMP, LPMX_F_SIGMAY, LPMX_F_NB, LPMX_F_PL_FLAG, LPMX_F_DL_AVE, LPMX_F_J2_BETA_EQ0,
         LPMX_F_J2_BETA_EQ2, LPMX_F_DAMAGE_D0);
}

static void law_outputs(void)
{
    PULL(LPMX_F_F, LPMX_F_PIN, LPMX_F_DL, LPMX_F_DL_AVE, LPMX_F_DDLP, LPMX_F_DDL, LPMX_F_CSX, LPMX_F_CSY, LPMX_F_CSZ,
         LPMX_F_DL_TOTAL, LPMX_F_TDL_TOTAL, LPMX_F_DDL_TOTAL, LPMX_F_TDDL_TOTAL, LPMX_F_DLP2, LPMX_F_J2_BETA2,
         LPMX_F_J2_ALPHA2, LPMX_F_J2_DLAMBDA, LPMX_F_PL_FLAG, LPMX_F_J2_BETA_EQ2);
}

/* crystal-plasticity state around a plmode 1 evaluation */
static void cp_inputs(void)
{
    PUSH(LPMX_F_CP_GY0, LPMX_F_CP_A_SINGLE0, LPMX_F_CP_A0);
}

static void cp_outputs(void)
{
    PULL(LPMX_F_CP_GY2, LPMX_F_CP_A_SINGLE2, LPMX_F_CP_A2, LPMX_F_CP_DA, LPMX_F_CP_DGY, LPMX_F_CP_DA_SINGLE,
         LPMX_F_CP_RSS, LPMX_F_CP_JACT);
    int fails = 0;
    CHK("crystal plasticity", lpmx_cp_failures(&fails));
    if (fails > 0)
    {
        /* the reference's LAPACKE_dgesv check exits here (constitutive.c:1235-1241) */
        printf("The diagonal element of the triangular factor of Jacobian matrix for cp_gamma \n");
        printf("is zero in %d return maps, so that it is singular.\n", fails);
        exit(1);
    }
}

void lpm_dropin_slip_systems(void)
{
    if (!d_have_model)
        return;
    double *t = dtmp(6L * nslipSys);
    for (int s = 0; s < nslipSys; s++)
        memcpy(t + 6 * s, schmid_tensor[s], sizeof(double) * 6);
    lpmx_cp_params p = {cp_tau0[0], cp_taus[0], cp_h0, cp_q, cp_eta, cp_p, dtime, (int)ceil(cp_maxloop)};
    CHK("slipSysDefine3D", lpmx_set_slip_systems(nslipSys, t, &p));
}

static void stress_outputs(void)
{
    PULL(LPMX_F_STRESS, LPMX_F_J2_STRESSEQ, LPMX_F_J2_STRESSM, LPMX_F_J2_TRIAXIALITY, LPMX_F_BOND_STRESS);
}

/* computeBondForceGeneral (reference src/constitutive.c:88-146). */
void computeBondForceGeneral(int pm, int t)
{
    if (pm != 0 && pm != 1 && pm != 3 && pm != 4 && pm != 5 && pm != 6)
        return; /* the reference computes no force for other values, and still forms the stress */
    need_model("computeBondForceGeneral");
    push_params();
    law_inputs();
    if (pm == 1)
        cp_inputs();
    CHK("computeBondForceGeneral", lpmx_bond_force_ex(pm, t));
    law_outputs();
    if (pm == 1)
        cp_outputs();
    stress_outputs();
    PULL(LPMX_F_DLP0, LPMX_F_J2_BETA0, LPMX_F_J2_ALPHA0, LPMX_F_CP_GY0, LPMX_F_CP_A_SINGLE0, LPMX_F_CP_A0,
         LPMX_F_J2_BETA_EQ0);
}

static void one_particle(int pm, int i)
{
    need_model("computeBondForce");
    law_inputs();
    CHK("computeBondForce", lpmx_bond_force_range(pm, i, 1));
    law_outputs();
}

void computeBondForceElastic(int i) { one_particle(6, i); }
void computeBondForceJ2mixedLinear3D(int ii) { one_particle(0, ii); }
void computeBondForceIncrementalUpdating(int ii) { one_particle(4, ii); }
/* The per-particle entry points of plmode 3 and 5 evaluate every particle (each
 * particle's final force needs its partners' updates); particle ii's force and state
 * are the reference's. */
void computeBondForceJ2nonlinearIso(int ii)
{
    (void)ii;
    need_model("computeBondForceJ2nonlinearIso");
    push_params();
    law_inputs();
    CHK("computeBondForceJ2nonlinearIso", lpmx_bond_force_range(5, 0, nparticle));
    law_outputs();
    PULL(LPMX_F_DLP0, LPMX_F_J2_BETA0, LPMX_F_J2_ALPHA0);
}
void computeBondForceCPMiehe(int ii)
{
    need_model("computeBondForceCPMiehe");
    law_inputs();
    cp_inputs();
    CHK("computeBondForceCPMiehe", lpmx_bond_force_range(1, ii, 1));
    law_outputs();
    cp_outputs();
}
void computeBondForceJ2energyReturnMap(int ii, int load_indicator)
{
    (void)ii;
    need_model("computeBondForceJ2energyReturnMap");
    law_inputs();
    CHK("computeBondForceJ2energyReturnMap", lpmx_bond_force_range_ex(3, load_indicator, 0, nparticle));
    law_outputs();
}

/* computeCab (reference src/constitutive.c:1864-1917) on the device. */
void computeCab()
{
    if (nslipSys <= 0)
        return;
    need_model("computeCab");
    PUSH(LPMX_F_CSX, LPMX_F_CSY, LPMX_F_CSZ, LPMX_F_DISTANCE, LPMX_F_KN, LPMX_F_TV, LPMX_F_NB, LPMX_F_DAMAGE_BROKEN);
    CHK("computeCab", lpmx_compute_cab());
    PULL(LPMX_F_CP_CAB);
}

/* computeStress (reference src/lpm_basic.c:53-125). */
void computeStress()
{
    need_model("computeStress");
    PUSH(LPMX_F_F, LPMX_F_CSX, LPMX_F_CSY, LPMX_F_CSZ, LPMX_F_NB, LPMX_F_DAMAGE_BROKEN);
    CHK("computeStress", lpmx_compute_stress());
    stress_outputs();
}

/* computedL (reference src/lpm_basic.c:252-291). */
void computedL()
{
    need_model("computedL");
    PUSH(LPMX_F_XYZ, LPMX_F_DLP0, LPMX_F_TV);
    CHK("computedL", lpmx_computedL());
    PULL(LPMX_F_DISTANCE, LPMX_F_DL, LPMX_F_DL_TOTAL, LPMX_F_TDL_TOTAL, LPMX_F_CSX, LPMX_F_CSY, LPMX_F_CSZ);
}

/* computeStrain (reference src/lpm_basic.c:127-249): the weighted least-squares
 * strain of every particle, on the device from the bond stretches dL. */
void computeStrain()
{
    need_model("computeStrain");
    PUSH(LPMX_F_DL);
    double *s = malloc(sizeof(double) * 6 * (size_t)nparticle);
    for (int i = 0; i < nparticle; i++)
        for (int c = 0; c < 6; c++)
            s[6 * i + c] = strain_tensor[i][c];
    CHK("computeStrain", lpmx_compute_strain(s));
    for (int i = 0; i < nparticle; i++)
        for (int c = 0; c < 6; c++)
            strain_tensor[i][c] = s[6 * i + c];
    free(s);
}

/* updateRR (reference src/stiffness.c:519-534). */
void updateRR()
{
    need_model("updateRR");
    PUSH(LPMX_F_PEX, LPMX_F_PIN, LPMX_F_DISPBC_INDEX);
    int nr = 0;
    CHK("updateRR", lpmx_update_rr(NULL, NULL, &nr));
    lpm_dropin_pull(LPMX_F_RESIDUAL);
    if (reaction_force && nr > 0)
        CHK("reaction", lpmx_get_reaction_force(reaction_force, nr));
}

/* ------------------------------------------------------------------- damage --- */
static FILE *bond_file(const char *name, int tstep)
{
    FILE *fp = fopen(name, "a+");
    if (fp)
        fprintf(fp, "TIMESTEP %d\n", tstep);
    return fp;
}

/* updateDuctileDamagePwiseNonlocal (reference src/constitutive.c:1757-1862). */
int updateDuctileDamagePwiseNonlocal(const char *dataName, int tstep)
{
    need_model("updateDuctileDamagePwiseNonlocal");
    FILE *fp = bond_file(dataName, tstep);
    PUSH(LPMX_F_DAMAGE_NONLOCAL0, LPMX_F_J2_DLAMBDA, LPMX_F_J2_TRIAXIALITY, LPMX_F_DAMAGE_BROKEN, LPMX_F_DAMAGE_D0,
         LPMX_F_DAMAGE_W);
    /* the broken-bond file lists newly broken bonds in (i, slot) order */
    const long nbn = (long)nparticle * nneighbors;
    unsigned char *was = malloc(nbn > 0 ? nbn : 1);
    host_read(LPMX_F_DAMAGE_BROKEN);
    for (long i = 0; i < nparticle; i++)
        for (long j = 0; j < nneighbors; j++)
            was[i * nneighbors + j] = damage_broken[i][j] > EPS || damage_broken[i][j] < -EPS;
    int k = 0;
    CHK("damage", lpmx_damage(0, &k));
    PULL(LPMX_F_DAMAGE_NONLOCAL0, LPMX_F_DAMAGE_BROKEN, LPMX_F_DAMAGE_D0, LPMX_F_DAMAGE_W);
    if (fp)
    {
        host_read(LPMX_F_DAMAGE_BROKEN);
        host_read(LPMX_F_NB_INITIAL);
        host_read(LPMX_F_NEIGHBORS);
        for (long i = 0; i < nparticle; i++)
            for (long j = 0; j < nb_initial[i]; j++)
                if (was[i * nneighbors + j] && damage_broken[i][j] == 0.0)
                    fprintf(fp, "%ld %d \n", i, neighbors[i][j]);
        fclose(fp);
    }
    free(was);
    return k;
}

/* updateBrittleDamage (reference src/constitutive.c:1437-1526). */
int updateBrittleDamage(const char *dataName, int tstep, int nbreak_)
{
    need_model("updateBrittleDamage");
    FILE *fp = bond_file(dataName, tstep);
    lpmx_params p = {plmode, J2_H, J2_xi, damage_L, damage_threshold, damagec_A, critical_bstrain, nbreak_, J2_C};
    CHK("params", lpmx_set_params(&p));
    if (ext_crit)
        CHK("critical stretch", lpmx_set_bond_critical(ext_crit));
    PUSH(LPMX_F_DL, LPMX_F_DAMAGE_BROKEN, LPMX_F_DAMAGE_W, LPMX_F_DAMAGE_D0);
    int k = 0;
    CHK("damage", lpmx_damage(6, &k));
    PULL(LPMX_F_DAMAGE_BROKEN, LPMX_F_DAMAGE_W, LPMX_F_DAMAGE_D0);
    int cnt = 0;
    CHK("broken", lpmx_get_broken(NULL, NULL, 0, &cnt));
    if (fp && cnt > 0)
    {
        int *bi = malloc(sizeof(int) * cnt), *bj = malloc(sizeof(int) * cnt);
        CHK("broken", lpmx_get_broken(bi, bj, cnt, &cnt));
        for (int t = 0; t < cnt; t++)
            fprintf(fp, "%d %d \n", bi[t], bj[t]);
        free(bi), free(bj);
    }
    if (fp)
        fclose(fp);
    return k;
}

/* The other ductile damage updates (reference src/constitutive.c:1529-1755) on the
 * device, with the bond-file lines the reference writes as it breaks in index order:
 * each newly broken entry "i neighbour" (bond-wise nonlocal), each newly broken pair
 * once, from its lower-index end (bond-wise local), each newly failed particle "i"
 * (point-wise local). */
static int ductile_variant(const char *dataName, int tstep, int law, const char *what)
{
    need_model(what);
    push_params();
    FILE *fp = bond_file(dataName, tstep);
    PUSH(LPMX_F_DAMAGE_NONLOCAL0, LPMX_F_DAMAGE_LOCAL0, LPMX_F_J2_DLAMBDA, LPMX_F_J2_TRIAXIALITY,
         LPMX_F_DAMAGE_BROKEN, LPMX_F_DAMAGE_D0, LPMX_F_DAMAGE_W, LPMX_F_NB);
    const long nbn = (long)nparticle * nneighbors;
    unsigned char *was = malloc(nbn > 0 ? nbn : 1);
    unsigned char *failed_before = malloc(nparticle > 0 ? nparticle : 1);
    host_read(LPMX_F_DAMAGE_BROKEN);
    host_read(LPMX_F_DAMAGE_LOCAL0);
    for (long i = 0; i < nparticle; i++)
    {
        for (long j = 0; j < nneighbors; j++)
            was[i * nneighbors + j] = damage_broken[i][j] > EPS || damage_broken[i][j] < -EPS;
        failed_before[i] = fabs(damage_local[i][0] - 1.0) <= EPS;
    }
    int k = 0;
    CHK(what, lpmx_damage_law(law, &k));
    PULL(LPMX_F_DAMAGE_NONLOCAL0, LPMX_F_DAMAGE_LOCAL0, LPMX_F_DAMAGE_BROKEN, LPMX_F_DAMAGE_D0, LPMX_F_DAMAGE_W,
         LPMX_F_NB);
    if (fp)
    {
        host_read(LPMX_F_DAMAGE_LOCAL0);
        host_read(LPMX_F_DAMAGE_BROKEN);
        host_read(LPMX_F_NB_INITIAL);
        host_read(LPMX_F_NEIGHBORS);
        for (long i = 0; i < nparticle; i++)
        {
            if (law == LPMX_DMG_PWISE_LOCAL)
            {
                if (!failed_before[i] && damage_local[i][0] == 1.0)
                    fprintf(fp, "%ld \n", i);
                continue;
            }
            for (long j = 0; j < nb_initial[i]; j++)
                if (was[i * nneighbors + j] && damage_broken[i][j] == 0.0 &&
                    (law != LPMX_DMG_BWISE_LOCAL || i < neighbors[i][j]))
                    fprintf(fp, "%ld %d \n", i, neighbors[i][j]);
        }
        fclose(fp);
    }
    free(was);
    free(failed_before);
    return k;
}

/* updateDuctileDamageBwiseLocal (reference src/constitutive.c:1607-1695). */
int updateDuctileDamageBwiseLocal(const char *dataName, int tstep)
{
    return ductile_variant(dataName, tstep, LPMX_DMG_BWISE_LOCAL, "updateDuctileDamageBwiseLocal");
}

/* updateDuctileDamagePwiseLocal (reference src/constitutive.c:1698-1755). */
int updateDuctileDamagePwiseLocal(const char *dataName, int tstep)
{
    return ductile_variant(dataName, tstep, LPMX_DMG_PWISE_LOCAL, "updateDuctileDamagePwiseLocal");
}

/* updateDuctileDamageBwiseNonlocal (reference src/constitutive.c:1529-1605). */
int updateDuctileDamageBwiseNonlocal(const char *dataName, int tstep)
{
    return ductile_variant(dataName, tstep, LPMX_DMG_BWISE_NONLOCAL, "updateDuctileDamageBwiseNonlocal");
}

/* updateDamageGeneral (reference src/constitutive.c:149-164). */
int updateDamageGeneral(const char *dataName, int tstep, int pm)
{
    if (pm == 0)
        return updateDuctileDamagePwiseNonlocal(dataName, tstep);
    if (pm == 5)
        return updateDuctileDamageBwiseLocal(dataName, tstep);
    if (pm == 6)
        return updateBrittleDamage(dataName, tstep, nbreak);
    return 0;
}

/* updateCrack (reference src/constitutive.c:1399-1434). */
void updateCrack()
{
    need_model("updateCrack");
    PUSH(LPMX_F_DAMAGE_BROKEN, LPMX_F_DAMAGE_W, LPMX_F_F, LPMX_F_CSX, LPMX_F_CSY, LPMX_F_CSZ, LPMX_F_FIX_INDEX);
    CHK("updateCrack", lpmx_update_crack());
    PULL(LPMX_F_NB, LPMX_F_DAMAGE_VISUAL, LPMX_F_F, LPMX_F_PIN, LPMX_F_FIX_INDEX);
}

/* ------------------------------------------------------------------ extensions --- */
/* A copy of count doubles, or NULL for a NULL source. */
static int ext_copy(double **dst, const double *src, long count, const char *what)
{
    free(*dst);
    *dst = NULL;
    if (!src)
        return LPMX_OK;
    if (!d_have_model || d_n != nparticle || count <= 0)
    {
        fprintf(stderr, "%s: no lattice yet; call searchNormalNeighbor() first\n", what);
        return LPMX_ERR_STATE;
    }
    *dst = malloc(sizeof(double) * count);
    if (!*dst)
        return LPMX_ERR_ARG;
    memcpy(*dst, src, sizeof(double) * count);
    return LPMX_OK;
}

int lpm_ext_set_modulus(const double *factor)
{
    return ext_copy(&ext_modulus, factor, nparticle, "lpm_ext_set_modulus");
}

int lpm_ext_set_bond_critical(const double *crit)
{
    int rc = ext_copy(&ext_crit, crit, (long)nparticle * nneighbors, "lpm_ext_set_bond_critical");
    if (rc != LPMX_OK)
        return rc;
    return lpmx_set_bond_critical(ext_crit);
}

int lpm_ext_set_solver(const lpmx_solver *s)
{
    if (!s)
    {
        ext_solver_set = 0;
        return LPMX_OK;
    }
    ext_solver = *s;
    ext_solver_set = 1;
    return LPMX_OK;
}

static const int all_fields[] = {
    LPMX_F_XYZ, LPMX_F_XYZ_TEMP, LPMX_F_NB, LPMX_F_PL_FLAG, LPMX_F_DL, LPMX_F_DL_AVE, LPMX_F_DDL, LPMX_F_DDLP,
    LPMX_F_F, LPMX_F_F_TEMP, LPMX_F_CSX, LPMX_F_CSY, LPMX_F_CSZ, LPMX_F_BOND_STRESS, LPMX_F_DAMAGE_BROKEN,
    LPMX_F_DAMAGE_W, LPMX_F_DLP0, LPMX_F_DLP1, LPMX_F_DLP2, LPMX_F_DAMAGE_D0, LPMX_F_DAMAGE_D1, LPMX_F_DL_TOTAL,
    LPMX_F_TDL_TOTAL, LPMX_F_DDL_TOTAL, LPMX_F_TDDL_TOTAL, LPMX_F_J2_BETA0, LPMX_F_J2_BETA1, LPMX_F_J2_BETA2,
    LPMX_F_STRESS, LPMX_F_J2_ALPHA0, LPMX_F_J2_ALPHA1, LPMX_F_J2_ALPHA2, LPMX_F_J2_DLAMBDA, LPMX_F_J2_STRESSEQ,
    LPMX_F_J2_STRESSM, LPMX_F_J2_TRIAXIALITY, LPMX_F_DAMAGE_NONLOCAL0, LPMX_F_DAMAGE_NONLOCAL1,
    LPMX_F_DAMAGE_VISUAL, LPMX_F_SIGMAY, LPMX_F_PIN, LPMX_F_PEX, LPMX_F_PEX_TEMP, LPMX_F_RESIDUAL, LPMX_F_DISP,
    LPMX_F_DISPBC_INDEX, LPMX_F_FIX_INDEX};

/* Load steps first..first+nsteps-1 of the driver's BC tables, run with every field
 * resident on the device (the loop of src/lpmc_project.c:382-546); the host
 * globals are pushed before and pulled after. */
int lpm_ext_run_steps(const lpmx_solver *s, int first, int nsteps, const int *nbd, struct dispBCPara (*dBP)[MAXSMALL],
                      const int *nbf, struct forceBCPara (*fBP)[MAXSMALL], char axis, int out_type,
                      lpmx_step_stats *out)
{
    need_model("lpm_ext_run_steps");
    push_list(all_fields, (int)(sizeof(all_fields) / sizeof(int)));
    lpmx_dispbc *db = calloc((size_t)nsteps * MAXSMALL + 1, sizeof(lpmx_dispbc));
    lpmx_forcebc *fb = calloc((size_t)nsteps * MAXSMALL + 1, sizeof(lpmx_forcebc));
    int *nd = calloc(nsteps + 1, sizeof(int)), *nf = calloc(nsteps + 1, sizeof(int));
    for (int s2 = 0; s2 < nsteps; s2++)
    {
        nd[s2] = nbd[first + s2], nf[s2] = nbf[first + s2];
        for (int k = 0; k < MAXSMALL; k++)
        {
            const struct dispBCPara *a = &dBP[first + s2][k];
            const struct forceBCPara *b = &fBP[first + s2][k];
            db[s2 * MAXSMALL + k] = (lpmx_dispbc){a->type, a->flag, a->step};
            fb[s2 * MAXSMALL + k] = (lpmx_forcebc){b->type, {b->step1, b->step2, b->step3}};
        }
    }
    int rc = lpmx_run_steps(s, nsteps, nd, db, nf, fb, MAXSMALL, axis, out_type, out);
    free(db), free(fb), free(nd), free(nf);
    pull_list(all_fields, (int)(sizeof(all_fields) / sizeof(int)));
    return rc;
}
