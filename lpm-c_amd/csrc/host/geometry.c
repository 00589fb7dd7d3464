/*
 * Setup of the drop-in API: the lattice description of createCuboid, whose points
 * the device generates (lpmx_lattice_points), geometry cuts, crack definition and
 * the allocation of every per-particle array (reference src/initialization.c).  The
 * particle positions must come out bit for bit as the reference's, since every later
 * list and force depends on them, so the lattice rules (candidate grid over [-a, a],
 * a the box diagonal, generation order, closed box test) are kept.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lpm.h"
#include "dropin.h"

static void rotation_matrix(void)
{
    const double a1 = angle1, a2 = angle2, a3 = angle3;
    double *R = R_matrix;
    if (eulerflag == 0) /* X1-Y2-Z3 */
    {
        R[0] = cos(a2) * cos(a3);
        R[1] = -cos(a2) * sin(a3);
        R[2] = sin(a2);
        R[3] = cos(a1) * sin(a3) + sin(a1) * sin(a2) * cos(a3);
        R[4] = cos(a1) * cos(a3) - sin(a1) * sin(a2) * sin(a3);
        R[5] = -sin(a1) * cos(a2);
        R[6] = sin(a1) * sin(a3) - cos(a1) * sin(a2) * cos(a3);
        R[7] = sin(a1) * cos(a3) + cos(a1) * sin(a2) * sin(a3);
        R[8] = cos(a1) * cos(a2);
    }
    else if (eulerflag == 1) /* Z1-Y2-Z3, Kocks */
    {
        R[0] = -cos(a1) * cos(a2) * cos(a3) - sin(a1) * sin(a3);
        R[1] = cos(a3) * sin(a1) - sin(a3) * cos(a2) * cos(a1);
        R[2] = sin(a2) * cos(a1);
        R[3] = sin(a3) * cos(a1) - cos(a3) * cos(a2) * sin(a1);
        R[4] = -cos(a1) * cos(a3) - sin(a1) * cos(a2) * sin(a3);
        R[5] = sin(a2) * sin(a1);
        R[6] = cos(a3) * sin(a2);
        R[7] = sin(a3) * sin(a2);
        R[8] = cos(a2);
    }
    else /* Z1-X2-Z3, Bunge */
    {
        R[0] = -sin(a1) * cos(a2) * sin(a3) + cos(a1) * cos(a3);
        R[1] = -sin(a3) * cos(a1) - cos(a3) * cos(a2) * sin(a1);
        R[2] = sin(a1) * sin(a2);
        R[3] = cos(a3) * sin(a1) + sin(a3) * cos(a2) * cos(a1);
        R[4] = -sin(a1) * sin(a3) + cos(a1) * cos(a2) * cos(a3);
        R[5] = -cos(a1) * sin(a2);
        R[6] = sin(a2) * sin(a3);
        R[7] = sin(a2) * cos(a3);
        R[8] = cos(a2);
    }
}

static void rotate(const double *t, double *p)
{
    for (int r = 0; r < 3; r++)
        p[r] = R_matrix[3 * r] * t[0] + R_matrix[3 * r + 1] * t[1] + R_matrix[3 * r + 2] * t[2];
}

/* createCuboid (reference src/initialization.c:6-560). */
void createCuboid()
{
    rotation_matrix();
    double a = dim == 2 ? sqrt(box_x * box_x + box_y * box_y) : sqrt(box_x * box_x + box_y * box_y + box_z * box_z);
    double bt[6] = {-a, a, -a, a, -a, a};
    switch (lattice)
    {
    case 0:
        hx = 2 * radius, hy = hx;
        particles_first_row = (int)floor((bt[1] - bt[0]) / hx);
        rows = (int)floor((bt[3] - bt[2]) / hy);
        nneighbors = 8, nneighbors1 = 4, nneighbors2 = 4;
        neighbor1_cutoff = 2.0 * radius, neighbor2_cutoff = 2.0 * sqrt(2.0) * radius;
        nneighbors_AFEM = 16, nneighbors_AFEM1 = 12, nneighbors_AFEM2 = 12;
        particle_volume = 4 * (radius * radius) * box_z;
        break;
    case 1:
        hx = 2 * radius, hy = hx * sqrt(3.0) / 2.0;
        particles_first_row = 1 + (int)floor((bt[1] - bt[0]) / hx);
        rows = 1 + (int)floor((bt[3] - bt[2]) / hy);
        nneighbors = 12, nneighbors1 = 6, nneighbors2 = 6;
        neighbor1_cutoff = 2.0 * radius, neighbor2_cutoff = 2.0 * sqrt(3.0) * radius;
        nneighbors_AFEM = 30, nneighbors_AFEM1 = 18, nneighbors_AFEM2 = 18;
        particle_volume = 2 * sqrt(3) * (radius * radius) * box_z;
        break;
    case 2:
        hx = 2 * radius, hy = hx, hz = hx;
        particles_first_row = 1 + (int)floor((bt[1] - bt[0]) / hx);
        rows = 1 + (int)floor((bt[3] - bt[2]) / hy);
        layers = 1 + (int)floor((bt[5] - bt[4]) / hz);
        nneighbors = 18, nneighbors1 = 6, nneighbors2 = 12;
        neighbor1_cutoff = 2.0 * radius, neighbor2_cutoff = 2.0 * sqrt(2.0) * radius;
        nneighbors_AFEM = 60, nneighbors_AFEM1 = 24, nneighbors_AFEM2 = 54;
        particle_volume = pow(2 * radius, 3);
        break;
    case 3:
        hx = 4 * radius / sqrt(2.0), hy = hx, hz = hx;
        particles_first_row = 1 + (int)floor((bt[1] - bt[0]) / hx);
        rows = 1 + (int)floor((bt[3] - bt[2]) / (hy / 2.0));
        layers = 1 + (int)floor((bt[5] - bt[4]) / (hz / 2.0));
        nneighbors = 18, nneighbors1 = 12, nneighbors2 = 6;
        neighbor1_cutoff = 2.0 * radius, neighbor2_cutoff = 2.0 * sqrt(2.0) * radius;
        nneighbors_AFEM = 60, nneighbors_AFEM1 = 54, nneighbors_AFEM2 = 24;
        particle_volume = 4.0 * sqrt(2.0) * pow(radius, 3);
        break;
    default:
        hx = 4.0 / sqrt(3.0) * radius, hy = hx, hz = hx;
        particles_first_row = 1 + (int)floor((bt[1] - bt[0]) / hx);
        rows = 1 + (int)floor((bt[3] - bt[2]) / hy);
        layers = 1 + (int)floor((bt[5] - bt[4]) / (hz / 2.0));
        nneighbors = 14, nneighbors1 = 8, nneighbors2 = 6;
        neighbor1_cutoff = 2.0 * radius, neighbor2_cutoff = 4.0 / sqrt(3.0) * radius;
        nneighbors_AFEM = 40, nneighbors_AFEM1 = 34, nneighbors_AFEM2 = 24;
        particle_volume = 32.0 * sqrt(3.0) / 9.0 * pow(radius, 3);
        break;
    }
    lpmx_lattice_spec sp;
    memset(&sp, 0, sizeof(sp));
    sp.lattice = lattice, sp.dim = dim, sp.pfr = particles_first_row, sp.rows = rows;
    sp.layers = lattice >= 2 ? layers : 1;
    sp.hx = hx, sp.hy = hy, sp.hz = hz, sp.radius = radius;
    for (int c = 0; c < 6; c++)
        sp.bt[c] = bt[c], sp.box[c] = box[c];
    memcpy(sp.R, R_matrix, sizeof(sp.R));
    int count = 0, rc = lpmx_lattice_points(&sp, NULL, 0, &count);
    if (rc != LPMX_OK)
        lpm_dropin_fatal("createCuboid", rc);
    nparticle = count;
    xyz = allocDouble2D(nparticle, NDIM, 0.0);
    double *pts = malloc(sizeof(double) * 3 * (size_t)(count > 0 ? count : 1));
    rc = lpmx_lattice_points(&sp, pts, count, &count);
    if (rc != LPMX_OK)
        lpm_dropin_fatal("createCuboid", rc);
    for (int i = 0; i < nparticle; i++)
        for (int c = 0; c < 3; c++)
            xyz[i][c] = pts[3 * i + c];
    free(pts);
    if (lattice == 2)
    {
        /* the 18 reference bond directions of the SC lattice, rotated */
        const double s = 1 / sqrt(2);
        const double ref[18][3] = {{1, 0, 0},  {0, 1, 0},   {0, 0, 1},   {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
                                   {s, s, 0},  {s, 0, s},   {0, s, s},   {-s, -s, 0}, {-s, 0, -s}, {0, -s, -s},
                                   {s, -s, 0}, {s, 0, -s},  {0, s, -s},  {-s, s, 0}, {-s, 0, s}, {0, -s, s}};
        bond_vector = allocDouble2D(nneighbors, dim, 0.0);
        for (int i = 0; i < nneighbors; i++)
            rotate(ref[i], bond_vector[i]);
    }
}

/* Keep the particles for which keep() holds, in order (the cut-out helpers). */
static void filter_particles(int (*keepf)(const double *, const void *), const void *arg)
{
    int m = 0;
    for (int i = 0; i < nparticle; i++)
        if (keepf(xyz[i], arg))
        {
            if (m != i)
                memcpy(xyz[m], xyz[i], sizeof(double) * NDIM);
            m++;
        }
    for (int i = m; i < nparticle; i++)
        free(xyz[i]);
    nparticle = m;
}

static int keep_in_box(const double *p, const void *arg)
{
    (void)arg;
    return p[0] >= box[0] && p[0] <= box[1] && p[1] >= box[2] && p[1] <= box[3] && p[2] >= box[4] && p[2] <= box[5];
}

/* moveParticle (reference src/initialization.c:830-866). */
void moveParticle(double *movexyz)
{
    lpm_dropin_host_access(xyz, 1);
    for (int i = 0; i < nparticle; i++)
        for (int j = 0; j < 3; j++)
            xyz[i][j] = xyz[i][j] + movexyz[j];
    filter_particles(keep_in_box, NULL);
}

struct circle
{
    const double *pc;
    double ra;
    char c;
};

static int keep_out_circle(const double *p, const void *arg)
{
    const struct circle *a = arg;
    double d2 = 0.0;
    if (a->c == 'x')
        d2 = (p[1] - a->pc[1]) * (p[1] - a->pc[1]) + (p[2] - a->pc[2]) * (p[2] - a->pc[2]);
    else if (a->c == 'y')
        d2 = (p[0] - a->pc[0]) * (p[0] - a->pc[0]) + (p[2] - a->pc[2]) * (p[2] - a->pc[2]);
    else if (a->c == 'z')
        d2 = (p[0] - a->pc[0]) * (p[0] - a->pc[0]) + (p[1] - a->pc[1]) * (p[1] - a->pc[1]);
    return d2 > a->ra * a->ra;
}

void removeCircle(double *pc, double ra, char c)
{
    struct circle a = {pc, ra, c};
    filter_particles(keep_out_circle, &a);
}

struct arc
{
    const double *pc;
    double ra, t1, t2;
};

static int keep_out_arc(const double *p, const void *arg)
{
    const struct arc *a = arg;
    double dx = p[0] - a->pc[0], dy = p[1] - a->pc[1];
    if (fabs(dx) < EPS && fabs(dy) < EPS)
        return 0;
    double th = atan2(dy, dx);
    return dx * dx + dy * dy > a->ra * a->ra || (th >= a->t1 && th <= a->t2);
}

void removeCirclePartZ(double *pc, double ra, double theta1, double theta2)
{
    struct arc a = {pc, ra, theta1, theta2};
    filter_particles(keep_out_arc, &a);
}

static int keep_out_block(const double *p, const void *arg)
{
    const double *r = arg;
    return p[0] < r[0] || p[0] > r[1] || p[1] < r[2] || p[1] > r[3] || p[2] < r[4] || p[2] > r[5];
}

void removeBlock(double r0, double r1, double r2, double r3, double r4, double r5)
{
    double r[6] = {r0, r1, r2, r3, r4, r5};
    filter_particles(keep_out_block, r);
}

static int keep_in_cylinder(const double *p, const void *arg)
{
    const struct circle *a = arg;
    double dx = p[0] - a->pc[0], dy = p[1] - a->pc[1];
    return dx * dx + dy * dy < a->ra * a->ra;
}

void createCylinderz(double *pc, double ra)
{
    struct circle a = {pc, ra, 'z'};
    filter_particles(keep_in_cylinder, &a);
}

struct ring
{
    const double *pc;
    double R, r;
};

static int keep_out_ring(const double *p, const void *arg)
{
    const struct ring *a = arg;
    double h = p[2] - a->pc[2];
    if (!(fabs(h) < a->r))
        return 1;
    double dx = p[0] - a->pc[0], dy = p[1] - a->pc[1], w = a->R - sqrt(a->r * a->r - h * h);
    return dx * dx + dy * dy < w * w;
}

void removeRingz(double *pc, double R, double r)
{
    struct ring a = {pc, R, r};
    filter_particles(keep_out_ring, &a);
}

struct crack
{
    double a1, a2, w, h;
};

static int keep_out_crack(const double *p, const void *arg)
{
    const struct crack *a = arg;
    if (!(p[0] < a->a1 || p[0] > a->a2 || p[1] > a->h + a->w || p[1] < a->h - a->w))
        return 0;
    double dx = p[0] - a->a2, dy = p[1] - a->h;
    return dx * dx + dy * dy > a->w * a->w;
}

void createCrack(double a1, double a2, double w, double h)
{
    struct crack a = {a1, a2, w, h};
    filter_particles(keep_out_crack, &a);
}

/* defineCrack (reference src/initialization.c:1097-1117): bonds crossing y = h
 * between x = a1 and a2 start broken. */
void defineCrack(double a1, double a2, double h)
{
    for (int i = 0; i < nparticle; i++)
    {
        damage_visual[i] = 0.0;
        for (int j = 0; j < nb_initial[i]; j++)
        {
            int q = neighbors[i][j];
            if ((xyz[q][1] - h) * (xyz[i][1] - h) < 0 && xyz[i][0] > a1 && xyz[i][0] < a2 && xyz[q][0] > a1 &&
                xyz[q][0] < a2)
            {
                damage_D[i][j][0] = 1.0;
                damage_D[i][j][1] = 1.0;
                damage_broken[i][j] = 0.0;
                damage_w[i][j] = 0.0;
                nb[i] -= 1;
            }
            damage_visual[i] += damage_broken[i][j];
        }
        damage_visual[i] = 1 - damage_visual[i] / nb_initial[i];
    }
}

/* initMatrices (reference src/initialization.c:1120-1194): every per-particle
 * array with its initial value. */
void initMatrices()
{
    const int n = nparticle, nb_ = nneighbors;
    disp = allocDouble1D(dim * n, 0);
    xyz_initial = allocDouble2D(n, NDIM, 0);
    xyz_temp = allocDouble2D(n, NDIM, 0);
    distance = allocDouble2D(n, nb_, 0.0);
    distance_initial = allocDouble2D(n, nb_, 0.0);
    csx = allocDouble2D(n, nb_, 0.0);
    csy = allocDouble2D(n, nb_, 0.0);
    csz = allocDouble2D(n, nb_, 0.0);
    csx_initial = allocDouble2D(n, nb_, 0.0);
    csy_initial = allocDouble2D(n, nb_, 0.0);
    csz_initial = allocDouble2D(n, nb_, 0.0);
    dL = allocDouble2D(n, nb_, 0.0);
    dL_total = allocDouble2D(n, 2, 0);
    TdL_total = allocDouble2D(n, 2, 0);
    dL_ave = allocDouble2D(n, nb_, 0.);
    dLp = allocDouble3D(n, nb_, 3, 0.);
    ddL = allocDouble2D(n, nb_, 0.0);
    ddL_total = allocDouble2D(n, 2, 0);
    TddL_total = allocDouble2D(n, 2, 0);
    ddLp = allocDouble2D(n, nb_, 0.);
    bond_stress = allocDouble2D(n, nb_, 0);
    bond_stretch = allocDouble2D(n, nb_, 0);
    state_v = allocInt1D(n, 0);
    pl_flag = allocInt1D(n, 0);
    nb = allocInt1D(n, -1);
    nb_initial = allocInt1D(n, -1);
    nb_conn = allocInt1D(n, -1);
    neighbors = allocInt2D(n, nb_, -1);
    neighbors1 = allocInt2D(n, nneighbors1, -1);
    neighbors2 = allocInt2D(n, nneighbors2, -1);
    nsign = allocInt2D(n, nb_, -1);
    conn = allocInt2D(n, nneighbors_AFEM + 1, -1);
    K_pointer = allocInt2D(n + 1, 2, 0);
    residual = allocDouble1D(dim * n, 0);
    Pin = allocDouble1D(NDIM * n, 0);
    Pex = allocDouble1D(dim * n, 0);
    Pex_temp = allocDouble1D(dim * n, 0);
    dispBC_index = allocInt1D(dim * n, 1);
    fix_index = allocInt1D(dim * n, 1);
    Kn = allocDouble2D(n, nb_, 0.0);
    Tv = allocDouble2D(n, nb_, 0.0);
    stress_tensor = allocDouble2D(n, 2 * NDIM, 0.);
    strain_tensor = allocDouble2D(n, 2 * NDIM, 0.);
    damage_broken = allocDouble2D(n, nb_, 1.);
    damage_D = allocDouble3D(n, nb_, 2, 0.);
    damage_w = allocDouble2D(n, nb_, 1.);
    damage_visual = allocDouble1D(n, 0.0);
    damage_local = allocDouble2D(n, 2, 0.0);
    damage_nonlocal = allocDouble2D(n, 2, 0.0);
    F = allocDouble2D(n, nb_, 0.);
    F_temp = allocDouble2D(n, nb_, 0.);
    bond_force = allocDouble2D(n, nb_, 0.);
    J2_beta = allocDouble3D(n, 2 * NDIM, 3, 0.);
    J2_alpha = allocDouble2D(n, 3, 0.);
    J2_beta_eq = allocDouble2D(n, 3, 0.);
    J2_stresseq = allocDouble1D(n, 0.);
    J2_stressm = allocDouble1D(n, 0.);
    J2_triaxiality = allocDouble1D(n, 0.);
    J2_dlambda = allocDouble1D(n, 0.);
}

/* slipSysDefine3D (reference src/initialization.c:563-827).  FCC (lattice 3): the 4
 * {111} planes with 6 <110> directions each (both senses), in the reference's order,
 * which decides ties in the active-set search; normals and directions rotated by
 * R_matrix (y = R x); Schmid tensors s11 s22 s33 s23 s13 s12 with engineering shear
 * terms.  Other lattices get no slip systems (BCC's are not part of this build).
 * Then the crystal-plasticity arrays, and the slip systems go to the device. */
void slipSysDefine3D()
{
    static const int pl[4][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1}};
    static const int dr[4][6][3] = {
        {{1, -1, 0}, {-1, 1, 0}, {-1, 0, 1}, {1, 0, -1}, {0, 1, -1}, {0, -1, 1}},
        {{1, 0, 1}, {-1, 0, -1}, {-1, -1, 0}, {1, 1, 0}, {0, 1, -1}, {0, -1, 1}},
        {{-1, 0, 1}, {1, 0, -1}, {0, -1, -1}, {0, 1, 1}, {1, 1, 0}, {-1, -1, 0}},
        {{-1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {-1, 0, -1}, {0, -1, -1}, {0, 1, 1}}};
    if (lattice == 3)
    {
        nslip_face = 4;
        max_nslip_vector = 6;
        nslipSys = 24;
        cp_gy = allocDouble3D(nparticle, nslipSys, 3, cp_tau0[0]);
        nslip_vector = allocInt1D(nslip_face, 6);
        const double a = 1 / sqrt(3), b = 1 / sqrt(2);
        double nloc[4][3], dloc[4][6][3];
        slip_normal = allocDouble2D(nslip_face, 3, 0);
        slip_vector = allocDouble3D(nslip_face, max_nslip_vector, 3, 0.);
        for (int p = 0; p < 4; p++)
        {
            for (int c = 0; c < 3; c++)
                nloc[p][c] = pl[p][c] * a;
            for (int r = 0; r < 3; r++)
                slip_normal[p][r] = R_matrix[3 * r] * nloc[p][0] + R_matrix[3 * r + 1] * nloc[p][1] +
                                    R_matrix[3 * r + 2] * nloc[p][2];
            for (int k = 0; k < 6; k++)
            {
                for (int c = 0; c < 3; c++)
                    dloc[p][k][c] = dr[p][k][c] * b;
                for (int r = 0; r < 3; r++)
                    slip_vector[p][k][r] = R_matrix[3 * r] * dloc[p][k][0] + R_matrix[3 * r + 1] * dloc[p][k][1] +
                                           R_matrix[3 * r + 2] * dloc[p][k][2];
            }
        }
        schmid_tensor_local = allocDouble2D(nslipSys, 6, 0);
        schmid_tensor = allocDouble2D(nslipSys, 6, 0);
        int s = 0;
        for (int p = 0; p < nslip_face; p++)
            for (int k = 0; k < nslip_vector[p]; k++, s++)
            {
                const double *n0 = nloc[p], *d0 = dloc[p][k], *n1 = slip_normal[p], *d1 = slip_vector[p][k];
                double *tl = schmid_tensor_local[s], *t = schmid_tensor[s];
                tl[0] = n0[0] * d0[0], tl[1] = n0[1] * d0[1], tl[2] = n0[2] * d0[2];
                tl[3] = n0[1] * d0[2] + n0[2] * d0[1];
                tl[4] = n0[0] * d0[2] + n0[2] * d0[0];
                tl[5] = n0[0] * d0[1] + n0[1] * d0[0];
                t[0] = n1[0] * d1[0], t[1] = n1[1] * d1[1], t[2] = n1[2] * d1[2];
                t[3] = n1[1] * d1[2] + n1[2] * d1[1];
                t[4] = n1[0] * d1[2] + n1[2] * d1[0];
                t[5] = n1[0] * d1[1] + n1[1] * d1[0];
            }
    }
    else
    {
        if (lattice == 4)
            fprintf(stderr, "slipSysDefine3D: BCC slip systems are not part of this build; none defined\n");
        nslipSys = 0;
        cp_gy = allocDouble3D(nparticle, 0, 3, 0.0);
    }
    cp_RSS = allocDouble2D(nparticle, nslipSys, 0.);
    cp_Jact = allocInt2D(nparticle, nslipSys, 0);
    cp_A = allocDouble2D(nparticle, 3, 0.);
    cp_A_single = allocDouble3D(nparticle, nslipSys, 3, 0.);
    cp_Cab = allocDouble2D(nparticle, nslipSys * nslipSys, 0.);
    cp_dA_single = allocDouble2D(nparticle, nslipSys, 0.0);
    cp_dA = allocDouble1D(nparticle, 0.0);
    cp_dgy = allocDouble2D(nparticle, nslipSys, 0.0);
    if (nslipSys > 0)
        lpm_dropin_slip_systems();
}
