/*
 * Text outputs of the drop-in API in the reference's file formats
 * (reference src/data_handler.c), written from the host globals, which the
 * drop-in layer keeps current after every call.  Files are opened for append as
 * the reference does, so reruns append.
 */
#include <stdio.h>
#include <stdlib.h>

#include "lpm.h"
#include "dropin.h"

/* writeDump (reference src/data_handler.c:42-84): LAMMPS dump of the particles. */
void writeDump(const char *dataName, int step, char flag)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, flag == 's' ? "w+" : "a+");
    if (!f)
        return;
    fprintf(f, "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n", step, nparticle);
    for (int c = 0; c < 3; c++)
        fprintf(f, "%8.8f %8.8f\n", box[2 * c], box[2 * c + 1]);
    /* crystal plasticity (plmode 1) reports the accumulated slips of eight of the 24
     * systems (0-based 1, 4, 10, 19, 9, 14, 17, 22) and the total slip; the other
     * laws the J2 and damage columns (data_handler.c:59-83) */
    static const int slip_cols[8] = {1, 4, 10, 19, 9, 14, 17, 22};
    const int cp = plmode == 1;
    if (cp)
        fprintf(f, "ITEM: ATOMS id type x y z dx dy dz s11 s22 s33 s23 s13 s12 gam11 gam12 gam13 gam14 gam21 gam22 "
                   "gam23 gam24 pl gam_total\n");
    else
        fprintf(f, "ITEM: ATOMS id type x y z dx dy dz s11 s22 s33 s23 s13 s12 stress_eq strain_eq stress_m "
                   "stress_tri damage_local damage_nonlocal pl damage\n");
    for (int i = 0; i < nparticle; i++)
    {
        fprintf(f, "%d %d", i, type[i]);
        for (int c = 0; c < 3; c++)
            fprintf(f, " %.4e", xyz[i][c]);
        for (int c = 0; c < 3; c++)
            fprintf(f, " %.4e", xyz[i][c] - xyz_initial[i][c]);
        for (int c = 0; c < 6; c++)
            fprintf(f, " %.4e", stress_tensor[i][c]);
        if (cp)
        {
            for (int c = 0; c < 8; c++)
                fprintf(f, " %.4e", cp_A_single[i][slip_cols[c]][0]);
            fprintf(f, " %d %.4e\n", pl_flag[i], cp_A[i][0]);
        }
        else
            fprintf(f, " %.4e %.4e %.4e %.4e %.4e %.4e %d %.4e\n", J2_stresseq[i], J2_alpha[i][0], J2_stressm[i],
                    J2_triaxiality[i], damage_local[i][0], damage_nonlocal[i][0], pl_flag[i], damage_visual[i]);
    }
    fclose(f);
}

/* writeRSS (reference src/data_handler.c:87-107). */
void writeRSS(const char *dataName, int step)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    fprintf(f, "TIMESTEP %d\nid type ", step);
    for (int i = 0; i < nslipSys; i++)
        fprintf(f, "slip%u ", i + 1);
    fprintf(f, "\n");
    for (int i = 0; i < nparticle; i++)
    {
        fprintf(f, "%u %d ", i, type[i]);
        for (int j = 0; j < nslipSys; j++)
            fprintf(f, "%.6e ", cp_RSS[i][j]);
        fprintf(f, "\n");
    }
    fclose(f);
}

static void type_average(const char *dataName, double **a, int t1, int tstep)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    int num = 0;
    double s[2 * NDIM] = {0.0};
    for (int i = 0; i < nparticle; i++)
        if (type[i] == t1)
        {
            num++;
            for (int j = 0; j < 2 * NDIM; j++)
                s[j] += a[i][j];
        }
    fprintf(f, "%d ", tstep);
    for (int j = 0; j < 2 * NDIM; j++)
        fprintf(f, "%.6e ", s[j] / num);
    fprintf(f, "\n");
    fclose(f);
}

/* writeStress / writeStrain (reference src/data_handler.c:110-161). */
void writeStress(const char *dataName, int t1, int tstep) { type_average(dataName, stress_tensor, t1, tstep); }
void writeStrain(const char *dataName, int t1, int tstep) { type_average(dataName, strain_tensor, t1, tstep); }

void writeCab(const char *dataName, int ii)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    fprintf(f, "Particle ID: %d\n", ii);
    for (int m = 0; m < nslipSys; m++)
    {
        for (int n = 0; n < nslipSys; n++)
            fprintf(f, "%.4e ", cp_Cab[ii][m * nslipSys + n]);
        fprintf(f, "\n");
    }
    fprintf(f, "\n");
    fclose(f);
}

void writeInternalForce(const char *dataName, int step)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    fprintf(f, "TIMESTEP %d\nid type fx fy fz\n", step);
    for (int i = 0; i < nparticle; i++)
    {
        double fx = 0., fy = 0., fz = 0.;
        for (int j = 0; j < nb_initial[i]; j++)
        {
            fx += csx[i][j] * F[i][j];
            fy += csy[i][j] * F[i][j];
            fz += csz[i][j] * F[i][j];
        }
        fprintf(f, "%u %d %.6e %.6e %.6e\n", i, type[i], fx, fy, fz);
    }
    fclose(f);
}

void writeJact(const char *dataName, int step)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    fprintf(f, "timestep %d\nid type actived_slip_systems\n", step);
    for (int i = 0; i < nparticle; i++)
    {
        fprintf(f, "%u %d ", i, type[i]);
        for (int m = 0; m < nslipSys; m++)
            if (cp_Jact[i][m] == 1)
                fprintf(f, "%d ", m);
        fprintf(f, "\n");
    }
    fclose(f);
}

static void bond_rows(const char *dataName, int step, double **a, double **den)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    fprintf(f, "TIMESTEP %d\n", step);
    for (int i = 0; i < nparticle; i++)
    {
        fprintf(f, "%d %d ", i, type[i]);
        for (int m = 0; m < nb_initial[i]; m++)
            fprintf(f, "%.4e ", den ? a[i][m] / den[i][m] : a[i][m]);
        fprintf(f, "\n");
    }
    fclose(f);
}

/* writeBondstretch / writeBondforce (reference src/data_handler.c:234-275). */
void writeBondstretch(const char *dataName, int step) { bond_rows(dataName, step, dL, distance_initial); }

void writeBondforce(const char *dataName, int step)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    fprintf(f, "TIMESTEP %d\n", step);
    for (int i = 0; i < nparticle; i++)
    {
        fprintf(f, "%d %d ", i, type[i]);
        for (int m = 0; m < nneighbors; m++)
            fprintf(f, "%.4e ", bond_force[i][m]);
        fprintf(f, "\n");
    }
    fclose(f);
}

/* writeDisp (reference src/data_handler.c:278-306): mean displacement of type t1. */
void writeDisp(const char *dataName, char c, int t1, int tstep)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    int k = c == 'x' ? 0 : (c == 'y' ? 1 : 2), num = 0;
    double p = 0, p0 = 0;
    for (int i = 0; i < nparticle; i++)
        if (type[i] == t1)
        {
            p += xyz[i][k];
            p0 += xyz_initial[i][k];
            num++;
        }
    fprintf(f, "%d %8.8e\n", tstep, (p - p0) / num);
    fclose(f);
}

/* writeReaction (reference src/data_handler.c:309-335): bond force of type t1
 * projected on the axis. */
void writeReaction(const char *dataName, char c, int t1, int tstep)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    /* an axis character other than x, y, z adds nothing (data_handler.c:320-327) */
    double **cs = c == 'x' ? csx : (c == 'y' ? csy : (c == 'z' ? csz : NULL));
    double force = 0.;
    for (int i = 0; cs && i < nparticle; i++)
        if (type[i] == t1)
            for (int j = 0; j < nb_initial[i]; j++)
                force += -cs[i][j] * F[i][j];
    fprintf(f, "%d %8.8e \n", tstep, force);
    fclose(f);
}

/* writeForce (reference src/data_handler.c:338-393): force on the lower fraction p
 * of the body along c. */
void writeForce(const char *dataname, char c, double p, int tstep)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataname, "a+");
    if (!f)
        return;
    /* an axis character other than x, y, z selects no particle (data_handler.c:346-373) */
    int k = c == 'x' ? 0 : (c == 'y' ? 1 : (c == 'z' ? 2 : -1));
    double lim = k < 0 ? 0.0 : box[2 * k] + p * (k == 0 ? box_x : (k == 1 ? box_y : box_z));
    double **cs = k == 0 ? csx : (k == 1 ? csy : csz);
    double force = 0.;
    for (int i = 0; k >= 0 && i < nparticle; i++)
        if (xyz_initial[i][k] < lim)
            for (int j = 0; j < nb_initial[i]; j++)
                force += -cs[i][j] * F[i][j];
    fprintf(f, "%d %8.8e \n", tstep, force);
    fclose(f);
}

void writeK_global(const char *dataName, int l)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "w+");
    if (!f)
        return;
    for (int i = 0; i < l; i++)
        fprintf(f, " %.5e\n", K_global[i]);
    fclose(f);
}

void writeDlambda(const char *dataName, int m, int n, int globalStep, int iterStep)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    fprintf(f, "Step (Global-Iteration) %d-%d\n", globalStep, iterStep);
    for (int i = m; i < n; i++)
        fprintf(f, "%d %.4e %.4e %.4e\n", i, J2_dlambda[i], J2_alpha[i][0], J2_alpha[i][1]);
    fclose(f);
}

void writeDamage(const char *dataName, int step)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    fprintf(f, "TIMESTEP %d\n", step);
    for (int i = 0; i < nparticle; i++)
    {
        fprintf(f, "%d %d ", i, type[i]);
        for (int m = 0; m < nb_initial[i]; m++)
            fprintf(f, "%.4e ", damage_D[i][m][0]);
        fprintf(f, "\n");
    }
    fclose(f);
}

void writeNeighbor(const char *dataName)
{
    lpm_dropin_sync_host();
    FILE *f = fopen(dataName, "a+");
    if (!f)
        return;
    for (int i = 0; i < nparticle; i++)
    {
        fprintf(f, "%d %d ", i, type[i]);
        for (int m = 0; m < nb_initial[i]; m++)
            fprintf(f, "%d ", neighbors[i][m]);
        fprintf(f, "\n");
    }
    fclose(f);
}
