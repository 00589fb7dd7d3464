/*
 * Host helpers of the drop-in API: allocation, copies, counts, the type setters,
 * the two MKL routines drivers call directly, and the output-only strain and bond
 * re-ordering (reference src/lpm_basic.c).  None of these is on the load-step hot
 * path; they keep drivers linking and running unchanged.  Those that read or change
 * a model array tell the drop-in layer first (lpm_dropin_host_access), so that in
 * its resident mode they see the device's current values.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lpm.h"
#include "dropin.h"

/* ---------------------------------------------------------------- allocation --- */
int *allocInt1D(int num, int a)
{
    int *r = malloc(sizeof(int) * (num > 0 ? num : 1));
    for (int i = 0; i < num; i++)
        r[i] = a;
    return r;
}

double *allocDouble1D(int num, double a)
{
    double *r = malloc(sizeof(double) * (num > 0 ? num : 1));
    for (int i = 0; i < num; i++)
        r[i] = a;
    return r;
}

int **allocInt2D(int row, int column, int a)
{
    int **r = malloc(sizeof(int *) * (row > 0 ? row : 1));
    for (int i = 0; i < row; i++)
    {
        r[i] = malloc(sizeof(int) * (column > 0 ? column : 1));
        for (int j = 0; j < column; j++)
            r[i][j] = a;
    }
    return r;
}

int ***allocInt3D(int row, int column, int layer, int a)
{
    int ***r = malloc(sizeof(int **) * (row > 0 ? row : 1));
    for (int i = 0; i < row; i++)
    {
        r[i] = malloc(sizeof(int *) * (column > 0 ? column : 1));
        for (int j = 0; j < column; j++)
        {
            r[i][j] = malloc(sizeof(int) * (layer > 0 ? layer : 1));
            for (int k = 0; k < layer; k++)
                r[i][j][k] = a;
        }
    }
    return r;
}

double **allocDouble2D(int row, int column, double a)
{
    double **r = malloc(sizeof(double *) * (row > 0 ? row : 1));
    for (int i = 0; i < row; i++)
    {
        r[i] = malloc(sizeof(double) * (column > 0 ? column : 1));
        for (int j = 0; j < column; j++)
            r[i][j] = a;
    }
    return r;
}

double ***allocDouble3D(int row, int column, int layer, double a)
{
    double ***r = malloc(sizeof(double **) * (row > 0 ? row : 1));
    for (int i = 0; i < row; i++)
    {
        r[i] = malloc(sizeof(double *) * (column > 0 ? column : 1));
        for (int j = 0; j < column; j++)
        {
            r[i][j] = malloc(sizeof(double) * (layer > 0 ? layer : 1));
            for (int k = 0; k < layer; k++)
                r[i][j][k] = a;
        }
    }
    return r;
}

void freeInt2D(int **arr, int row)
{
    for (int i = 0; i < row; i++)
        free(arr[i]);
    free(arr);
}

void freeDouble2D(double **arr, int row)
{
    for (int i = 0; i < row; i++)
        free(arr[i]);
    free(arr);
}

void freeDouble3D(double ***arr, int row, int column)
{
    for (int i = 0; i < row; i++)
    {
        for (int j = 0; j < column; j++)
            free(arr[i][j]);
        free(arr[i]);
    }
    free(arr);
}

void copyInt2D(int **target, int **source, int row, int col)
{
    lpm_dropin_host_access(source, 0);
    lpm_dropin_host_access(target, 1);
    for (int i = 0; i < row; i++)
        memcpy(target[i], source[i], sizeof(int) * col);
}

void copyDouble1D(double *target, double *source, int row)
{
    lpm_dropin_host_access(source, 0);
    lpm_dropin_host_access(target, 1);
    memcpy(target, source, sizeof(double) * row);
}

void copyDouble2D(double **target, double **source, int row, int col)
{
    lpm_dropin_host_access(source, 0);
    lpm_dropin_host_access(target, 1);
    for (int i = 0; i < row; i++)
        memcpy(target[i], source[i], sizeof(double) * col);
}

void copyDouble3D(double ***target, double ***source, int row, int col, int deep)
{
    lpm_dropin_host_access(source, 0);
    lpm_dropin_host_access(target, 1);
    for (int i = 0; i < row; i++)
        for (int j = 0; j < col; j++)
            memcpy(target[i][j], source[i][j], sizeof(double) * deep);
}

/* ------------------------------------------------------------------- counting --- */
int findMaxDouble(double *arr, int len)
{
    lpm_dropin_host_access(arr, 0);
    int t = 0;
    for (int i = 0; i < len; i++)
        if (arr[i] > arr[t])
            t = i;
    return t;
}

int findMaxInt(int *arr, int len)
{
    lpm_dropin_host_access(arr, 0);
    int t = 0;
    for (int i = 0; i < len; i++)
        if (arr[i] > arr[t])
            t = i;
    return t;
}

int countEqual(int *arr, int len, int b)
{
    lpm_dropin_host_access(arr, 0);
    int c = 0;
    for (int i = 0; i < len; i++)
        c += arr[i] == b;
    return c;
}

int countNEqual(int *arr, int len, int b)
{
    lpm_dropin_host_access(arr, 0);
    int c = 0;
    for (int i = 0; i < len; i++)
        c += arr[i] != b;
    return c;
}

int countLarger(int *arr, int len, int b)
{
    lpm_dropin_host_access(arr, 0);
    int c = 0;
    for (int i = 0; i < len; i++)
        c += arr[i] > b;
    return c;
}

int findNElement(int *in, int *out, int len, int a)
{
    lpm_dropin_host_access(in, 0);
    int k = 0;
    for (int i = 0; i < len; i++)
        if (in[i] != a)
            out[k++] = in[i];
    return k;
}

double sumDouble(double *arr, int len)
{
    lpm_dropin_host_access(arr, 0);
    double s = 0.0;
    for (int i = 0; i < len; i++)
        s += arr[i];
    return s;
}

int sumInt(int *arr, int len)
{
    lpm_dropin_host_access(arr, 0);
    int s = 0;
    for (int i = 0; i < len; i++)
        s += arr[i];
    return s;
}

/* Shell sort, gaps len/2, len/4, ..., 1 (the order sortInt produces for ints is
 * the ascending order whatever the algorithm). */
void sortInt(int *arr, int len)
{
    lpm_dropin_host_access(arr, 1);
    for (int r = len / 2; r >= 1; r /= 2)
        for (int i = r; i < len; i++)
        {
            int t = arr[i], j = i - r;
            while (j >= 0 && arr[j] > t)
            {
                arr[j + r] = arr[j];
                j -= r;
            }
            arr[j + r] = t;
        }
}

int getLine(char *line, int max, FILE *fp)
{
    int c = 0, c0 = 0, i = 0;
    while (--max > 0 && (c = fgetc(fp)) != EOF && c != '\n')
    {
        if (i == 0)
            c0 = c;
        line[i++] = c0 == '#' ? '#' : c;
    }
    if (c == '\n')
        line[i++] = c;
    line[i] = '\0';
    return i;
}

void printDouble(double *target, int m, int n)
{
    lpm_dropin_host_access(target, 0);
    for (int i = m - 1; i < n; i++)
        printf("%8.8e \n", target[i]);
    printf("\n");
}

void printInt(int *target, int n)
{
    lpm_dropin_host_access(target, 0);
    for (int i = 0; i < n; i++)
        printf("%d ", target[i]);
    printf("\n");
}

/* --------------------------------------------------------------- type setters --- */
void setTypeRect(double r0, double r1, double r2, double r3, double r4, double r5, int t)
{
    lpm_dropin_host_access(xyz, 0);
    lpm_dropin_host_access(type, 1);
    for (int i = 0; i < nparticle; i++)
        if (xyz[i][0] >= r0 && xyz[i][0] <= r1 && xyz[i][1] >= r2 && xyz[i][1] <= r3 && xyz[i][2] >= r4 &&
            xyz[i][2] <= r5)
            type[i] = t;
}

void setTypeCircle(double x, double y, double r, int t)
{
    lpm_dropin_host_access(xyz, 0);
    lpm_dropin_host_access(type, 1);
    for (int i = 0; i < nparticle; i++)
    {
        double dx = xyz[i][0] - x, dy = xyz[i][1] - y;
        if (dx * dx + dy * dy <= r * r)
            type[i] = t;
    }
}

void setTypeFullNeighbor(int k)
{
    lpm_dropin_host_access(nb_initial, 0);
    lpm_dropin_host_access(type, 1);
    for (int i = 0; i < nparticle; i++)
        if (nb_initial[i] == nneighbors)
            type[i] = k;
}

/* ------------------------------------------------------------- MKL routines --- */
/* Euclidean norm with the usual scaling against overflow. */
double cblas_dnrm2(const int n, const double *x, const int incx)
{
    lpm_dropin_host_access(x, 0);
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++)
    {
        double a = fabs(x[(long)i * incx]);
        if (a == 0.0)
            continue;
        if (scale < a)
        {
            ssq = 1.0 + ssq * (scale / a) * (scale / a);
            scale = a;
        }
        else
            ssq += (a / scale) * (a / scale);
    }
    return scale * sqrt(ssq);
}

void mkl_free_buffers(void) {}

/* ------------------------------------------------ output-only post-processing --- */
/* computeBondForce / computeBondStretch (reference src/lpm_basic.c:8-50): bond
 * values re-ordered by the lattice's reference bond vectors.  Output only. */
static void reorder_bonds(double **dst, int stretch)
{
    if (!bond_vector || !dst)
        return;
    lpm_dropin_sync_host();
    for (int i = 0; i < nparticle; i++)
        for (int k = 0; k < nneighbors; k++)
            for (int j = 0; j < nb_initial[i]; j++)
            {
                double p[3] = {csx[i][j] - bond_vector[k][0], csy[i][j] - bond_vector[k][1],
                               csz[i][j] - bond_vector[k][2]};
                if (sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) < TOL)
                {
                    dst[i][k] = stretch ? dL[i][j] / distance_initial[i][j] : F[i][j];
                    break;
                }
            }
}

void computeBondForce() { reorder_bonds(bond_force, 0); }
void computeBondStretch() { reorder_bonds(bond_stretch, 1); }
