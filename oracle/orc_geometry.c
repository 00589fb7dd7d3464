/*
 * Oracle: lattice generation and neighbour / connectivity lists.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Lattice constants per type (reference src/initialization.c:96-560, the
 * "model parameters" block of each lattice). */
void orc_lattice_params(int lattice, double radius, double box_z, orc_lattice *o)
{
    memset(o, 0, sizeof(*o));
    o->lattice = lattice;
    o->radius = radius;
    o->dim = lattice < 2 ? 2 : 3;
    switch (lattice)
    {
    case 0: /* square: initialization.c:104-112 */
        o->nbmax = 8, o->nb1max = 4, o->nb2max = 4, o->nafem = 16;
        o->cut1 = 2.0 * radius, o->cut2 = 2.0 * sqrt(2.0) * radius;
        o->volume = 4 * (radius * radius) * box_z;
        break;
    case 1: /* hexagon: initialization.c:171-179 */
        o->nbmax = 12, o->nb1max = 6, o->nb2max = 6, o->nafem = 30;
        o->cut1 = 2.0 * radius, o->cut2 = 2.0 * sqrt(3.0) * radius;
        o->volume = 2 * sqrt(3) * (radius * radius) * box_z;
        break;
    case 2: /* simple cubic: initialization.c:251-259 */
        o->nbmax = 18, o->nb1max = 6, o->nb2max = 12, o->nafem = 60;
        o->cut1 = 2.0 * radius, o->cut2 = 2.0 * sqrt(2.0) * radius;
        o->volume = pow(2 * radius, 3);
        break;
    case 3: /* FCC: initialization.c:336-345 */
        o->nbmax = 18, o->nb1max = 12, o->nb2max = 6, o->nafem = 60;
        o->cut1 = 2.0 * radius, o->cut2 = 2.0 * sqrt(2.0) * radius;
        o->volume = 4.0 * sqrt(2.0) * pow(radius, 3);
        break;
    default: /* BCC: initialization.c:471-479 */
        o->nbmax = 14, o->nb1max = 8, o->nb2max = 6, o->nafem = 40;
        o->cut1 = 2.0 * radius, o->cut2 = 4.0 / sqrt(3.0) * radius;
        o->volume = 32.0 * sqrt(3.0) / 9.0 * pow(radius, 3);
        break;
    }
}

/* Rotation matrix of the three Euler conventions (initialization.c:42-94). */
static void rotation(int eulerflag, double a1, double a2, double a3, double R[9])
{
    if (eulerflag == 0)
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
    else if (eulerflag == 1)
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
    else
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

/* Accept a candidate point if its rotated image lies in the (closed) box.
 * The reference rotates with cblas_dgemv; for the identity rotation used by every
 * benchmark configuration the product is exact, whatever the summation order. */
static int keep(const double R[9], const double *box, int dim, const double t[3], double p[3])
{
    for (int r = 0; r < 3; r++)
        p[r] = R[3 * r] * t[0] + R[3 * r + 1] * t[1] + R[3 * r + 2] * t[2];
    if (!(p[0] >= box[0] && p[0] <= box[1] && p[1] >= box[2] && p[1] <= box[3]))
        return 0;
    if (dim == 3 && !(p[2] >= box[4] && p[2] <= box[5]))
        return 0;
    return 1;
}

static void emit(double *x, int cap, int *n, const double p[3], int dim)
{
    if (x && *n < cap)
    {
        x[3 * *n + 0] = p[0];
        x[3 * *n + 1] = p[1];
        x[3 * *n + 2] = dim == 3 ? p[2] : 0.0;
    }
    (*n)++;
}

/* createCuboid (reference src/initialization.c:6-560): candidate lattice points
 * over [-a, a] with a the box diagonal, rotated, clipped to the box, kept in
 * generation order.  Returns the particle count; writes at most cap points. */
int orc_create_lattice(int lattice, double radius, const double *box, int eulerflag,
                       double a1, double a2, double a3, double *x, int cap)
{
    double R[9];
    rotation(eulerflag, a1, a2, a3, R);
    int dim = lattice < 2 ? 2 : 3;
    double bx = box[1] - box[0], by = box[3] - box[2], bz = box[5] - box[4];
    /* pow(x, 2) is written x*x, as gcc -O2 folds it in the reference build */
    double a = dim == 2 ? sqrt(bx * bx + by * by) : sqrt(bx * bx + by * by + bz * bz);
    double bt[6] = {-a, a, -a, a, -a, a};
    int n = 0;
    double t[3] = {0, 0, 0}, p[3];

    if (lattice == 0) /* initialization.c:96-161 */
    {
        double hx = 2 * radius, hy = hx;
        int pfr = (int)floor((bt[1] - bt[0]) / hx);
        int rows = (int)floor((bt[3] - bt[2]) / hy);
        for (int j = 1; j <= rows; j++)
        {
            t[1] = bt[2] + hy * (j - 1) + radius;
            for (int i = 1; i <= pfr; i++)
            {
                t[0] = bt[0] + hx * (i - 1) + radius;
                if (keep(R, box, 2, t, p))
                    emit(x, cap, &n, p, 2);
            }
        }
    }
    else if (lattice == 1) /* initialization.c:163-239 */
    {
        double hx = 2 * radius, hy = hx * sqrt(3.0) / 2.0;
        int pfr = 1 + (int)floor((bt[1] - bt[0]) / hx);
        int rows = 1 + (int)floor((bt[3] - bt[2]) / hy);
        for (int j = 1; j <= rows; j++)
        {
            t[1] = bt[2] + hy * (j - 1);
            if (j % 2 == 1)
                for (int i = 1; i <= pfr; i++)
                {
                    t[0] = bt[0] + hx * (i - 1);
                    if (keep(R, box, 2, t, p))
                        emit(x, cap, &n, p, 2);
                }
            else
                for (int i = 1; i <= pfr - 1; i++)
                {
                    t[0] = bt[0] + hx * (2 * i - 1) / 2.0;
                    if (keep(R, box, 2, t, p))
                        emit(x, cap, &n, p, 2);
                }
        }
    }
    else if (lattice == 2) /* initialization.c:241-315; x starts at box[0], y/z at -a */
    {
        double h = 2 * radius;
        int pfr = 1 + (int)floor((bt[1] - bt[0]) / h);
        int rows = 1 + (int)floor((bt[3] - bt[2]) / h);
        int layers = 1 + (int)floor((bt[5] - bt[4]) / h);
        for (int k = 1; k <= layers; k++)
        {
            t[2] = bt[4] + h * (k - 1);
            for (int j = 1; j <= rows; j++)
            {
                t[1] = bt[2] + h * (j - 1);
                for (int i = 1; i <= pfr; i++)
                {
                    t[0] = box[0] + h * (i - 1);
                    if (keep(R, box, 3, t, p))
                        emit(x, cap, &n, p, 3);
                }
            }
        }
    }
    else if (lattice == 3) /* initialization.c:326-459 */
    {
        double h = 4 * radius / sqrt(2.0);
        int pfr = 1 + (int)floor((bt[1] - bt[0]) / h);
        int rows = 1 + (int)floor((bt[3] - bt[2]) / (h / 2.0));
        int layers = 1 + (int)floor((bt[5] - bt[4]) / (h / 2.0));
        for (int k = 1; k <= layers; k++)
        {
            t[2] = bt[4] + h / 2.0 * (k - 1);
            for (int j = 1; j <= rows; j++)
            {
                t[1] = bt[2] + h / 2.0 * (j - 1);
                int full = (k % 2 == 1) == (j % 2 == 1);
                if (full)
                    for (int i = 1; i <= pfr; i++)
                    {
                        t[0] = bt[0] + h * (i - 1);
                        if (keep(R, box, 3, t, p))
                            emit(x, cap, &n, p, 3);
                    }
                else
                    for (int i = 1; i <= pfr - 1; i++)
                    {
                        t[0] = bt[0] + h * (i - 1) + h / 2.0;
                        if (keep(R, box, 3, t, p))
                            emit(x, cap, &n, p, 3);
                    }
            }
        }
    }
    else /* BCC: initialization.c:461-559 */
    {
        double h = 4.0 / sqrt(3.0) * radius;
        int pfr = 1 + (int)floor((bt[1] - bt[0]) / h);
        int rows = 1 + (int)floor((bt[3] - bt[2]) / h);
        int layers = 1 + (int)floor((bt[5] - bt[4]) / (h / 2.0));
        for (int k = 1; k <= layers; k++)
        {
            t[2] = bt[4] + h / 2.0 * (k - 1);
            if (k % 2 == 1)
            {
                for (int j = 1; j <= rows; j++)
                {
                    t[1] = bt[2] + h * (j - 1);
                    for (int i = 1; i <= pfr; i++)
                    {
                        t[0] = bt[0] + h * (i - 1);
                        if (keep(R, box, 3, t, p))
                            emit(x, cap, &n, p, 3);
                    }
                }
            }
            else
            {
                for (int j = 1; j <= rows - 1; j++)
                {
                    t[1] = bt[2] + h * (j - 1) + h / 2.0;
                    for (int i = 1; i <= pfr - 1; i++)
                    {
                        t[0] = bt[0] + h * (i - 1) + h / 2.0;
                        if (keep(R, box, 3, t, p))
                            emit(x, cap, &n, p, 3);
                    }
                }
            }
        }
    }
    return n;
}

/* Distance exactly as the reference writes it: sqrt of the sum of the three
 * squared differences, x*x for pow(x, 2) (neighbor.c:18, constitutive.c:249). */
static double dist3(const double *a, const double *b)
{
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* Append j to i's lists if it passes the shell predicate of neighbor.c:20-41. */
static int try_pair(orc_model *m, int i, int j, int *i1, int *i2, int *i3)
{
    const double *xi = m->x + 3 * i, *xj = m->x + 3 * j;
    double dis = dist3(xj, xi);
    int shell;
    if ((dis < 1.01 * m->cut1) && (j != i))
        shell = 0;
    else if ((dis > 1.01 * m->cut1) && (dis < 1.01 * m->cut2))
        shell = 1;
    else
        return 0;
    if (*i3 >= m->nbmax)
        return -1;
    if (shell == 0)
    {
        if (*i1 >= m->nb1max)
            return -1;
        m->nbr1[i * m->nb1max + (*i1)++] = j;
    }
    else
    {
        if (*i2 >= m->nb2max)
            return -1;
        m->nbr2[i * m->nb2max + (*i2)++] = j;
    }
    int b = i * m->nbmax + *i3;
    m->cs0[3 * b + 0] = (xi[0] - xj[0]) / dis;
    m->cs0[3 * b + 1] = (xi[1] - xj[1]) / dis;
    m->cs0[3 * b + 2] = (xi[2] - xj[2]) / dis;
    m->L0[b] = dis;
    m->nsign[b] = shell;
    m->nbr[b] = j;
    (*i3)++;
    return 1;
}

static void reset_lists(orc_model *m)
{
    for (int64_t k = 0; k < (int64_t)m->n * m->nbmax; k++)
        m->nbr[k] = -1, m->nsign[k] = -1;
    for (int64_t k = 0; k < (int64_t)m->n * m->nb1max; k++)
        m->nbr1[k] = -1;
    for (int64_t k = 0; k < (int64_t)m->n * m->nb2max; k++)
        m->nbr2[k] = -1;
}

/* searchNormalNeighbor (reference src/neighbor.c:9-46), brute force over all j in
 * increasing index.  Returns 0, or -1 if a list overflows its capacity. */
int orc_neighbor_search(orc_model *m)
{
    reset_lists(m);
    for (int i = 0; i < m->n; i++)
    {
        int i1 = 0, i2 = 0, i3 = 0;
        for (int j = 0; j < m->n; j++)
            if (try_pair(m, i, j, &i1, &i2, &i3) < 0)
                return -1;
        m->nb[i] = i3;
        m->nb_initial[i] = i3;
    }
    return 0;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Same lists from a cell list: candidates from the 27 (9 in 2D) surrounding cells
 * are sorted by index before the predicate is applied, so that the lists equal the
 * brute-force ones bit for bit.  Used for oracle runs too large for O(N^2). */
int orc_neighbor_search_cells(orc_model *m)
{
    reset_lists(m);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < m->n; i++)
        for (int c = 0; c < 3; c++)
        {
            lo[c] = fmin(lo[c], m->x[3 * i + c]);
            hi[c] = fmax(hi[c], m->x[3 * i + c]);
        }
    double cell = 1.01 * m->cut2 * 1.0001;
    int nc[3];
    for (int c = 0; c < 3; c++)
        nc[c] = 1 + (int)floor((hi[c] - lo[c]) / cell);
    int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    int *head = malloc(sizeof(int) * (ncell + 1));
    int *list = malloc(sizeof(int) * m->n);
    int *cid = malloc(sizeof(int) * m->n);
    memset(head, 0, sizeof(int) * (ncell + 1));
    for (int i = 0; i < m->n; i++)
    {
        int g[3];
        for (int c = 0; c < 3; c++)
            g[c] = (int)floor((m->x[3 * i + c] - lo[c]) / cell);
        cid[i] = (g[2] * nc[1] + g[1]) * nc[0] + g[0];
        head[cid[i] + 1]++;
    }
    for (int64_t k = 0; k < ncell; k++)
        head[k + 1] += head[k];
    int *fill = malloc(sizeof(int) * ncell);
    memcpy(fill, head, sizeof(int) * ncell);
    for (int i = 0; i < m->n; i++)
        list[fill[cid[i]]++] = i;
    free(fill);
    int cand[4096];
    int rc = 0;
    for (int i = 0; i < m->n && rc == 0; i++)
    {
        int g[3];
        for (int c = 0; c < 3; c++)
            g[c] = (int)floor((m->x[3 * i + c] - lo[c]) / cell);
        int nca = 0;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++)
                {
                    int q[3] = {g[0] + dx, g[1] + dy, g[2] + dz};
                    if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= nc[0] || q[1] >= nc[1] || q[2] >= nc[2])
                        continue;
                    int c = (q[2] * nc[1] + q[1]) * nc[0] + q[0];
                    for (int s = head[c]; s < head[c + 1] && nca < 4096; s++)
                        cand[nca++] = list[s];
                }
        qsort(cand, nca, sizeof(int), cmp_int);
        int i1 = 0, i2 = 0, i3 = 0;
        for (int s = 0; s < nca; s++)
            if (try_pair(m, i, cand[s], &i1, &i2, &i3) < 0)
            {
                rc = -1;
                break;
            }
        m->nb[i] = i3;
        m->nb_initial[i] = i3;
    }
    free(head);
    free(list);
    free(cid);
    return rc;
}

/* searchAFEMNeighbor (reference src/neighbor.c:49-141): conn[i] is the sorted set
 * of the bond partners of i and the same-shell bond partners of those partners;
 * K_pointer[i][0] counts conn entries >= i, K_pointer[i+1][1] = K_pointer[i][1] +
 * dim*dim*K_pointer[i][0] - (3 in 3D, 1 in 2D).  Returns len(K_global), or -1 if a
 * conn row overflows. */
int64_t orc_afem_connectivity(orc_model *m)
{
    int *col = malloc(sizeof(int) * (m->nbmax * (m->nbmax + 1) + 1));
    int64_t rc = 0;
    for (int i = 0; i < m->n; i++)
    {
        int nc = 0;
        for (int a = 0; a < m->nb[i]; a++)
        {
            int b = i * m->nbmax + a;
            int j = m->nbr[b], s = m->nsign[b];
            col[nc++] = j;
            const int *lst = s == 0 ? m->nbr1 + (int64_t)j * m->nb1max : m->nbr2 + (int64_t)j * m->nb2max;
            int cap = s == 0 ? m->nb1max : m->nb2max;
            for (int k = 0; k < cap; k++)
                if (lst[k] != -1)
                    col[nc++] = lst[k];
        }
        /* first-occurrence de-duplication, then ascending order */
        int nu = 0;
        for (int k = 0; k < nc; k++)
        {
            int dup = 0;
            for (int u = 0; u < nu; u++)
                if (col[u] == col[k])
                {
                    dup = 1;
                    break;
                }
            if (!dup)
                col[nu++] = col[k];
        }
        qsort(col, nu, sizeof(int), cmp_int);
        if (nu > m->connmax)
        {
            rc = -1;
            break;
        }
        for (int k = 0; k < m->connmax; k++)
            m->conn[(int64_t)i * m->connmax + k] = k < nu ? col[k] : -1;
        m->nb_conn[i] = nu;
    }
    free(col);
    if (rc < 0)
        return rc;
    m->kptr[1] = 0;
    for (int i = 0; i < m->n; i++)
    {
        int k = 0;
        for (int c = 0; c < m->nb_conn[i]; c++)
            if (m->conn[(int64_t)i * m->connmax + c] >= i)
                k++;
        m->kptr[2 * i] = k;
        m->kptr[2 * (i + 1) + 1] = m->kptr[2 * i + 1] + m->dim * m->dim * k - (m->dim == 2 ? 1 : 3);
    }
    return m->kptr[2 * m->n + 1];
}
