/*
 * A load-step driver written against the reference API only (lpm.h): SC lattice,
 * J2 mixed-linear material, uniaxial tension with the bottom face held in z and
 * 3-2-1 corner constraints.  It makes the calls the reference driver makes
 * (src/lpmc_project.c:99-470): setup, per-step tangent, BCs, predictor, then its
 * own Newton loop over switchStateV / setDispBC_stiffnessUpdate3D / solverCG /
 * computeBondForceGeneral / updateRR, damage and crack updates, switchStateV(1).
 *
 * Usage: sc_uniaxial OUT L HEIGHT NEWTON_TOL NEWTON_MAXIT STRAIN1 [STRAIN2 ...]
 * (box [0, L]^3; the top face moves by strain increment * HEIGHT per step)
 * OUT receives one line per step ("step reaction mean_disp newton_iterations")
 * followed by the final positions, all in %.17g.
 */
#include <stdio.h>
#include <stdlib.h>

#include "lpm.h"
#include "lpm_ext.h"

/* The globals are defined by liblpm.so; a driver that defines them itself (as the
 * reference driver does, src/lpmc_project.c:17-45) takes precedence at link time. */

int main(int argc, char **argv)
{
    if (argc < 7)
    {
        fprintf(stderr, "usage: %s OUT L HEIGHT NEWTON_TOL NEWTON_MAXIT STRAIN...\n", argv[0]);
        return 2;
    }
    const char *out = argv[1];
    const double L = atof(argv[2]), height = atof(argv[3]), newton_tol = atof(argv[4]);
    const int newton_maxit = atoi(argv[5]), nsteps = argc - 6;

    /* lattice and state (src/lpmc_project.c:60-106) */
    lattice = 2;
    dim = 3;
    radius = 0.5;
    eulerflag = 0;
    angle1 = angle2 = angle3 = 0.0;
    double b[6] = {0.0, L, 0.0, L, 0.0, L};
    for (int k = 0; k < 6; k++)
        box[k] = b[k];
    box_x = box_y = box_z = L;
    createCuboid();
    double mv[3] = {0.0, 0.0, 0.0};
    moveParticle(mv);
    initMatrices();
    copyDouble2D(xyz_initial, xyz, nparticle, NDIM);
    searchNormalNeighbor();
    searchAFEMNeighbor();

    /* types: 1 top face, 2 bottom face, 3 bottom corner at min x and y,
     * 4 bottom corner at max x and min y */
    double zmin = 1e300, zmax = -1e300;
    for (int i = 0; i < nparticle; i++)
        zmin = xyz[i][2] < zmin ? xyz[i][2] : zmin, zmax = xyz[i][2] > zmax ? xyz[i][2] : zmax;
    type = allocInt1D(nparticle, 0);
    int c3 = -1, c4 = -1;
    for (int i = 0; i < nparticle; i++)
    {
        if (xyz[i][2] >= zmax - 0.6 * radius)
            type[i] = 1;
        if (xyz[i][2] <= zmin + 0.6 * radius)
        {
            type[i] = 2;
            /* lexicographic on (y, x) and (y, -x) */
            if (c3 < 0 || xyz[i][1] < xyz[c3][1] || (xyz[i][1] == xyz[c3][1] && xyz[i][0] < xyz[c3][0]))
                c3 = i;
            if (c4 < 0 || xyz[i][1] < xyz[c4][1] || (xyz[i][1] == xyz[c4][1] && xyz[i][0] > xyz[c4][0]))
                c4 = i;
        }
    }
    type[c3] = 3;
    type[c4] = 4;
    ntype = 5;

    /* material (src/lpmc_project.c:186-244) */
    const double E = 200e3, mu = 0.3;
    Ce = allocDouble2D(ntype, 3, 0.0);
    for (int t = 0; t < ntype; t++)
    {
        Ce[t][0] = E * (1.0 - mu) / (1.0 + mu) / (1.0 - 2.0 * mu);
        Ce[t][1] = E * mu / (1.0 + mu) / (1.0 - 2.0 * mu);
        Ce[t][2] = E / 2.0 / (1.0 + mu);
    }
    sigmay = allocDouble1D(nparticle, 200.0);
    plmode = 0;
    J2_H = 2000.0;
    J2_xi = 0.0;
    damage_L = 0.5;
    damage_threshold = 0.9;
    damagec_A = 0.0;
    critical_bstrain = 1e-2;
    nbreak = 20;
    calcKnTv();
    computedL();

    /* tight CG so that the converged state is compared, not the iterates */
    lpmx_solver s = {LPMX_FP64, 1e-24, 0.0, 100 * dim * nparticle, newton_tol, newton_maxit, 1e-2, 100, 0};
    lpm_ext_set_solver(&s);

    FILE *f = fopen(out, "w");
    if (!f)
        return 3;
    double prev = 0.0;
    for (int i = 0; i < nsteps; i++)
    {
        const double e = atof(argv[6 + i]);
        struct dispBCPara dbp[7] = {{2, 'z', 0.0}, {3, 'x', 0.0}, {3, 'y', 0.0}, {3, 'z', 0.0},
                                    {4, 'y', 0.0}, {4, 'z', 0.0}, {1, 'z', (e - prev) * height}};
        prev = e;
        copyDouble2D(xyz_temp, xyz, nparticle, NDIM);
        copyDouble2D(F_temp, F, nparticle, nneighbors);
        copyDouble1D(Pex_temp, Pex, dim * nparticle);
        calcStiffness3DFiniteDifference(6);
        setDispBC(7, dbp);
        setForceBC(0, NULL);
        computeBondForceGeneral(4, 0);
        int total_ni = 0;
    broken:
        updateRR();
        double nres = cblas_dnrm2(dim * nparticle, residual, 1);
        double nrea = cblas_dnrm2(countNEqual(dispBC_index, nparticle * dim, 1), reaction_force, 1);
        const double tolm = nres > nrea ? nres : nrea;
        int ni = 0;
        while (nres > newton_tol * tolm && ni < newton_maxit)
        {
            switchStateV(0);
            setDispBC_stiffnessUpdate3D();
            solverCG();
            computeBondForceGeneral(plmode, 0);
            updateRR();
            ni++;
            nres = cblas_dnrm2(dim * nparticle, residual, 1);
        }
        total_ni += ni;
        int broke = updateDamageGeneral("bond_driver.txt", i + 1, plmode);
        updateCrack();
        switchStateV(1);
        if (broke > 0)
        {
            calcStiffness3DFiniteDifference(6);
            goto broken;
        }
        double force = 0.0, dsum = 0.0;
        int cnt = 0;
        for (int p = 0; p < nparticle; p++)
            if (type[p] == 1)
            {
                for (int j = 0; j < nb_initial[p]; j++)
                    force += -csz[p][j] * F[p][j];
                dsum += xyz[p][2] - xyz_initial[p][2];
                cnt++;
            }
        fprintf(f, "%d %.17g %.17g %d\n", i + 1, force, dsum / cnt, total_ni);
    }
    for (int p = 0; p < nparticle; p++)
        fprintf(f, "%.17g %.17g %.17g\n", xyz[p][0], xyz[p][1], xyz[p][2]);
    fclose(f);
    return 0;
}
