/*
 * A load-step driver written against the reference API only (lpm.h): 2D hexagonal
 * elastic-brittle lattice in shear, with the reference driver's loop
 * (src/lpmc_project.c:382-546): per step the tangent, the BCs and the predictor,
 * the Newton loop over switchStateV / setDispBC_stiffnessUpdate2D / solverCG /
 * computeBondForceGeneral / updateRR, computeStrain, updateDamageGeneral (which
 * appends the broken bonds to the bond file), updateCrack, switchStateV(1), the
 * writeDisp / writeReaction outputs, and after bond breakage a new tangent and a
 * re-solve of the same step.  A writeDump of the last step closes the run.
 *
 * Usage: hex_brittle OUTDIR W H CRITICAL NBREAK NEWTON_TOL NEWTON_MAXIT SHEAR1 [SHEAR2 ...]
 * (box [0, W] x [0, H]; bottom edge held, top edge held in y and moved in x by the
 * shear-strain increment times the height).  OUTDIR receives the reference's files
 * result_bond.txt, result_disp.txt, result_reaction.txt, result_dump.txt, and
 * positions.txt with the final positions in %.17g.
 */
#include <stdio.h>
#include <stdlib.h>

#include "lpm.h"
#include "lpm_ext.h"

int main(int argc, char **argv)
{
    if (argc < 9)
    {
        fprintf(stderr, "usage: %s OUTDIR W H CRITICAL NBREAK NEWTON_TOL NEWTON_MAXIT SHEAR...\n", argv[0]);
        return 2;
    }
    const char *dir = argv[1];
    const double W = atof(argv[2]), Hb = atof(argv[3]), crit = atof(argv[4]), newton_tol = atof(argv[6]);
    const int nbr_cap = atoi(argv[5]), newton_maxit = atoi(argv[7]), nsteps = argc - 8;
    char bondFile[512], dispFile[512], forceFile[512], dumpFile[512], posFile[512];
    snprintf(bondFile, sizeof(bondFile), "%s/result_bond.txt", dir);
    snprintf(dispFile, sizeof(dispFile), "%s/result_disp.txt", dir);
    snprintf(forceFile, sizeof(forceFile), "%s/result_reaction.txt", dir);
    snprintf(dumpFile, sizeof(dumpFile), "%s/result_dump.txt", dir);
    snprintf(posFile, sizeof(posFile), "%s/positions.txt", dir);

    /* lattice and state (src/lpmc_project.c:60-106) */
    lattice = 1;
    dim = 2;
    radius = 0.5;
    eulerflag = 0;
    angle1 = angle2 = angle3 = 0.0;
    double b[6] = {0.0, W, 0.0, Hb, 0.0, 1.0};
    for (int k = 0; k < 6; k++)
        box[k] = b[k];
    box_x = W, box_y = Hb, box_z = 1.0;
    createCuboid();
    double mv[3] = {0.0, 0.0, 0.0};
    moveParticle(mv);
    initMatrices();
    copyDouble2D(xyz_initial, xyz, nparticle, NDIM);
    searchNormalNeighbor();
    searchAFEMNeighbor();

    /* types: 1 top edge, 2 bottom edge */
    double ymin = 1e300, ymax = -1e300;
    for (int i = 0; i < nparticle; i++)
        ymin = xyz[i][1] < ymin ? xyz[i][1] : ymin, ymax = xyz[i][1] > ymax ? xyz[i][1] : ymax;
    type = allocInt1D(nparticle, 0);
    for (int i = 0; i < nparticle; i++)
    {
        if (xyz[i][1] >= ymax - 0.6 * radius)
            type[i] = 1;
        if (xyz[i][1] <= ymin + 0.6 * radius)
            type[i] = 2;
    }
    ntype = 3;
    const double height = ymax - ymin;

    /* elastic-brittle material, plane strain (src/lpmc_project.c:186-244) */
    const double E = 70e3, mu = 0.25;
    Ce = allocDouble2D(ntype, 3, 0.0);
    for (int t = 0; t < ntype; t++)
    {
        Ce[t][0] = E * (1.0 - mu) / (1.0 + mu) / (1.0 - 2.0 * mu);
        Ce[t][1] = E * mu / (1.0 + mu) / (1.0 - 2.0 * mu);
        Ce[t][2] = E / 2.0 / (1.0 + mu);
    }
    sigmay = allocDouble1D(nparticle, 1e30);
    plmode = 6;
    critical_bstrain = crit;
    nbreak = nbr_cap;
    calcKnTv();
    computedL();

    /* tight CG so that the converged state is compared, not the iterates */
    lpmx_solver s = {LPMX_FP64, 1e-24, 0.0, 200 * dim * nparticle, newton_tol, newton_maxit, 1e-2, 100, 0, 0};
    lpm_ext_set_solver(&s);

    double prev = 0.0;
    for (int i = 0; i < nsteps; i++)
    {
        const double g = atof(argv[8 + i]);
        struct dispBCPara dbp[4] = {{2, 'x', 0.0}, {2, 'y', 0.0}, {1, 'y', 0.0}, {1, 'x', (g - prev) * height}};
        prev = g;
        copyDouble2D(xyz_temp, xyz, nparticle, NDIM);
        copyDouble2D(F_temp, F, nparticle, nneighbors);
        copyDouble1D(Pex_temp, Pex, dim * nparticle);
        calcStiffness2DFiniteDifference(6);
        setDispBC(4, dbp);
        setForceBC(0, NULL);
        computeBondForceGeneral(4, 0);
    label_broken_bond:
        updateRR();
        double nres = cblas_dnrm2(dim * nparticle, residual, 1);
        double nrea = cblas_dnrm2(countNEqual(dispBC_index, nparticle * dim, 1), reaction_force, 1);
        const double tolm = nres > nrea ? nres : nrea;
        int ni = 0;
        while (nres > newton_tol * tolm && ni < newton_maxit)
        {
            switchStateV(0);
            setDispBC_stiffnessUpdate2D();
            solverCG();
            computeBondForceGeneral(plmode, 0);
            updateRR();
            ni++;
            nres = cblas_dnrm2(dim * nparticle, residual, 1);
        }
        if (ni >= newton_maxit)
        {
            fprintf(stderr, "step %d: Newton cap\n", i + 1);
            return 4;
        }
        computeStrain();
        double smax = 0.0;
        for (int p = 0; p < nparticle; p++)
            for (int j = 0; j < nb_initial[p]; j++)
                if (dL[p][j] / distance_initial[p][j] > smax)
                    smax = dL[p][j] / distance_initial[p][j];
        int broke = updateDamageGeneral(bondFile, i + 1, plmode);
        printf("step %d: Newton %d, largest bond stretch %.6e, candidates %d\n", i + 1, ni, smax, broke);
        updateCrack();
        switchStateV(1);
        writeDisp(dispFile, 'x', 1, i + 1);
        writeReaction(forceFile, 'x', 1, i + 1);
        if (broke > 0)
        {
            calcStiffness2DFiniteDifference(6);
            goto label_broken_bond;
        }
    }
    writeDump(dumpFile, nsteps, 's');
    FILE *f = fopen(posFile, "w");
    if (!f)
        return 3;
    for (int p = 0; p < nparticle; p++)
        fprintf(f, "%.17g %.17g %.17g\n", xyz[p][0], xyz[p][1], xyz[p][2]);
    fclose(f);
    mkl_free_buffers();
    return 0;
}
