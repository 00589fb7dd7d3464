"""Writes refdriver_sc9261.npz: the oracle's run of the reference driver's own problem.

The reference's driver, src/lpmc_project.c, hard-codes its problem (lines 69-356):
SC lattice, radius 0.2499999944120646, box [-0.2, 10.2]^3 -> 21^3 = 9,261 particles;
types 1 (top layer, z >= 10 - 1.2 r), 2 (bottom layer, z <= 1.2 r), 3 (full
neighbour list), 0 elsewhere; E = 146 GPa, nu = 0.3 for all four types; J2
mixed-linear hardening (plmode 0, yield 200 MPa, H = 38.714 GPa, xi = 0) with the
default damage settings (damage_L 0.5, threshold 0.9, damagec_A 0); 91 load steps:
the top layer held in z, the bottom layer loaded by a z force of -2000 per step for
steps 1-14, +2000 for steps 15-48 and -2000 for steps 49-91.  Its loop
(src/lpmc_project.c:382-546) uses the finite-difference tangent, unpreconditioned
CG with the dcg rule |r|^2 <= 1e-8 |r0|^2 + 1e-12 capped at dim*n, Newton to
TOLITER = 1e-4 of max(|R0|, |reaction|) capped at MAXITER = 100.

This script runs the oracle (the serial C restatement of that loop) with the same
settings and records, per step, the values the driver writes to result_disp.txt
(writeDisp 'z' of type 2) and result_force.txt (writeReaction 'z' of type 2), and
the Newton iterations.

Run:  python tests/golden/make_refdriver_sc9261.py   (about 30 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_py as O  # noqa: E402

RADIUS = 0.2499999944120646
BOX = [-0.2, 10.2, -0.2, 10.2, -0.2, 10.2]
NSTEPS = 91


def model():
    x = O.create_lattice(2, RADIUS, BOX)
    n = x.shape[0]
    c = O.elastic_constants(146e3, 0.3)
    m = O.Model(x, 2, RADIUS, np.zeros(n, np.int32), [c] * 4, plmode=0, sigmay=200.0, J2_H=38.714e3, J2_xi=0.0,
                damage_L=0.5, damage_threshold=0.9, damagec_A=0.0, critical_bstrain=1e-2, nbreak=20)
    m.neighbors(cells=True)
    m.connectivity()
    # setTypeRect (src/lpm_basic.c:330-337) twice, then setTypeFullNeighbor (:351-358)
    t = np.zeros(n, np.int32)
    z = x[:, 2]
    t[(z >= 10 - 1.2 * RADIUS) & (z <= 100.0)] = 1
    t[(z >= -100.0) & (z <= 1.2 * RADIUS)] = 2
    t[m.nb_initial == m.nbmax] = 3
    m.type[:] = t
    O.lib().orc_kntv(m.p)
    O.lib().orc_computedL(m.p)
    return m, x, t


def steps():
    out = []
    for i in range(NSTEPS):
        f = -2000.0 if (i < 14 or i >= 48) else 2000.0
        out.append(([(1, "z", 0.0)], [(2, (0.0, 0.0, f))]))
    return out


def main(out=os.path.join(HERE, "refdriver_sc9261.npz")):
    m, x, t = model()
    n = m.n
    assert n == 21 ** 3
    s = O.solver(rtol2=1e-8, atol2=1e-12, maxit=3 * n, newton_tol=1e-4, newton_maxit=100, tangent=0,
                 nthreads=min(8, os.cpu_count() or 1))
    rec = {k: [] for k in ("disp", "reaction", "newton_iters")}
    t0 = time.time()
    for k, st in enumerate(steps()):
        rc, o = O.run_steps(m, [st], s, out_axis="z", out_type=2)
        rec["disp"].append(o[0].disp)
        rec["reaction"].append(o[0].reaction)
        rec["newton_iters"].append(o[0].newton_iters)
        print(f"step {k + 1}: disp {o[0].disp:.8e} reaction {o[0].reaction:.8e} newton {o[0].newton_iters} "
              f"({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(out, types=t, **{k: np.array(v) for k, v in rec.items()})
    print("wrote", out)


if __name__ == "__main__":
    main()
