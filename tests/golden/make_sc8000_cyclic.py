"""Writes sc8000_cyclic.npz: the oracle's trajectory of the cyclic SC workload.

Workload: SC lattice, radius 0.5 mm, box [0, 19.035]^3 -> 20^3 = 8,000 particles;
J2 mixed-linear hardening (E = 200 GPa, nu = 0.3, yield 200 MPa, H = 2 GPa, xi = 0)
with the reference's nonlocal ductile damage (L = 0.5, threshold 0.9, A = 0;
reference src/lpmc_project.c:207-244); bottom face held in z, corners 3-2-1; the
top face moved in z along 4 symmetric cycles of +/-0.5% strain, 20 load steps per
half-cycle (160 steps); the load-step loop of src/lpmc_project.c:382-546 with
Newton to 1e-12 of max(|R0|, |reaction|) (Anderson depth 8 to get there) and CG
to 1e-14 relative.

Run:  python tests/golden/make_sc8000_cyclic.py   (about 40 minutes on 8 cores)
The oracle is the checker; the parity test compares the device loop with this file.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_py as O  # noqa: E402
from lpm_harness import P  # noqa: E402

L_BOX = 19.035
NEWTON_TOL = 1e-12


def model():
    x = O.create_lattice(2, 0.5, [0, L_BOX, 0, L_BOX, 0, L_BOX])
    t = P.sc_types(x)
    c = O.elastic_constants(200e3, 0.3)
    m = O.Model(x, 2, 0.5, t, [c] * 5, plmode=0, sigmay=200.0, J2_H=2000.0, J2_xi=0.0, damage_L=0.5,
                damage_threshold=0.9, damagec_A=0.0)
    m.setup(cells=True)
    return m, x, t


def steps_for(x):
    height = float(x[:, 2].max() - x[:, 2].min())
    return P.sc_uniaxial_steps(height, P.cyclic_strains(5e-3, 20, 4))


def solver(n, nthreads=8):
    return O.solver(rtol2=1e-28, atol2=0.0, maxit=100 * 3 * n, newton_tol=NEWTON_TOL, newton_maxit=2000, tangent=1,
                    fixed_diag=1, accel=8, nthreads=nthreads)


def main(out=os.path.join(HERE, "sc8000_cyclic.npz")):
    m, x, t = model()
    n = m.n
    steps = steps_for(x)
    s = solver(n, min(8, os.cpu_count() or 1))
    rec = {k: [] for k in ("reaction", "disp", "newton_iters")}
    snap = {}
    t0 = time.time()
    for k, st in enumerate(steps):
        rc, o = O.run_steps(m, [st], s)
        assert rc == 0, f"step {k + 1} hit the Newton cap"
        rec["reaction"].append(o[0].reaction)
        rec["disp"].append(o[0].disp)
        rec["newton_iters"].append(o[0].newton_iters)
        print(f"step {k + 1}: reaction {o[0].reaction:.6e} disp {o[0].disp:.6e} newton {o[0].newton_iters} "
              f"({time.time() - t0:.0f} s)", flush=True)
        if k + 1 in (40, len(steps)):
            tag = f"s{k + 1}"
            snap[tag + "_u"] = m.x - x
            snap[tag + "_alpha"] = m.alpha[:, 1].copy()
            snap[tag + "_dnl"] = m.dnl[:, 1].copy()
    np.savez_compressed(out, **{k: np.array(v) for k, v in rec.items()}, **snap)
    print("wrote", out)


if __name__ == "__main__":
    main()
