"""The cyclic SC workload on the device against the oracle's trajectory.

SC 20^3 = 8,000 particles, J2 mixed-linear hardening with the nonlocal ductile
damage of the reference's default path, 4 symmetric cycles of +/-0.5% strain with
20 load steps per half-cycle (160 steps: loading, unloading, reverse yield at
every peak).  The oracle's trajectory (serial C restatement, Newton to 1e-12) is
the committed fixture tests/golden/sc8000_cyclic.npz, written by
tests/golden/make_sc8000_cyclic.py: at this size the oracle needs about 40
minutes, the device loop seconds.  Reactions and mean displacements of every
step, and the displacement field, hardening variable and nonlocal damage after
the first cycle and after the last, are compared: FP64 to 1e-10 relative, the
mixed form to its stated 1e-8.
"""
import os

import numpy as np
import pytest

from lpm_harness import P, assert_rel, get, tol_for

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sc8000_cyclic.npz")
L_BOX = 19.035


@pytest.mark.parametrize("prec", [P.LPMX_FP64, P.LPMX_MIXED])
def test_cyclic_sc8000_matches_oracle(prec):
    g = np.load(FIXTURE)
    n, x, t, c = P.build_sc_model(L_BOX)
    assert n == 8000
    height = float(x[:, 2].max() - x[:, 2].min())
    strains = P.cyclic_strains(5e-3, 20, 4)
    steps = P.sc_uniaxial_steps(height, strains)
    assert len(steps) == len(g["reaction"]) == 160
    s = P.solver(precision=prec, rtol2=1e-8, atol2=0.0, maxit=300 * n, newton_tol=1e-12, newton_maxit=4000,
                 accel_depth=32, precond=2)
    tol = tol_for(prec)
    print(f"cyclic SC 8000, precision {prec}: 160 steps", flush=True)
    rc, out = P.run_steps(steps[:40], s, out_axis="z", out_type=1)
    assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters) for o in out])
    u = get("XYZ", (n, 3)) - x
    assert_rel("displacement after cycle 1", u, g["s40_u"], tol)
    assert_rel("hardening variable after cycle 1", get("J2_ALPHA1", (n,)), g["s40_alpha"], tol)
    assert_rel("nonlocal damage after cycle 1", get("DAMAGE_NONLOCAL1", (n,)), g["s40_dnl"], tol)
    rc, out2 = P.run_steps(steps[40:], s, out_axis="z", out_type=1)
    assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters) for o in out2])
    out = out + out2
    print("Newton iterations per step:", [o.newton_iters for o in out], flush=True)
    assert_rel("reaction per step", [o.reaction for o in out], g["reaction"], tol)
    assert_rel("mean top displacement per step", [o.disp for o in out], g["disp"], tol)
    u = get("XYZ", (n, 3)) - x
    assert_rel("displacement after cycle 4", u, g["s160_u"], tol)
    assert_rel("hardening variable after cycle 4", get("J2_ALPHA1", (n,)), g["s160_alpha"], tol)
    assert_rel("nonlocal damage after cycle 4", get("DAMAGE_NONLOCAL1", (n,)), g["s160_dnl"], tol)
    # the path unloads and yields in reverse: the reaction changes sign each half-cycle
    r = g["reaction"]
    assert r.max() > 0 > r.min()
