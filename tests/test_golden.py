"""The committed oracle fixtures are what the oracle computes (CPU).

tests/golden/sc8000_cyclic.npz holds the oracle's trajectory of the cyclic SC
workload (tests/golden/make_sc8000_cyclic.py).  Its first two load steps are
recomputed here with the oracle and compared, and the prescribed strain path is
checked against the file's mean top-face displacements.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "sc8000_cyclic.npz")


def test_cyclic_fixture_is_the_oracles_trajectory(oracle):
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_sc8000_cyclic as mk
    from lpm_harness import P

    g = np.load(FIXTURE)
    assert g["reaction"].shape == (160,) and g["disp"].shape == (160,)
    for tag in ("s40", "s160"):
        assert g[tag + "_u"].shape == (8000, 3)
        assert g[tag + "_alpha"].shape == (8000,) and g[tag + "_dnl"].shape == (8000,)
    # the top face follows the strain path: +0.5% at step 10, -0.5% at step 30, 0 after every cycle
    m, x, t = mk.model()
    height = float(x[:, 2].max() - x[:, 2].min())
    e = P.cyclic_strains(5e-3, 20, 4)
    np.testing.assert_allclose(g["disp"], e * height, rtol=0, atol=1e-12)
    np.testing.assert_allclose(e[[9, 29, 39, 159]], [5e-3, -5e-3, 0.0, 0.0], rtol=0, atol=1e-17)
    # plastic flow in both directions, and damage accumulating from it
    assert g["s40_alpha"].max() > 0 and g["s160_alpha"].max() > g["s40_alpha"].max()
    assert g["s160_dnl"].max() > g["s40_dnl"].max() > 0
    # recompute the first two steps (elastic, then first yield)
    s = mk.solver(m.n, min(8, os.cpu_count() or 1))
    steps = mk.steps_for(x)
    rc, out = oracle.run_steps(m, steps[:2], s)
    assert rc == 0
    r = np.array([o.reaction for o in out])
    assert np.abs(r - g["reaction"][:2]).max() <= 1e-10 * np.abs(g["reaction"][:2]).max()
