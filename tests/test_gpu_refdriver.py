"""The reference's own driver, unchanged, on the device library.

oracle/_ref/lpmc_project is src/lpmc_project.c of the reference compiled from the
reference tree against include/ and liblpm.so with no edit (oracle/Makefile,
target ref-driver; __graft_entry__.build() makes it where the reference sources
are mounted).  It runs its hard-coded problem: SC 21^3 = 9,261 particles, J2
mixed-linear hardening, 91 force-controlled load steps through plastic loading,
reverse yield and reloading, with the reference's solver settings (FD tangent
call, solverCG, Newton to TOLITER 1e-4, MAXITER 100), and writes its result files.
Its mean displacement and reaction of the loaded layer after every step are
compared with the oracle's run of the same problem and settings
(tests/golden/refdriver_sc9261.npz, written by
tests/golden/make_refdriver_sc9261.py).  Both runs stop each step at the
reference's TOLITER from different CG iterates (block-Jacobi PCG on the device,
unpreconditioned CG in the oracle, both to the reference's dcg rule), so they agree
to round-off of the accepted iterates, not bit for bit: the same Newton count in
every step, displacements and reactions within 1e-6 of their largest values (1.2e-7
and 2.8e-7 measured).
"""
import os
import re
import subprocess

import numpy as np
import pytest

from lpm_harness import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "oracle", "_ref", "lpmc_project")
FIXTURE = os.path.join(ROOT, "tests", "golden", "refdriver_sc9261.npz")


def read_series(path):
    rows = [ln.split() for ln in open(path).read().splitlines() if ln.strip()]
    steps = np.array([int(r[0]) for r in rows])
    vals = np.array([float(r[1]) for r in rows])
    return steps, vals


@pytest.mark.skipif(not os.path.exists(BIN), reason="the reference driver is built from the reference sources, "
                                                     "which are not mounted where this tree was built")
def test_reference_driver_runs_unchanged(tmp_path):
    g = np.load(FIXTURE)
    print("running the reference driver (91 load steps)", flush=True)
    r = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    iters = [int(m.group(2)) for m in re.finditer(r"Loading step (\d+) has finished in (\d+) iterations", r.stdout)]
    assert len(iters) == 91
    sd, disp = read_series(tmp_path / "result_disp.txt")
    sf, force = read_series(tmp_path / "result_force.txt")
    np.testing.assert_array_equal(sd, np.arange(92))
    np.testing.assert_array_equal(sf, np.arange(92))
    for name in ("result_strain.txt", "result_stress.txt", "result_position.dump", "result_bforce.txt",
                 "result_stretch.txt", "result_neighbor.txt"):
        assert (tmp_path / name).stat().st_size > 0, name
    ed = np.abs(disp[1:] - g["disp"]).max() / np.abs(g["disp"]).max()
    ef = np.abs(force[1:] - g["reaction"]).max() / np.abs(g["reaction"]).max()
    print(f"Newton iterations: device {sum(iters)}, oracle {int(g['newton_iters'].sum())}; largest difference "
          f"of the mean displacement {ed:.2e}, of the reaction {ef:.2e} (of their largest values)", flush=True)
    # same iterate sequence: the Newton count of every step is the oracle's
    np.testing.assert_array_equal(iters, g["newton_iters"])
    assert ed <= 1e-6 and ef <= 1e-6, (ed, ef)
