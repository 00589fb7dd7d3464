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
import filecmp
import os
import re
import subprocess
import time

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


RESULT_FILES = ("result_disp.txt", "result_force.txt", "result_strain.txt", "result_stress.txt",
                "result_position.dump", "result_bforce.txt", "result_stretch.txt", "result_neighbor.txt",
                "result_brokenbonds.txt")


def run_driver(path, resident):
    env = dict(os.environ)
    env.pop("LPMX_DROPIN_RESIDENT", None)
    if resident:
        env["LPMX_DROPIN_RESIDENT"] = "1"
    path.mkdir()
    print(f"running the reference driver (91 load steps), resident mode {int(resident)}", flush=True)
    t0 = time.perf_counter()
    r = subprocess.run([BIN], cwd=path, env=env, capture_output=True, text=True, timeout=900)
    wall = time.perf_counter() - t0
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r, wall


@pytest.mark.skipif(not os.path.exists(BIN), reason="the reference driver is built from the reference sources, "
                                                     "which are not mounted where this tree was built")
def test_reference_driver_runs_unchanged(tmp_path):
    """The driver as built, and again with the drop-in layer in its resident mode
    (LPMX_DROPIN_RESIDENT=1: the device holds the model, host copies come back only
    when library code reads them): every result file is the same byte for byte."""
    g = np.load(FIXTURE)
    r, t_eager = run_driver(tmp_path / "eager", False)
    iters = [int(m.group(2)) for m in re.finditer(r"Loading step (\d+) has finished in (\d+) iterations", r.stdout)]
    assert len(iters) == 91
    sd, disp = read_series(tmp_path / "eager" / "result_disp.txt")
    sf, force = read_series(tmp_path / "eager" / "result_force.txt")
    np.testing.assert_array_equal(sd, np.arange(92))
    np.testing.assert_array_equal(sf, np.arange(92))
    for name in RESULT_FILES:
        assert (tmp_path / "eager" / name).stat().st_size > 0, name
    ed = np.abs(disp[1:] - g["disp"]).max() / np.abs(g["disp"]).max()
    ef = np.abs(force[1:] - g["reaction"]).max() / np.abs(g["reaction"]).max()
    print(f"Newton iterations: device {sum(iters)}, oracle {int(g['newton_iters'].sum())}; largest difference "
          f"of the mean displacement {ed:.2e}, of the reaction {ef:.2e} (of their largest values)", flush=True)
    # same iterate sequence: the Newton count of every step is the oracle's
    np.testing.assert_array_equal(iters, g["newton_iters"])
    assert ed <= 1e-6 and ef <= 1e-6, (ed, ef)
    _, t_res = run_driver(tmp_path / "resident", True)
    print(f"wall time of the driver: {t_eager:.1f} s, resident mode {t_res:.1f} s", flush=True)
    for name in RESULT_FILES:
        assert filecmp.cmp(tmp_path / "eager" / name, tmp_path / "resident" / name, shallow=False), name
