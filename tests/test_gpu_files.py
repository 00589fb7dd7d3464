"""File-level parity of the drop-in's outputs with the oracle (GPU).

A driver written against the reference API only (tests/drivers/hex_brittle.c: 2D
hexagonal elastic-brittle shear with the reference driver's loop, bond breakage and
re-solves) runs on the device through liblpm.so and writes the reference's files:
the broken-bond file of updateDamageGeneral (src/constitutive.c:1440-1442, 1481,
1519), writeDisp and writeReaction after every Newton loop (src/data_handler.c:
278-335) and writeDump of the last step (src/data_handler.c:42-84).  The oracle runs
the same steps and writes the bond, displacement and reaction files in the same
format (orc_set_output); the dump is checked against the oracle's arrays.
- bond file: identical text (which bonds break, in which order, per solve);
- displacement and reaction files: the same lines and step numbers, values equal
  to the 9 significant digits the format prints;
- final positions (printed in %.17g by the driver): FP64 displacement to 1e-10;
- dump: header, ids, types, flags identical; the %.4e columns to their resolution.
"""
import os
import subprocess

import numpy as np
import pytest

from lpm_harness import ROOT, TOL_FP64, P, assert_rel

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "tests", "drivers", "hex_brittle")


def read_lines(path):
    with open(path) as f:
        return f.read().splitlines()


def compare_step_file(got_path, want_path):
    got, want = read_lines(got_path), read_lines(want_path)
    assert len(got) == len(want), (got_path, len(got), len(want))
    steps_g = [int(r.split()[0]) for r in got]
    steps_w = [int(r.split()[0]) for r in want]
    assert steps_g == steps_w
    vg = np.array([float(r.split()[1]) for r in got])
    vw = np.array([float(r.split()[1]) for r in want])
    # "%8.8e" keeps 9 significant digits: values 1e-10 apart print within one unit
    assert np.all(np.abs(vg - vw) <= 1.5e-8 * np.abs(vw) + 1e-300), (got_path, np.abs(vg - vw).max())
    return steps_g, vw


def test_hex_brittle_driver_files_match_oracle(oracle, tmp_path):
    assert os.path.exists(DRIVER), "tests/drivers/hex_brittle is built by __graft_entry__.build()"
    W, H, crit, nbreak, tol, maxit = 14.0, 12.0, 1e-3, 20, 1e-11, 600
    shears = np.linspace(4e-4, 4e-3, 10)
    gdir, odir = tmp_path / "device", tmp_path / "oracle"
    gdir.mkdir()
    odir.mkdir()
    print(f"hex brittle driver: {W}x{H}, critical stretch {crit}, {len(shears)} shear steps", flush=True)
    r = subprocess.run([DRIVER, str(gdir), "%.17g" % W, "%.17g" % H, "%.17g" % crit, str(nbreak), "%.17g" % tol, str(maxit)] +
                       ["%.17g" % g for g in shears], capture_output=True, text=True, timeout=600, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    # the oracle on the same steps, writing the same files
    x0 = oracle.create_lattice(1, 0.5, [0.0, W, 0.0, H, 0.0, 1.0])
    n = len(x0)
    t = P.hex_types(x0)
    c = oracle.elastic_constants(70e3, 0.25)
    m = oracle.Model(x0, 1, 0.5, t, [c] * 3, plmode=6, sigmay=1e30, critical_bstrain=crit, nbreak=nbreak)
    m.setup()
    height = float(x0[:, 1].max() - x0[:, 1].min())
    steps = P.hex_shear_steps(height, shears)
    s = oracle.solver(rtol2=1e-24, atol2=0.0, maxit=200 * 2 * n, newton_tol=tol, newton_maxit=maxit, tangent=1,
                      fixed_diag=0)
    oracle.set_output(bond=odir / "result_bond.txt", reaction=odir / "result_reaction.txt",
                      disp=odir / "result_disp.txt", axis="x", t1=1)
    try:
        rc, out = oracle.run_steps(m, steps, s, out_axis="x", out_type=1)
    finally:
        oracle.set_output()
    assert rc == 0
    assert sum(o.broken for o in out) > 0, "the path breaks bonds"
    print(r.stdout[-3000:], flush=True)
    print("oracle steps:", [(o.broken, o.resolves, o.newton_iters) for o in out], flush=True)
    # full-precision positions
    xg = np.loadtxt(gdir / "positions.txt")
    assert_rel("displacement", xg - x0, m.x - x0, TOL_FP64)
    steps_seen, _ = compare_step_file(gdir / "result_reaction.txt", odir / "result_reaction.txt")
    compare_step_file(gdir / "result_disp.txt", odir / "result_disp.txt")
    assert len(steps_seen) > len(shears), "re-solves after breakage write their own lines"
    # broken-bond file: identical text
    bond_g, bond_w = read_lines(gdir / "result_bond.txt"), read_lines(odir / "result_bond.txt")
    assert bond_g == bond_w, (bond_g[:12], bond_w[:12])
    assert sum(1 for row in bond_w if not row.startswith("TIMESTEP")) > 0
    # dump of the last step
    rows = read_lines(gdir / "result_dump.txt")
    assert rows[0] == "ITEM: TIMESTEP" and int(rows[1]) == len(shears)
    assert rows[2] == "ITEM: NUMBER OF ATOMS" and int(rows[3]) == n
    assert rows[4] == "ITEM: BOX BOUNDS pp pp pp"
    for k, (lo, hi) in enumerate(((0.0, W), (0.0, H), (0.0, 1.0))):
        assert rows[5 + k] == f"{lo:8.8f} {hi:8.8f}"
    assert rows[8].startswith("ITEM: ATOMS id type x y z dx dy dz s11 s22 s33 s23 s13 s12 stress_eq")
    body = [r.split() for r in rows[9:9 + n]]
    assert len(body) == n and len(rows) == 9 + n
    ids = [int(b[0]) for b in body]
    assert ids == list(range(n))
    assert [int(b[1]) for b in body] == list(t)
    assert [int(b[20]) for b in body] == list(m.pl_flag)
    vals = np.array([[float(v) for v in b[2:20]] + [float(b[21])] for b in body])
    want = np.concatenate([m.x, m.x - x0, m.stress, m.seq[:, None], m.alpha[:, :1], m.sm[:, None],
                           m.triax[:, None], m.dloc[:, :1], m.dnl[:, :1], m.damage_visual[:, None]], axis=1)
    scale = np.abs(want).max(axis=0) + 1e-300
    # "%.4e" keeps 5 significant digits
    assert np.all(np.abs(vals - want) <= 6e-5 * np.abs(want) + 1e-12 * scale), np.abs(vals - want).max(axis=0)
