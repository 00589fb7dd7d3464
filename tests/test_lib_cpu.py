"""CPU checks of liblpm.so: it loads, exports every entry point its headers declare,
its host writers produce the reference's layouts, and its device entry points
(lattice generation included) fail loudly when no GPU is present."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from lpm_harness import P, ROOT, host_lattice

HEADERS = ["lpm.h", "boundary.h", "constitutive.h", "data_handler.h", "initialization.h", "lpm_basic.h",
           "neighbor.h", "solver.h", "stiffness.h", "lpmx.h", "lpm_ext.h"]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def test_library_exports_every_declared_function():
    L = P.lib()
    names = declared_functions()
    assert len(names) > 100
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_globals_are_exported():
    for g in ("nparticle", "xyz", "K_global", "dLp", "J2_beta", "damage_D", "box", "R_matrix", "pbc"):
        C.c_int.in_dll(P.lib(), g)


def test_sc_box_sizes_give_the_configured_counts(oracle):
    for k, n in ((20, 8000), (60, 216000), (100, 1000000)):
        L = P.SC_BOX[n]
        assert oracle.create_lattice(2, 0.5, [0, L, 0, L, 0, L]).shape[0] == n
    for k in (6, 8, 10):
        L = P.sc_box_for(k)
        assert oracle.create_lattice(2, 0.5, [0, L, 0, L, 0, L]).shape[0] == k ** 3


def test_cblas_dnrm2_and_helpers():
    rng = np.random.default_rng(0)
    v = rng.standard_normal(1000) * 1e150
    got = P.lib().cblas_dnrm2(1000, P.ptr(v), 1)
    assert abs(got - np.linalg.norm(v)) <= 1e-14 * np.linalg.norm(v)
    assert P.lib().cblas_dnrm2(3, P.ptr(np.array([3.0, 0.0, 4.0])), 1) == 5.0


def test_modulus_field_statistics():
    f = P.uniform_modulus(200000, 0)
    assert f.min() >= 0.95 and f.max() <= 1.05
    assert abs(f.mean() - 1.0) < 1e-3
    np.testing.assert_array_equal(f, P.uniform_modulus(200000, 0))
    assert not np.array_equal(f, P.uniform_modulus(200000, 2))


def test_device_calls_fail_loudly_without_gpu():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    desc = P.ModelDesc(3, 2, 8, 18, 6, 12, 60, 0.5, 1.0, 1.41, 1.0)
    x = np.zeros((8, 3))
    assert P.lib().lpmx_model_create(C.byref(desc), P.ptr(x)) == -5
    assert b"device" in P.lib().lpmx_last_error()
    spec = P.LatticeSpec(lattice=2, dim=3, pfr=2, rows=2, layers=2, hx=1.0, hy=1.0, hz=1.0, radius=0.5)
    cnt = C.c_int(-1)
    assert P.lib().lpmx_lattice_points(C.byref(spec), None, 0, C.byref(cnt)) == -5
    # the drop-in entry point exits with status 1, as the reference's fatal paths do
    code = ("import sys; sys.path.insert(0, %r); from lpm_harness import P; "
            "P.make_lattice(2, 0.5, [0, 5.125, 0, 5.125, 0, 5.125]); P.lib().searchNormalNeighbor()"
            % os.path.join(ROOT, "tests"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "device" in r.stderr


REF_SRC = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="the reference sources are not mounted here")
@pytest.mark.parametrize("driver", ["src/lpmc_project.c", "examples/3_point_bending_sq_brittle.c",
                                    "examples/shear_hex_brittle.c", "examples/CT_sc_ductile_nonlocal.c"])
def test_reference_drivers_compile_and_link_unchanged(tmp_path, driver):
    """The reference's own drivers build against include/ and liblpm.so with no edit
    (the drop-in claim at link level; nothing is run)."""
    lib_dir = os.path.join(ROOT, "lpm-c_amd")
    out = tmp_path / "drv"
    r = subprocess.run(["gcc", "-std=gnu11", "-O2", "-fopenmp", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(REF_SRC, driver), "-L" + lib_dir, "-llpm", "-Wl,-rpath," + lib_dir, "-lm",
                        "-o", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert out.exists()


def test_write_disp_reports_the_mean_displacement(oracle, tmp_path):
    """writeDisp (host, src/data_handler.c:278-306) appends "step mean-displacement" of
    one particle type along an axis: under a uniform translation every type reports
    the translation, a line per call."""
    L = P.sc_box_for(5)
    n, x = host_lattice(oracle, 2, 0.5, [0, L, 0, L, 0, L])
    t = P.sc_types(x)
    P.set_types(t, 5)
    u = np.array([0.125, -0.5, 0.25])
    xyz = P.gptr("xyz", 2)
    for i in range(n):
        for c in range(3):
            xyz[i][c] = x[i, c] + u[c]
    f = str(tmp_path / "disp.txt")
    fn = P.lib().writeDisp
    fn.argtypes = [C.c_char_p, C.c_char, C.c_int, C.c_int]
    fn.restype = None
    for step, (axis, typ) in enumerate([(b"z", 1), (b"x", 2), (b"y", 0)], start=3):
        fn(f.encode(), axis, typ, step)
    lines = open(f).read().split("\n")
    assert lines[-1] == ""
    got = [(int(a), float(b)) for a, b in (ln.split() for ln in lines[:-1])]
    assert got == [(3, 0.25), (4, 0.125), (5, -0.5)]


def test_write_dump_lammps_layout(oracle, tmp_path):
    """writeDump (host, src/data_handler.c:42-84) writes a LAMMPS dump: the timestep,
    the particle count, the box, the column header, then one row per particle with
    its id, type, position and displacement."""
    L = P.sc_box_for(4)
    n, x = host_lattice(oracle, 2, 0.5, [0, L, 0, L, 0, L])
    P.set_types(P.sc_types(x), 5)
    u = np.array([0.0, 0.0, 0.03125])
    xyz = P.gptr("xyz", 2)
    for i in range(n):
        xyz[i][2] = x[i, 2] + u[2]
    f = str(tmp_path / "dump.txt")
    fn = P.lib().writeDump
    fn.argtypes = [C.c_char_p, C.c_int, C.c_char]
    fn.restype = None
    fn(f.encode(), 7, b"s")
    lines = open(f).read().splitlines()
    assert lines[:4] == ["ITEM: TIMESTEP", "7", "ITEM: NUMBER OF ATOMS", str(n)]
    assert lines[4] == "ITEM: BOX BOUNDS pp pp pp"
    assert lines[8].startswith("ITEM: ATOMS id type x y z dx dy dz")
    rows = [ln.split() for ln in lines[9:]]
    assert len(rows) == n
    t = P.sc_types(x)
    for i, r in enumerate(rows):
        assert int(r[0]) == i and int(r[1]) == t[i]
        np.testing.assert_allclose([float(v) for v in r[2:8]], list(x[i] + u) + list(u), rtol=1e-4, atol=1e-4)


def test_write_dump_crystal_plasticity_columns(oracle, tmp_path):
    """writeDump with plmode 1 (host, src/data_handler.c:59-70): the slip header and,
    per particle, the stress, the accumulated slips of systems 1, 4, 10, 19, 9, 14,
    17 and 22 (cp_A_single, current slot), pl_flag and the total slip cp_A."""
    L = P.lib()
    Lb = P.sc_box_for(4)
    n, x = host_lattice(oracle, 2, 0.5, [0, Lb, 0, Lb, 0, Lb])
    P.set_types(P.sc_types(x), 5)
    a3 = L.allocDouble3D
    a3.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
    a3.restype = C.c_void_p
    a2 = L.allocDouble2D
    P.set_gptr("cp_A_single", C.cast(a3(n, 24, 3, 0.0), C.POINTER(C.POINTER(C.POINTER(C.c_double)))),
               C.c_double, 3)
    P.set_gptr("cp_A", a2(n, 3, 0.0), C.c_double, 2)
    cas = P.gptr("cp_A_single", 3)
    ca = P.gptr("cp_A", 2)
    st = P.gptr("stress_tensor", 2)
    pl = P.gptr("pl_flag", 1, C.c_int)
    for i in range(n):
        for k in range(24):
            cas[i][k][0] = 1e-3 * (i + 1) + 1e-5 * k
            cas[i][k][1] = -1.0  # other slots are not written
        ca[i][0] = 0.5 * (i + 1)
        for c in range(6):
            st[i][c] = 10.0 * c + i
        pl[i] = i % 2
    P.gint("plmode").value = 1
    try:
        f = str(tmp_path / "dump_cp.txt")
        fn = L.writeDump
        fn.argtypes = [C.c_char_p, C.c_int, C.c_char]
        fn.restype = None
        fn(f.encode(), 3, b"s")
    finally:
        P.gint("plmode").value = 0
    lines = open(f).read().splitlines()
    assert lines[8] == ("ITEM: ATOMS id type x y z dx dy dz s11 s22 s33 s23 s13 s12 gam11 gam12 gam13 gam14 "
                        "gam21 gam22 gam23 gam24 pl gam_total")
    rows = [ln.split() for ln in lines[9:]]
    assert len(rows) == n and all(len(r) == 24 for r in rows)
    sys_cols = [1, 4, 10, 19, 9, 14, 17, 22]
    for i, r in enumerate(rows):
        assert [float(v) for v in r[8:14]] == [float("%.4e" % (10.0 * c + i)) for c in range(6)]
        assert [float(v) for v in r[14:22]] == [float("%.4e" % (1e-3 * (i + 1) + 1e-5 * k)) for k in sys_cols]
        assert int(r[22]) == i % 2 and float(r[23]) == float("%.4e" % (0.5 * (i + 1)))
