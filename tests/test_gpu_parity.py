"""Parity of the HIP path (through the drop-in API and lpmx) with the CPU oracle on
the same inputs.  Integer and index results are compared bit for bit; the
floating-point laws are the reference's evaluation order compiled without
contraction on both sides and are compared bit for bit too, except where a
device library function (exp) or a different summation order enters, which is
stated per check."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from lpm_harness import ROOT, TOL_CP, TOL_FP64, P, assert_rel, get, oracle_model, push, sc_pair, tol_for

pytestmark = pytest.mark.gpu


def host2d(name, n, w, dtype=np.float64):
    return P.read2d(name, n, w, dtype)


@pytest.mark.parametrize("kind", ["sc", "hex", "fcc"])
def test_neighbor_lists_and_pattern_bit_exact(oracle, kind):
    L = P.lib()
    if kind == "sc":
        lat, box = 2, [0, 7.321, 0, 7.321, 0, 7.321]
    elif kind == "hex":
        lat, box = 1, [0, 12.0, 0, 10.0, 0, 1.0]
    else:
        lat, box = 3, [0, 6.5, 0, 6.5, 0, 6.5]
    n = P.make_lattice(lat, 0.5, box)
    L.searchNormalNeighbor()
    lenK = L.searchAFEMNeighbor()
    x = P.positions()
    m = oracle_model(oracle, x, lat, np.zeros(n, np.int32), [(1.0, 0.5, 0.25)])
    nb = P.gint("nneighbors").value
    np.testing.assert_array_equal(host2d("neighbors", n, nb, np.int32), m.nbr)
    np.testing.assert_array_equal(host2d("nsign", n, nb, np.int32), m.nsign)
    np.testing.assert_array_equal(P.read1d("nb_initial", n, np.int32), m.nb_initial)
    np.testing.assert_array_equal(host2d("distance_initial", n, nb), m.L0)
    for c, name in enumerate(("csx_initial", "csy_initial", "csz_initial")):
        np.testing.assert_array_equal(host2d(name, n, nb), m.cs0[:, :, c])
    cm = P.gint("nneighbors_AFEM").value + 1
    np.testing.assert_array_equal(host2d("conn", n, cm, np.int32), m.conn)
    np.testing.assert_array_equal(P.read1d("nb_conn", n, np.int32), m.nb_conn)
    np.testing.assert_array_equal(host2d("K_pointer", n + 1, 2, np.int32), m.kptr)
    assert lenK == m.s.nK
    nb1 = P.gint("nneighbors1").value
    np.testing.assert_array_equal(host2d("neighbors1", n, nb1, np.int32), m.nbr1)


def test_modulus_field_belongs_to_its_lattice():
    """lpm_ext_set_modulus copies the field and a new lattice drops it: a model built
    after one with a modulus field has the plain constants (the harness's earlier
    builds once left the library a pointer into freed memory)."""
    L = P.lib()
    box = P.sc_box_for(5)
    emod = 1.0 + 0.05 * np.random.default_rng(3).uniform(-1, 1, 125)
    n, _, _, _ = P.build_sc_model(box, emod=emod)
    kn_field = host2d("Kn", n, 18)
    del emod  # the library holds its own copy
    L.calcKnTv()
    np.testing.assert_array_equal(host2d("Kn", n, 18), kn_field)
    # a new lattice without a field, against one built with an all-ones field
    P.make_lattice(2, 0.5, [0.0, box, 0.0, box, 0.0, box])
    L.searchNormalNeighbor()
    L.searchAFEMNeighbor()
    P.set_types(P.sc_types(P.positions()), 5)
    P.set_materials([P.elastic_constants(200e3, 0.3)] * 5, 200.0)
    L.calcKnTv()
    kn_plain = host2d("Kn", n, 18)
    assert not np.array_equal(kn_plain, kn_field)
    P.build_sc_model(box, emod=np.ones(n))
    np.testing.assert_array_equal(host2d("Kn", n, 18), kn_plain)


def test_bond_constants_and_tangent(oracle):
    L = P.lib()
    _, n, x, t, m = sc_pair(oracle, 6)
    np.testing.assert_array_equal(host2d("Kn", n, 18), m.Kn)
    np.testing.assert_array_equal(host2d("Tv", n, 18), m.Tv)
    rng = np.random.default_rng(4)
    xd = x + 2e-3 * rng.standard_normal(x.shape)
    dlp = 1e-4 * rng.standard_normal((n, 18)) * (m.nbr >= 0)
    push("XYZ", xd)
    push("DLP0", dlp)
    P.check(L.lpmx_tangent(), "tangent")
    nK = m.s.nK
    IK = np.zeros(3 * n + 1, np.int32)
    JK = np.zeros(nK, np.int32)
    K = np.zeros(nK)
    P.check(L.lpmx_export_upper_csr(P.ptr(IK), P.ptr(JK), P.ptr(K)), "export")
    m.x[:] = xd
    m.dLp[:, :, 0] = dlp
    oracle.lib().orc_tangent_analytic(m.p)
    np.testing.assert_array_equal(IK, m.IK)
    np.testing.assert_array_equal(JK, m.JK)
    # same Jacobian, different summation order of the contributions
    assert np.abs(K - m.K).max() <= 1e-12 * np.abs(m.K).max()
    oracle.lib().orc_tangent_fd(m.p)
    np.testing.assert_array_equal(IK, m.IK)
    np.testing.assert_array_equal(JK, m.JK)
    assert np.abs(K - m.K).max() <= 1e-5 * np.abs(m.K).max()
    # the device SpMV equals the upper-CSR product of the same values
    oracle.lib().orc_tangent_analytic(m.p)
    v = rng.standard_normal(3 * n)
    y = np.zeros(3 * n)
    P.check(L.lpmx_spmv(P.LPMX_FP64, P.ptr(v), P.ptr(y)), "spmv")
    m.A["K"][:] = K
    yo = np.zeros(3 * n)
    oracle.lib().orc_spmv_upper(m.p, oracle.ptr(v), oracle.ptr(yo))
    assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()
    P.check(L.lpmx_spmv(P.LPMX_MIXED, P.ptr(v), P.ptr(y)), "spmv32")
    assert np.abs(y - yo).max() <= 1e-5 * np.abs(yo).max()


@pytest.mark.parametrize("kind", ["sc", "hex"])
def test_tangent_with_broken_bonds(oracle, kind):
    """The tangent rows (bond records, eight rows per work-group, the last slice
    partly filled) with a share of broken bonds (damage_broken 0 on one side of a
    bond, or both): against the oracle's analytic tangent to 1e-12, SC and the 2D
    hexagonal lattice."""
    L = P.lib()
    if kind == "sc":
        _, n, x, t, m = sc_pair(oracle, 7)
        d = 3
    else:
        n, x, t, c = P.build_hex_model([0.0, 12.0, 0.0, 10.0, 0.0, 1.0], critical=1.0)
        m = oracle_model(oracle, x, 1, t, [c] * 3, plmode=6, sigmay=1e30, critical_bstrain=1.0, nbreak=20)
        d = 2
    nb = m.nbr.shape[1]
    rng = np.random.default_rng(11)
    xd = x + 2e-3 * rng.standard_normal(x.shape) * (np.arange(3) < d)
    dlp = 1e-4 * rng.standard_normal((n, nb)) * (m.nbr >= 0)
    broken = np.where((rng.random((n, nb)) < 0.1) & (m.nbr >= 0), 0.0, 1.0)
    push("XYZ", xd)
    push("DLP0", dlp)
    push("DAMAGE_BROKEN", broken)
    P.check(L.lpmx_tangent(), "tangent")
    nK = m.s.nK
    IK = np.zeros(d * n + 1, np.int32)
    JK = np.zeros(nK, np.int32)
    K = np.zeros(nK)
    P.check(L.lpmx_export_upper_csr(P.ptr(IK), P.ptr(JK), P.ptr(K)), "export")
    m.x[:] = xd
    m.dLp[:, :, 0] = dlp
    m.dmg_broken[:] = broken
    oracle.lib().orc_tangent_analytic(m.p)
    np.testing.assert_array_equal(IK, m.IK)
    np.testing.assert_array_equal(JK, m.JK)
    err = np.abs(K - m.K).max() / np.abs(m.K).max()
    print(f"{kind}: {n} particles, {int((broken == 0).sum())} broken bond sides, tangent rel diff {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("mode", [6, 0, 4])
def test_constitutive_laws_bit_exact(oracle, mode):
    L = P.lib()
    _, n, x, t, m = sc_pair(oracle, 6)
    rng = np.random.default_rng(9 + mode)
    eps = 2.5e-3 * np.array([[1.0, 0.1, 0.0], [0.1, -0.4, 0.05], [0.0, 0.05, -0.3]])
    xd = x + x @ eps.T + 2e-4 * rng.standard_normal(x.shape)
    xt = x + 0.5 * (xd - x)
    ft = 10.0 * rng.standard_normal((n, 18)) * (m.nbr >= 0)
    dlp = 2e-5 * rng.standard_normal((n, 18)) * (m.nbr >= 0)
    beta = 3.0 * rng.standard_normal((n, 6))
    alpha = 1e-3 * rng.uniform(size=n)
    for name, v in (("XYZ", xd), ("XYZ_TEMP", xt), ("F_TEMP", ft), ("DLP0", dlp), ("J2_BETA0", beta),
                    ("J2_ALPHA0", alpha)):
        push(name, v)
    # predictor uses the direction cosines of the last evaluation: give both the same
    P.check(L.lpmx_computedL(), "computedL")
    P.check(L.lpmx_bond_force(mode), "bond force")
    m.x[:] = xd
    m.xt[:] = xt
    m.Ftemp[:] = ft
    m.dLp[:, :, 0] = dlp
    m.beta[:, :, 0] = beta
    m.alpha[:, 0] = alpha
    oracle.lib().orc_computedL(m.p)
    oracle.lib().orc_bond_force_general(m.p, mode)
    nb = (n, 18)
    np.testing.assert_array_equal(get("F", nb), m.F)
    np.testing.assert_array_equal(get("PIN", (3 * n,)), m.Pin)
    np.testing.assert_array_equal(get("STRESS", (n, 6)), m.stress)
    np.testing.assert_array_equal(get("BOND_STRESS", nb), m.bond_stress)
    if mode != 4:
        np.testing.assert_array_equal(get("DL", nb), m.dL)
        for c, f in enumerate(("CSX", "CSY", "CSZ")):
            np.testing.assert_array_equal(get(f, nb), m.cs[:, :, c])
    if mode == 0:
        dl = get("J2_DLAMBDA", (n,))
        assert (dl > 0).sum() > n // 4, "the test deformation yields a good share of particles"
        np.testing.assert_array_equal(dl, m.dlambda)
        np.testing.assert_array_equal(get("DLP0", nb), m.dLp[:, :, 0])
        np.testing.assert_array_equal(get("J2_BETA0", (n, 6)), m.beta[:, :, 0])
        np.testing.assert_array_equal(get("J2_ALPHA0", (n,)), m.alpha[:, 0])
        np.testing.assert_array_equal(get("DL_AVE", nb), m.dL_ave)
        np.testing.assert_array_equal(get("PL_FLAG", (n,)), m.pl_flag)


def test_bc_elimination_and_solve(oracle):
    L = P.lib()
    _, n, x, t, m = sc_pair(oracle, 6)
    L.calcStiffness3DFiniteDifference(6)
    # bottom face fully fixed, top face fixed in z: no rigid-body mode left
    dbc = np.ones(3 * n, np.int32)
    dbc.reshape(n, 3)[t == 2, :] = 0
    dbc.reshape(n, 3)[t == 3, :] = 0
    dbc.reshape(n, 3)[t == 1, 2] = 0
    P.write1d("dispBC_index", dbc)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(3 * n) * dbc
    P.write1d("residual", r)
    L.setDispBC_stiffnessUpdate3D()
    nK = m.s.nK
    K = P.read1d("K_global", nK)
    oracle.lib().orc_tangent_analytic(m.p)
    m.dispBC[:] = dbc
    m.residual[:] = r
    oracle.lib().orc_bc_stiffness_update(m.p)
    np.testing.assert_array_equal(P.read1d("residual", 3 * n), m.residual)
    assert np.abs(K - m.K).max() <= 1e-12 * np.abs(m.K).max()
    dense = oracle.upper_csr_dense(m)
    want = np.linalg.solve(dense, m.residual)
    for prec, tol in ((P.LPMX_FP64, 1e-10), (P.LPMX_MIXED, 1e-10)):
        s = P.solver(precision=prec, rtol2=1e-26, atol2=0.0, maxit=50 * 3 * n)
        L.lpm_ext_set_solver(C.byref(s))
        P.write1d("residual", m.residual)
        x0 = P.positions()
        L.setDispBC_stiffnessUpdate3D()
        L.solverCG()
        d = P.read1d("disp", 3 * n)
        assert np.abs(d - want).max() <= tol * np.abs(want).max(), prec
        np.testing.assert_allclose(P.positions() - x0, d.reshape(n, 3), rtol=0, atol=1e-15)
    L.lpm_ext_set_solver(None)


def test_nonlocal_damage_and_crack(oracle):
    L = P.lib()
    _, n, x, t, m = sc_pair(oracle, 6)
    rng = np.random.default_rng(5)
    dlam = rng.uniform(0, 1e-2, n)
    dlam[rng.uniform(size=n) < 0.05] = 2.0  # a few particles far past the threshold
    tri = rng.uniform(-0.5, 0.5, n)
    for name, v in (("J2_DLAMBDA", dlam), ("J2_TRIAXIALITY", tri)):
        push(name, v)
    P.set_params(plmode=0, damage_L=0.5, damage_threshold=0.9, damagec_A=0.3)
    P.check(L.lpmx_set_params(C.byref(P.Params(0, 2000.0, 0.0, 0.5, 0.9, 0.3, 1e-2, 20))), "params")
    broken = C.c_int()
    P.check(L.lpmx_damage(0, C.byref(broken)), "damage")
    m.dlambda[:] = dlam
    m.triax[:] = tri
    m.s.damagec_A = 0.3
    k = oracle.lib().orc_damage_pwise_nonlocal(m.p)
    assert broken.value == k and k > 0
    # Gaussian weights use the device exp: agreement to rounding
    np.testing.assert_allclose(get("DAMAGE_NONLOCAL0", (n,)), m.dnl[:, 0], rtol=1e-14, atol=0)
    np.testing.assert_array_equal(get("DAMAGE_BROKEN", (n, 18)), m.dmg_broken)
    np.testing.assert_allclose(get("DAMAGE_W", (n, 18)), m.dmg_w, rtol=1e-13, atol=1e-15)
    # crack bookkeeping on forces of a deformed state
    ft = rng.standard_normal((n, 18)) * (m.nbr >= 0)
    push("F", ft)
    push("DAMAGE_W", m.dmg_w)
    P.check(L.lpmx_computedL(), "computedL")
    P.check(L.lpmx_update_crack(), "crack")
    m.F[:] = ft
    oracle.lib().orc_computedL(m.p)
    oracle.lib().orc_update_crack(m.p)
    np.testing.assert_array_equal(get("NB", (n,)), m.nb)
    np.testing.assert_array_equal(get("F", (n, 18)), m.F)
    np.testing.assert_array_equal(get("PIN", (3 * n,)), m.Pin)
    np.testing.assert_array_equal(get("FIX_INDEX", (3 * n,)), m.fixidx)


def test_brittle_selection_bit_exact(oracle):
    L = P.lib()
    n = P.make_lattice(1, 0.5, [0, 14.0, 0, 12.0, 0, 1.0])
    L.searchNormalNeighbor()
    L.searchAFEMNeighbor()
    x = P.positions()
    P.set_types(np.zeros(n, np.int32), 1)
    c = P.elastic_constants(70e3, 0.25)
    P.set_materials([c], 200.0)
    P.set_params(plmode=6, critical_bstrain=1e-3, nbreak=7)
    L.calcKnTv()
    m = oracle_model(oracle, x, 1, np.zeros(n, np.int32), [c], plmode=6, critical_bstrain=1e-3, nbreak=7)
    rng = np.random.default_rng(8)
    live = m.nbr >= 0
    # symmetric stretches with many ties, as (i, j) and (j, i) carry the same dL
    s = np.round(rng.uniform(0, 2e-3, (n, 12)), 5)
    for i in range(n):
        for a in range(m.nb_initial[i]):
            j = m.nbr[i, a]
            if j < i:
                b = list(m.nbr[j]).index(i)
                s[i, a] = s[j, b]
    dl = np.where(live, s * m.L0, 0.0)
    push("DL", dl)
    P.check(L.lpmx_set_params(C.byref(P.Params(6, 0.0, 0.0, 0.5, 0.9, 0.0, 1e-3, 7))), "params")
    k = C.c_int()
    P.check(L.lpmx_damage(6, C.byref(k)), "brittle")
    m.dL[:] = dl
    bi = np.zeros(64, np.int32)
    bj = np.zeros(64, np.int32)
    ko = oracle.lib().orc_damage_brittle(m.p, oracle.ptr(bi), oracle.ptr(bj), 64)
    assert k.value == ko and ko > 7
    np.testing.assert_array_equal(get("DAMAGE_BROKEN", (n, 12)), m.dmg_broken)
    gi = np.zeros(64, np.int32)
    gj = np.zeros(64, np.int32)
    cnt = C.c_int()
    P.check(L.lpmx_get_broken(P.ptr(gi), P.ptr(gj), 64, C.byref(cnt)), "broken")
    assert cnt.value == 7
    np.testing.assert_array_equal(gi[:7], bi[:7])
    np.testing.assert_array_equal(gj[:7], bj[:7])


@pytest.mark.parametrize("prec", [P.LPMX_FP64, P.LPMX_MIXED])
def test_load_steps_match_oracle(oracle, prec):
    """An elastic step and three plastic steps of uniaxial tension (0.05% .. 0.2%).
    Newton runs to 1e-11 on both sides; the FP64 CG to 1e-12 relative residual, the
    mixed form's refinement to 1e-8 (its FP64 residual cannot be driven further
    through FP32 corrections), which leaves the converged Newton solution unchanged."""
    k = 6
    _, n, x, t, m = sc_pair(oracle, k)
    steps = P.sc_uniaxial_steps(k - 1.0, np.array([0.5e-3, 1e-3, 1.5e-3, 2e-3]))
    s = P.solver(precision=prec, rtol2=1e-24 if prec == P.LPMX_FP64 else 1e-16, atol2=0.0, maxit=100 * 3 * n,
                 newton_tol=1e-11, newton_maxit=600)
    rc, out = P.run_steps(steps, s, out_axis="z", out_type=1)
    assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters, o.cg_iters, o.ir_passes) for o in out])
    so = oracle.solver(rtol2=1e-24, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-11, newton_maxit=600, tangent=1,
                       fixed_diag=1)
    rco, outo = oracle.run_steps(m, steps, so, out_axis="z", out_type=1)
    assert rco == 0
    xg = get("XYZ", (n, 3))
    u, uo = xg - x, m.x - x
    assert_rel("displacement", u, uo, tol_for(prec))
    r = np.array([o.reaction for o in out])
    ro = np.array([o.reaction for o in outo])
    np.testing.assert_allclose(r, ro, rtol=tol_for(prec))
    dl = get("J2_ALPHA1", (n,))
    assert dl.max() > 0, "the load path reaches the plastic regime"
    np.testing.assert_allclose(dl, m.alpha[:, 1], rtol=0, atol=1e-10 * dl.max())
    np.testing.assert_allclose(get("DAMAGE_NONLOCAL1", (n,)), m.dnl[:, 1], rtol=0, atol=1e-9 * m.dnl[:, 1].max())


@pytest.mark.parametrize("prec", [P.LPMX_FP64, P.LPMX_MIXED])
def test_accelerated_newton_same_solution(oracle, prec):
    """Anderson-accelerated Newton reaches the oracle's (plain modified-Newton)
    converged solution on the elastic + plastic path, in fewer iterations."""
    k = 6
    _, n, x, t, m = sc_pair(oracle, k)
    steps = P.sc_uniaxial_steps(k - 1.0, np.array([0.5e-3, 1e-3, 1.5e-3, 2e-3]))
    # Newton to 1e-12: at 1e-11 the plastic steps fix the displacements to ~1.5e-10 only
    s = P.solver(precision=prec, rtol2=1e-24 if prec == P.LPMX_FP64 else 1e-16, atol2=0.0, maxit=100 * 3 * n,
                 newton_tol=1e-12, newton_maxit=600, accel_depth=5)
    rc, out = P.run_steps(steps, s, out_axis="z", out_type=1)
    assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters, o.cg_iters) for o in out])
    so = oracle.solver(rtol2=1e-26, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-12, newton_maxit=1200, tangent=1,
                       fixed_diag=1)
    rco, outo = oracle.run_steps(m, steps, so, out_axis="z", out_type=1)
    assert rco == 0
    u, uo = get("XYZ", (n, 3)) - x, m.x - x
    assert_rel("displacement", u, uo, tol_for(prec))
    np.testing.assert_allclose([o.reaction for o in out], [o.reaction for o in outo], rtol=tol_for(prec))
    ni = [o.newton_iters for o in out]
    nio = [o.newton_iters for o in outo]
    assert sum(ni[2:]) < 0.5 * sum(nio[2:]), (ni, nio)
    assert all(o.accel_mixed > 0 for o in out[1:])


DRIVER = os.path.join(ROOT, "tests", "drivers", "sc_uniaxial")


def test_reference_api_driver_matches_oracle(oracle, tmp_path):
    """A C driver that uses only the reference API (lpm.h) and its own Newton loop,
    linked to liblpm.so, reproduces the oracle's load steps (drop-in at run time)."""
    assert os.path.exists(DRIVER), "tests/drivers/sc_uniaxial is built by __graft_entry__.build()"
    k = 6
    L, height = P.sc_box_for(k), k - 1.0
    strains = [0.5e-3, 1e-3]
    out = tmp_path / "driver.txt"
    print(f"reference-API driver: k={k}, strains {strains}", flush=True)
    r = subprocess.run([DRIVER, str(out), repr(L), repr(height), "1e-11", "600"] + [repr(e) for e in strains],
                       capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = open(out).read().split("\n")
    steps = [row.split() for row in rows[:len(strains)]]
    xg = np.array([[float(v) for v in row.split()] for row in rows[len(strains):] if row.strip()])
    x0 = oracle.create_lattice(2, 0.5, [0.0, L, 0.0, L, 0.0, L])
    n = len(x0)
    assert xg.shape == (n, 3)
    c = P.elastic_constants(200e3, 0.3)
    m = oracle_model(oracle, x0, 2, P.sc_types(x0), [c] * 5, plmode=0, sigmay=200.0, J2_H=2000.0)
    so = oracle.solver(rtol2=1e-24, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-11, newton_maxit=600, tangent=1,
                       fixed_diag=0)
    rco, outo = oracle.run_steps(m, P.sc_uniaxial_steps(height, np.array(strains)), so, out_axis="z", out_type=1)
    assert rco == 0
    assert_rel("reaction", [float(s[1]) for s in steps], [o.reaction for o in outo], TOL_FP64)
    u, uo = xg - x0, m.x - x0
    assert_rel("displacement", u, uo, TOL_FP64)
    assert int(steps[-1][3]) > 1


@pytest.mark.parametrize("prec", [P.LPMX_FP64, P.LPMX_MIXED])
def test_brittle_shear_steps_same_broken_bonds(oracle, prec):
    """Elastic-brittle hexagonal lattice in shear with a Weibull per-bond critical
    stretch: bonds break, the tangent is rebuilt and the step re-solved; the broken
    bond set, the per-step break counts, reactions and displacements match the
    oracle (broken sets bit-exact)."""
    box = [0.0, 14.0, 0.0, 12.0, 0.0, 1.0]
    n, x, t, c = P.build_hex_model(box, critical=1e-3, nbreak=20)
    nb = P.gint("nneighbors").value
    crit = P.weibull_bond_critical(get("NEIGHBORS", (n, nb)), 1, shape=10.0, scale=1e-3)
    P.build_hex_model(box, critical=1e-3, nbreak=20, crit=crit)
    m = oracle_model(oracle, x, 1, t, [c] * 3, plmode=6, sigmay=1e30, critical_bstrain=1e-3, nbreak=20, crit=crit)
    height = float(x[:, 1].max() - x[:, 1].min())
    steps = P.hex_shear_steps(height, np.linspace(4e-4, 4e-3, 10))
    s = P.solver(precision=prec, rtol2=1e-24 if prec == P.LPMX_FP64 else 1e-16, atol2=0.0, maxit=100 * 2 * n,
                 newton_tol=1e-11, newton_maxit=600)
    rc, out = P.run_steps(steps, s, out_axis="x", out_type=1)
    assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters, o.broken) for o in out])
    so = oracle.solver(rtol2=1e-24, atol2=0.0, maxit=100 * 2 * n, newton_tol=1e-11, newton_maxit=600, tangent=1,
                       fixed_diag=1)
    rco, outo = oracle.run_steps(m, steps, so, out_axis="x", out_type=1)
    assert rco == 0
    assert [o.broken for o in out] == [o.broken for o in outo]
    assert sum(o.broken for o in outo) > 0, "the load path breaks bonds"
    np.testing.assert_array_equal(get("DAMAGE_BROKEN", (n, nb)), m.dmg_broken)
    r, ro = np.array([o.reaction for o in out]), np.array([o.reaction for o in outo])
    assert_rel("reaction", r, ro, tol_for(prec))
    u, uo = get("XYZ", (n, 3)) - x, m.x - x
    assert_rel("displacement", u, uo, tol_for(prec))


CP = dict(tau0=10.0, taus=1e20, h0=100.0, q=1.0, eta=1.0, p=1.0, dtime=1.0, maxloop=20)


def cp_pair(oracle, L):
    n, x, t, c = P.build_fcc_cp_model(L, CP)
    m = oracle_model(oracle, x, 3, t, [c] * 5, plmode=1, sigmay=1e30, cp=CP)
    oracle.lib().orc_compute_cab(m.p)
    return n, x, t, c, m


@pytest.mark.parametrize("case", ["single", "multi", "random"])
def test_crystal_plasticity_return_map_bit_exact(oracle, case):
    """computeCab and one plmode-1 force evaluation: slip stiffness, active slip sets,
    slips, hardening, plastic stretch, forces and stress bit-exact with the oracle,
    for single slip, [001] multi-slip (eight tied systems) and a random strain."""
    L = 6.5
    n, x, t, c, m = cp_pair(oracle, L)
    np.testing.assert_array_equal(get("CP_CAB", (n, 24 * 24)), m.cp_Cab.reshape(n, -1))
    mu = c[2]
    if case == "single":
        d = np.array([-1.0, 2.0, 3.0]) / np.sqrt(14.0)
        eps = np.outer(d, d)
        fac = max(2 * mu * (s[0] * d[0] ** 2 + s[1] * d[1] ** 2 + s[2] * d[2] ** 2 + s[3] * d[1] * d[2] +
                            s[4] * d[0] * d[2] + s[5] * d[0] * d[1]) for s in m.schmid)
        eps *= 1.02 * CP["tau0"] / fac
    elif case == "multi":
        eps = np.diag([-0.3, -0.3, 1.0]) * 1.5 * CP["tau0"] / (mu * 0.8)
    else:
        rng = np.random.default_rng(5)
        a = rng.standard_normal((3, 3))
        eps = 2e-3 * (a + a.T) / 2
    xd = x + x @ eps.T
    if case == "random":
        xd = xd + 1e-4 * np.random.default_rng(6).standard_normal(x.shape)
    push("XYZ", xd)
    P.check(P.lib().lpmx_bond_force(1), "plmode 1")
    fails = C.c_int()
    P.check(P.lib().lpmx_cp_failures(C.byref(fails)), "failures")
    m.x[:] = xd
    oracle.lib().orc_bond_force_general(m.p, 1)
    assert fails.value == 0 and m.s.cp_fail == 0
    jact = get("CP_JACT", (n, 24))
    np.testing.assert_array_equal(jact, m.cp_Jact)
    assert jact.any(), "the case reaches plastic slip"
    if case == "multi":
        assert jact.sum(axis=1).max() > 1
    for name, ref in (("CP_DA_SINGLE", m.cp_dA_single), ("CP_DGY", m.cp_dgy), ("CP_RSS", m.cp_RSS),
                      ("CP_GY0", m.cp_gy[:, :, 0]), ("CP_A_SINGLE0", m.cp_A_single[:, :, 0])):
        np.testing.assert_array_equal(get(name, ref.shape), ref, err_msg=name)
    np.testing.assert_array_equal(get("CP_A0", (n,)), m.cp_A[:, 0])
    np.testing.assert_array_equal(get("CP_DA", (n,)), m.cp_dA)
    np.testing.assert_array_equal(get("DLP0", (n, 18)), m.dLp[:, :, 0])
    np.testing.assert_array_equal(get("F", (n, 18)), m.F)
    np.testing.assert_array_equal(get("PIN", (3 * n,)), m.Pin)
    np.testing.assert_array_equal(get("STRESS", (n, 6)), m.stress)
    np.testing.assert_array_equal(get("PL_FLAG", (n,)), m.pl_flag)


@pytest.mark.parametrize("prec", [P.LPMX_FP64, P.LPMX_MIXED])
def test_crystal_plasticity_load_steps_match_oracle(oracle, prec):
    """FCC crystal under [001] tension through first slip: the fused device loop and
    the oracle reach the same state (active slip sets bit-exact at the end).  Both
    accelerate the modified-Newton iteration (its contraction is ~0.996 once slip
    starts); the equilibrium reached does not depend on that."""
    L = 4.3
    n, x, t, c, m = cp_pair(oracle, L)
    height = float(x[:, 2].max() - x[:, 2].min())
    e_y = CP["tau0"] / (c[2] * 0.8)
    steps = P.sc_uniaxial_steps(height, np.array([0.5, 1.0, 1.3]) * e_y)
    s = P.solver(precision=prec, rtol2=1e-24 if prec == P.LPMX_FP64 else 1e-16, atol2=0.0, maxit=100 * 3 * n,
                 newton_tol=1e-11, newton_maxit=800, accel_depth=5)
    rc, out = P.run_steps(steps, s, out_axis="z", out_type=1)
    assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters) for o in out])
    so = oracle.solver(rtol2=1e-24, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-11, newton_maxit=800, tangent=1,
                       fixed_diag=1, accel=5)
    rco, outo = oracle.run_steps(m, steps, so, out_axis="z", out_type=1)
    assert rco == 0 and m.s.cp_fail == 0
    jact = get("CP_JACT", (n, 24))
    assert jact.any(), "the path reaches slip"
    np.testing.assert_array_equal(jact, m.cp_Jact)
    r, ro = np.array([o.reaction for o in out]), np.array([o.reaction for o in outo])
    assert_rel("reaction", r, ro, tol_for(prec))
    u, uo = get("XYZ", (n, 3)) - x, m.x - x
    assert_rel("displacement", u, uo, TOL_CP)


def test_j2_vjp_matches_oracle(oracle):
    """The device reverse sweep of one J2 force evaluation against the oracle's
    (itself pinned by central differences): same seeds, same state."""
    rng = np.random.default_rng(21)
    k = 6
    L = P.sc_box_for(k)
    n0 = k ** 3
    emod = 1.0 + 0.05 * rng.uniform(-1, 1, n0)
    sy = 200.0 + 20 * rng.uniform(-1, 1, n0)
    n, x, t, c = P.build_sc_model(L, sigmay=sy, H=2000.0, emod=emod)
    P.check(P.lib().lpmx_set_params(C.byref(P.Params(0, 2000.0, 0.3, 0.5, 0.9, 0.0, 1e-2, 20))), "params")
    m = oracle_model(oracle, x, 2, t, [c] * 5, plmode=0, sigmay=sy, J2_H=2000.0, J2_xi=0.3, emod=emod)
    nb = m.nbmax
    live = m.nbr >= 0
    eps = 1.2e-3 * np.array([[1.0, 0.1, 0], [0.1, -0.4, 0.05], [0, 0.05, -0.3]])
    xd = x + x @ eps.T + 2e-5 * rng.standard_normal((n, 3))
    dlp = 1e-5 * rng.standard_normal((n, nb)) * live
    beta = 5.0 * rng.standard_normal((n, 6))
    alpha = 1e-3 * rng.uniform(0, 1, n)
    push("XYZ", xd)
    push("DLP0", dlp)
    push("J2_BETA0", beta)
    push("J2_ALPHA0", alpha)
    m.x[:] = xd
    m.dLp[:, :, 0] = dlp
    m.beta[:, :, 0] = beta
    m.alpha[:, 0] = alpha
    seeds = [rng.standard_normal(3 * n), np.ascontiguousarray(rng.standard_normal((n, nb)) * live * 1e5),
             np.ascontiguousarray(rng.standard_normal((n, 6))), rng.standard_normal(n) * 1e3]
    dev = [np.zeros(3 * n), np.zeros((n, nb)), np.zeros((n, 6)), np.zeros(n), np.zeros(n), np.zeros(n)]
    hb = np.zeros(1)
    P.check(P.lib().lpmx_vjp(*[P.ptr(s) for s in seeds], *[P.ptr(o) for o in dev], P.ptr(hb)), "vjp")
    ref = [np.zeros(3 * n), np.zeros((n, nb)), np.zeros((n, 6)), np.zeros(n), np.zeros(n), np.zeros(n)]
    hr = C.c_double()
    oracle.lib().orc_eval_j2(m.p)
    assert m.pl_flag.sum() > 0, "the state is past yield"
    oracle.lib().orc_vjp_j2(m.p, *[oracle.ptr(s) for s in seeds], *[oracle.ptr(o) for o in ref], C.byref(hr))
    for name, a, b in zip(["positions", "dLp0", "beta0", "alpha0", "modulus", "sigmay"], dev, ref):
        assert np.abs(b).max() > 0, name
        assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max(), name
    assert abs(hb[0] - hr.value) <= 1e-11 * abs(hr.value)


# ------------------------------------------------------------------ history adjoint --
def _history_case(k, seed, strains):
    rng = np.random.default_rng(seed)
    n0 = k ** 3
    emod = 1.0 + 0.05 * rng.uniform(-1, 1, n0)
    sy = 200.0 + 30.0 * rng.uniform(-1, 1, n0)
    L = P.sc_box_for(k)
    return L, emod, sy, P.sc_uniaxial_steps(L - 1.0, np.array(strains)), rng


def _dev_solver(n, tol=1e-12):
    return P.solver(precision=P.LPMX_FP64, rtol2=1e-26, atol2=0.0, maxit=100 * 3 * n, newton_tol=tol,
                    newton_maxit=800, accel_depth=5)


def test_history_adjoint_matches_oracle(oracle):
    """Device discrete adjoint of a three-step J2 tension history (damage off) against
    the oracle's, whose adjoint systems are dense LU solves (and which is pinned by
    central differences of its own forward history).  With a checkpoint every step
    the reverse sweep re-runs two of the three steps: the recomputed trajectory is
    bitwise the recorded one, so the gradients are identical."""
    L, emod, sy, steps, _ = _history_case(5, 5, [0.6e-3, 1.1e-3, 1.5e-3])
    n, x, t, c = P.build_sc_model(L, sigmay=sy, H=2000.0, emod=emod)
    s = _dev_solver(n)
    res = {}
    for ck in (0, 1):
        if ck:
            P.build_sc_model(L, sigmay=sy, H=2000.0, emod=emod)
        rc, R, out = P.history_forward(steps, s, ckpt_every=ck)
        assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters) for o in out])
        rbar = 2.0 * (R - 0.9 * R)
        x_after = get("XYZ", (n, 3)).copy()
        a_after = get("J2_ALPHA1", (n,)).copy()
        o = P.adjoint_opts(tol=1e-12, maxit=400, depth=8, rtol2=1e-20, n=n)
        rc, ge, gy, gh, st = P.history_backward(o, rbar)
        assert rc == 0, P.lib().lpmx_last_error()
        print(f"ckpt {ck}: adj iters {st.adj_iters} (max {st.adj_iters_max}), cg {st.adj_cg_iters}, "
              f"recomputed {st.recomputed_steps}, rel res {st.max_rel_res:.2e}, free {st.max_rel_free:.2e}")
        assert st.recomputed_steps == (2 if ck else 0)
        assert st.max_rel_free <= 1e-10
        res[ck] = (R, ge, gy, gh)
        # the model is left as the forward left it
        np.testing.assert_array_equal(get("XYZ", (n, 3)), x_after)
        np.testing.assert_array_equal(get("J2_ALPHA1", (n,)), a_after)
    for a, b in zip(res[0], res[1]):
        np.testing.assert_array_equal(a, b)
    R, ge, gy, gh = res[0]
    m = oracle_model(oracle, x, 2, t, [c] * 5, plmode=0, sigmay=sy, J2_H=2000.0, emod=emod)
    so = oracle.solver(rtol2=1e-26, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-12, newton_maxit=800, tangent=1,
                       fixed_diag=1, accel=5, skip_damage=1)
    rc, J, Ro, geo, gyo, gho = oracle.adjoint_run(m, steps, so, 0.9 * R)
    assert rc == 0
    assert m.pl_flag.sum() > 0
    assert_rel("reaction history", R, Ro, TOL_FP64)
    for name, a, b in (("modulus", ge, geo), ("sigmay", gy, gyo)):
        assert np.abs(b).max() > 0, name
        assert np.abs(a - b).max() <= 1e-7 * np.abs(b).max(), (name, np.abs(a - b).max(), np.abs(b).max())
    assert abs(gh - gho) <= 1e-7 * abs(gho), (gh, gho)


def test_history_adjoint_central_differences_and_autograd():
    """Gradients of J = sum (R_n - t_n)^2 over a six-step J2 tension history on the
    device (checkpoint every 2 steps), through the torch.autograd harness, against
    central differences of the device forward: the global modulus E, a random
    direction of the per-particle modulus field, the yield stress field along a random
    direction, and H, each to 1e-6."""
    import torch

    k = 7
    L, emod, sy, steps, rng = _history_case(k, 8, [0.5e-3, 1.0e-3, 1.4e-3, 1.8e-3, 2.2e-3, 2.6e-3])
    n = k ** 3
    E0, H0 = 200e3, 2000.0
    spec = dict(L=L, nu=0.3, steps=steps, solver=_dev_solver(n), ckpt_every=2,
                opts=P.adjoint_opts(tol=1e-11, maxit=400, depth=8, rtol2=1e-16, n=n))
    F = P.reaction_history()

    def reactions(E, em, s_, H):
        with torch.no_grad():
            return F.apply(torch.tensor(E, dtype=torch.float64), torch.from_numpy(em), torch.from_numpy(s_),
                           torch.tensor(H, dtype=torch.float64), spec).numpy()

    target = 0.9 * reactions(E0, emod, sy, H0)
    tE = torch.tensor(E0, dtype=torch.float64, requires_grad=True)
    tm = torch.from_numpy(emod.copy()).requires_grad_(True)
    ts = torch.from_numpy(sy.copy()).requires_grad_(True)
    tH = torch.tensor(H0, dtype=torch.float64, requires_grad=True)
    R = F.apply(tE, tm, ts, tH, spec)
    J = ((R - torch.from_numpy(target)) ** 2).sum()
    J.backward()
    assert (P.field("PL_FLAG", n) > 0).sum() > 0, "the history reaches the plastic regime"

    def Jof(E, em, s_, H):
        return float(np.sum((reactions(E, em, s_, H) - target) ** 2))

    d1 = rng.standard_normal(n)
    d2 = rng.standard_normal(n)

    # Each forward history carries its Newton tolerance (1e-12 of the reaction) as
    # noise, which a central difference divides by 2h; the truncation error is O(h^2).
    # h = 1e-5 keeps both near 1e-7 for E and the two fields; J moves least with H, so
    # its step is 1e-4 (noise ~1e-8 of J, truncation ~1e-8).
    def fd(h, plus, minus):
        return (Jof(*plus) - Jof(*minus)) / (2 * h)

    hs = dict(E=1e-5, field=1e-5, H=1e-4)
    checks = [
        ("E", hs["E"], fd(hs["E"], (E0 * (1 + hs["E"]), emod, sy, H0), (E0 * (1 - hs["E"]), emod, sy, H0)) / E0,
         tE.grad.item()),
        ("modulus field", hs["field"],
         fd(hs["field"], (E0, emod * (1 + hs["field"] * d1), sy, H0), (E0, emod * (1 - hs["field"] * d1), sy, H0)),
         float(tm.grad.numpy() @ (emod * d1))),
        ("yield stress", hs["field"],
         fd(hs["field"], (E0, emod, sy * (1 + hs["field"] * d2), H0), (E0, emod, sy * (1 - hs["field"] * d2), H0)),
         float(ts.grad.numpy() @ (sy * d2))),
        ("H", hs["H"], fd(hs["H"], (E0, emod, sy, H0 * (1 + hs["H"])), (E0, emod, sy, H0 * (1 - hs["H"]))) / H0,
         tH.grad.item()),
    ]
    for name, h, f, a in checks:
        print(f"h {h:g} {name}: FD {f:.10e} adjoint {a:.10e} rel {abs(f - a) / abs(a):.2e}", flush=True)
    for name, h, f, a in checks:
        assert abs(f - a) <= 1e-6 * abs(a), (h, name, f, a)


# ------------------------------------------------------------------ AMG ------------
@pytest.mark.parametrize("prec", [P.LPMX_FP64, P.LPMX_MIXED])
def test_amg_solve_matches_block_jacobi(capfd, prec):
    """The smoothed-aggregation AMG preconditioner (20^3 lattice: a hierarchy of three
    levels) gives the block-Jacobi solution of the same constrained system, in far
    fewer CG iterations; FP64 and the mixed form."""
    import re

    L = P.lib()
    k = 20
    n, x, t, c = P.build_sc_model(P.sc_box_for(k))
    L.calcStiffness3DFiniteDifference(6)
    dbc = np.ones(3 * n, np.int32)
    dbc.reshape(n, 3)[t == 2, :] = 0
    dbc.reshape(n, 3)[t == 3, :] = 0
    dbc.reshape(n, 3)[t == 1, 2] = 0
    P.write1d("dispBC_index", dbc)
    r = np.random.default_rng(12).standard_normal(3 * n) * dbc
    sols, iters = {}, {}
    for pc in (0, 1):
        s = P.solver(precision=prec, rtol2=1e-26, atol2=0.0, maxit=50 * 3 * n, precond=pc)
        L.lpm_ext_set_solver(C.byref(s))
        P.write1d("residual", r)
        L.setDispBC_stiffnessUpdate3D()
        capfd.readouterr()
        L.solverCG()
        out, err = capfd.readouterr()
        m = re.search(r"solved after (\d+) iterations", out)
        assert m, out + err
        iters[pc] = int(m.group(1))
        sols[pc] = P.read1d("disp", 3 * n)
    L.lpm_ext_set_solver(None)
    print("iterations block-Jacobi", iters[0], "AMG", iters[1])
    assert_rel("AMG against block-Jacobi solution", sols[1], sols[0], tol_for(prec))
    assert iters[1] * 4 < iters[0]


def test_amg_hierarchy_follows_constraint_set():
    """The loop keeps the AMG hierarchy across load steps whose constraint set is the
    same (the device fingerprint of the constrained DOFs matches) and rebuilds it when
    the set changes (the top face also held in x)."""
    L = P.lib()
    k = 12
    n, x, t, c = P.build_sc_model(P.sc_box_for(k))
    steps = P.sc_uniaxial_steps(k - 1.0, np.array([2e-4, 4e-4, 6e-4]))
    s = P.solver(precision=P.LPMX_FP64, rtol2=1e-8, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-10,
                 newton_maxit=600, accel_depth=8, precond=2)

    def setup_time():
        lev, ts, dev = C.c_int(), C.c_double(), C.c_int64()
        P.check(L.lpmx_amg_info(C.byref(lev), C.byref(ts), C.byref(dev)), "amg info")
        assert lev.value >= 2
        return ts.value

    rc, _ = P.run_steps(steps[:1], s)
    assert rc == 0, P.lib().lpmx_last_error()
    t1 = setup_time()
    rc, _ = P.run_steps(steps[1:2], s)
    assert rc == 0, P.lib().lpmx_last_error()
    assert setup_time() == t1, "same constraint set: the hierarchy is kept"
    d, f = steps[2]
    rc, _ = P.run_steps([(d + [(1, "x", 0.0)], f)], s)
    assert rc == 0, P.lib().lpmx_last_error()
    assert setup_time() != t1, "another constraint set: the hierarchy is rebuilt"


@pytest.mark.parametrize("prec", [P.LPMX_FP64, P.LPMX_MIXED])
def test_load_steps_amg_match_oracle(oracle, prec):
    """The fused loop with the AMG preconditioner (12^3: two levels) reaches the
    oracle's converged states over an elastic and three plastic steps."""
    k = 12
    _, n, x, t, m = sc_pair(oracle, k)
    steps = P.sc_uniaxial_steps(k - 1.0, np.array([0.5e-3, 1e-3, 1.5e-3, 2e-3]))
    s = P.solver(precision=prec, rtol2=1e-24 if prec == P.LPMX_FP64 else 1e-16, atol2=0.0, maxit=100 * 3 * n,
                 newton_tol=1e-11, newton_maxit=600, accel_depth=8, precond=1)
    rc, out = P.run_steps(steps, s, out_axis="z", out_type=1)
    assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters, o.cg_iters) for o in out])
    print("AMG: Newton", [o.newton_iters for o in out], "CG", [o.cg_iters for o in out])
    so = oracle.solver(rtol2=1e-24, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-11, newton_maxit=600, tangent=1,
                       fixed_diag=1, accel=8, nthreads=8)
    rco, outo = oracle.run_steps(m, steps, so, out_axis="z", out_type=1)
    assert rco == 0
    u, uo = get("XYZ", (n, 3)) - x, m.x - x
    assert_rel("displacement", u, uo, tol_for(prec))
    np.testing.assert_allclose([o.reaction for o in out], [o.reaction for o in outo], rtol=tol_for(prec))


# ----------------------------------------------- one V-cycle per Newton iteration --
@pytest.mark.parametrize("prec", [P.LPMX_FP64, P.LPMX_MIXED])
@pytest.mark.parametrize("case", ["j2", "hex", "brittle", "cp"])
def test_vcycle_newton_matches_oracle(oracle, case, prec):
    """Preconditioner 2 of the fused loop: one AMG V-cycle per Newton iteration in
    place of the linear solve, the map u + B R(u) mixed by Anderson(32).  It reaches
    the oracle's equilibria: J2 tension (12^3, one elastic and three plastic steps),
    2D elastic hexagonal shear (900 particles), the brittle hexagonal shear with
    Weibull bonds (broken-bond sets bit-exact) and FCC crystal plasticity through first
    slip (active-slip sets bit-exact).  All but the brittle case (196 particles) have
    hierarchies of at least two levels."""
    s = None
    if case == "j2":
        k = 12
        _, n, x, t, m = sc_pair(oracle, k)
        steps = P.sc_uniaxial_steps(k - 1.0, np.array([0.5e-3, 1e-3, 1.5e-3, 2e-3]))
        axis, dim, accel = "z", 3, 8
    elif case == "hex":
        box = [0.0, 30.0, 0.0, 26.0, 0.0, 1.0]  # 900 particles: a two-level 2D hierarchy
        n, x, t, c = P.build_hex_model(box, critical=1.0, nbreak=20)  # no bond reaches 1.0
        m = oracle_model(oracle, x, 1, t, [c] * 3, plmode=6, sigmay=1e30, critical_bstrain=1.0, nbreak=20)
        height = float(x[:, 1].max() - x[:, 1].min())
        steps = P.hex_shear_steps(height, np.linspace(4e-4, 1.2e-3, 3))
        axis, dim, accel = "x", 2, 0
    elif case == "brittle":
        box = [0.0, 14.0, 0.0, 12.0, 0.0, 1.0]  # 196 particles (the hierarchy is its dense level)
        n, x, t, c = P.build_hex_model(box, critical=1e-3, nbreak=20)
        nb = P.gint("nneighbors").value
        crit = P.weibull_bond_critical(get("NEIGHBORS", (n, nb)), 1, shape=10.0, scale=1e-3)
        P.build_hex_model(box, critical=1e-3, nbreak=20, crit=crit)
        m = oracle_model(oracle, x, 1, t, [c] * 3, plmode=6, sigmay=1e30, critical_bstrain=1e-3, nbreak=20,
                         crit=crit)
        height = float(x[:, 1].max() - x[:, 1].min())
        steps = P.hex_shear_steps(height, np.linspace(4e-4, 4e-3, 10))
        axis, dim, accel = "x", 2, 0
    else:
        n, x, t, c, m = cp_pair(oracle, 6.5)  # 500 particles: a two-level hierarchy
        height = float(x[:, 2].max() - x[:, 2].min())
        steps = P.sc_uniaxial_steps(height, np.array([0.5, 1.0, 1.3]) * (CP["tau0"] / (c[2] * 0.8)))
        axis, dim, accel = "z", 3, 5
    s = P.solver(precision=prec, rtol2=1e-8, atol2=0.0, maxit=100 * dim * n, newton_tol=1e-11, newton_maxit=4000,
                 accel_depth=32, precond=2)
    rc, out = P.run_steps(steps, s, out_axis=axis, out_type=1)
    assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters) for o in out])
    print(case, "V-cycle Newton iterations", [o.newton_iters for o in out])
    so = oracle.solver(rtol2=1e-24, atol2=0.0, maxit=100 * dim * n, newton_tol=1e-11, newton_maxit=800, tangent=1,
                       fixed_diag=1, accel=accel)
    rco, outo = oracle.run_steps(m, steps, so, out_axis=axis, out_type=1)
    assert rco == 0
    if case == "brittle":
        nb = P.gint("nneighbors").value
        assert [o.broken for o in out] == [o.broken for o in outo]
        assert sum(o.broken for o in outo) > 0
        np.testing.assert_array_equal(get("DAMAGE_BROKEN", (n, nb)), m.dmg_broken)
    if case == "cp":
        assert m.s.cp_fail == 0
        jact = get("CP_JACT", (n, 24))
        assert jact.any()
        np.testing.assert_array_equal(jact, m.cp_Jact)
    r, ro = np.array([o.reaction for o in out]), np.array([o.reaction for o in outo])
    assert_rel("reaction", r, ro, tol_for(prec))
    u, uo = get("XYZ", (n, 3)) - x, m.x - x
    assert_rel("displacement", u, uo, TOL_CP if case == "cp" else tol_for(prec))
