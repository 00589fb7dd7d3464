"""Parity at the configured full sizes: the 1,000,000-particle SC lattice of the
load-step workload, the 1,000,000-particle hexagonal brittle lattice and the
256,000-particle FCC crystal (40^3 cells).  One force evaluation each, bit-exact
against the oracle run at the same size (serial C: about 20 s for the SC case, 30 s
for the FCC one), plus a size-independent property for the FCC case."""
import ctypes as C

import numpy as np
import pytest

from lpm_harness import P, get, push

pytestmark = pytest.mark.gpu

CP = dict(tau0=10.0, taus=1e20, h0=100.0, q=1.0, eta=1.0, p=1.0, dtime=1.0, maxloop=20)


def full_model(O, x, lattice, types, Ce, **kw):
    """The oracle's model with the cell-list neighbour search (the brute-force one is
    quadratic in n)."""
    m = O.Model(x, lattice, 0.5, types, Ce, **kw)
    m.setup(cells=True)
    return m


def test_sc_1m_lists_and_j2_force_bit_exact(oracle):
    """SC 100^3 (box 99.195, modulus field of seed 0): neighbour lists and shells,
    then one J2 evaluation at a 0.3% uniaxial strain with noise, where every particle
    yields: bond forces, stretches, internal forces, stress and the whole J2 state
    bit-exact with the oracle."""
    L = P.lib()
    emod = P.uniform_modulus(100 ** 3, 0)
    n, x, t, c = P.build_sc_model(99.195, emod=emod)
    assert n == 1_000_000
    print(f"SC full size: {n} particles", flush=True)
    m = full_model(oracle, x, 2, t, [c] * 5, plmode=0, sigmay=200.0, J2_H=2000.0, emod=emod)
    nb = P.gint("nneighbors").value
    np.testing.assert_array_equal(get("NEIGHBORS", (n, nb)), m.nbr)
    np.testing.assert_array_equal(get("NSIGN", (n, nb)), m.nsign)
    np.testing.assert_array_equal(get("KN", (n, nb)), m.Kn)
    np.testing.assert_array_equal(get("TV", (n, nb)), m.Tv)
    eps = np.diag([-0.3 * 3e-3, -0.3 * 3e-3, 3e-3])
    xd = x + x @ eps.T + 1e-5 * np.random.default_rng(21).standard_normal(x.shape)
    push("XYZ", xd)
    P.check(L.lpmx_bond_force(0), "bond force")
    m.x[:] = xd
    oracle.lib().orc_bond_force_general(m.p, 0)
    sh = (n, nb)
    np.testing.assert_array_equal(get("F", sh), m.F)
    np.testing.assert_array_equal(get("DL", sh), m.dL)
    np.testing.assert_array_equal(get("PIN", (3 * n,)), m.Pin)
    np.testing.assert_array_equal(get("STRESS", (n, 6)), m.stress)
    dl = get("J2_DLAMBDA", (n,))
    assert (dl > 0).all(), "every particle yields"
    np.testing.assert_array_equal(dl, m.dlambda)
    np.testing.assert_array_equal(get("DLP0", sh), m.dLp[:, :, 0])
    np.testing.assert_array_equal(get("J2_BETA0", (n, 6)), m.beta[:, :, 0])
    np.testing.assert_array_equal(get("J2_ALPHA0", (n,)), m.alpha[:, 0])
    np.testing.assert_array_equal(get("PL_FLAG", (n,)), m.pl_flag)


def test_hex_1m_brittle_breaks_bit_exact(oracle):
    """Hexagonal 1,000,000 particles (box 999.6 x 865.77) with Weibull critical
    stretches (shape 10, scale 1e-3, seed 1) under a uniform 9e-4 shear: the elastic
    stretches, the candidate count (296) and the broken-bond set of the 20 largest,
    ties by the reference's order, bit-exact with the oracle."""
    L = P.lib()
    box = [0.0, 999.6, 0.0, 865.77, 0.0, 1.0]
    n, x, t, c = P.build_hex_model(box, critical=1e-3, nbreak=20)
    assert n == 1_000_000
    print(f"hex full size: {n} particles", flush=True)
    nb = P.gint("nneighbors").value
    crit = P.weibull_bond_critical(get("NEIGHBORS", (n, nb)), 1, shape=10.0, scale=1e-3)
    P.build_hex_model(box, critical=1e-3, nbreak=20, crit=crit)
    m = full_model(oracle, x, 1, t, [c] * 3, plmode=6, sigmay=1e30, critical_bstrain=1e-3, nbreak=20, crit=crit)
    np.testing.assert_array_equal(get("NEIGHBORS", (n, nb)), m.nbr)
    xd = x.copy()
    xd[:, 0] += 9e-4 * x[:, 1]
    push("XYZ", xd)
    P.check(L.lpmx_bond_force(6), "elastic force")
    k = C.c_int()
    P.check(L.lpmx_damage(6, C.byref(k)), "brittle")
    m.x[:] = xd
    oracle.lib().orc_bond_force_general(m.p, 6)
    np.testing.assert_array_equal(get("F", (n, nb)), m.F)
    bi = np.zeros(64, np.int32)
    bj = np.zeros(64, np.int32)
    ko = oracle.lib().orc_damage_brittle(m.p, oracle.ptr(bi), oracle.ptr(bj), 64)
    assert k.value == ko and ko > 20, (k.value, ko)
    np.testing.assert_array_equal(get("DAMAGE_BROKEN", (n, nb)), m.dmg_broken)
    assert int((m.dmg_broken == 0).sum()) == 20, "the 20 largest directed entries"
    gi = np.zeros(64, np.int32)
    gj = np.zeros(64, np.int32)
    cnt = C.c_int()
    P.check(L.lpmx_get_broken(P.ptr(gi), P.ptr(gj), 64, C.byref(cnt)), "broken")
    assert cnt.value == 20
    np.testing.assert_array_equal(gi[:20], bi[:20])
    np.testing.assert_array_equal(gj[:20], bj[:20])


def test_fcc_256k_single_slip_return_map(oracle):
    """FCC 40^3 cells (box 55.907, 256,000 particles) under a uniform strain along a
    single-slip direction just past first yield: slip stiffness, active sets,
    resolved shear stresses, slips and forces bit-exact with the oracle at this size;
    no return map fails; every particle with a full list activates exactly the most
    loaded of the 24 systems, with the slip of the full-list particles of a 500-
    particle model under the same strain (a particle's response depends only on its
    own bonds) to 1e-8 relative: at coordinates up to 56 the bond vectors round at
    ~1e-14, which the 2% overstress amplifies to a 1e-9 spread of the slips."""
    L = P.lib()
    n, x, t, c = P.build_fcc_cp_model(55.907, CP)
    assert n == 256_000
    print(f"FCC full size: {n} particles", flush=True)
    sch = np.zeros((64, 6))
    ns = oracle.lib().orc_slip_systems_fcc(oracle.ptr(np.eye(3).reshape(9).copy()), oracle.ptr(sch))
    assert ns == 24
    d = np.array([-1.0, 2.0, 3.0]) / np.sqrt(14.0)
    mu = c[2]
    fac = np.array([2 * mu * (s[0] * d[0] ** 2 + s[1] * d[1] ** 2 + s[2] * d[2] ** 2 + s[3] * d[1] * d[2] +
                              s[4] * d[0] * d[2] + s[5] * d[0] * d[1]) for s in sch[:ns]])
    a = int(np.argmax(fac))
    assert np.sort(fac)[-2] < 0.95 * fac[a]
    ev = 1.02 * CP["tau0"] / fac[a] * np.outer(d, d)
    xd = x + x @ ev.T
    push("XYZ", xd)
    P.check(L.lpmx_bond_force(1), "plmode 1")
    fails = C.c_int()
    P.check(L.lpmx_cp_failures(C.byref(fails)), "failures")
    assert fails.value == 0
    nb = P.gint("nneighbors").value
    m = full_model(oracle, x, 3, t, [c] * 5, plmode=1, sigmay=1e30, cp=CP)
    oracle.lib().orc_compute_cab(m.p)
    np.testing.assert_array_equal(get("CP_CAB", (n, ns * ns)), m.cp_Cab.reshape(n, -1))
    m.x[:] = xd
    oracle.lib().orc_bond_force_general(m.p, 1)
    assert m.s.cp_fail == 0
    np.testing.assert_array_equal(get("CP_JACT", (n, ns)), m.cp_Jact)
    np.testing.assert_array_equal(get("CP_RSS", (n, ns)), m.cp_RSS)
    np.testing.assert_array_equal(get("CP_DA_SINGLE", (n, ns)), m.cp_dA_single)
    np.testing.assert_array_equal(get("F", (n, nb)), m.F)
    np.testing.assert_array_equal(get("STRESS", (n, 6)), m.stress)
    del m
    full = get("NB_INITIAL", (n,)) == nb
    assert full.sum() > n // 2
    jact = get("CP_JACT", (n, ns))
    one = np.zeros(ns, np.int32)
    one[a] = 1
    assert (jact[full] == one).all(), "exactly the most loaded system, at every full-list particle"
    slip = get("CP_DA_SINGLE", (n, ns))[full, a]
    # the same strain on a small oracle model
    xs = oracle.create_lattice(3, 0.5, [0.0, 6.5, 0.0, 6.5, 0.0, 6.5])
    ts = P.sc_types(xs)
    ms = oracle.Model(xs, 3, 0.5, ts, [c] * 5, plmode=1, sigmay=1e30, cp=CP)
    ms.setup()
    oracle.lib().orc_compute_cab(ms.p)
    ms.x[:] = xs + xs @ ev.T
    oracle.lib().orc_bond_force_general(ms.p, 1)
    assert ms.s.cp_fail == 0
    fs = ms.nb_initial == nb
    ref = ms.cp_dA_single[fs, a]
    assert (ms.cp_Jact[fs] == one).all() and ref.min() > 0
    g0 = np.median(ref)
    assert np.abs(ref - g0).max() <= 1e-8 * g0
    assert np.abs(slip - g0).max() <= 1e-8 * g0, (slip.min(), slip.max(), g0)


@pytest.mark.skip(reason="does not converge yet: on the 256,000-particle FCC crystal the first 0.05% [001] step "
                         "(eight tied slip systems activate) brings the residual from 1.5e5 to 0.53 in 100 Newton "
                         "iterations with one AMG V-cycle each, then stalls between 2e-2 and 1e-1 for 300 more "
                         "(0.075 with a PCG solve per correction after 100); a 1e-10 tolerance needs 1.5e-5.  The "
                         "elastic tangent of the reference contracts the slip modes by only ~1 - h0/mu per "
                         "iteration, and the return map's slip Newton stops at TOLITER, which leaves a residual "
                         "floor near 1e-12 of the reaction on small crystals")
def test_fcc_256k_001_tension_converges():
    """FCC 40^3 cells (256,000 particles), crystal plasticity with CRSS 10 MPa and
    hardening 100 MPa, tension along [001] to 3% in 60 steps: the first steps past
    first slip converge with the fused loop (one V-cycle per iteration, Anderson 32)
    in both precision forms, with the eight tied systems active."""
    n, x, t, c = P.build_fcc_cp_model(55.907, CP)
    assert n == 256_000
    height = float(x[:, 2].max() - x[:, 2].min())
    steps = P.sc_uniaxial_steps(height, np.arange(1, 61) * (0.03 / 60))
    for prec in (P.LPMX_MIXED, P.LPMX_FP64):
        P.build_fcc_cp_model(55.907, CP)
        s = P.solver(precision=prec, rtol2=1e-8, atol2=0.0, maxit=300 * n, newton_tol=1e-10, newton_maxit=1000,
                     accel_depth=32, precond=2)
        rc, out = P.run_steps(steps[:3], s, out_axis="z", out_type=1)
        assert rc == 0, [(o.status, o.newton_iters, o.res_norm / o.tol_mult) for o in out]
        assert get("CP_JACT", (n, 24)).sum(axis=1).max() >= 8
