"""The reference's remaining laws on the device against the oracle (GPU).

- the energy-based J2 return map (plmode 3, reference src/constitutive.c:286-463):
  one evaluation from a random slot-0 state under a random deformation, in tension
  and in compression, bit for bit (the reference's serial order and the per-particle
  form are identical, which the oracle's CPU tests check);
- the nonlinear-isotropic J2 law (plmode 5, :689-863) in its one-update-per-particle
  form, against the oracle's restatement of that form;
- the point-wise local, bond-wise nonlocal and bond-wise local ductile damage
  updates (:1529-1755), from random multipliers, triaxialities, damage and broken
  bonds: broken sets, counts and damage bit for bit (the nonlocal Gaussian weights
  call exp: to rounding);
- the fused load-step loop with plmode 5 and its bond-wise local damage, against the
  oracle's loop in the same form.
"""
import ctypes as C

import numpy as np
import pytest

from lpm_harness import TOL_FP64, P, assert_rel, get, oracle_model, push, sc_pair

pytestmark = pytest.mark.gpu


def random_state(m, rng, alpha_scale=1e-3):
    live = m.nbr >= 0
    n = m.n
    lp = np.zeros(live.shape)
    # plastic stretch symmetric over the two directions of a bond
    for i in range(n):
        for a in range(m.nb_initial[i]):
            q = m.nbr[i, a]
            if q > i:
                v = 2e-5 * rng.standard_normal()
                lp[i, a] = v
                lp[q, list(m.nbr[q]).index(i)] = v
    return lp, alpha_scale * rng.uniform(size=n), 20.0 * rng.standard_normal(n), 10.0 * rng.standard_normal((n, 6))


def set_state(m, x, lp, alpha, beq, beta):
    push("XYZ", x)
    push("DLP0", lp)
    push("J2_ALPHA0", alpha)
    push("J2_BETA_EQ0", beq)
    push("J2_BETA0", beta)
    m.x[:] = x
    m.dLp[:, :, 0] = lp
    m.alpha[:, 0] = alpha
    m.beta_eq[:, 0] = beq
    m.beta[:, :, 0] = beta


@pytest.mark.parametrize("t", [1, -1])
def test_energy_j2_law_bit_exact(oracle, t):
    L = P.lib()
    rng = np.random.default_rng(40 + t)
    sy = rng.uniform(100.0, 1500.0, 343)  # per-particle yield stress: some particles yield, some do not
    _, n, x0, types, m = sc_pair(oracle, 7, sigmay=sy, H=2000.0)
    P.check(L.lpmx_set_params(C.byref(P.Params(3, 2000.0, 0.3, 0.5, 0.9, 0.0, 1e-2, 20, 0.0))), "params")
    m.s.plmode, m.s.J2_xi = 3, 0.3
    lp, alpha, beq, beta = random_state(m, rng)
    eps = t * 4e-3 * np.array([[1.0, 0.2, 0.0], [0.2, -0.5, 0.1], [0.0, 0.1, -0.5]])
    x = x0 + x0 @ eps.T + 1e-4 * rng.standard_normal(x0.shape)
    set_state(m, x, lp, alpha, beq, beta)
    P.check(L.lpmx_bond_force_ex(3, t), "plmode 3")
    oracle.lib().orc_bond_force_j2energy(m.p, t, 0)
    pl = get("PL_FLAG", (n,))
    assert 0 < pl.sum() < n, "some particles yield, some do not"
    np.testing.assert_array_equal(pl, m.pl_flag)
    for name, ref in (("F", m.F), ("DL", m.dL), ("PIN", m.Pin.reshape(n, 3)), ("DDLP", m.ddLp),
                      ("DLP0", m.dLp[:, :, 0]), ("J2_ALPHA0", m.alpha[:, 0]), ("J2_BETA_EQ0", m.beta_eq[:, 0]),
                      ("J2_DLAMBDA", m.dlambda), ("STRESS", m.stress)):
        np.testing.assert_array_equal(get(name, ref.shape), ref, err_msg=name)


def test_nonlinear_j2_law_matches_its_oracle(oracle):
    L = P.lib()
    _, n, x0, types, m = sc_pair(oracle, 7, sigmay=200.0, H=2000.0)
    C_kin = 1500.0
    P.check(L.lpmx_set_params(C.byref(P.Params(5, 2000.0, 0.0, 0.5, 0.9, 0.0, 1e-2, 20, C_kin))), "params")
    m.s.plmode, m.s.J2_C = 5, C_kin
    rng = np.random.default_rng(51)
    # hardening variables up to 1.5: the law's yield stress spans 620 .. 2100 MPa
    lp, alpha, beq, beta = random_state(m, rng, alpha_scale=1.5)
    eps = 6e-3 * np.array([[1.0, 0.3, 0.0], [0.3, -0.6, 0.2], [0.0, 0.2, -0.4]])
    x = x0 + x0 @ eps.T + 2e-4 * rng.standard_normal(x0.shape)
    set_state(m, x, lp, alpha, beq, 30.0 * beta)
    # computeBondForceGeneral(5): the law, the stress, then switchStateV(2), which
    # restores slot 0 from slot 2 (the law advanced slot 0 in place and writes no slot 2)
    P.check(L.lpmx_bond_force_ex(5, 1), "plmode 5")
    oracle.lib().orc_bond_force_j2nonlinear(m.p, C_kin, 1)
    dlam = get("J2_DLAMBDA", (n,))
    assert 0 < (dlam > 0).sum() < n, "some particles yield, some do not"
    # the hardening law calls exp: its values agree to rounding, the bisection's
    # decisions (and so every multiplier) exactly
    np.testing.assert_array_equal(dlam, m.dlambda)
    assert np.abs(get("DDLP", (n, 18))).max() > 0, "plastic stretch increments"
    for name, ref in (("F", m.F), ("DL", m.dL), ("PIN", m.Pin.reshape(n, 3)), ("DDLP", m.ddLp),
                      ("STRESS", m.stress), ("DLP0", m.dLp[:, :, 0]), ("J2_ALPHA0", m.alpha[:, 0]),
                      ("J2_BETA0", m.beta[:, :, 0])):
        assert_rel(name, get(name, ref.shape), ref, 1e-13)


@pytest.mark.parametrize("law", ["pwise_local", "bwise_nonlocal", "bwise_local"])
def test_ductile_damage_variants(oracle, law):
    L = P.lib()
    _, n, x0, types, m = sc_pair(oracle, 7)
    rng = np.random.default_rng({"pwise_local": 61, "bwise_nonlocal": 62, "bwise_local": 63}[law])
    thr, A = 0.9, 0.4
    live = m.nbr >= 0
    dlam = rng.uniform(0, 0.3, n)
    tri = rng.uniform(-0.6, 0.6, n)
    dloc = rng.uniform(0, 0.85, n)
    dloc[rng.uniform(size=n) < 0.05] = 1.0
    dnl = rng.uniform(0, 0.85, n)
    dnl[rng.uniform(size=n) < 0.05] = 0.95
    br = np.ones(live.shape)
    for i in range(n):  # a few broken pairs
        for a in range(m.nb_initial[i]):
            q = m.nbr[i, a]
            if q > i and rng.uniform() < 0.03:
                br[i, a] = 0.0
                br[q, list(m.nbr[q]).index(i)] = 0.0
    br *= live
    D = np.where(br == 0, 1.0, rng.uniform(0, 0.5, live.shape)) * live
    nb = (br > 0).sum(1).astype(np.int32)
    for name, v in (("J2_DLAMBDA", dlam), ("J2_TRIAXIALITY", tri), ("DAMAGE_LOCAL0", dloc),
                    ("DAMAGE_NONLOCAL0", dnl), ("DAMAGE_BROKEN", br), ("DAMAGE_D0", D), ("DAMAGE_W", 1 - D),
                    ("NB", nb)):
        push(name, v)
    P.check(L.lpmx_set_params(C.byref(P.Params(5, 2000.0, 0.0, 0.5, thr, A, 1e-2, 20, 0.0))), "params")
    m.dlambda[:], m.triax[:], m.dloc[:, 0], m.dnl[:, 0] = dlam, tri, dloc, dnl
    m.dmg_broken[:], m.dmgD[:, :, 0], m.dmg_w[:], m.nb[:] = br, D, 1 - D, nb
    m.s.damagec_A, m.s.damage_threshold = A, thr
    code = {"pwise_local": 1, "bwise_nonlocal": 2, "bwise_local": 3}[law]
    cnt = C.c_int()
    P.check(L.lpmx_damage_law(code, C.byref(cnt)), law)
    if law == "pwise_local":
        k = oracle.lib().orc_damage_pwise_local(m.p, 0)
    elif law == "bwise_nonlocal":
        k = oracle.lib().orc_damage_bwise_nonlocal(m.p)
    else:
        k = oracle.lib().orc_damage_bwise_local(m.p, 0)
    print(law, "count", cnt.value, "oracle", k)
    assert cnt.value == k and k > 0
    np.testing.assert_array_equal(get("DAMAGE_BROKEN", (n, 18)), m.dmg_broken)
    if law == "bwise_nonlocal":
        np.testing.assert_allclose(get("DAMAGE_NONLOCAL0", (n,)), m.dnl[:, 0], rtol=1e-14, atol=0)
        np.testing.assert_allclose(get("DAMAGE_D0", (n, 18)), m.dmgD[:, :, 0], rtol=1e-14, atol=1e-16)
        np.testing.assert_allclose(get("DAMAGE_W", (n, 18)), m.dmg_w, rtol=1e-13, atol=1e-15)
    else:
        np.testing.assert_array_equal(get("DAMAGE_LOCAL0", (n,)), m.dloc[:, 0])
        np.testing.assert_array_equal(get("DAMAGE_D0", (n, 18)), m.dmgD[:, :, 0])
        np.testing.assert_array_equal(get("DAMAGE_W", (n, 18)), m.dmg_w)
    if law == "bwise_local":
        np.testing.assert_array_equal(get("NB", (n,)), m.nb)


@pytest.mark.parametrize("prec", [P.LPMX_FP64, P.LPMX_MIXED])
def test_plmode5_load_steps_match_oracle(oracle, prec):
    """Tension past the nonlinear law's yield (620 MPa) with its bond-wise local
    damage after every step, through the fused loop and the oracle's loop, both in
    the one-update-per-particle form of plmode 5."""
    L = P.lib()
    k = 6
    _, n, x, t, m = sc_pair(oracle, k, sigmay=620.0, H=2000.0)
    C_kin = 1000.0
    P.set_params(plmode=5, J2_H=2000.0, damage_threshold=0.9, damagec_A=0.0)
    P.check(L.lpmx_set_params(C.byref(P.Params(5, 2000.0, 0.0, 0.5, 0.9, 0.0, 1e-2, 20, C_kin))), "params")
    m.s.plmode, m.s.J2_C, m.s.j2nl_jacobi = 5, C_kin, 1
    steps = P.sc_uniaxial_steps(k - 1.0, np.array([2e-3, 4e-3, 5e-3]))
    s = P.solver(precision=prec, rtol2=1e-24 if prec == P.LPMX_FP64 else 1e-16, atol2=0.0, maxit=100 * 3 * n,
                 newton_tol=1e-12, newton_maxit=800, accel_depth=5)
    rc, out = P.run_steps(steps, s, out_axis="z", out_type=1)
    assert rc == 0, (P.lib().lpmx_last_error(), [(o.status, o.newton_iters) for o in out])
    so = oracle.solver(rtol2=1e-26, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-12, newton_maxit=800, tangent=1,
                       fixed_diag=1, accel=5)
    rco, outo = oracle.run_steps(m, steps, so, out_axis="z", out_type=1)
    assert rco == 0
    assert get("J2_DLAMBDA", (n,)).max() > 0, "the path yields"
    tol = TOL_FP64 if prec == P.LPMX_FP64 else 1e-8
    assert_rel("reaction", [o.reaction for o in out], [o.reaction for o in outo], tol)
    assert_rel("displacement", get("XYZ", (n, 3)) - x, m.x - x, tol)
    np.testing.assert_array_equal(get("DAMAGE_BROKEN", (n, 18)), m.dmg_broken)
