"""Pins of the oracle's load-step functions by what the mathematics fixes.

Each test checks one oracle function against a closed form, a physical invariant
or the defining property of the reference routine it restates, written here
independently of the oracle's code, so that a dropped term, a wrong factor, index
or slot in the oracle fails a CPU test:

- the J2 mixed-linear return map (reference src/constitutive.c:466-686) along a
  loading, unloading, reverse-yield and non-proportional path against the textbook
  radial return with mixed hardening; below yield against the elastic law bit for
  bit; under non-uniform plastic flow against action-reaction (F_ij = F_ji);
- updateCrack (src/constitutive.c:1399-1434), updateRR (src/stiffness.c:519-534),
  setDispBC / setForceBC (src/boundary.c:12-70), switchStateV
  (src/constitutive.c:10-85), writeReaction / writeDisp (src/data_handler.c:278-335)
  against their defining properties;
- the load-step loop (src/lpmc_project.c:382-546): a converged step is an
  equilibrium of the pinned law, its new state is the law's update at the converged
  positions, the prescribed displacements are met and the reactions balance.
"""
import numpy as np
import pytest

from test_oracle import E0, NU0, deep_interior, hexm, sc, voigt_C


# ---------------------------------------------------------------- J2 return map --
def to_voigt_strain(e):
    return np.array([e[0, 0], e[1, 1], e[2, 2], 2 * e[1, 2], 2 * e[0, 2], 2 * e[0, 1]])


def tensor_from_voigt_stress(s):
    return np.array([[s[0], s[5], s[4]], [s[5], s[1], s[3]], [s[4], s[3], s[2]]])


def voigt_from_tensor(t):
    """Tensor components in the order xx, yy, zz, yz, xz, xy (no factor 2)."""
    return np.array([t[0, 0], t[1, 1], t[2, 2], t[1, 2], t[0, 2], t[0, 1]])


def radial_return(c, sy, H, xi, eps, ep, beta, alpha):
    """Textbook small-strain radial return with linear mixed hardening (isotropic
    part (1 - xi) H, kinematic part 2/3 xi H): trial stress C:(eps - ep), relative
    deviator s - beta, yield function sqrt(3/2 s:s) - (sy + (1 - xi) H alpha),
    multiplier f / (3 mu + H), flow direction 3/2 s / seq.  Tensors are 3x3."""
    mu = c[2]
    sig = voigt_C(*c) @ to_voigt_strain(eps - ep)
    st = tensor_from_voigt_stress(sig)
    s = st - np.trace(st) / 3.0 * np.eye(3) - beta
    seq = np.sqrt(1.5 * np.sum(s * s))
    f = seq - (sy + (1.0 - xi) * H * alpha)
    dlam = f / (3.0 * mu + H) if f > 0 else 0.0
    dep = dlam * 1.5 * s / seq
    ep2, beta2, alpha2 = ep + dep, beta + 2.0 / 3.0 * xi * H * dep, alpha + dlam
    sig2 = voigt_C(*c) @ to_voigt_strain(eps - ep2)
    return dict(sig=sig2, ep=ep2, beta=beta2, alpha=alpha2, dlam=dlam, f=f)


def test_j2_return_map_follows_the_textbook_radial_return(oracle):
    """A uniform strain path at deep-interior particles: first yield under a strain
    with all three shear components, elastic unloading, reverse yield against the
    back stress, then a non-proportional plastic increment.  The yield stress is
    small (yield at strains near 1e-6), so the geometric terms the lattice carries
    beyond the small-strain closed form are ~1e-6 relative, and H is comparable to
    the shear modulus, so the hardening terms weigh in the multiplier.  Checked at
    every stage: the stress from the final bond forces (the second stretch pass and
    the plastic stretch of each bond, engineering shear included), the multiplier,
    the hardening variable, the back stress, and the plastic bond stretch
    L0 n.eps_p.n stored for the next step."""
    sy, H, xi = 0.2, 3.0e4, 0.4
    m, c = sc(oracle, 7.1, sigmay=sy, J2_H=H, J2_xi=xi)
    deep = deep_interior(m)
    assert len(deep) >= 27
    D1 = np.array([[1.0, 0.6, -0.3], [0.6, -0.2, 0.5], [-0.3, 0.5, 0.4]])
    D2 = np.array([[-0.3, 0.1, 0.7], [0.1, 0.8, -0.2], [0.7, -0.2, -0.1]])
    zero = np.zeros((3, 3))
    a = 1.6 * sy / radial_return(c, 0.0, 0.0, 0.0, D1, zero, zero, 0.0)["f"]  # seq(D1) from a zero yield stress
    e1 = a * D1
    path = [("first yield", e1, True), ("elastic unloading", 0.5 * e1, False),
            ("reverse yield", -0.8 * e1, True), ("non-proportional", -0.8 * e1 + 0.9 * a * D2, True)]
    ep, beta, alpha = zero.copy(), zero.copy(), 0.0
    live = m.nbr >= 0
    n0 = m.cs0  # bond directions at the reference configuration
    for label, eps, plastic in path:
        want = radial_return(c, sy, H, xi, eps, ep, beta, alpha)
        assert (want["f"] > 0) == plastic, (label, want["f"])
        m.x[:] = m.x0 + m.x0 @ eps.T
        m.pl_flag[:] = 0
        oracle.lib().orc_bond_force_general(m.p, 0)  # ends with slot 0 = slot 2
        oracle.lib().orc_switch_state(m.p, 1)  # the converged state of the step
        scale = np.abs(want["sig"]).max()
        err = np.abs(m.stress[deep] - want["sig"]).max() / scale
        assert err <= 2e-5, (label, err)
        assert np.all(m.pl_flag[deep] == int(plastic)), label
        np.testing.assert_allclose(m.dlambda[deep], want["dlam"], rtol=2e-5, atol=1e-12 * a, err_msg=label)
        np.testing.assert_allclose(m.alpha[deep, 0], want["alpha"], rtol=2e-5, err_msg=label)
        bscale = np.abs(want["beta"]).max()
        assert np.abs(m.beta[deep, :, 0] - voigt_from_tensor(want["beta"])).max() <= 2e-5 * bscale, label
        if plastic:
            # the final stress lies on the hardened, shifted yield surface
            for i in deep:
                st = tensor_from_voigt_stress(m.stress[i])
                s = st - np.trace(st) / 3.0 * np.eye(3) - tensor_from_voigt_stress(m.beta[i, :, 0])
                seq = np.sqrt(1.5 * np.sum(s * s))
                surf = sy + (1.0 - xi) * H * m.alpha[i, 0]
                assert abs(seq - surf) <= 1e-4 * surf, (label, seq, surf)
        # plastic bond stretch: L0 n.eps_p.n on every bond of a deep particle
        nn = n0[deep]
        q = np.einsum("pjk,kl,pjl->pj", nn, want["ep"], nn)
        lp = m.L0[deep] * q
        got = m.dLp[deep, :, 0]
        assert np.abs(got - lp).max() <= 2e-5 * np.abs(lp).max(), label
        assert np.all(live[deep])
        ep, beta, alpha = want["ep"], want["beta"], want["alpha"]


def test_j2_below_yield_is_the_elastic_law_bit_for_bit(oracle):
    """Below yield the J2 law's forces, stretches, internal forces and stress are
    the elastic law's (reference src/constitutive.c:228-283) with the same plastic
    bond stretch, bit for bit, for any deformation and any plastic stretch that is
    the same on both directions of a bond; the plastic state is carried over
    unchanged."""
    rng = np.random.default_rng(5)
    models = []
    for plmode in (0, 6):
        m, c = sc(oracle, 6.2, sigmay=1e12, J2_H=2000.0, J2_xi=0.3, plmode=plmode)
        models.append(m)
    m0, m6 = models
    n, nb = m0.n, m0.nbmax
    live = m0.nbr >= 0
    lp = np.zeros((n, nb))
    for i in range(n):
        for j in range(nb):
            q = m0.nbr[i, j]
            if q > i:
                v = 1e-4 * rng.standard_normal()
                lp[i, j] = v
                lp[q, list(m0.nbr[q]).index(i)] = v
    beta0 = 50.0 * rng.standard_normal((n, 6))
    alpha0 = 1e-3 * rng.uniform(size=n)
    x = m0.x0 + m0.x0 @ (2e-3 * rng.standard_normal((3, 3))).T + 1e-3 * rng.standard_normal(m0.x0.shape)
    for m in models:
        m.dLp[:, :, 0] = lp
        m.beta[:, :, 0] = beta0
        m.alpha[:, 0] = alpha0
        m.x[:] = x
    oracle.lib().orc_bond_force_general(m0.p, 0)
    oracle.lib().orc_bond_force_general(m6.p, 6)
    assert np.abs(m6.F[live]).max() > 0
    assert m0.pl_flag.sum() == 0
    for name in ("F", "dL", "Pin", "stress", "cs", "dLtot", "TdLtot"):
        np.testing.assert_array_equal(getattr(m0, name), getattr(m6, name), err_msg=name)
    np.testing.assert_array_equal(m0.dLp[:, :, 0], lp)
    np.testing.assert_array_equal(m0.beta[:, :, 0], beta0)
    np.testing.assert_array_equal(m0.alpha[:, 0], alpha0)
    assert not m0.dlambda.any()


def reverse_slots(m):
    rev = np.full(m.nbr.shape, -1)
    for i in range(m.n):
        for j in range(m.nb_initial[i]):
            rev[i, j] = list(m.nbr[m.nbr[i, j]]).index(i)
    return rev


def test_j2_forces_obey_action_reaction_under_nonuniform_flow(oracle):
    """Under a non-affine deformation with a random plastic state, where neighbours
    return to the yield surface by different amounts, each bond force is formed from
    the mean elastic stretch of its two directions and both endpoints' shell sums,
    so F_ij = F_ji bit for bit and the internal forces sum to zero.  Forming it from
    one direction's stretch alone breaks both."""
    rng = np.random.default_rng(8)
    m, c = sc(oracle, 6.2, sigmay=rng.uniform(100.0, 600.0, 252), J2_H=2000.0, J2_xi=0.3)
    live = m.nbr >= 0
    m.dLp[:, :, 0] = 2e-5 * rng.standard_normal(live.shape) * live
    m.beta[:, :, 0] = 10.0 * rng.standard_normal((m.n, 6))
    m.alpha[:, 0] = 1e-3 * rng.uniform(size=m.n)
    eps = 2e-3 * np.array([[1.0, 0.3, 0.0], [0.3, -0.4, 0.2], [0.0, 0.2, -0.5]])
    m.x[:] = m.x0 + m.x0 @ eps.T + 2e-4 * rng.standard_normal(m.x.shape)
    oracle.lib().orc_bond_force_general(m.p, 0)
    assert 0 < m.pl_flag.sum() < m.n, "some particles yield, some do not"
    dlp2 = m.dLp[:, :, 0]
    rev = reverse_slots(m)
    ii, jj = np.nonzero(live)
    q = m.nbr[ii, jj]
    assert np.abs(dlp2[ii, jj] - dlp2[q, rev[ii, jj]]).max() > 0, "the plastic stretches differ across bonds"
    np.testing.assert_array_equal(m.F[ii, jj], m.F[q, rev[ii, jj]])
    tot = m.Pin.reshape(-1, 3).sum(0)
    assert np.abs(tot).max() <= 1e-12 * np.abs(m.F).max() * m.n


# ----------------------------------------------------------------- updateCrack --
def test_update_crack_properties(oracle):
    """updateCrack: the bond count is the number of unbroken bonds, the visual damage
    1 - sum(broken)/nb_initial, every bond force is multiplied by its weight once more
    (F_out = w F_in, bit for bit; weights of 1 leave F unchanged), Pin is the sum of
    n F over the bonds, and a particle with no unbroken bond is fixed (fix_index 0 on
    all of its DOFs) with no internal force; no other particle's fix_index changes."""
    m, _ = sc(oracle, 5.2)
    rng = np.random.default_rng(3)
    live = m.nbr >= 0
    n = m.n
    m.cs[:] = rng.standard_normal(m.cs.shape)
    F0 = rng.standard_normal(live.shape) * live
    # weights of one: forces unchanged, counts full
    m.F[:] = F0
    oracle.lib().orc_update_crack(m.p)
    np.testing.assert_array_equal(m.F, F0)
    np.testing.assert_array_equal(m.nb, m.nb_initial)
    assert not m.damage_visual.any()
    assert np.all(m.fixidx == 1)
    np.testing.assert_allclose(m.Pin.reshape(n, 3), np.einsum("ijk,ij->ik", m.cs, F0), rtol=0,
                               atol=1e-14 * np.abs(F0).max() * m.nbmax)
    # partial damage, some broken bonds, one particle with every bond broken
    w = rng.uniform(0.2, 1.0, live.shape) * live
    br = (rng.uniform(size=live.shape) > 0.3).astype(np.float64) * live
    p = 17
    br[p, :] = 0.0
    w[p, :] = 0.0
    m.dmg_w[:] = w
    m.dmg_broken[:] = br
    m.F[:] = F0
    oracle.lib().orc_update_crack(m.p)
    np.testing.assert_array_equal(m.F, F0 * w)
    np.testing.assert_array_equal(m.nb, (br[:, :] > 0).sum(1))
    np.testing.assert_allclose(m.damage_visual, 1.0 - br.sum(1) / m.nb_initial, rtol=0, atol=1e-15)
    fix = m.fixidx.reshape(n, 3)
    assert np.all(fix[p] == 0)
    assert np.all(np.delete(fix, p, axis=0) == 1)
    assert not m.Pin.reshape(n, 3)[p].any()
    np.testing.assert_allclose(m.Pin.reshape(n, 3), np.einsum("ijk,ij->ik", m.cs, F0 * w), rtol=0,
                               atol=1e-14 * np.abs(F0).max() * m.nbmax)


# -------------------------------------------------------------------- updateRR --
@pytest.mark.parametrize("dim", [3, 2])
def test_update_rr_residual_and_reaction(oracle, dim):
    """updateRR: residual = Pex - Pin on free DOFs and 0 on prescribed ones (Pex and
    the residual dim-strided, Pin always 3-strided); the reaction list is Pin on the
    prescribed DOFs in DOF order."""
    m, _ = sc(oracle, 4.3) if dim == 3 else hexm(oracle, 6.0, 5.0)
    n = m.n
    rng = np.random.default_rng(4)
    pin = rng.standard_normal((n, 3))
    pex = rng.standard_normal((n, dim))
    mask = (rng.uniform(size=(n, dim)) > 0.25).astype(np.int32)
    m.Pin[:] = pin.reshape(-1)
    m.Pex[:] = pex.reshape(-1)
    m.dispBC[:] = mask.reshape(-1)
    oracle.lib().orc_update_rr(m.p)
    np.testing.assert_array_equal(m.residual.reshape(n, dim), mask * (pex - pin[:, :dim]))
    nr = int((mask == 0).sum())
    assert m.s.nreaction == nr
    np.testing.assert_array_equal(m.reaction[:nr], pin[:, :dim][mask == 0])


# ------------------------------------------------------------------ BC setters --
def test_set_disp_bc_moves_the_type_by_the_step(oracle):
    """setDispBC: every particle of a listed type moves by exactly the step along the
    flag's axis and that DOF becomes prescribed; entries for one type accumulate;
    other particles and DOFs are untouched; a flag that names no axis does nothing."""
    m, _ = sc(oracle, 5.2)
    n = m.n
    t = np.zeros(n, np.int32)
    t[::3] = 1
    t[1::7] = 2
    m.type[:] = t
    x0 = m.x.copy()
    bcs = [(1, "z", 0.125), (1, "x", -0.25), (2, "y", 0.5), (2, "y", 0.25), (1, "q", 9.0)]
    steps = [(bcs, [])]
    maxbc, nbd, dbc, nbf, fbc = oracle.bc_tables(steps)
    oracle.lib().orc_set_disp_bc(m.p, int(nbd[0]), dbc)
    want = x0.copy()
    want[t == 1, 2] += 0.125
    want[t == 1, 0] += -0.25
    want[t == 2, 1] += 0.5
    want[t == 2, 1] += 0.25
    np.testing.assert_array_equal(m.x, want)
    mask = np.ones((n, 3), np.int32)
    mask[t == 1, 2] = mask[t == 1, 0] = 0
    mask[t == 2, 1] = 0
    np.testing.assert_array_equal(m.dispBC.reshape(n, 3), mask)


@pytest.mark.parametrize("dim", [3, 2])
def test_set_force_bc_spreads_the_step_over_the_type(oracle, dim):
    """setForceBC: each entry's force step is spread evenly over the particles of its
    type, so Pex summed over the type grows by the step (the z component only in
    3D); other particles' Pex is untouched."""
    m, _ = sc(oracle, 5.2) if dim == 3 else hexm(oracle, 6.0, 5.0)
    n = m.n
    t = np.zeros(n, np.int32)
    t[::5] = 3
    m.type[:] = t
    rng = np.random.default_rng(6)
    pex0 = rng.standard_normal(n * dim)
    m.Pex[:] = pex0
    f = (1.5, -2.25, 7.0)
    maxbc, nbd, dbc, nbf, fbc = oracle.bc_tables([([], [(3, f)])])
    oracle.lib().orc_set_force_bc(m.p, int(nbf[0]), fbc)
    d = (m.Pex - pex0).reshape(n, dim)
    cnt = int((t == 3).sum())
    for c in range(dim):
        np.testing.assert_allclose(d[t == 3, c], f[c] / cnt, rtol=1e-13)
        assert abs(d[t == 3, c].sum() - f[c]) <= 1e-12 * abs(f[c])
    assert not d[t != 3].any()


# ----------------------------------------------------------------- switchStateV --
def test_switch_state_slot_copies(oracle):
    """switchStateV: flag 0 restores the converged slot (slot 0 = slot 1) of the
    plastic stretch, bond damage, back stress, hardening variable, equivalent back
    stress and the local and nonlocal damage; flag 1 stores it (slot 1 = slot 0);
    flag 2 takes the law's update (slot 0 = slot 2) of the plastic stretch, back
    stress, hardening variable and equivalent back stress only."""
    m, _ = sc(oracle, 4.3)
    rng = np.random.default_rng(9)
    names = {"dLp": 3, "dmgD": 2, "beta": 3, "alpha": 3, "beta_eq": 3, "dloc": 2, "dnl": 2}
    law = {"dLp", "beta", "alpha", "beta_eq"}

    def fill():
        for k in names:
            getattr(m, k)[...] = rng.standard_normal(getattr(m, k).shape)
        return {k: getattr(m, k).copy() for k in names}

    for flag, (dst, src) in ((0, (0, 1)), (1, (1, 0)), (2, (0, 2))):
        before = fill()
        oracle.lib().orc_switch_state(m.p, flag)
        for k, ns in names.items():
            a, b = getattr(m, k), before[k]
            for slot in range(ns):
                if slot == dst and (flag != 2 or k in law):
                    np.testing.assert_array_equal(a[..., slot], b[..., src], err_msg=f"{k} flag {flag}")
                else:
                    np.testing.assert_array_equal(a[..., slot], b[..., slot], err_msg=f"{k} flag {flag}")


# ------------------------------------------------------- writeReaction / writeDisp --
def test_mean_disp_and_reaction_balance(oracle):
    """writeDisp: a uniform translation of a type is reported as that translation on
    each axis.  writeReaction: minus the sum of n F over the type's bonds, i.e. minus
    its particles' internal force; the elastic bond forces are equal and opposite on
    the two directions of a bond, so the reactions of all types sum to zero; an axis
    character other than x, y, z adds nothing."""
    m, _ = sc(oracle, 5.2)
    n = m.n
    t = (np.arange(n) % 4).astype(np.int32)
    m.type[:] = t
    rng = np.random.default_rng(10)
    v = np.array([0.3, -0.7, 0.11])
    m.x[:] = m.x0 + 1e-3 * rng.standard_normal(m.x.shape)
    m.x[t == 2] = m.x0[t == 2] + v
    for c, ax in enumerate("xyz"):
        got = oracle.lib().orc_mean_disp(m.p, ax.encode(), 2)
        assert abs(got - v[c]) <= 1e-14 * 8, (ax, got)
        want = np.mean(m.x[t == 1, c] - m.x0[t == 1, c])
        assert abs(oracle.lib().orc_mean_disp(m.p, ax.encode(), 1) - want) <= 1e-13
    m.x[:] = m.x0 + m.x0 @ (1e-3 * rng.standard_normal((3, 3))).T + 1e-4 * rng.standard_normal(m.x.shape)
    oracle.lib().orc_bond_force_general(m.p, 6)
    pin = m.Pin.reshape(n, 3)
    scale = np.abs(m.F).max() * m.nbmax * n
    for c, ax in enumerate("xyz"):
        r = [oracle.lib().orc_reaction(m.p, ax.encode(), tt) for tt in range(4)]
        for tt in range(4):
            assert abs(r[tt] + pin[t == tt, c].sum()) <= 1e-13 * scale
        assert abs(sum(r)) <= 1e-13 * scale
        assert abs(r[1]) > 1e-6 * scale
    assert oracle.lib().orc_reaction(m.p, b"w", 1) == 0.0


# -------------------------------------------------------------- load-step loop --
def test_converged_steps_are_equilibria_of_the_law(oracle):
    """The load-step loop on a J2 cube (tension past yield, unloading, compression
    into reverse yield; damage update off): after each step the prescribed
    displacements are met exactly, the J2 law evaluated at the converged positions
    from the previous step's converged state leaves a residual within the Newton
    tolerance (Pex - Pin on the free DOFs), the state the step stores is that
    evaluation's update bit for bit, and the reactions of the loaded face and of the
    supports balance."""
    from lpm_harness import P

    k = 5
    L = P.sc_box_for(k)
    x = oracle.create_lattice(2, 0.5, [0, L, 0, L, 0, L])
    t = P.sc_types(x)
    c = oracle.elastic_constants(E0, NU0)
    m = oracle.Model(x, 2, 0.5, t, [c] * 5, plmode=0, sigmay=200.0, J2_H=2000.0, J2_xi=0.3)
    m.setup()
    n = m.n
    height = k - 1.0
    strains = np.array([0.6e-3, 1.6e-3, 0.8e-3, -1.2e-3])
    steps = P.sc_uniaxial_steps(height, strains)
    tol = 1e-11
    s = oracle.solver(rtol2=1e-26, atol2=0.0, maxit=100 * 3 * n, newton_tol=tol, newton_maxit=800, tangent=1,
                      fixed_diag=1, accel=5, skip_damage=1)
    top = t == 1
    for st, e in enumerate(strains):
        prev = {kk: getattr(m, kk)[..., 1].copy() for kk in ("dLp", "beta", "alpha", "beta_eq")}
        rc, out = oracle.run_steps(m, [steps[st]], s)
        assert rc == 0, st
        o = out[0]
        # prescribed displacements
        np.testing.assert_allclose(m.x[top, 2] - m.x0[top, 2], e * height, rtol=1e-12)
        for tt, axes in ((2, [2]), (3, [0, 1, 2]), (4, [1, 2])):
            np.testing.assert_array_equal(m.x[t == tt][:, axes], m.x0[t == tt][:, axes])
        conv = {kk: getattr(m, kk)[..., 1].copy() for kk in prev}
        xs = m.x.copy()
        # the law at the converged positions from the previous converged state
        for kk, v in prev.items():
            getattr(m, kk)[..., 0] = v
        oracle.lib().orc_bond_force_general(m.p, 0)
        np.testing.assert_array_equal(m.x, xs)
        for kk, v in conv.items():
            np.testing.assert_array_equal(getattr(m, kk)[..., 0], v, err_msg=kk)
        mask = m.dispBC.reshape(n, 3)
        R = mask * (m.Pex.reshape(n, 3) - m.Pin.reshape(n, 3))
        assert np.linalg.norm(R) <= tol * o.tol_mult, (st, np.linalg.norm(R), o.tol_mult)
        # reaction balance along z: loaded face against the supports
        rz = [oracle.lib().orc_reaction(m.p, b"z", tt) for tt in range(5)]
        assert abs(rz[0]) <= 1e-6 * abs(rz[1]), "interior particles carry no net force"
        assert abs(rz[1] + rz[2] + rz[3] + rz[4]) <= 1e-9 * abs(rz[1])
        assert abs(o.reaction - rz[1]) <= 1e-12 * abs(rz[1])
        assert np.sign(rz[1]) == -np.sign(e), (st, rz[1])
    assert m.alpha[:, 1].max() > 0, "the path is plastic"


# ---------------------------------------------------------------- computeStrain --
def packed_to_reference(x, dim):
    """Least-squares unknowns in the reference's packing (j <= k, row-major) to its
    strain components xx yy zz yz xz xy."""
    s = np.zeros(6)
    if dim == 2:
        s[0], s[5], s[1] = x
    else:
        s[0], s[5], s[4], s[1], s[3], s[2] = x
    return s


def bond_rows(m, i):
    """Unit bond vectors (initial configuration) and weights of particle i's bonds."""
    d = m.dim
    e, w, b = [], [], []
    for j in range(m.nb_initial[i]):
        q = m.nbr[i, j]
        u = (m.x0[q] - m.x0[i])[:d] / m.L0[i, j]
        e.append(u)
        w.append(0.1 if m.nsign[i, j] == 0 else 0.9)
        b.append(j)
    return np.array(e), np.array(w), b


def test_compute_strain_recovers_a_stretch_field_linear_in_the_strain(oracle):
    """computeStrain (reference src/lpm_basic.c:127-249) fits a strain to the bond
    stretches: when every bond's stretch is exactly e.eps.e (e its unit vector) the fit
    returns eps itself, with engineering shears, in the order xx yy zz yz xz xy, at
    every particle whose system is regular (3D SC and 2D hexagonal lattices)."""
    rng = np.random.default_rng(11)
    for m, d in ((sc(oracle, 5.2)[0], 3), (hexm(oracle, 8.0, 7.0)[0], 2)):
        a = rng.standard_normal((d, d)) * 1e-3
        eps = 0.5 * (a + a.T)
        for i in range(m.n):
            e, _, bs = bond_rows(m, i)
            for k, j in enumerate(bs):
                m.dL[i, j] = m.L0[i, j] * (e[k] @ eps @ e[k])
        st = oracle_strain = m.compute_strain()
        if d == 3:
            want = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[1, 2], 2 * eps[0, 2], 2 * eps[0, 1]])
        else:
            want = np.array([eps[0, 0], eps[1, 1], 0.0, 0.0, 0.0, 2 * eps[0, 1]])
        err = np.abs(oracle_strain - want).max() / np.abs(want).max()
        assert err < 1e-10, (d, err)
        assert st.shape == (m.n, 6)


def test_compute_strain_is_the_weighted_least_squares_fit(oracle):
    """For arbitrary stretches the result is the minimiser of
    sum_b w_b (dL_b / L_b - sum_{m<=l} x_ml e_m e_l)^2, w = 0.1 on the first shell and
    0.9 on the second, as numpy's least-squares solver finds it."""
    rng = np.random.default_rng(12)
    for m in (sc(oracle, 5.2)[0], hexm(oracle, 8.0, 7.0)[0]):
        d = m.dim
        live = m.nbr >= 0
        m.dL[live] = 1e-3 * rng.standard_normal(int(live.sum())) * m.L0[live]
        st = m.compute_strain()
        pairs = [(a, b) for a in range(d) for b in range(d) if a <= b]
        for i in range(m.n):
            e, w, bs = bond_rows(m, i)
            F = np.array([[u[a] * u[b] for a, b in pairs] for u in e])
            y = np.array([m.dL[i, j] / m.L0[i, j] for j in bs])
            sw = np.sqrt(w)
            x, *_ = np.linalg.lstsq(sw[:, None] * F, sw * y, rcond=None)
            np.testing.assert_allclose(st[i], packed_to_reference(x, d), rtol=0, atol=1e-9 * np.abs(x).max())


def test_compute_strain_singular_system_stores_its_right_hand_side(oracle):
    """A particle whose bonds all lie on one line has a singular system: dgesv leaves
    the right-hand side unsolved and the reference stores it as the strain (its copy
    of the previous particle's strain is overwritten), except at particle 0, which
    keeps what it had.  A chain of three particles along x, placed first, ahead of a
    cubic block."""
    blk = oracle.create_lattice(2, 0.5, [0, 4.2, 0, 4.2, 0, 4.2])
    chain = np.array([[100.0, 100.0, 100.0], [101.0, 100.0, 100.0], [102.0, 100.0, 100.0]])
    x = np.vstack([chain, blk])
    m = oracle.Model(x, 2, 0.5, np.zeros(x.shape[0], np.int32), [oracle.elastic_constants(E0, NU0)])
    m.setup()
    assert list(m.nb_initial[:3]) == [1, 2, 1]
    rng = np.random.default_rng(13)
    live = m.nbr >= 0
    m.dL[live] = 1e-3 * rng.standard_normal(int(live.sum()))
    init = np.full((m.n, 6), 7.0)
    st = m.compute_strain(init)
    np.testing.assert_array_equal(st[0], init[0])
    for i in (1, 2):
        rhs = sum(0.1 * m.dL[i, j] / m.L0[i, j] * ((m.x0[m.nbr[i, j], 0] - m.x0[i, 0]) / m.L0[i, j]) ** 2
                  for j in range(m.nb_initial[i]))
        np.testing.assert_allclose(st[i], [rhs, 0, 0, 0, 0, 0], rtol=1e-15, atol=0)
    # the block's particles are regular and fitted
    assert np.all(np.abs(st[3:] - 7.0) > 1.0)


# ------------------------------------------------------------------- calcKnTv --
def test_kntv_endpoint_average_over_types_and_modulus_field(oracle):
    """calcKnTv for SC (reference src/stiffness.c:178-203): per type the bond constants
    r (C11 - C12 - C44), r C44 and r (C12 - C44) / 18, and per bond the mean of its two
    endpoints' values (the first-shell constant for a first-shell bond, the second-shell
    one otherwise; Tv likewise), each endpoint's value scaled by its modulus factor
    first (the configurations' per-particle modulus).  Three types at random and a
    factor field in [0.95, 1.05]; every bond's Kn and Tv to round-off, Kn and Tv
    symmetric in (i, j)."""
    x = oracle.create_lattice(2, 0.5, [0, 5.2, 0, 5.2, 0, 5.2])
    n = x.shape[0]
    rng = np.random.default_rng(21)
    types = rng.integers(0, 3, n).astype(np.int32)
    Ce = [oracle.elastic_constants(E, nu) for E, nu in ((200e3, 0.3), (120e3, 0.25), (70e3, 0.33))]
    em = 1.0 + 0.05 * rng.uniform(-1, 1, n)
    m = oracle.Model(x, 2, 0.5, types, Ce, emod=em)
    m.setup()
    r = 0.5
    kt = np.array([[r * (c11 - c12 - c44), r * c44, r * (c12 - c44) / 18.0] for c11, c12, c44 in Ce])
    live = m.nbr >= 0
    i_idx = np.broadcast_to(np.arange(n)[:, None], m.nbr.shape)[live]
    j_idx = m.nbr[live]
    s = m.nsign[live]
    kn_i = em[i_idx] * kt[types[i_idx], s]
    kn_j = em[j_idx] * kt[types[j_idx], s]
    tv_i = em[i_idx] * kt[types[i_idx], 2]
    tv_j = em[j_idx] * kt[types[j_idx], 2]
    np.testing.assert_allclose(m.Kn[live], 0.5 * (kn_i + kn_j), rtol=1e-15, atol=0)
    np.testing.assert_allclose(m.Tv[live], 0.5 * (tv_i + tv_j), rtol=1e-14, atol=0)
    # symmetry: bond (i, j) and its reverse carry the same constants
    for i, j in zip(i_idx[:500], j_idx[:500]):
        k = int(np.where(m.nbr[j] == i)[0][0])
        row = int(np.where(m.nbr[i] == j)[0][0])
        assert m.Kn[i, row] == m.Kn[j, k] and m.Tv[i, row] == m.Tv[j, k]
