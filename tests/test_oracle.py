"""Pins of the CPU oracle against closed forms, invariants and brute force.

The reference ships no tests and no golden vectors, and it cannot be built here
(its Intel MKL headers are absent), so each oracle function is pinned by what the
mathematics fixes, chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails a check.
"""
import itertools

import ctypes as C

import numpy as np
import pytest

E0, NU0 = 200e3, 0.3


def sc(O, L, ntype=1, **kw):
    x = O.create_lattice(2, 0.5, [0, L, 0, L, 0, L])
    c = O.elastic_constants(E0, NU0)
    m = O.Model(x, 2, 0.5, np.zeros(x.shape[0], np.int32), [c] * ntype, **kw)
    m.setup()
    return m, c


def hexm(O, W, H, **kw):
    x = O.create_lattice(1, 0.5, [0, W, 0, H, 0, 1.0])
    c = O.elastic_constants(70e3, 0.25)
    m = O.Model(x, 1, 0.5, np.zeros(x.shape[0], np.int32), [c], **kw)
    m.setup()
    return m, c


def fcc(O, L, **kw):
    x = O.create_lattice(3, 0.5, [0, L, 0, L, 0, L])
    c = O.elastic_constants(70e3, 0.3)
    m = O.Model(x, 3, 0.5, np.zeros(x.shape[0], np.int32), [c], **kw)
    m.setup()
    return m, c


# ---------------------------------------------------------------- geometry -------
@pytest.mark.parametrize("L,n", [(19.035, 8000), (59.3, 216000), (99.195, 1000000)])
def test_sc_particle_counts(oracle, L, n):
    x = oracle.create_lattice(2, 0.5, [0, L, 0, L, 0, L])
    assert x.shape[0] == n
    k = round(n ** (1 / 3))
    # x fastest, then y, then z; unit spacing; x starts at box[0]
    g = x.reshape(k, k, k, 3)
    assert np.all(np.diff(g[..., 0], axis=2) == 1.0)
    assert np.all(np.diff(g[..., 1], axis=1) == 1.0)
    assert np.all(np.diff(g[..., 2], axis=0) == 1.0)
    assert g[0, 0, 0, 0] == 0.0


def test_hex_and_fcc_counts(oracle):
    assert oracle.create_lattice(1, 0.5, [0, 999.6, 0, 865.77, 0, 1]).shape[0] == 1000000
    assert oracle.create_lattice(3, 0.5, [0, 55.907, 0, 55.907, 0, 55.907]).shape[0] == 256000


# -------------------------------------------------------------- neighbours -------
def shells(lattice):
    if lattice == 2:
        o1 = [v for v in itertools.product((-1, 0, 1), repeat=3) if sum(map(abs, v)) == 1]
        o2 = [v for v in itertools.product((-1, 0, 1), repeat=3) if sum(map(abs, v)) == 2]
        return np.array(o1, float), np.array(o2, float)
    if lattice == 3:
        h = 1.0 / np.sqrt(2.0)  # a/2 with a = 4r/sqrt(2), r = 0.5
        o1 = [np.array(v) * h for v in itertools.product((-1, 0, 1), repeat=3) if sum(map(abs, v)) == 2]
        o2 = [np.array(v) * 2 * h for v in itertools.product((-1, 0, 1), repeat=3) if sum(map(abs, v)) == 1]
        return np.array(o1), np.array(o2)
    ang = np.arange(6) * np.pi / 3
    o1 = np.stack([np.cos(ang), np.sin(ang), 0 * ang], 1)
    o2 = np.stack([np.sqrt(3) * np.cos(ang + np.pi / 6), np.sqrt(3) * np.sin(ang + np.pi / 6), 0 * ang], 1)
    return o1, o2


@pytest.mark.parametrize("kind", ["sc", "fcc", "hex"])
def test_neighbor_shells_closed_form(oracle, kind):
    m = {"sc": lambda: sc(oracle, 7.1)[0], "fcc": lambda: fcc(oracle, 4.0)[0],
         "hex": lambda: hexm(oracle, 9.0, 8.0)[0]}[kind]()
    lat = {"sc": 2, "fcc": 3, "hex": 1}[kind]
    o1, o2 = shells(lat)
    full = np.where(m.nb_initial == m.nbmax)[0]
    assert len(full) > 0
    for i in full[:: max(1, len(full) // 25)]:
        nb = m.nb_initial[i]
        js = m.nbr[i, :nb]
        assert np.all(np.diff(js) > 0), "lists are in increasing particle index"
        d = m.x0[js] - m.x0[i]
        for o, s in ((o1, 0), (o2, 1)):
            sel = m.nsign[i, :nb] == s
            got = sorted(map(tuple, np.round(d[sel], 6)))
            assert got == sorted(map(tuple, np.round(o, 6)))
        np.testing.assert_array_equal(m.cs0[i, :nb], (m.x0[i] - m.x0[js]) / m.L0[i, :nb, None])
    # every bond is mutual and has a mutual reverse slot
    for i in range(m.n):
        for j in m.nbr[i, :m.nb_initial[i]]:
            assert i in m.nbr[j, :m.nb_initial[j]]


@pytest.mark.parametrize("kind", ["sc", "fcc", "hex"])
def test_cell_list_equals_brute_force(oracle, kind):
    m = {"sc": lambda: sc(oracle, 6.3)[0], "fcc": lambda: fcc(oracle, 3.5)[0],
         "hex": lambda: hexm(oracle, 7.0, 6.0)[0]}[kind]()
    ref = {k: m.A[k].copy() for k in ("nbr", "nsign", "L0", "cs0", "nb_initial", "nbr1", "nbr2")}
    m.neighbors(cells=True)
    for k, v in ref.items():
        np.testing.assert_array_equal(m.A[k], v, err_msg=k)


@pytest.mark.parametrize("kind,interior", [("sc", 61), ("fcc", 61), ("hex", 31)])
def test_conn_sets_and_pointers(oracle, kind, interior):
    m = {"sc": lambda: sc(oracle, 7.1)[0], "fcc": lambda: fcc(oracle, 6.5)[0],
         "hex": lambda: hexm(oracle, 9.0, 8.0)[0]}[kind]()
    assert m.nb_conn.max() == interior
    d2 = m.dim * m.dim - (1 if m.dim == 2 else 3)
    for i in range(m.n):
        # definition: i's bond partners and the same-shell partners of each partner
        want = set()
        for a in range(m.nb_initial[i]):
            j, s = m.nbr[i, a], m.nsign[i, a]
            want.add(j)
            want |= {q for b, q in enumerate(m.nbr[j, :m.nb_initial[j]]) if m.nsign[j, b] == s}
        got = m.conn[i, :m.nb_conn[i]]
        assert list(got) == sorted(want)
        k = int(np.sum(got >= i))
        assert m.kptr[i, 0] == k
        assert m.kptr[i + 1, 1] == m.kptr[i, 1] + m.dim * m.dim * k - (1 if m.dim == 2 else 3)
    assert m.s.nK == m.kptr[m.n, 1]
    assert d2 > 0


def test_kntv_closed_form(oracle):
    m, (c11, c12, c44) = sc(oracle, 5.2)
    r = 0.5
    s0 = m.nsign == 0
    s1 = m.nsign == 1
    assert np.allclose(m.Kn[s0], r * (c11 - c12 - c44), rtol=1e-15)
    assert np.allclose(m.Kn[s1], r * c44, rtol=1e-15)
    assert np.allclose(m.Tv[s0 | s1], r * (c12 - c44) / 18.0, rtol=1e-14)


# --------------------------------------------------------- constitutive laws -----
def deep_interior(m):
    full = m.nb_initial == m.nbmax
    return np.array([i for i in np.where(full)[0] if np.all(full[m.nbr[i, :m.nbmax]])])


def voigt_C(c11, c12, c44):
    C = np.zeros((6, 6))
    C[:3, :3] = c12
    np.fill_diagonal(C[:3, :3], c11)
    C[3, 3] = C[4, 4] = C[5, 5] = c44
    return C


@pytest.mark.parametrize("kind", ["sc", "fcc"])
def test_reference_config_and_affine_patch(oracle, kind):
    m, c = sc(oracle, 7.1, sigmay=1e9) if kind == "sc" else fcc(oracle, 6.5, sigmay=1e9)
    oracle.lib().orc_bond_force_general(m.p, 6)
    assert np.abs(m.Pin).max() == 0.0, "no force at the reference configuration"
    eps = 1e-7 * np.array([[1.0, 0.2, -0.3], [0.2, -0.5, 0.1], [-0.3, 0.1, 0.7]])
    m.x[:] = m.x0 + m.x0 @ eps.T
    oracle.lib().orc_bond_force_general(m.p, 6)
    deep = deep_interior(m)
    assert len(deep) > 0
    ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[1, 2], 2 * eps[0, 2], 2 * eps[0, 1]])
    sig = voigt_C(*c) @ ev
    np.testing.assert_allclose(m.stress[deep], np.broadcast_to(sig, (len(deep), 6)), rtol=0,
                               atol=2e-6 * np.abs(sig).max())
    # interior equilibrium, up to the cancellation in dis - L0 at strain 1e-7
    assert np.abs(m.Pin.reshape(-1, 3)[deep]).max() < 1e-6 * np.abs(m.F).max()


@pytest.mark.parametrize("kind", ["sc", "hex"])
def test_predictor_continues_the_elastic_law(oracle, kind):
    """The incremental predictor (plmode 4) adds the elastic law's response to the
    stretch increment dis(x) - dis(x_temp) to F_temp.  The elastic law is linear in
    the stretch, so with F_temp the elastic forces at x_temp the predicted forces at x
    are the elastic forces at x, to round-off, for any two deformed states."""
    m, _ = sc(oracle, 7.1, sigmay=1e9) if kind == "sc" else hexm(oracle, 12.0, 10.0, sigmay=1e9)
    rng = np.random.default_rng(31)
    e1 = 2e-3 * rng.standard_normal((3, 3))
    e2 = 1e-3 * rng.standard_normal((3, 3))
    if kind == "hex":
        e1[2, :] = e1[:, 2] = e2[2, :] = e2[:, 2] = 0.0
    x0 = m.x0.copy()
    noise = 1e-4 * rng.standard_normal(x0.shape) * (1 if kind == "sc" else np.array([1.0, 1.0, 0.0]))
    xt = x0 + x0 @ e1.T + noise
    x1 = xt + x0 @ e2.T + 0.5 * noise[::-1]
    m.x[:] = xt
    oracle.lib().orc_bond_force_general(m.p, 6)
    m.Ftemp[:] = m.F
    m.xt[:] = xt
    m.x[:] = x1
    oracle.lib().orc_bond_force_general(m.p, 4)
    fpred = m.F.copy()
    oracle.lib().orc_bond_force_general(m.p, 6)
    live = m.nbr >= 0
    scale = np.abs(m.F[live]).max()
    assert scale > 0
    assert np.abs(fpred[live] - m.F[live]).max() <= 1e-10 * scale
    # and a plausible slip (the increment's sign) is visible at that tolerance
    assert np.abs(m.Ftemp[live] - m.F[live]).max() > 1e-3 * scale


def test_hex_affine_patch_plane(oracle):
    m, (c11, c12, c44) = hexm(oracle, 12.0, 10.0, sigmay=1e9)
    eps = 1e-7 * np.array([[1.0, 0.4, 0], [0.4, -0.6, 0], [0, 0, 0]])
    m.x[:] = m.x0 + m.x0 @ eps.T
    oracle.lib().orc_bond_force_general(m.p, 6)
    deep = deep_interior(m)
    sig = np.array([c11 * eps[0, 0] + c12 * eps[1, 1], c12 * eps[0, 0] + c11 * eps[1, 1], 2 * c44 * eps[0, 1]])
    got = m.stress[deep][:, [0, 1, 5]] * 1.0  # box_z = 1
    np.testing.assert_allclose(got, np.broadcast_to(sig, got.shape), rtol=0, atol=2e-6 * np.abs(sig).max())


def test_radial_return_closed_form(oracle):
    H = 2000.0
    m, c = sc(oracle, 7.1, sigmay=200.0, J2_H=H)
    eps = 4e-3 * np.array([[1.0, 0, 0], [0, -0.5, 0], [0, 0, -0.5]])
    m.x[:] = m.x0 + m.x0 @ eps.T
    oracle.lib().orc_bond_force_general(m.p, 0)
    deep = deep_interior(m)
    ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 0, 0, 0])
    s = voigt_C(*c) @ ev
    s[:3] -= s[:3].mean()
    seq = np.sqrt(1.5 * (np.sum(s[:3] ** 2) + 2 * np.sum(s[3:] ** 2)))
    dlam = (seq - 200.0) / (3 * c[2] + H)
    assert dlam > 0
    np.testing.assert_allclose(m.dlambda[deep], dlam, rtol=5e-3)
    # plastic flow reduces the deviatoric stress to the hardened yield surface
    assert np.all(m.pl_flag[deep] == 1)


def j2energy_pair(oracle, sigmay, seed):
    """Two identical SC models with a random slot-0 state and a random deformation."""
    out = []
    for _ in range(2):
        m, c = sc(oracle, 7.1, sigmay=sigmay, J2_H=2000.0, J2_xi=0.3)
        rng = np.random.default_rng(seed)
        live = m.nbr >= 0
        if sigmay < 1e6:
            m.dLp[:, :, 0] = 2e-5 * rng.standard_normal(live.shape) * live
            m.alpha[:, 0] = 1e-3 * rng.uniform(size=m.n)
            m.beta_eq[:, 0] = 20.0 * rng.standard_normal(m.n)
        eps = 4e-3 * np.array([[1.0, 0.2, 0.0], [0.2, -0.5, 0.1], [0.0, 0.1, -0.5]])
        m.x[:] = m.x0 + m.x0 @ eps.T + 1e-4 * rng.standard_normal(m.x.shape)
        out.append(m)
    return out


def test_j2energy_serial_order_equals_jacobi_form(oracle):
    """The energy-based J2 law (plmode 3) recomputes each particle's list from slot 0
    and writes only the particle's own slot 2, so the reference's serial sweep gives
    the values of the Jacobi form (every update once, all second stretch passes,
    then every force), bit for bit, with a good share of the particles yielding."""
    ms, mj = j2energy_pair(oracle, 200.0, 41)
    oracle.lib().orc_bond_force_j2energy(ms.p, 1, 0)
    oracle.lib().orc_bond_force_j2energy(mj.p, 1, 1)
    assert ms.pl_flag.sum() > ms.n // 4
    assert 0 < ms.dlambda.max() < 1
    for k in ("F", "Pin", "dL", "dL_ave", "ddLp", "dLtot", "TdLtot", "cs", "stress", "dLp", "alpha", "beta_eq",
              "dlambda", "pl_flag"):
        np.testing.assert_array_equal(getattr(ms, k), getattr(mj, k), err_msg=k)


def j2nl_model(oracle, strain, noise):
    m, c = sc(oracle, 7.1, sigmay=1e9)
    eps = strain * np.array([[1.0, 0.2, 0.0], [0.2, -0.5, 0.1], [0.0, 0.1, -0.5]])
    m.x[:] = m.x0 + m.x0 @ eps.T + noise * np.random.default_rng(5).standard_normal(m.x.shape)
    return m, c, eps


@pytest.mark.parametrize("jacobi", [0, 1])
def test_j2nonlinear_elastic_branch_is_the_elastic_law(oracle, jacobi):
    """Below SY(0) = 620 MPa the nonlinear-isotropic J2 law (plmode 5) leaves the state
    alone and its internal forces are the elastic law's."""
    m5, _, _ = j2nl_model(oracle, 5e-4, 1e-5)
    m6, _, _ = j2nl_model(oracle, 5e-4, 1e-5)
    oracle.lib().orc_bond_force_j2nonlinear(m5.p, 0.0, jacobi)
    oracle.lib().orc_bond_force_general(m6.p, 6)
    assert np.all(m5.dlambda == 0)
    np.testing.assert_array_equal(m5.Pin, m6.Pin)


def test_j2nonlinear_return_map_closed_form(oracle):
    """Under a uniform deviatoric strain a deep-interior particle's trial stress is
    C:eps, and the multiplier of the Jacobi form is the root of
    sigma_eq - 3 mu dl - SY(dl) = 0 (SY(a) = 620 + 3300 (1 - exp(-0.4 a)), no
    kinematic hardening) to the bisection's bracket TOLITER = 1e-4."""
    from scipy.optimize import brentq

    m, c, eps = j2nl_model(oracle, 4e-3, 0.0)
    oracle.lib().orc_bond_force_j2nonlinear(m.p, 0.0, 1)
    deep = deep_interior(m)
    ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[1, 2], 2 * eps[0, 2], 2 * eps[0, 1]])
    s = voigt_C(*c) @ ev
    s[:3] -= s[:3].mean()
    seq = np.sqrt(1.5 * (np.sum(s[:3] ** 2) + 2 * np.sum(s[3:] ** 2)))
    assert seq > 620.0
    root = brentq(lambda dl: seq - 1.5 * dl * 2.0 * c[2] - (620.0 + 3300.0 * (1.0 - np.exp(-0.4 * dl))), 0.0, 1.0,
                  xtol=1e-15)
    assert np.all(np.abs(m.dlambda[deep] - root) <= 1.5e-4), (m.dlambda[deep].min(), m.dlambda[deep].max(), root)


def test_j2nonlinear_keeps_no_plastic_state(oracle):
    """plmode 5 advances slot 0 in place and writes no slot 2; the switchStateV(2) that
    ends every evaluation copies slot 2 over slot 0, so a yielding evaluation leaves no
    plastic state behind (the reference's behaviour, kept by the restatement).  In the
    serial order a particle's multiplier is the one of its last recomputation, from a
    later neighbour's list, which starts on the yield surface and finds none."""
    mj, _, _ = j2nl_model(oracle, 4e-3, 1e-5)
    ms, _, _ = j2nl_model(oracle, 4e-3, 1e-5)
    oracle.lib().orc_bond_force_j2nonlinear(mj.p, 0.0, 1)
    oracle.lib().orc_bond_force_j2nonlinear(ms.p, 0.0, 0)
    assert (mj.dlambda > 0).sum() > mj.n // 4, "the deformation yields"
    for m in (mj, ms):
        assert np.all(m.alpha[:, 0] == 0) and np.all(m.dLp[:, :, 0] == 0)
    # the serial sweep did advance the particles (its forces are not the elastic law's)
    me, _, _ = j2nl_model(oracle, 4e-3, 1e-5)
    oracle.lib().orc_bond_force_general(me.p, 6)
    assert np.abs(ms.Pin - me.Pin).max() > 1e-3 * np.abs(me.Pin).max()
    assert np.all(ms.dlambda[deep_interior(ms)] == 0)


def bwise_model(oracle, seed):
    """SC model with a random local damage field, multipliers, triaxialities and a few
    bonds broken beforehand (both entries)."""
    m, _ = sc(oracle, 7.1, damage_threshold=0.9, damagec_A=0.5)
    rng = np.random.default_rng(seed)
    m.dloc[:, 0] = rng.uniform(0.6, 1.05, m.n)
    m.dlambda[:] = 0.05 * rng.uniform(size=m.n)
    m.triax[:] = rng.uniform(-3.0, 1.0, m.n)
    for _ in range(10):
        i = int(rng.integers(m.n))
        j = int(rng.integers(m.nb_initial[i]))
        q = m.nbr[i, j]
        m.dmg_broken[i, j] = 0.0
        m.dmg_broken[q, list(m.nbr[q]).index(i)] = 0.0
    return m


def test_bwise_local_damage_does_not_depend_on_the_order(oracle):
    """The bond-wise local damage (the damage of plmode 5) decides each bond from its
    two ends only: visiting the particles in reverse gives the same state and count."""
    a, b = bwise_model(oracle, 51), bwise_model(oracle, 51)
    ka = oracle.lib().orc_damage_bwise_local(a.p, 0)
    kb = oracle.lib().orc_damage_bwise_local(b.p, 1)
    assert ka == kb and ka > 10
    for k in ("dloc", "dmgD", "dmg_broken", "nb", "dmg_w"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


def test_bwise_local_damage_brute_force(oracle):
    """Against a pairwise restatement: the local damage rule, a bond broken exactly when
    the mean damage of its unbroken ends passes the threshold, the counts, and the
    weights 1 - D with D = 1 on broken bonds and at isolated particles."""
    m = bwise_model(oracle, 52)
    thr, A = 0.9, 0.5
    d0, dl, tx = m.dloc[:, 0].copy(), m.dlambda.copy(), m.triax.copy()
    live = m.nbr >= 0
    broken0 = (m.dmg_broken == 0) & live
    k = oracle.lib().orc_damage_bwise_local(m.p, 0)
    f = 1.0 + A * tx
    d1 = np.where((f > 0) & (d0 <= thr), d0 + f * dl, np.where(d0 > 1.0, 1.0, d0))
    np.testing.assert_allclose(m.dloc[:, 0], d1, rtol=0, atol=1e-15)
    q = np.where(live, m.nbr, 0)
    mean = 0.5 * (d1[:, None] + d1[q])
    newly = live & ~broken0 & (mean > thr)
    broken = broken0 | newly
    np.testing.assert_array_equal(m.dmg_broken == 0, broken)
    assert k == int(newly.sum()) // 2
    nb = m.nb_initial - broken.sum(axis=1)
    np.testing.assert_array_equal(m.nb, nb)
    iso = (nb[:, None] == 0) | (nb[q] == 0)
    D = np.where(broken | iso, 1.0, mean)
    np.testing.assert_allclose(m.dmg_w[live], 1.0 - D[live], rtol=0, atol=1e-15)


def test_pwise_local_damage_order_and_brute_force(oracle):
    """Point-wise local damage: a particle whose damage passes the threshold fails, all
    its bonds break both ways, and a bond's damage is the larger of its ends; the
    reverse particle order gives the same state."""
    a, b = bwise_model(oracle, 53), bwise_model(oracle, 53)
    thr, A = 0.9, 0.5
    d0, dl, tx = a.dloc[:, 0].copy(), a.dlambda.copy(), a.triax.copy()
    live = a.nbr >= 0
    broken0 = (a.dmg_broken == 0) & live
    ka = oracle.lib().orc_damage_pwise_local(a.p, 0)
    kb = oracle.lib().orc_damage_pwise_local(b.p, 1)
    assert ka == kb
    for k in ("dloc", "dmgD", "dmg_broken", "dmg_w"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    f = 1.0 + A * tx
    d1 = np.where((f > 0) & (d0 <= thr), d0 + f * dl, d0)
    fail = (d1 > thr) & (np.abs(d1 - 1.0) > 1e-6)
    assert ka == int(fail.sum()) and ka > 10
    d1 = np.where(fail, 1.0, d1)
    np.testing.assert_allclose(a.dloc[:, 0], d1, rtol=0, atol=1e-15)
    q = np.where(live, a.nbr, 0)
    broken = broken0 | (live & (fail[:, None] | fail[q]))
    np.testing.assert_array_equal(a.dmg_broken == 0, broken)
    D = np.maximum(d1[:, None], d1[q])
    np.testing.assert_array_equal(a.dmg_w[live], 1.0 - D[live])


def test_bwise_nonlocal_damage_brute_force(oracle):
    """Bond-wise nonlocal damage: a particle below the threshold gains the Gaussian-
    weighted mean of the local rates of itself and its list, and every directed bond
    whose ends average above the threshold breaks."""
    m = bwise_model(oracle, 54)
    thr, A, Ld = 0.9, 0.5, m.s.damage_L
    rng = np.random.default_rng(55)
    m.dnl[:, 0] = rng.uniform(0.6, 1.05, m.n)
    live = m.nbr >= 0
    broken0 = (m.dmg_broken == 0) & live
    m.dmgD[:, :, 0] = np.where(broken0, 1.0, 0.0)
    d0, dl, tx = m.dnl[:, 0].copy(), m.dlambda.copy(), m.triax.copy()
    k = oracle.lib().orc_damage_bwise_nonlocal(m.p)
    f = 1.0 + A * tx
    r = np.where(f > 0, dl * f, 0.0)
    q = np.where(live, m.nbr, 0)
    phi = np.where(live, 1.0 / Ld / np.sqrt(2 * np.pi) * np.exp(-0.5 * m.L0 ** 2 / Ld ** 2), 0.0)
    V = m.s.volume
    Ddot = V * r + (r[q] * phi * V).sum(axis=1)
    Aw = V + (phi * V).sum(axis=1)
    d1 = np.where(d0 > thr, np.minimum(d0, 1.0), np.where(Ddot > 0, d0 + Ddot / Aw, d0))
    np.testing.assert_allclose(m.dnl[:, 0], d1, rtol=1e-14, atol=0)
    mean = 0.5 * (d1[:, None] + d1[q])
    newly = live & ~broken0 & (mean > thr)
    assert k == int(newly.sum()) and k > 10
    np.testing.assert_array_equal(m.dmg_broken == 0, broken0 | newly)
    D = np.where(broken0 | newly, 1.0, mean)
    np.testing.assert_allclose(m.dmg_w[live], 1.0 - D[live], rtol=0, atol=1e-14)


def test_j2energy_elastic_branch_is_the_elastic_law(oracle):
    """Below yield the energy-based J2 law has no plastic stretch, the half-averaged
    stretch of a bond equals its own, and its forces are the elastic law's."""
    m3, m6 = j2energy_pair(oracle, 1e9, 43)
    oracle.lib().orc_bond_force_j2energy(m3.p, 1, 0)
    oracle.lib().orc_bond_force_general(m6.p, 6)
    assert m3.pl_flag.sum() == 0 and np.all(m3.dlambda == 0)
    np.testing.assert_array_equal(m3.F, m6.F)
    np.testing.assert_array_equal(m3.Pin, m6.Pin)
    np.testing.assert_array_equal(m3.stress, m6.stress)


def test_j2energy_plastic_return_lands_on_the_yield_surface(oracle):
    """The plastic branch of the energy-based J2 law (constitutive.c:286-463) returns
    to the hardened yield surface.  From the updated state alone (slot 2): the
    particle's elastic stretches with its new plastic stretch, their deviatoric energy
    U_d = sum_shells 1/2 Kn (dl - shell mean)^2, sigma = sqrt(6 c44 U_d / V) with
    V = volume nb / nbmax, and the updated back stress and hardening satisfy
    |sigma - beta_eq| = sigmay + (1 - xi) H alpha to the bisection's tolerance
    (dlambda within TOLITER = 1e-4 of the root).  This ties the plastic stretch
    increment (dlambda f_d / 4 Kn with f_d scaled by 1 / (1 + dlambda/2)), the
    energy-stress relation and the alpha and beta_eq updates together without
    restating the return equation; a factor of 2 off in any of them moves sigma by
    about dlambda sigma / 2, over 1% at the particles checked.  The strain is small
    enough that the root lies inside the bisection's bracket [0, 1] (past it the
    return stops at dlambda = 1, short of the surface); the update ends with the
    slot 2 -> 0 copy, so slot 2 is the updated state."""
    m, _ = sc(oracle, 7.1, sigmay=200.0, J2_H=2000.0, J2_xi=0.3)
    rng = np.random.default_rng(41)
    live = m.nbr >= 0
    m.dLp[:, :, 0] = 2e-5 * rng.standard_normal(live.shape) * live
    m.alpha[:, 0] = 1e-3 * rng.uniform(size=m.n)
    m.beta_eq[:, 0] = 20.0 * rng.standard_normal(m.n)
    eps = 1.1e-3 * np.array([[1.0, 0.2, 0.0], [0.2, -0.5, 0.1], [0.0, 0.1, -0.5]])
    m.x[:] = m.x0 + m.x0 @ eps.T + 3e-5 * rng.standard_normal(m.x.shape)
    oracle.lib().orc_bond_force_j2energy(m.p, 1, 0)
    s = m.s
    H, xi, nbm = s.J2_H, s.J2_xi, s.nbmax
    Ce = np.asarray(m.Ce).reshape(-1, 3)
    sel = np.flatnonzero((m.pl_flag == 1) & (m.dlambda > 1e-2) & (m.dlambda < 0.99))
    assert len(sel) > 100
    worst = 0.0
    for i in sel:
        nb = int(m.nb[i])
        assert nb == m.nb_initial[i]  # no broken bonds here
        q = m.nbr[i, :nb]
        dl = np.linalg.norm(m.x[i] - m.x[q], axis=1) - m.L0[i, :nb] - m.dLp[i, :nb, 2]
        ud = 0.0
        for shell in (0, 1):
            k = m.nsign[i, :nb] == shell
            if k.any():
                ud += 0.5 * np.sum(m.Kn[i, :nb][k] * (dl[k] - dl[k].mean()) ** 2)
        sig = np.sqrt(6.0 * Ce[m.type[i], 2] * ud / (s.volume * nb / nbm))
        f = abs(sig - m.beta_eq[i, 2]) - (m.sigmay[i] + (1.0 - xi) * H * m.alpha[i, 2])
        worst = max(worst, abs(f) / m.sigmay[i])
    assert worst < 5e-4, worst


# ---------------------------------------------------------------- tangent --------
def test_analytic_jacobian_matches_central_differences(oracle):
    m, _ = sc(oracle, 3.2)
    rng = np.random.default_rng(3)
    m.dLp[:, :, 0] = 1e-4 * rng.standard_normal(m.dLp[:, :, 0].shape) * (m.nbr >= 0)
    m.x[:] = m.x0 + 2e-3 * rng.standard_normal(m.x.shape)
    x = m.x.copy()
    n = m.n
    Jn = np.zeros((3 * n, 3 * n))
    h = 1e-6
    f1, f2 = np.zeros(3 * n), np.zeros(3 * n)
    for col in range(3 * n):
        xp = x.copy().reshape(-1)
        xp[col] += h
        oracle.lib().orc_internal_force(m.p, oracle.ptr(xp.reshape(n, 3).copy()), oracle.ptr(f1))
        xp[col] -= 2 * h
        oracle.lib().orc_internal_force(m.p, oracle.ptr(xp.reshape(n, 3).copy()), oracle.ptr(f2))
        Jn[:, col] = (f1 - f2) / (2 * h)
    oracle.lib().orc_tangent_analytic(m.p)
    Ka = oracle.upper_csr_dense(m)
    Jsym = Jn.copy()
    up = np.triu(np.ones_like(Jn), 1)
    # off-diagonal blocks: 0.5 (J + J^T); diagonal blocks: the upper triangle of J
    Jsym = 0.5 * (Jn + Jn.T)
    for i in range(n):
        b = Jn[3 * i:3 * i + 3, 3 * i:3 * i + 3]
        Jsym[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.triu(b) + np.triu(b, 1).T
    assert up.sum() > 0
    err = np.abs(Ka - Jsym).max() / np.abs(Jsym).max()
    assert err < 1e-7, err


def test_fd_tangent_pattern_symmetry_and_translation(oracle):
    m, _ = sc(oracle, 4.3)
    rng = np.random.default_rng(5)
    m.x[:] = m.x0 + 1e-3 * rng.standard_normal(m.x.shape)
    oracle.lib().orc_tangent_fd(m.p)
    Kfd = m.K.copy()
    ik, jk = m.IK.copy(), m.JK.copy()
    oracle.lib().orc_tangent_analytic(m.p)
    np.testing.assert_array_equal(m.IK, ik)
    np.testing.assert_array_equal(m.JK, jk)
    assert np.abs(Kfd - m.K).max() / np.abs(m.K).max() < 1e-5
    # 1-based upper CSR, ascending columns in each row, diagonal first
    for r in range(m.dim * m.n):
        cols = jk[ik[r] - 1:ik[r + 1] - 1] - 1
        assert cols[0] == r and np.all(np.diff(cols) > 0)
    D = oracle.upper_csr_dense(m)
    for c in range(3):
        t = np.zeros(3 * m.n)
        t[c::3] = 1.0
        assert np.abs(D @ t).max() < 1e-9 * np.abs(D).max()


def test_bc_elimination_and_cg_vs_dense(oracle):
    m, _ = sc(oracle, 4.3, ntype=3)
    n = m.n
    types = np.zeros(n, np.int32)
    z = m.x0[:, 2]
    types[z < z.min() + 0.1] = 2
    types[z > z.max() - 0.1] = 1
    m.type[:] = types
    oracle.lib().orc_kntv(m.p)
    oracle.lib().orc_tangent_analytic(m.p)
    m.dispBC.reshape(n, 3)[types == 2, :] = 0  # bottom fully fixed
    m.dispBC.reshape(n, 3)[types == 1, 2] = 0  # top z prescribed
    rng = np.random.default_rng(1)
    m.residual[:] = rng.standard_normal(3 * n) * m.dispBC
    D0 = oracle.upper_csr_dense(m)
    oracle.lib().orc_bc_stiffness_update(m.p)
    D = oracle.upper_csr_dense(m)
    bc = np.where(m.dispBC == 0)[0]
    nd = np.linalg.norm(np.diag(D0))
    for r in bc:
        row = D[r].copy()
        assert abs(row[r] - nd) <= 1e-14 * nd
        row[r] = 0
        assert not row.any() and not D[:, r][np.arange(3 * n) != r].any()
    want = np.linalg.solve(D, m.residual)
    x0 = m.x.copy()
    s = oracle.solver(rtol2=1e-28, atol2=0.0, maxit=20 * 3 * n)
    it = oracle.lib().orc_solve_cg(m.p, s, None)
    assert it > 0
    np.testing.assert_allclose(m.disp, want, rtol=0, atol=1e-10 * np.abs(want).max())
    np.testing.assert_allclose(m.x - x0, m.disp.reshape(n, 3), rtol=0, atol=1e-15)
    # the threaded timing baseline runs the same CG recurrence on the full CSR
    for k in (1, 7, 30):
        s1 = oracle.solver(rtol2=0.0, atol2=0.0, maxit=k)
        m.residual[:] = rng.standard_normal(3 * n) * m.dispBC
        rr_ser = C.c_double()
        oracle.lib().orc_solve_cg(m.p, s1, C.byref(rr_ser))
        rr_par = C.c_double()
        oracle.lib().orc_cg_baseline.restype = C.c_double
        t = oracle.lib().orc_cg_baseline(m.p, 4, k, C.byref(rr_par))
        assert t >= 0.0
        assert abs(rr_par.value - rr_ser.value) <= 1e-9 * rr_ser.value


# ---------------------------------------------------------------- damage ---------
def test_nonlocal_damage_uniform_field(oracle):
    m, _ = sc(oracle, 6.3, damage_L=0.5)
    m.dlambda[:] = 3e-3
    m.triax[:] = 0.25
    oracle.lib().orc_damage_pwise_nonlocal(m.p)
    # a uniform rate field is reproduced exactly by the normalised Gaussian average
    np.testing.assert_allclose(m.dnl[:, 0], 3e-3, rtol=1e-14)
    live = m.nbr >= 0
    np.testing.assert_allclose(m.dmg_w[live], 1 - 3e-3, rtol=1e-14)


def test_nonlocal_damage_weights_brute_force(oracle):
    m, _ = sc(oracle, 5.2, damage_L=0.7, damagec_A=0.5)
    rng = np.random.default_rng(7)
    m.dlambda[:] = rng.uniform(0, 1e-2, m.n)
    m.triax[:] = rng.uniform(-0.5, 0.5, m.n)
    oracle.lib().orc_damage_pwise_nonlocal(m.p)
    L = 0.7
    X = m.x0
    d = np.sqrt(((X[None] - X[:, None]) ** 2).sum(-1))
    w = np.where(d < 3 * L, np.exp(-0.5 * d * d / L / L), 0.0)
    rate = m.dlambda * (1 + 0.5 * m.triax)
    np.testing.assert_allclose(m.dnl[:, 0], (w @ rate) / w.sum(1), rtol=1e-12)


def test_shell_sort_selection(oracle):
    rng = np.random.default_rng(11)
    for k in (1, 2, 7, 40, 401):
        v = rng.uniform(0, 1, k)
        order = np.zeros(k, np.int32)
        oracle.lib().orc_shell_sort_select(k, oracle.ptr(v), 5, oracle.ptr(order))
        assert np.all(np.diff(v[order]) >= 0)
        assert sorted(order) == list(range(k))
    # with ties the order of equal keys is the shell sort's, not a stable sort's:
    # check against an independent restatement of the reference's gap sequence
    v = np.repeat(rng.uniform(0, 1, 30), 2)[rng.permutation(60)]
    order = np.zeros(60, np.int32)
    oracle.lib().orc_shell_sort_select(60, oracle.ptr(v), 5, oracle.ptr(order))
    idx = list(range(60))
    gap = 30
    while gap >= 1:
        for i in range(gap, 60):
            t = idx[i]
            j = i - gap
            while j >= 0 and v[idx[j]] > v[t]:
                idx[j + gap] = idx[j]
                j -= gap
            idx[j + gap] = t
        gap //= 2
    assert list(order) == idx


def test_brittle_breaks_largest(oracle):
    m, _ = hexm(oracle, 8.0, 7.0, plmode=6, nbreak=4, critical_bstrain=1e-3)
    rng = np.random.default_rng(2)
    live = m.nbr >= 0
    m.dL[:] = 0
    m.dL[live] = m.L0[live] * rng.uniform(0, 2e-3, live.sum())
    strain = m.dL / np.where(live, m.L0, 1)
    cand = np.argwhere(live & (strain >= 1e-3))
    k = oracle.lib().orc_damage_brittle(m.p, None, None, 0)
    assert k == len(cand)
    broken = np.argwhere(live & (m.dmg_broken == 0))
    assert len(broken) == 4
    top = np.sort(strain[live])[-4:]
    assert np.allclose(np.sort(strain[tuple(broken.T)]), top)


def test_weibull_bond_critical_is_symmetric_with_the_weibull_mean(oracle):
    from lpm_harness import P

    x = oracle.create_lattice(1, 0.5, [0.0, 40.0, 0.0, 40.0, 0.0, 1.0])
    t = np.zeros(len(x), np.int32)
    m = oracle.Model(x, 1, 0.5, t, [oracle.elastic_constants(70e3, 0.25)], plmode=6)
    m.setup()
    crit = P.weibull_bond_critical(m.nbr, 1, shape=10.0, scale=1e-3)
    vals = []
    for i in range(m.n):
        for a in range(m.nb_initial[i]):
            j = m.nbr[i, a]
            b = list(m.nbr[j]).index(i)
            assert crit[i, a] == crit[j, b]
            vals.append(crit[i, a])
    vals = np.array(vals)
    import math

    mean = 1e-3 * math.gamma(1.0 + 1.0 / 10.0)
    sd = 1e-3 * math.sqrt(math.gamma(1.0 + 2.0 / 10.0) - math.gamma(1.0 + 1.0 / 10.0) ** 2)
    # each undirected bond appears twice
    assert abs(vals.mean() - mean) <= 4 * sd / math.sqrt(len(vals) / 2)
    assert not np.array_equal(crit, P.weibull_bond_critical(m.nbr, 2))


# ----------------------------------------------------------- crystal plasticity ----
CP = dict(tau0=10.0, taus=1e20, h0=100.0, q=1.0, eta=1.0, p=1.0, dtime=1.0, maxloop=20)


def fcc_cp(O, L, cp=CP):
    x = O.create_lattice(3, 0.5, [0, L, 0, L, 0, L])
    c = O.elastic_constants(70e3, 0.3)
    m = O.Model(x, 3, 0.5, np.zeros(x.shape[0], np.int32), [c], plmode=1, cp=cp)
    m.setup()
    O.lib().orc_compute_cab(m.p)
    return m, c


def test_fcc_slip_systems_closed_form(oracle):
    m, _ = fcc_cp(oracle, 4.3)
    t = m.schmid
    assert t.shape == (24, 6)
    # n . d = 0: the Schmid tensor is traceless; |sym(n x d)|^2 = 1/2 for unit n, d
    assert np.abs(t[:, :3].sum(axis=1)).max() < 1e-15
    np.testing.assert_allclose((t[:, :3] ** 2).sum(1) + (t[:, 3:] ** 2).sum(1) / 2, 0.5, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(t[1::2], -t[0::2])  # both senses of each direction
    zz = t[:, 2]
    big = np.isclose(np.abs(zz), 1 / np.sqrt(6), rtol=0, atol=1e-15)
    assert big.sum() == 16 and (zz[big] > 0).sum() == 8 and np.all(zz[~big] == 0)
    # a rotation of the crystal rotates every Schmid tensor
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    sr = np.zeros((24, 6))
    oracle.lib().orc_slip_systems_fcc(oracle.ptr(np.ascontiguousarray(R.reshape(9))), oracle.ptr(sr))

    def full(v):
        return np.array([[v[0], v[5] / 2, v[4] / 2], [v[5] / 2, v[1], v[3] / 2], [v[4] / 2, v[3] / 2, v[2]]])
    for s in range(24):
        np.testing.assert_allclose(full(sr[s]), R @ full(t[s]) @ R.T, rtol=0, atol=1e-15)


def test_lu_solve_matches_dense_solve(oracle):
    rng = np.random.default_rng(3)
    for n in (5, 24):
        a = rng.standard_normal((n, n)) + n * np.eye(n) * rng.choice([-1, 1], n)[:, None] * 0.1
        a = a[rng.permutation(n)]  # pivoting is needed
        b = rng.standard_normal(n)
        aa, bb = np.ascontiguousarray(a.copy()), b.copy()
        assert oracle.lib().orc_lu_solve(n, oracle.ptr(aa), oracle.ptr(bb)) == 0
        np.testing.assert_allclose(bb, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)
        # the factors reproduce a row permutation of a
        L = np.tril(aa, -1) + np.eye(n)
        U = np.triu(aa)
        LU = L @ U
        rows = [int(np.argmin(np.abs(a - LU[k]).sum(1))) for k in range(n)]
        assert sorted(rows) == list(range(n))
        np.testing.assert_allclose(LU, a[rows], rtol=0, atol=1e-12 * np.abs(a).max())
    s = np.zeros((3, 3))
    s[0] = [1.0, 2.0, 3.0]
    assert oracle.lib().orc_lu_solve(3, oracle.ptr(s), oracle.ptr(np.ones(3))) > 0


def test_cab_is_minus_rss_derivative(oracle):
    """C_ab = -dRSS_a/dgamma_b of the return map's own stress path (linear)."""
    m, _ = fcc_cp(oracle, 6.5)
    i = deep_interior(m)[0]
    cab = m.cp_Cab[i]
    g0 = np.zeros(24)
    r0 = np.zeros(24)
    oracle.lib().orc_cp_rss(m.p, int(i), oracle.ptr(g0), oracle.ptr(r0))
    h = 1e-4
    for b in range(24):
        g = np.zeros(24)
        g[b] = h
        r = np.zeros(24)
        oracle.lib().orc_cp_rss(m.p, int(i), oracle.ptr(g), oracle.ptr(r))
        np.testing.assert_allclose(-(r - r0) / h, cab[:, b], rtol=0, atol=1e-8 * np.abs(cab).max())
    assert np.abs(cab - cab.T).max() <= 1e-9 * np.abs(cab).max()


def test_single_slip_return_map_closed_form(oracle):
    """Uniaxial strain along a single-slip direction just past first yield: exactly
    the most loaded system activates, and its slip solves the scalar residual
    RSS_tr - C'aa g - (tau0 + h0 g)(1 + g eta/dt) = 0 in closed form."""
    m, c = fcc_cp(oracle, 6.5)
    t = np.array([-1.0, 2.0, 3.0]) / np.sqrt(14.0)
    mu = c[2]
    ev = np.outer(t, t)
    fac = np.array([2 * mu * (s[0] * t[0] ** 2 + s[1] * t[1] ** 2 + s[2] * t[2] ** 2 + s[3] * t[1] * t[2] +
                              s[4] * t[0] * t[2] + s[5] * t[0] * t[1]) for s in m.schmid])
    a = int(np.argmax(fac))
    second = np.sort(fac)[-2]
    assert second < 0.95 * fac[a]
    e = 1.02 * CP["tau0"] / fac[a]
    m.x[:] = m.x0 + m.x0 @ (e * ev).T
    i = deep_interior(m)[0]
    # trial RSS and the slip stiffness at the deformed geometry, from the stress path
    oracle.lib().orc_computedL(m.p)
    r0, r1 = np.zeros(24), np.zeros(24)
    g = np.zeros(24)
    oracle.lib().orc_cp_rss(m.p, int(i), oracle.ptr(g), oracle.ptr(r0))
    g[a] = 1e-4
    oracle.lib().orc_cp_rss(m.p, int(i), oracle.ptr(g), oracle.ptr(r1))
    caa = (r0[a] - r1[a]) / 1e-4
    assert abs(r0[a] - 1.02 * CP["tau0"]) < 1e-3 * CP["tau0"]  # the bond model's stress is C:eps
    oracle.lib().orc_bond_force_general(m.p, 1)
    assert m.s.cp_fail == 0
    assert list(np.nonzero(m.cp_Jact[i])[0]) == [a]
    k = CP["eta"] / CP["dtime"]
    qa, qb, qc = k * CP["h0"], caa + CP["h0"] + CP["tau0"] * k, -(r0[a] - CP["tau0"])
    gam = (-qb + np.sqrt(qb * qb - 4 * qa * qc)) / (2 * qa)
    assert gam > 0
    np.testing.assert_allclose(m.cp_dA_single[i, a], gam, rtol=1e-6)
    # the accumulated slip and the hardening are formed from the slips at the start
    # of the last inner iteration, before its Newton update (constitutive.c:1111-1131)
    np.testing.assert_allclose(m.cp_A[i, 2], gam, rtol=1e-3)
    np.testing.assert_allclose(m.cp_dgy[i, a], CP["h0"] * m.cp_A[i, 2], rtol=1e-12)
    gy2 = m.cp_gy[i, :, 2]
    assert abs(gy2[a] - (CP["tau0"] + m.cp_dgy[i, a])) <= 1e-12 * CP["tau0"]
    np.testing.assert_array_equal(np.delete(gy2, a), CP["tau0"])
    assert m.pl_flag[i] == 1


def test_accelerated_loop_reaches_the_plain_equilibrium(oracle):
    """The oracle's optional Anderson mixing changes the path, not the equilibrium:
    J2 tension through yield, plain modified Newton against the accelerated one."""
    from lpm_harness import P

    out = {}
    for acc in (0, 5):
        L = P.sc_box_for(6)
        x = oracle.create_lattice(2, 0.5, [0, L, 0, L, 0, L])
        c = oracle.elastic_constants(E0, NU0)
        m = oracle.Model(x, 2, 0.5, P.sc_types(x), [c] * 5, plmode=0, sigmay=200.0, J2_H=2000.0)
        m.setup()
        n = m.n
        steps = P.sc_uniaxial_steps(5.0, np.array([0.5e-3, 1e-3, 1.5e-3]))
        s = oracle.solver(rtol2=1e-24, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-11, newton_maxit=600,
                          tangent=1, fixed_diag=1, accel=acc)
        rc, o = oracle.run_steps(m, steps, s, out_axis="z", out_type=1)
        assert rc == 0
        out[acc] = (m.x - x, [k.newton_iters for k in o], [k.reaction for k in o])
    u0, u5 = out[0][0], out[5][0]
    assert np.abs(u0 - u5).max() <= 1e-9 * np.abs(u0).max()
    np.testing.assert_allclose(out[5][2], out[0][2], rtol=1e-9)
    assert sum(out[5][1]) < 0.5 * sum(out[0][1])


# ------------------------------------------------------------------ adjoint --------
def j2_state_model(oracle, seed, xi=0.3):
    """SC model past yield with a non-trivial slot-0 state and modulus field."""
    rng = np.random.default_rng(seed)
    x = oracle.create_lattice(2, 0.5, [0, 5.125, 0, 5.125, 0, 5.125])
    n = x.shape[0]
    c = oracle.elastic_constants(E0, NU0)
    emod = 1.0 + 0.05 * rng.uniform(-1, 1, n)
    m = oracle.Model(x, 2, 0.5, np.zeros(n, np.int32), [c], plmode=0, sigmay=200.0 + 20 * rng.uniform(-1, 1, n),
                     J2_H=2000.0, J2_xi=xi, emod=emod)
    m.setup()
    eps = 1.2e-3 * np.array([[1.0, 0.1, 0], [0.1, -0.4, 0.05], [0, 0.05, -0.3]])
    m.x[:] = m.x0 + m.x0 @ eps.T + 2e-5 * rng.standard_normal((n, 3))
    live = m.nbr >= 0
    m.dLp[:, :, 0] = 1e-5 * rng.standard_normal(m.nbr.shape) * live
    m.beta[:, :, 0] = 5.0 * rng.standard_normal((n, 6))
    m.alpha[:, 0] = 1e-3 * rng.uniform(0, 1, n)
    return m, rng


def j2_objective(oracle, m, seeds):
    Pb, dLp2b, beta2b, alpha2b = seeds
    oracle.lib().orc_kntv(m.p)
    oracle.lib().orc_eval_j2(m.p)
    return (Pb @ m.Pin + np.sum(dLp2b * m.dLp[:, :, 2]) + np.sum(beta2b * m.beta[:, :, 2]) +
            alpha2b @ m.alpha[:, 2])


def test_j2_vjp_matches_central_differences(oracle):
    """Reverse mode of one J2 force evaluation against central differences of the
    oracle's own forward law, along random directions of every input: positions,
    the slot-0 plastic stretch, back stress and hardening variable, the
    per-particle modulus factor and yield stress, and H."""
    m, rng = j2_state_model(oracle, 11)
    n, nb = m.n, m.nbmax
    live = m.nbr >= 0
    seeds = (rng.standard_normal(3 * n), rng.standard_normal((n, nb)) * live * 1e5,
             rng.standard_normal((n, 6)), rng.standard_normal(n) * 1e3)
    j2_objective(oracle, m, seeds)
    assert 0 < m.pl_flag.sum() < n, "some particles yield, some do not"
    out = [np.zeros(3 * n), np.zeros((n, nb)), np.zeros((n, 6)), np.zeros(n), np.zeros(n), np.zeros(n)]
    Hb = C.c_double()
    P = oracle.ptr
    oracle.lib().orc_vjp_j2(m.p, P(seeds[0]), P(np.ascontiguousarray(seeds[1])), P(np.ascontiguousarray(seeds[2])),
                            P(seeds[3]), *[P(o) for o in out], C.byref(Hb))
    xb, dLp0b, beta0b, alpha0b, eb, syb = out

    def fd(setter, direction, h):
        setter(+h)
        jp = j2_objective(oracle, m, seeds)
        setter(-2 * h)
        jm = j2_objective(oracle, m, seeds)
        setter(+h)
        return (jp - jm) / (2 * h)

    checks = []
    d = rng.standard_normal((n, 3))

    def sx(h):
        m.x[:] += h * d
    checks.append(("positions", fd(sx, d, 1e-7), xb @ d.reshape(-1)))
    d2 = rng.standard_normal((n, nb)) * live

    def sp(h):
        m.dLp[:, :, 0] += h * d2
    checks.append(("dLp0", fd(sp, d2, 1e-9), np.sum(dLp0b * d2)))
    d3 = rng.standard_normal((n, 6))

    def sb(h):
        m.beta[:, :, 0] += h * d3
    checks.append(("beta0", fd(sb, d3, 1e-3), np.sum(beta0b * d3)))
    d4 = rng.uniform(0, 1, n)

    def sa(h):
        m.alpha[:, 0] += h * d4
    checks.append(("alpha0", fd(sa, d4, 1e-7), alpha0b @ d4))
    d5 = rng.standard_normal(n)

    def se(h):
        m.emod[:] += h * d5
    checks.append(("modulus", fd(se, d5, 1e-7), eb @ d5))
    d6 = rng.standard_normal(n)

    def ss(h):
        m.sigmay[:] += h * d6
    checks.append(("sigmay", fd(ss, d6, 1e-3), syb @ d6))

    def sh(h):
        m.s.J2_H += h
    checks.append(("H", fd(sh, None, 1e-2), Hb.value))
    for name, f, a in checks:
        assert abs(f - a) <= 1e-6 * max(abs(f), abs(a)) + 1e-12, (name, f, a)


def j2_history_model(oracle, emod, sy, H):
    from lpm_harness import P

    L = P.sc_box_for(5)
    x = oracle.create_lattice(2, 0.5, [0, L, 0, L, 0, L])
    c = oracle.elastic_constants(E0, NU0)
    m = oracle.Model(x, 2, 0.5, P.sc_types(x), [c] * 5, plmode=0, sigmay=sy, J2_H=H, emod=emod)
    m.setup()
    steps = P.sc_uniaxial_steps(L - 1.0, np.array([0.6e-3, 1.1e-3, 1.5e-3]))
    return m, steps


def history_solver(oracle, n):
    return oracle.solver(rtol2=1e-26, atol2=0.0, maxit=100 * 3 * n, newton_tol=1e-12, newton_maxit=800,
                         tangent=1, fixed_diag=1, accel=5, skip_damage=1)


def test_history_adjoint_matches_central_differences(oracle):
    """Discrete adjoint of a three-step J2 tension history (damage off) against
    central differences of the whole forward history, for the per-particle modulus
    factor and yield stress and for H.  J = sum_n (R_n - t_n)^2 over the reactions
    of the moved face.  The adjoint solves are dense LU on the free DOFs."""
    from lpm_harness import P

    n = oracle.create_lattice(2, 0.5, [0, P.sc_box_for(5)] * 3).shape[0]
    rng = np.random.default_rng(5)
    e0 = 1.0 + 0.05 * rng.uniform(-1, 1, n)
    y0 = 200.0 + 30.0 * rng.uniform(-1, 1, n)
    H0 = 2000.0
    s = history_solver(oracle, n)

    def forward(emod, sy, H):
        m, steps = j2_history_model(oracle, emod, sy, H)
        rc, o = oracle.run_steps(m, steps, s)
        assert rc == 0
        return m, np.array([k.reaction for k in o])

    m, R0 = forward(e0, y0, H0)
    assert 0 < m.pl_flag.sum() < n
    target = 0.9 * R0
    m, steps = j2_history_model(oracle, e0, y0, H0)
    rc, J, reac, ge, gy, gh = oracle.adjoint_run(m, steps, s, target)
    assert rc == 0
    np.testing.assert_allclose(reac, R0, rtol=1e-10)
    assert abs(J - np.sum((R0 - target) ** 2)) <= 1e-10 * J

    def Jof(*a):
        return float(np.sum((forward(*a)[1] - target) ** 2))

    d1 = rng.standard_normal(n)
    d2 = rng.standard_normal(n)
    checks = []
    h = 1e-4
    checks.append(("modulus", (Jof(e0 + h * d1, y0, H0) - Jof(e0 - h * d1, y0, H0)) / (2 * h), ge @ d1))
    h = 1e-3
    checks.append(("sigmay", (Jof(e0, y0 + h * d2, H0) - Jof(e0, y0 - h * d2, H0)) / (2 * h), gy @ d2))
    h = 5.0
    checks.append(("H", (Jof(e0, y0, H0 + h) - Jof(e0, y0, H0 - h)) / (2 * h), gh))
    for name, f, a in checks:
        assert abs(f - a) <= 1e-6 * max(abs(f), abs(a)), (name, f, a)
