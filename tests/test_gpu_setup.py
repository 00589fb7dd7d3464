"""Setup and post-processing on the device against the oracle.

createCuboid's lattice points come from the device generator (lpmx_lattice_points,
one thread per candidate row): compared bit for bit with the oracle's restatement
of the lattice rules (reference src/initialization.c:6-560) for all five lattices,
rotated by both Euler conventions, and at 1M SC particles.  computeStrain
(reference src/lpm_basic.c:127-249) runs on the device (lpmx_compute_strain, one
thread per particle) and is compared bit for bit with the oracle's, whose pins are
the exact recovery of a linear stretch field, numpy's weighted least squares and
the stored right-hand side of a singular system: on SC and hexagonal lattices and
on a collinear chain placed ahead of a block.
"""
import numpy as np
import pytest

from lpm_harness import P, host_lattice, oracle_model

pytestmark = pytest.mark.gpu

LATTICES = [(2, [0, 9.151, 0, 9.151, 0, 9.151]), (1, [0, 20.0, 0, 17.0, 0, 1.0]), (3, [0, 8.0, 0, 8.0, 0, 8.0]),
            (0, [0, 10.0, 0, 8.0, 0, 1.0]), (4, [0, 6.0, 0, 6.0, 0, 6.0])]


@pytest.mark.parametrize("lat,box", LATTICES)
@pytest.mark.parametrize("rot", [None, (0, (0.3, -0.2, 0.7)), (1, (0.4, 0.5, -0.3))])
def test_create_cuboid_matches_oracle(oracle, lat, box, rot):
    eul, ang = (0, (0.0, 0.0, 0.0)) if rot is None else rot
    if lat < 2:
        ang = (0.0, 0.0, ang[2])  # in-plane rotation of the 2D lattices
    n = P.make_lattice(lat, 0.5, box, angles=ang, eulerflag=eul)
    x = P.positions()
    xo = oracle.create_lattice(lat, 0.5, box, eulerflag=eul, angles=ang)
    print(f"lattice {lat} rotation {rot}: {n} particles", flush=True)
    assert n == xo.shape[0] > 0
    np.testing.assert_array_equal(x, xo)
    lp = oracle.lattice_params(lat, 0.5, box[5] - box[4])
    assert P.gint("nneighbors").value == lp.nbmax
    assert P.gdouble("particle_volume").value == lp.volume
    assert P.gdouble("neighbor2_cutoff").value == lp.cut2


def test_create_cuboid_1m_sc_matches_oracle(oracle):
    L = P.SC_BOX[1000000]
    n = P.make_lattice(2, 0.5, [0, L, 0, L, 0, L])
    assert n == 1000000
    xo = oracle.create_lattice(2, 0.5, [0, L, 0, L, 0, L])
    np.testing.assert_array_equal(P.positions(), xo)


def strain_case(oracle, kind):
    """A model on the device (through the reference API) and the same one in the
    oracle, with the same random bond stretches dL on both."""
    L = P.lib()
    if kind == "sc":
        lat, box = 2, [0, 6.2, 0, 6.2, 0, 6.2]
        n = P.make_lattice(lat, 0.5, box)
        x = P.positions()
    elif kind == "hex":
        lat, box = 1, [0, 12.0, 0, 10.0, 0, 1.0]
        n = P.make_lattice(lat, 0.5, box)
        x = P.positions()
    else:
        lat, box = 2, [0, 4.2, 0, 4.2, 0, 4.2]
        blk = oracle.create_lattice(lat, 0.5, box)
        chain = np.array([[100.0, 100.0, 100.0], [101.0, 100.0, 100.0], [102.0, 100.0, 100.0]])
        n, x = host_lattice(oracle, lat, 0.5, box, np.vstack([chain, blk]))
    L.searchNormalNeighbor()
    m = oracle_model(oracle, x, lat, np.zeros(n, np.int32), [(1.0, 0.5, 0.25)])
    nb = P.gint("nneighbors").value
    np.testing.assert_array_equal(P.read2d("neighbors", n, nb, np.int32), m.nbr)
    rng = np.random.default_rng({"sc": 1, "hex": 2, "chain": 3}[kind])
    live = m.nbr >= 0
    dl = np.zeros((n, nb))
    dl[live] = 1e-3 * rng.standard_normal(int(live.sum())) * m.L0[live]
    m.dL[:] = dl
    rows = P.gptr("dL", 2)
    st = P.gptr("strain_tensor", 2)
    for i in range(n):
        for j in range(nb):
            rows[i][j] = dl[i, j]
        for c in range(6):
            st[i][c] = 7.0
    return n, m


@pytest.mark.parametrize("kind", ["sc", "hex", "chain"])
def test_compute_strain_matches_oracle(oracle, kind):
    n, m = strain_case(oracle, kind)
    P.lib().computeStrain()
    got = P.read2d("strain_tensor", n, 6)
    want = m.compute_strain(np.full((n, 6), 7.0))
    print(f"{kind}: {n} particles, max |strain| {np.abs(want).max():.3e}", flush=True)
    np.testing.assert_array_equal(got, want)
    if kind == "chain":
        assert np.all(got[0] == 7.0)            # a singular particle 0 keeps its strain
        assert got[1, 0] != 7.0 and np.all(got[1, 1:] == 0.0)  # the others store the rhs


def test_solver_pardiso_gives_the_direct_solution(oracle):
    """solverPARDISO (reference src/solver.c:3-92) factorises the constrained tangent
    (symmetric positive definite, mtype 2) and solves K disp = residual, then moves
    the particles by disp.  The drop-in's answer must be the direct solution: against
    numpy's dense solve of the same upper-CSR system to 1e-10 relative, with the
    positions moved by exactly disp; with the extension solver settings left at their
    defaults, as a reference driver leaves them."""
    from lpm_harness import sc_pair
    L = P.lib()
    L.lpm_ext_set_solver(None)
    _, n, x, t, m = sc_pair(oracle, 6)
    L.calcStiffness3DFiniteDifference(6)
    dbc = np.ones(3 * n, np.int32)
    dbc.reshape(n, 3)[t == 2, :] = 0
    dbc.reshape(n, 3)[t == 3, :] = 0
    dbc.reshape(n, 3)[t == 1, 2] = 0
    P.write1d("dispBC_index", dbc)
    rng = np.random.default_rng(7)
    P.write1d("residual", rng.standard_normal(3 * n) * dbc)
    L.setDispBC_stiffnessUpdate3D()
    K = P.read1d("K_global", m.s.nK)
    r = P.read1d("residual", 3 * n)
    oracle.lib().orc_tangent_analytic(m.p)
    m.dispBC[:] = dbc
    oracle.lib().orc_bc_stiffness_update(m.p)
    m.K[:] = K  # the device's eliminated tangent, in the oracle's upper-CSR pattern
    want = np.linalg.solve(oracle.upper_csr_dense(m), r)
    x0 = P.positions()
    L.solverPARDISO()
    d = P.read1d("disp", 3 * n)
    err = np.abs(d - want).max() / np.abs(want).max()
    print(f"solverPARDISO: {3 * n} DOF, max relative error {err:.2e}", flush=True)
    assert err <= 1e-10
    np.testing.assert_array_equal(P.positions(), x0 + d.reshape(n, 3))
