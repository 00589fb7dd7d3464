"""Shared setup of the parity tests: the product harness (lpm-c_amd/__init__.py,
loaded by path because of the package directory's name) and model builders that
give the oracle exactly the product's inputs."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def product():
    spec = importlib.util.spec_from_file_location("lpm_c_amd", os.path.join(ROOT, "lpm-c_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


P = product()


def oracle_model(O, x, lattice, types, Ce, **kw):
    m = O.Model(x, lattice, 0.5, types, Ce, **kw)
    m.setup()
    return m


def sc_pair(O, k, E=200e3, nu=0.3, sigmay=200.0, H=2000.0, emod=None):
    """The same SC model in the product (device, through the drop-in API) and in
    the oracle: k^3 particles, uniaxial types, J2 mixed-linear material."""
    L = P.sc_box_for(k)
    n, x, t, c = P.build_sc_model(L, E=E, nu=nu, sigmay=sigmay, H=H, emod=emod)
    m = oracle_model(O, x, 2, t, [c] * 5, plmode=0, sigmay=sigmay, J2_H=H, emod=emod)
    return L, n, x, t, m


def host_lattice(O, lattice, radius, box, x=None):
    """The host state that createCuboid and initMatrices leave (lattice globals,
    positions, per-particle arrays), with the lattice description from the oracle's
    restatement of createCuboid and its positions (or the given ones, x [n][3]): for
    CPU tests of host code, since the product generates the points on the device, and
    for models on positions no lattice rule makes."""
    import ctypes as C
    L = P.lib()
    lp = O.lattice_params(lattice, radius, box[5] - box[4])
    if x is None:
        x = O.create_lattice(lattice, radius, box)
    n = x.shape[0]
    P.gint("lattice").value = lattice
    P.gint("dim").value = lp.dim
    P.gdouble("radius").value = radius
    b = (C.c_double * 6).in_dll(L, "box")
    for k in range(6):
        b[k] = box[k]
    P.gdouble("box_x").value = box[1] - box[0]
    P.gdouble("box_y").value = box[3] - box[2]
    P.gdouble("box_z").value = box[5] - box[4]
    P.gint("nneighbors").value = lp.nbmax
    P.gint("nneighbors1").value = lp.nb1max
    P.gint("nneighbors2").value = lp.nb2max
    P.gdouble("particle_volume").value = lp.volume
    P.gdouble("neighbor1_cutoff").value = lp.cut1
    P.gdouble("neighbor2_cutoff").value = lp.cut2
    P.gint("nneighbors_AFEM").value = lp.nafem
    P.gint("nparticle").value = n
    P.set_gptr("xyz", L.allocDouble2D(n, 3, 0.0), C.c_double, 2)
    xyz = P.gptr("xyz", 2)
    for i in range(n):
        for c in range(3):
            xyz[i][c] = x[i, c]
    L.initMatrices()
    L.copyDouble2D(P.gptr("xyz_initial", 2), xyz, n, 3)
    return n, x


def push(name, arr):
    a = np.ascontiguousarray(arr)
    P.check(P.lib().lpmx_put(P.F[name], a.ctypes.data, a.size), name)


def get(name, shape):
    return P.field(name, int(np.prod(shape))).reshape(shape)


# FP64 displacements and reactions must agree with the oracle to 1e-10 relative;
# the mixed form states 1e-8.
TOL_FP64 = 1e-10
TOL_MIXED = 1e-8
# Crystal plasticity: the reference's return map stops its slip Newton iteration at
# |r| <= TOLITER, so the internal force it returns is a piecewise function of the
# positions whose jumps leave a residual floor near 1e-12 of the reaction (the
# oracle's own loop stalls there), and the slip modes soften the tangent by E/h0;
# two converged solvers then agree on the displacements to a few 1e-10 .. 1e-9
# (measured: 6.8e-10 on 108 particles, 3.4e-9 on 500, at a Newton tolerance of
# 1e-11).  The slip sets are compared bit for bit; displacements to 1e-8 in both forms.
TOL_CP = 1e-8


def tol_for(prec):
    return TOL_FP64 if prec == P.LPMX_FP64 else TOL_MIXED


def assert_rel(label, a, ref, tol):
    """max|a - ref| <= tol * max|ref|, with the measured ratio printed first."""
    a, ref = np.asarray(a, np.float64), np.asarray(ref, np.float64)
    scale = float(np.abs(ref).max())
    err = float(np.abs(a - ref).max()) / scale if scale > 0 else float(np.abs(a - ref).max())
    print(f"{label}: relative error {err:.3e} (bound {tol:.0e})", flush=True)
    assert err <= tol, (label, err, tol)
