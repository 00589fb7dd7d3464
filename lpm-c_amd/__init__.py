"""Python harness over liblpm.so (the drop-in lpm.h API and the lpmx device layer).

Host code of the product is C; this module is the thin test/benchmark harness: it
loads the in-tree library, reads and writes the reference's globals through ctypes,
builds the benchmark configurations through the reference API (createCuboid,
initMatrices, searchNormalNeighbor, ...), and wraps the lpmx entry points.  It never
computes anything itself; without the built library or a gfx950 device the calls
raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblpm.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

LPMX_FP64, LPMX_MIXED = 0, 1
FIELDS = """XYZ XYZ_INITIAL XYZ_TEMP NEIGHBORS NSIGN NB NB_INITIAL NB_CONN TYPE PL_FLAG CONN K_POINTER
DISTANCE_INITIAL CSX_INITIAL CSY_INITIAL CSZ_INITIAL KN TV DL DL_AVE DDL DDLP F F_TEMP CSX CSY CSZ BOND_STRESS
DISTANCE DAMAGE_BROKEN DAMAGE_W DLP0 DLP1 DLP2 DAMAGE_D0 DAMAGE_D1 DL_TOTAL TDL_TOTAL DDL_TOTAL TDDL_TOTAL
J2_BETA0 J2_BETA1 J2_BETA2 STRESS J2_ALPHA0 J2_ALPHA1 J2_ALPHA2 J2_DLAMBDA J2_STRESSEQ J2_STRESSM
J2_TRIAXIALITY DAMAGE_NONLOCAL0 DAMAGE_NONLOCAL1 DAMAGE_VISUAL SIGMAY PIN PEX PEX_TEMP RESIDUAL DISP
DISPBC_INDEX FIX_INDEX CP_GY0 CP_GY1 CP_GY2 CP_A_SINGLE0 CP_A_SINGLE1 CP_A_SINGLE2 CP_A0 CP_A1 CP_A2 CP_DA
CP_DGY CP_DA_SINGLE CP_RSS CP_JACT CP_CAB J2_BETA_EQ0 J2_BETA_EQ1 J2_BETA_EQ2 DAMAGE_LOCAL0 DAMAGE_LOCAL1""".split()
F = {name: k for k, name in enumerate(FIELDS)}
INT_FIELDS = {"NEIGHBORS", "NSIGN", "NB", "NB_INITIAL", "NB_CONN", "TYPE", "PL_FLAG", "CONN", "K_POINTER",
              "DISPBC_INDEX", "FIX_INDEX", "CP_JACT"}


class LpmxError(RuntimeError):
    pass


class ModelDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("lattice", C.c_int), ("n", C.c_int), ("nneighbors", C.c_int),
                ("nneighbors1", C.c_int), ("nneighbors2", C.c_int), ("nneighbors_AFEM", C.c_int),
                ("radius", C.c_double), ("neighbor1_cutoff", C.c_double), ("neighbor2_cutoff", C.c_double),
                ("particle_volume", C.c_double)]


class Params(C.Structure):
    _fields_ = [("plmode", C.c_int), ("J2_H", C.c_double), ("J2_xi", C.c_double), ("damage_L", C.c_double),
                ("damage_threshold", C.c_double), ("damagec_A", C.c_double), ("critical_bstrain", C.c_double),
                ("nbreak", C.c_int), ("J2_C", C.c_double)]


class LatticeSpec(C.Structure):
    """lpmx_lattice_spec: createCuboid's candidate grid."""
    _fields_ = [("lattice", C.c_int), ("dim", C.c_int), ("pfr", C.c_int), ("rows", C.c_int), ("layers", C.c_int),
                ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double), ("radius", C.c_double),
                ("bt", C.c_double * 6), ("box", C.c_double * 6), ("R", C.c_double * 9)]


class DispBC(C.Structure):
    _fields_ = [("type", C.c_int), ("flag", C.c_char), ("step", C.c_double)]


class ForceBC(C.Structure):
    _fields_ = [("type", C.c_int), ("step", C.c_double * 3)]


class Solver(C.Structure):
    _fields_ = [("precision", C.c_int), ("rtol2", C.c_double), ("atol2", C.c_double), ("maxit", C.c_int),
                ("newton_tol", C.c_double), ("newton_maxit", C.c_int), ("inner_rtol2", C.c_double),
                ("ir_maxit", C.c_int), ("accel_depth", C.c_int), ("precond", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("reaction", C.c_double), ("disp", C.c_double), ("newton_iters", C.c_int), ("cg_iters", C.c_int),
                ("resolves", C.c_int), ("broken", C.c_int), ("ir_passes", C.c_int), ("status", C.c_int),
                ("res_norm", C.c_double), ("tol_mult", C.c_double), ("t_tangent", C.c_double),
                ("t_solve", C.c_double), ("t_force", C.c_double), ("t_other", C.c_double), ("t_total", C.c_double),
                ("accel_mixed", C.c_int), ("accel_reverts", C.c_int), ("accel_restarts", C.c_int)]


class AdjointOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int), ("depth", C.c_int), ("precision", C.c_int),
                ("rtol2", C.c_double), ("cg_maxit", C.c_int)]


class AdjointStats(C.Structure):
    _fields_ = [("t_forward", C.c_double), ("t_recompute", C.c_double), ("t_backward", C.c_double),
                ("checkpoints", C.c_int), ("recomputed_steps", C.c_int), ("adj_iters", C.c_int),
                ("adj_cg_iters", C.c_int), ("adj_iters_max", C.c_int), ("max_rel_res", C.c_double),
                ("max_rel_free", C.c_double), ("ckpt_bytes", C.c_double)]


# reference API entry points declared in include/*.h (argument types for ctypes)
_API = {
    "createCuboid": [], "initMatrices": [], "searchNormalNeighbor": [], "searchAFEMNeighbor": [],
    "calcKnTv": [], "computedL": [], "computeStress": [], "updateRR": [], "updateCrack": [],
    "calcStiffness3DFiniteDifference": [C.c_int], "calcStiffness2DFiniteDifference": [C.c_int],
    "setDispBC_stiffnessUpdate3D": [], "setDispBC_stiffnessUpdate2D": [], "solverCG": [],
    "switchStateV": [C.c_int], "computeBondForceGeneral": [C.c_int, C.c_int],
    "setTypeRect": [C.c_double] * 6 + [C.c_int], "setTypeFullNeighbor": [C.c_int],
    "moveParticle": [_dp], "copyDouble2D": [C.c_void_p, C.c_void_p, C.c_int, C.c_int],
    "allocInt1D": [C.c_int, C.c_int], "allocDouble1D": [C.c_int, C.c_double],
    "allocDouble2D": [C.c_int, C.c_int, C.c_double], "cblas_dnrm2": [C.c_int, _dp, C.c_int],
    "updateDamageGeneral": [C.c_char_p, C.c_int, C.c_int], "slipSysDefine3D": [],
}
_RES = {"searchAFEMNeighbor": C.c_int, "allocInt1D": _ip, "allocDouble1D": _dp,
        "allocDouble2D": C.POINTER(_dp), "cblas_dnrm2": C.c_double, "updateDamageGeneral": C.c_int}

_lib = None


def lib():
    """The loaded liblpm.so with argument types set; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LpmxError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    for name, args in _API.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = _RES.get(name)
    L.lpmx_last_error.restype = C.c_char_p
    sig = {
        "lpmx_model_create": [C.POINTER(ModelDesc), _dp], "lpmx_model_destroy": [],
        "lpmx_set_params": [C.POINTER(Params)], "lpmx_neighbor_build": [],
        "lpmx_connectivity_build": [C.POINTER(C.c_int64)],
        "lpmx_set_materials": [C.c_int, _ip, _dp, _dp, _dp], "lpmx_set_bond_critical": [_dp],
        "lpmx_get": [C.c_int, C.c_void_p, C.c_int64], "lpmx_put": [C.c_int, C.c_void_p, C.c_int64],
        "lpmx_tangent": [], "lpmx_export_upper_csr": [_ip, _ip, _dp], "lpmx_import_upper_csr": [_dp],
        "lpmx_bc_eliminate": [], "lpmx_solve": [C.c_int, C.c_double, C.c_double, C.c_int, _ip, _dp],
        "lpmx_computedL": [], "lpmx_bond_force": [C.c_int], "lpmx_bond_force_range": [C.c_int, C.c_int, C.c_int],
        "lpmx_compute_stress": [], "lpmx_switch_state": [C.c_int],
        "lpmx_update_rr": [_dp, _dp, _ip], "lpmx_get_reaction_force": [_dp, C.c_int],
        "lpmx_set_disp_bc": [C.c_int, C.POINTER(DispBC)], "lpmx_set_force_bc": [C.c_int, C.POINTER(ForceBC)],
        "lpmx_damage": [C.c_int, _ip], "lpmx_get_broken": [_ip, _ip, C.c_int, _ip], "lpmx_update_crack": [],
        "lpmx_reaction": [C.c_char, C.c_int, _dp, _dp],
        "lpmx_run_steps": [C.POINTER(Solver), C.c_int, _ip, C.POINTER(DispBC), _ip, C.POINTER(ForceBC), C.c_int,
                           C.c_char, C.c_int, C.POINTER(StepStats)],
        "lpmx_bench_spmv": [C.c_int, C.c_int, _dp, _dp],
        "lpmx_spmv_bytes": [C.c_int, _dp],
        "lpmx_compute_cab": [],
        "lpmx_lattice_points": [C.POINTER(LatticeSpec), _dp, C.c_int, _ip], "lpmx_compute_strain": [_dp],
        "lpmx_vjp": [_dp] * 11,
        "lpmx_history_forward": [C.POINTER(Solver), C.c_int, C.c_int, _ip, C.POINTER(DispBC), _ip,
                                 C.POINTER(ForceBC), C.c_int, C.c_char, C.c_int, _dp, C.POINTER(StepStats)],
        "lpmx_history_backward": [C.POINTER(AdjointOpts), _dp, _dp, _dp, _dp, C.POINTER(AdjointStats)],
        "lpmx_cp_failures": [_ip],
        "lpmx_matrix_info": [C.POINTER(C.c_int64), C.POINTER(C.c_int64), _ip],
        "lpmx_spmv": [C.c_int, _dp, _dp],
        "lpmx_set_preconditioner": [C.c_int],
        "lpmx_profile_amg_read": [C.POINTER(C.c_double), C.POINTER(C.c_int64)],
        "lpmx_amg_smoother_bytes": [C.c_int, C.POINTER(C.c_double)],
        "lpmx_amg_info": [_ip, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
        "lpmx_amg_export": [C.c_int, C.c_int, C.c_int64, _ip, _ip, _dp, _ip],
        "lpmx_profile_force_read": [C.POINTER(C.c_double), C.POINTER(C.c_int64)],
        "lpmx_j2_particle_bytes": [C.POINTER(C.c_double)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.lpmx_history_free.argtypes = []
    L.lpmx_history_free.restype = None
    L.lpmx_profile_spmv.argtypes = [C.c_int]
    L.lpmx_profile_spmv.restype = None
    L.lpmx_profile_spmv_read.argtypes = [_dp, C.POINTER(C.c_int64)]
    L.lpmx_profile_spmv_read.restype = None
    L.lpmx_profile_amg_read.restype = None
    L.lpmx_profile_force.argtypes = [C.c_int]
    L.lpmx_profile_force.restype = None
    L.lpm_ext_set_modulus.argtypes = [_dp]
    L.lpm_ext_set_bond_critical.argtypes = [_dp]
    L.lpm_ext_set_solver.argtypes = [C.POINTER(Solver)]
    _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        raise LpmxError(f"{what}: status {rc}: {lib().lpmx_last_error().decode()}")


def ptr(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


# ---------------------------------------------------------------- globals --------
def gint(name):
    return C.c_int.in_dll(lib(), name)


def gdouble(name):
    return C.c_double.in_dll(lib(), name)


def gptr(name, depth, ctype=C.c_double):
    t = ctype
    for _ in range(depth):
        t = C.POINTER(t)
    return t.in_dll(lib(), name)


def read1d(name, count, dtype=np.float64):
    ct = C.c_double if dtype == np.float64 else C.c_int
    p = gptr(name, 1, ct)
    return np.ctypeslib.as_array(p, shape=(count,)).copy()


def read2d(name, rows, cols, dtype=np.float64):
    ct = C.c_double if dtype == np.float64 else C.c_int
    p = gptr(name, 2, ct)
    out = np.empty((rows, cols), dtype)
    for i in range(rows):
        out[i] = np.ctypeslib.as_array(p[i], shape=(cols,))
    return out


def write1d(name, values):
    arr = np.asarray(values)
    ct = C.c_double if arr.dtype == np.float64 else C.c_int
    p = gptr(name, 1, ct)
    for k, v in enumerate(arr):
        p[k] = v


def field(name, count=None):
    """A device field in the reference layout (lpmx_get)."""
    dt = np.int32 if name in INT_FIELDS else np.float64
    if count is None:
        raise ValueError("count required")
    out = np.zeros(count, dt)
    check(lib().lpmx_get(F[name], out.ctypes.data_as(C.c_void_p), count), f"lpmx_get {name}")
    return out


# ----------------------------------------------------- model construction --------
def elastic_constants(E, nu, plane_stress=False):
    """C11, C12, C44 as the reference driver forms them (src/lpmc_project.c:188-194)."""
    if plane_stress:
        return (E / (1.0 - nu) / (1.0 + nu), E * nu / (1.0 - nu) / (1.0 + nu), E / 2.0 / (1.0 + nu))
    return (E * (1.0 - nu) / (1.0 + nu) / (1.0 - 2.0 * nu), E * nu / (1.0 + nu) / (1.0 - 2.0 * nu),
            E / 2.0 / (1.0 + nu))


def make_lattice(lattice, radius, box, angles=(0.0, 0.0, 0.0), eulerflag=0):
    """createCuboid + initMatrices + xyz_initial = xyz through the reference API."""
    L = lib()
    gint("lattice").value = lattice
    gint("dim").value = 2 if lattice < 2 else 3
    gdouble("radius").value = radius
    gint("eulerflag").value = eulerflag
    for k, a in enumerate(("angle1", "angle2", "angle3")):
        gdouble(a).value = angles[k]
    b = (C.c_double * 6).in_dll(L, "box")
    for k in range(6):
        b[k] = box[k]
    gdouble("box_x").value = box[1] - box[0]
    gdouble("box_y").value = box[3] - box[2]
    gdouble("box_z").value = box[5] - box[4]
    L.createCuboid()
    mv = np.zeros(3)
    L.moveParticle(ptr(mv))
    L.initMatrices()
    n = gint("nparticle").value
    L.copyDouble2D(gptr("xyz_initial", 2), gptr("xyz", 2), n, 3)
    return n


def positions():
    n = gint("nparticle").value
    return read2d("xyz", n, 3)


def set_gptr(name, value, ctype, depth):
    """Point the global `name` (a depth-level pointer to ctype) at `value`."""
    t = ctype
    for _ in range(depth):
        t = C.POINTER(t)
    g = t.in_dll(lib(), name)
    C.cast(C.addressof(g), C.POINTER(t))[0] = value


def set_types(types, ntype):
    """type[] (allocated through allocInt1D, as the reference driver does) and ntype."""
    n = gint("nparticle").value
    p = lib().allocInt1D(n, 0)
    tv = np.ascontiguousarray(types, np.int32)
    C.memmove(p, tv.ctypes.data, 4 * n)
    set_gptr("type", p, C.c_int, 1)
    gint("ntype").value = ntype


def set_materials(Ce_rows, sigmay):
    """Ce[ntype][3] and sigmay[n] through the reference allocators."""
    L = lib()
    nt = len(Ce_rows)
    ce = L.allocDouble2D(nt, 3, 0.0)
    for t, row in enumerate(Ce_rows):
        for c in range(3):
            ce[t][c] = row[c]
    set_gptr("Ce", ce, C.c_double, 2)
    n = gint("nparticle").value
    sy = L.allocDouble1D(n, 0.0)
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sigmay, np.float64), (n,)))
    C.memmove(sy, sv.ctypes.data, 8 * n)
    set_gptr("sigmay", sy, C.c_double, 1)


def set_params(plmode=0, J2_H=0.0, J2_xi=0.0, damage_L=0.5, damage_threshold=0.9, damagec_A=0.0,
               critical_bstrain=1e-2, nbreak=20):
    gint("plmode").value = plmode
    gdouble("J2_H").value = J2_H
    gdouble("J2_xi").value = J2_xi
    gdouble("damage_L").value = damage_L
    gdouble("damage_threshold").value = damage_threshold
    gdouble("damagec_A").value = damagec_A
    gdouble("critical_bstrain").value = critical_bstrain
    gint("nbreak").value = nbreak


def solver(precision=LPMX_FP64, rtol2=1e-8, atol2=1e-12, maxit=None, newton_tol=1e-4, newton_maxit=100,
           inner_rtol2=1e-2, ir_maxit=100, accel_depth=0, precond=0):
    n = gint("nparticle").value * gint("dim").value
    return Solver(precision, rtol2, atol2, maxit if maxit is not None else n, newton_tol, newton_maxit,
                  inner_rtol2, ir_maxit, accel_depth, precond)


def bc_tables(steps):
    """steps: list of (disp_entries, force_entries); disp (type, 'x'|'y'|'z', step),
    force (type, (fx, fy, fz))."""
    maxbc = max(1, max(max(len(d), len(f)) for d, f in steps))
    ns = len(steps)
    dbc = (DispBC * (ns * maxbc))()
    fbc = (ForceBC * (ns * maxbc))()
    nbd = np.zeros(ns, np.int32)
    nbf = np.zeros(ns, np.int32)
    for s, (dl, fl) in enumerate(steps):
        nbd[s], nbf[s] = len(dl), len(fl)
        for k, (t, flag, st) in enumerate(dl):
            dbc[s * maxbc + k] = DispBC(t, flag.encode(), st)
        for k, (t, st) in enumerate(fl):
            fbc[s * maxbc + k].type = t
            for c in range(3):
                fbc[s * maxbc + k].step[c] = st[c]
    return maxbc, nbd, dbc, nbf, fbc


def run_steps(steps, solv, out_axis="z", out_type=1):
    """lpmx_run_steps over the BC tables; returns the per-step statistics."""
    maxbc, nbd, dbc, nbf, fbc = bc_tables(steps)
    out = (StepStats * len(steps))()
    rc = lib().lpmx_run_steps(C.byref(solv), len(steps), ptr(nbd), dbc, ptr(nbf), fbc, maxbc, out_axis.encode(),
                              out_type, out)
    return rc, list(out)


# ------------------------------------------------------------------ adjoint -----
def history_forward(steps, solv, ckpt_every=5, out_axis="z", out_type=1):
    """lpmx_history_forward: the load steps with damage off, recorded for the reverse
    sweep (a checkpoint every ckpt_every steps).  Returns (rc, reactions, stats)."""
    maxbc, nbd, dbc, nbf, fbc = bc_tables(steps)
    reac = np.zeros(len(steps))
    out = (StepStats * len(steps))()
    rc = lib().lpmx_history_forward(C.byref(solv), ckpt_every, len(steps), ptr(nbd), dbc, ptr(nbf), fbc, maxbc,
                                    out_axis.encode(), out_type, ptr(reac), out)
    return rc, reac, list(out)


def adjoint_opts(tol=1e-10, maxit=400, depth=8, precision=LPMX_FP64, rtol2=1e-8, cg_maxit=None, n=None):
    o = AdjointOpts()
    o.tol, o.maxit, o.depth, o.precision, o.rtol2 = tol, maxit, depth, precision, rtol2
    o.cg_maxit = cg_maxit if cg_maxit is not None else 3 * (n if n is not None else gint("nparticle").value)
    return o


def history_backward(opts, rbar):
    """lpmx_history_backward for dJ/dR = rbar.  Returns (rc, g_emod [n], g_sigmay [n],
    g_H, stats)."""
    n = gint("nparticle").value
    rb = np.ascontiguousarray(rbar, np.float64)
    ge, gy, gh = np.zeros(n), np.zeros(n), np.zeros(1)
    st = AdjointStats()
    rc = lib().lpmx_history_backward(C.byref(opts), ptr(rb), ptr(ge), ptr(gy), ptr(gh), C.byref(st))
    return rc, ge, gy, gh[0], st


_rh_class = None


def reaction_history():
    """A torch.autograd.Function R = f(E, emod, sigmay, H): the reaction curve of an SC
    uniaxial J2 history on the device, differentiated by the discrete adjoint.

        R = reaction_history().apply(E, emod, sigmay, H, spec)

    E and H are 0-d float64 tensors, emod and sigmay [n] float64 tensors; spec is a
    dict with L, nu, steps (BC tables), solver, ckpt_every and opts (adjoint options).
    Each call builds the model on the device; backward runs the reverse sweep of the
    history that forward recorded, so a backward belongs to the latest forward."""
    global _rh_class
    if _rh_class is not None:
        return _rh_class
    import torch

    class ReactionHistory(torch.autograd.Function):
        @staticmethod
        def forward(ctx, E, emod, sigmay, H, spec):
            Ev, Hv = float(E), float(H)
            em = emod.detach().cpu().double().numpy().copy()
            sy = sigmay.detach().cpu().double().numpy().copy()
            build_sc_model(spec["L"], E=Ev, nu=spec.get("nu", 0.3), sigmay=sy, H=Hv, emod=em)
            rc, reac, _ = history_forward(spec["steps"], spec["solver"], spec.get("ckpt_every", 5),
                                          spec.get("axis", "z"), spec.get("out_type", 1))
            check(rc, "lpmx_history_forward")
            ctx.spec, ctx.E, ctx.em = spec, Ev, em
            return torch.from_numpy(reac)

        @staticmethod
        def backward(ctx, gR):
            opts = ctx.spec.get("opts") or adjoint_opts()
            rc, ge, gy, gh, _ = history_backward(opts, gR.detach().cpu().double().numpy())
            check(rc, "lpmx_history_backward")
            gE = float(np.dot(ctx.em, ge)) / ctx.E
            return (torch.tensor(gE, dtype=torch.float64), torch.from_numpy(ge), torch.from_numpy(gy),
                    torch.tensor(gh, dtype=torch.float64), None)

    _rh_class = ReactionHistory
    return _rh_class


# ------------------------------------------------------------ configurations -----
SC_BOX = {8000: 19.035, 216000: 59.3, 1000000: 99.195}


def sc_box_for(k):
    """Box edge L of an SC lattice (radius 0.5) with k particles per edge.  x starts
    at 0, y and z at -a + integer (a = sqrt(3) L, the box diagonal), so besides
    k-1 <= L < k the first y/z value ceil(a) - a must not exceed L - (k-1)."""
    for m in range(1, 1000):
        L = k - 1 + m / 1000.0
        a = np.sqrt(3.0) * L
        f = np.ceil(a) - a
        if 1e-6 < f <= L - (k - 1) - 1e-9:
            return L
    raise ValueError(k)


def sc_types(x, radius=0.5):
    """Types of the uniaxial SC configurations: 1 top face, 2 bottom face, 3 the
    bottom corner at minimum x and y (fully fixed), 4 the bottom corner at maximum x
    and minimum y (held in y, which removes the rigid rotation about z that a face
    fixed in z and one fixed corner leave free), 0 elsewhere."""
    z = x[:, 2]
    t = np.zeros(len(x), np.int32)
    t[z >= z.max() - 0.6 * radius] = 1
    t[z <= z.min() + 0.6 * radius] = 2
    bot = np.where(t == 2)[0]
    a = bot[np.lexsort((x[bot, 0], x[bot, 1]))[0]]
    b = bot[np.lexsort((-x[bot, 0], x[bot, 1]))[0]]
    t[a] = 3
    t[b] = 4
    return t


def sc_uniaxial_steps(height, strains):
    """Displacement steps: bottom (type 2) fixed in z, corner 3 fixed in x, y, z,
    corner 4 fixed in y and z, top (type 1) moved in z by the strain increments
    times the height."""
    steps = []
    prev = 0.0
    for e in strains:
        du = (e - prev) * height
        prev = e
        steps.append(([(2, "z", 0.0), (3, "x", 0.0), (3, "y", 0.0), (3, "z", 0.0), (4, "y", 0.0), (4, "z", 0.0),
                       (1, "z", du)], []))
    return steps


def cyclic_strains(amplitude, steps_per_half, cycles):
    """Total strain after each load step of a symmetric triangular strain cycle
    repeated `cycles` times: 0 -> +a -> -a -> 0, each half-cycle (a peak to the
    opposite peak) in `steps_per_half` equal increments 2a/steps_per_half.  Strains
    are integer multiples of the increment, so the path returns to 0 exactly."""
    q = steps_per_half // 2
    de = 2.0 * amplitude / steps_per_half
    count, out = 0, []
    for _ in range(cycles):
        for k in range(2 * steps_per_half):
            count += 1 if (k < q or k >= q + steps_per_half) else -1
            out.append(count * de)
    return np.array(out)


def uniform_modulus(n, seed, amp=0.05):
    """Per-particle modulus factor 1 + amp*U(-1, 1); U from the splitmix64 stream of
    `seed` (element i is the (i+1)-th output), 53-bit mantissa."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + gamma * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 1.0 + amp * (2.0 * u - 1.0)


def build_sc_model(L, E=200e3, nu=0.3, sigmay=200.0, H=2000.0, emod=None, damage=True):
    """SC elastoplastic model of the uniaxial configurations through the reference
    API, on the device: lattice, lists, pattern, types, materials, bond constants."""
    Lb = lib()
    n = make_lattice(2, 0.5, [0.0, L, 0.0, L, 0.0, L])
    Lb.searchNormalNeighbor()
    Lb.searchAFEMNeighbor()
    x = field("XYZ", 3 * n).reshape(n, 3)  # the device model's copy of xyz, made by the search
    t = sc_types(x)
    set_types(t, 5)
    c = elastic_constants(E, nu)
    set_materials([c] * 5, sigmay)
    set_params(plmode=0, J2_H=H, J2_xi=0.0, damage_L=0.5, damage_threshold=0.9, damagec_A=0.0,
               critical_bstrain=1e-2, nbreak=20)
    # the library copies the field; a new lattice has none until it is set
    check(Lb.lpm_ext_set_modulus(ptr(np.ascontiguousarray(emod, np.float64)) if emod is not None else None),
          "modulus field")
    Lb.calcKnTv()
    Lb.computedL()
    return n, x, t, c




def hex_types(x, radius=0.5):
    """Types of the 2D shear configurations: 1 top edge, 2 bottom edge, 0 elsewhere."""
    y = x[:, 1]
    t = np.zeros(len(x), np.int32)
    t[y >= y.max() - 0.6 * radius] = 1
    t[y <= y.min() + 0.6 * radius] = 2
    return t


def hex_shear_steps(height, strains):
    """Displacement steps: bottom (type 2) held in x and y, top (type 1) held in y and
    moved in x by the shear-strain increments times the height."""
    steps = []
    prev = 0.0
    for g in strains:
        du = (g - prev) * height
        prev = g
        steps.append(([(2, "x", 0.0), (2, "y", 0.0), (1, "y", 0.0), (1, "x", du)], []))
    return steps


def _splitmix64(z):
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def weibull_bond_critical(nbr, seed, shape=10.0, scale=1e-3):
    """Per-bond critical stretch [n][nb] from Weibull(shape, scale), one draw per
    undirected bond: the uniform of bond {i, j} is the splitmix64 hash of
    (seed, min(i, j) * n + max(i, j)), so (i, j) and (j, i) share it.  Empty slots
    get the scale."""
    nbr = np.asarray(nbr)
    n = nbr.shape[0]
    i = np.repeat(np.arange(n, dtype=np.int64)[:, None], nbr.shape[1], axis=1)
    j = nbr.astype(np.int64)
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    key = (lo * n + hi).astype(np.uint64)
    with np.errstate(over="ignore"):
        h = _splitmix64(_splitmix64(np.uint64(seed)) ^ key)
    u = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    crit = scale * (-np.log1p(-u)) ** (1.0 / shape)
    return np.where(nbr >= 0, crit, scale)


def build_hex_model(box, E=70e3, nu=0.25, critical=1e-3, nbreak=20, crit=None):
    """2D hexagonal elastic-brittle model (plmode 6) of the shear configurations, plane
    strain constants, through the reference API on the device."""
    Lb = lib()
    n = make_lattice(1, 0.5, box)
    Lb.searchNormalNeighbor()
    Lb.searchAFEMNeighbor()
    x = field("XYZ", 3 * n).reshape(n, 3)  # the device model's copy of xyz, made by the search
    t = hex_types(x)
    set_types(t, 3)
    c = elastic_constants(E, nu)
    set_materials([c] * 3, 1e30)
    set_params(plmode=6, critical_bstrain=critical, nbreak=nbreak)
    # the library copies the values (a new lattice drops them)
    check(Lb.lpm_ext_set_bond_critical(ptr(np.ascontiguousarray(crit, np.float64)) if crit is not None else None),
          "critical stretch")
    Lb.calcKnTv()
    Lb.computedL()
    return n, x, t, c


def set_cp_params(tau0, taus, h0, q, eta, p, dtime, maxloop):
    """The reference's crystal-plasticity globals (cp_tau0[0], cp_taus[0], cp_h0, cp_q,
    cp_eta, cp_p, dtime, cp_maxloop)."""
    L = lib()
    (C.c_double * 3).in_dll(L, "cp_tau0")[0] = tau0
    (C.c_double * 3).in_dll(L, "cp_taus")[0] = taus
    for name, v in (("cp_h0", h0), ("cp_q", q), ("cp_eta", eta), ("cp_p", p), ("dtime", dtime),
                    ("cp_maxloop", float(maxloop))):
        gdouble(name).value = v


def build_fcc_cp_model(L, cp, E=70e3, nu=0.3):
    """FCC crystal-plasticity model (plmode 1) through the reference API on the
    device: lattice, lists, uniaxial types, materials, bond constants, slip systems
    (slipSysDefine3D) and the slip stiffness (computeCab)."""
    Lb = lib()
    n = make_lattice(3, 0.5, [0.0, L, 0.0, L, 0.0, L])
    Lb.searchNormalNeighbor()
    Lb.searchAFEMNeighbor()
    x = field("XYZ", 3 * n).reshape(n, 3)  # the device model's copy of xyz, made by the search
    t = sc_types(x)
    set_types(t, 5)
    c = elastic_constants(E, nu)
    set_materials([c] * 5, 1e30)
    set_params(plmode=1)
    set_cp_params(**cp)
    Lb.calcKnTv()
    Lb.computedL()
    Lb.slipSysDefine3D()
    Lb.computeCab()
    return n, x, t, c
