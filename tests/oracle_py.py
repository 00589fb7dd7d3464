"""ctypes harness for the CPU oracle (oracle/liboracle.so).

Test infrastructure only: tests, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg use it as the checker.  Arrays are numpy-owned; the C side
only reads and writes them through the ``orc_model`` struct.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# LPM_ORACLE_SO names another build of the oracle (a copy with a planted mistake,
# to check that the pins catch it)
ORACLE_SO = os.environ.get("LPM_ORACLE_SO") or os.path.join(ROOT, "oracle", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcModel(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("lattice", C.c_int), ("n", C.c_int), ("nbmax", C.c_int),
        ("nb1max", C.c_int), ("nb2max", C.c_int), ("connmax", C.c_int), ("ntype", C.c_int),
        ("radius", C.c_double), ("cut1", C.c_double), ("cut2", C.c_double), ("volume", C.c_double),
        ("plmode", C.c_int), ("J2_H", C.c_double), ("J2_xi", C.c_double),
        ("damage_L", C.c_double), ("damage_threshold", C.c_double), ("damagec_A", C.c_double),
        ("critical_bstrain", C.c_double), ("nbreak", C.c_int),
        ("type", _ip), ("Ce", _dp), ("sigmay", _dp), ("emod", _dp), ("crit", _dp),
        ("x", _dp), ("x0", _dp), ("xt", _dp),
        ("nbr", _ip), ("nsign", _ip), ("nbr1", _ip), ("nbr2", _ip), ("nb", _ip), ("nb_initial", _ip),
        ("L0", _dp), ("cs0", _dp), ("conn", _ip), ("nb_conn", _ip), ("kptr", _ip),
        ("Kn", _dp), ("Tv", _dp), ("dL", _dp), ("dL_ave", _dp), ("ddL", _dp), ("ddLp", _dp),
        ("F", _dp), ("Ftemp", _dp), ("bond_stress", _dp), ("cs", _dp), ("dLp", _dp), ("dmgD", _dp),
        ("dmg_broken", _dp), ("dmg_w", _dp),
        ("dLtot", _dp), ("TdLtot", _dp), ("ddLtot", _dp), ("TddLtot", _dp),
        ("beta", _dp), ("alpha", _dp), ("beta_eq", _dp),
        ("dlambda", _dp), ("stress", _dp), ("seq", _dp), ("sm", _dp), ("triax", _dp),
        ("pl_flag", _ip), ("dnl", _dp), ("dloc", _dp), ("damage_visual", _dp),
        ("Pin", _dp), ("Pex", _dp), ("Pex_temp", _dp), ("residual", _dp), ("disp", _dp),
        ("dispBC", _ip), ("fixidx", _ip), ("reaction", _dp), ("nreaction", C.c_int),
        ("nK", C.c_int64), ("K", _dp), ("IK", _ip), ("JK", _ip),
        ("dist", _dp), ("nslip", C.c_int), ("schmid", _dp),
        ("cp_tau0", C.c_double), ("cp_taus", C.c_double), ("cp_h0", C.c_double), ("cp_q", C.c_double),
        ("cp_eta", C.c_double), ("cp_p", C.c_double), ("dtime", C.c_double), ("cp_maxloop", C.c_int),
        ("cp_gy", _dp), ("cp_A_single", _dp), ("cp_A", _dp), ("cp_dA", _dp),
        ("cp_dgy", _dp), ("cp_dA_single", _dp), ("cp_RSS", _dp), ("cp_Jact", _ip), ("cp_Cab", _dp),
        ("state_v", _ip), ("cp_fail", C.c_int), ("J2_C", C.c_double), ("j2nl_jacobi", C.c_int),
    ]


class OrcLattice(C.Structure):
    _fields_ = [("dim", C.c_int), ("lattice", C.c_int), ("nbmax", C.c_int), ("nb1max", C.c_int),
                ("nb2max", C.c_int), ("nafem", C.c_int), ("radius", C.c_double), ("cut1", C.c_double),
                ("cut2", C.c_double), ("volume", C.c_double)]


class OrcDispBC(C.Structure):
    _fields_ = [("type", C.c_int), ("flag", C.c_char), ("step", C.c_double)]


class OrcForceBC(C.Structure):
    _fields_ = [("type", C.c_int), ("step", C.c_double * 3)]


class OrcSolver(C.Structure):
    _fields_ = [("rtol2", C.c_double), ("atol2", C.c_double), ("maxit", C.c_int),
                ("newton_tol", C.c_double), ("newton_maxit", C.c_int), ("tangent", C.c_int),
                ("fixed_diag", C.c_int), ("nthreads", C.c_int), ("accel", C.c_int),
                ("skip_damage", C.c_int)]


class OrcStepOut(C.Structure):
    _fields_ = [("reaction", C.c_double), ("disp", C.c_double), ("newton_iters", C.c_int),
                ("cg_iters", C.c_int), ("resolves", C.c_int), ("broken", C.c_int),
                ("res_norm", C.c_double), ("tol_mult", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(ORACLE_SO)
        P = C.POINTER(OrcModel)
        L.orc_lattice_params.argtypes = [C.c_int, C.c_double, C.c_double, C.POINTER(OrcLattice)]
        L.orc_create_lattice.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_double, C.c_double,
                                         C.c_double, _dp, C.c_int]
        L.orc_create_lattice.restype = C.c_int
        for f in ("orc_neighbor_search", "orc_neighbor_search_cells", "orc_damage_pwise_nonlocal"):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = C.c_int
        L.orc_afem_connectivity.argtypes = [P]
        L.orc_afem_connectivity.restype = C.c_int64
        for f in ("orc_kntv", "orc_computedL", "orc_compute_stress", "orc_tangent_fd",
                  "orc_tangent_analytic", "orc_update_rr", "orc_bc_stiffness_update", "orc_update_crack",
                  "orc_compute_cab"):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = None
        L.orc_eval_j2.argtypes = [P]
        L.orc_eval_j2.restype = None
        L.orc_compute_strain.argtypes = [P, _dp]
        L.orc_compute_strain.restype = None
        L.orc_vjp_j2.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_vjp_j2.restype = None
        L.orc_adjoint_run.argtypes = [P, C.POINTER(OrcSolver), C.c_int, _ip, C.POINTER(OrcDispBC), _ip,
                                      C.POINTER(OrcForceBC), C.c_int, C.c_char, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_adjoint_run.restype = C.c_int
        L.orc_slip_systems_fcc.argtypes = [_dp, _dp]
        L.orc_slip_systems_fcc.restype = C.c_int
        L.orc_cp_rss.argtypes = [P, C.c_int, _dp, _dp]
        L.orc_cp_rss.restype = None
        L.orc_lu_solve.argtypes = [C.c_int, _dp, _dp]
        L.orc_lu_solve.restype = C.c_int
        for f in ("orc_bond_force_elastic", "orc_bond_force_j2", "orc_bond_force_incremental",
                  "orc_switch_state", "orc_bond_force_general", "orc_bond_force_cp"):
            getattr(L, f).argtypes = [P, C.c_int]
            getattr(L, f).restype = None
        L.orc_bond_force_j2energy.argtypes = [P, C.c_int, C.c_int]
        L.orc_bond_force_j2energy.restype = None
        L.orc_bond_force_j2nonlinear.argtypes = [P, C.c_double, C.c_int]
        L.orc_bond_force_j2nonlinear.restype = None
        L.orc_internal_force.argtypes = [P, _dp, _dp]
        L.orc_set_disp_bc.argtypes = [P, C.c_int, C.POINTER(OrcDispBC)]
        L.orc_set_force_bc.argtypes = [P, C.c_int, C.POINTER(OrcForceBC)]
        L.orc_nrm2.argtypes = [C.c_int, _dp]
        L.orc_nrm2.restype = C.c_double
        L.orc_spmv_upper.argtypes = [P, _dp, _dp]
        L.orc_solve_cg.argtypes = [P, C.POINTER(OrcSolver), _dp]
        L.orc_solve_cg.restype = C.c_int
        L.orc_damage_bwise_nonlocal.argtypes = [P]
        L.orc_damage_bwise_nonlocal.restype = C.c_int
        L.orc_damage_pwise_local.argtypes = [P, C.c_int]
        L.orc_damage_pwise_local.restype = C.c_int
        L.orc_damage_bwise_local.argtypes = [P, C.c_int]
        L.orc_damage_bwise_local.restype = C.c_int
        L.orc_damage_brittle.argtypes = [P, _ip, _ip, C.c_int]
        L.orc_damage_brittle.restype = C.c_int
        L.orc_shell_sort_select.argtypes = [C.c_int, _dp, C.c_int, _ip]
        L.orc_shell_sort_select.restype = C.c_int
        L.orc_reaction.argtypes = [P, C.c_char, C.c_int]
        L.orc_reaction.restype = C.c_double
        L.orc_mean_disp.argtypes = [P, C.c_char, C.c_int]
        L.orc_mean_disp.restype = C.c_double
        L.orc_run_steps.argtypes = [P, C.POINTER(OrcSolver), C.c_int, _ip, C.POINTER(OrcDispBC), _ip,
                                    C.POINTER(OrcForceBC), C.c_int, C.c_char, C.c_int, C.POINTER(OrcStepOut)]
        L.orc_run_steps.restype = C.c_int
        L.orc_set_output.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char, C.c_int, C.c_int]
        L.orc_set_output.restype = None
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def lattice_params(lattice, radius, box_z=1.0):
    o = OrcLattice()
    lib().orc_lattice_params(lattice, radius, box_z, C.byref(o))
    return o


def create_lattice(lattice, radius, box, eulerflag=0, angles=(0.0, 0.0, 0.0)):
    b = np.ascontiguousarray(box, dtype=np.float64)
    n = lib().orc_create_lattice(lattice, radius, ptr(b), eulerflag, *angles, None, 0)
    x = np.zeros((n, 3))
    lib().orc_create_lattice(lattice, radius, ptr(b), eulerflag, *angles, ptr(x), n)
    return x


def elastic_constants(E, nu, plane_stress=False):
    """C11, C12, C44 as the reference driver forms them (src/lpmc_project.c:188-194)."""
    if plane_stress:
        return (E / (1.0 - nu) / (1.0 + nu), E * nu / (1.0 - nu) / (1.0 + nu), E / 2.0 / (1.0 + nu))
    return (E * (1.0 - nu) / (1.0 + nu) / (1.0 - 2.0 * nu), E * nu / (1.0 + nu) / (1.0 - 2.0 * nu),
            E / 2.0 / (1.0 + nu))


class Model:
    """All state of one oracle model, numpy-backed."""

    def __init__(self, x, lattice, radius, types, Ce, box_z=1.0, plmode=0, sigmay=200.0,
                 J2_H=0.0, J2_xi=0.0, damage_L=0.5, damage_threshold=0.9, damagec_A=0.0,
                 critical_bstrain=1e-2, nbreak=20, emod=None, crit=None, cp=None, J2_C=0.0, j2nl_jacobi=0):
        lp = lattice_params(lattice, radius, box_z)
        n = x.shape[0]
        d = lp.dim
        nbm = lp.nbmax
        self.lp = lp
        self.n, self.dim, self.nbmax = n, d, nbm
        self.connmax = lp.nafem + 1
        Ce = np.ascontiguousarray(Ce, dtype=np.float64).reshape(-1, 3)
        f, i32 = np.float64, np.int32
        A = {}
        A["type"] = np.ascontiguousarray(types, dtype=i32)
        A["Ce"] = Ce.copy()
        A["sigmay"] = np.full(n, sigmay, f) if np.isscalar(sigmay) else np.ascontiguousarray(sigmay, f)
        A["emod"] = None if emod is None else np.ascontiguousarray(emod, f)
        A["crit"] = None if crit is None else np.ascontiguousarray(crit, f).reshape(n * nbm)
        A["x"] = np.ascontiguousarray(x, f).copy()
        A["x0"] = A["x"].copy()
        A["xt"] = np.zeros((n, 3), f)
        for k in ("nbr", "nsign"):
            A[k] = np.full((n, nbm), -1, i32)
        A["nbr1"] = np.full((n, lp.nb1max), -1, i32)
        A["nbr2"] = np.full((n, lp.nb2max), -1, i32)
        A["nb"] = np.full(n, -1, i32)
        A["nb_initial"] = np.full(n, -1, i32)
        A["L0"] = np.zeros((n, nbm), f)
        A["cs0"] = np.zeros((n, nbm, 3), f)
        A["conn"] = np.full((n, self.connmax), -1, i32)
        A["nb_conn"] = np.full(n, -1, i32)
        A["kptr"] = np.zeros((n + 1, 2), i32)
        for k in ("Kn", "Tv", "dL", "dL_ave", "ddL", "ddLp", "F", "Ftemp", "bond_stress"):
            A[k] = np.zeros((n, nbm), f)
        A["cs"] = np.zeros((n, nbm, 3), f)
        A["dLp"] = np.zeros((n, nbm, 3), f)
        A["dmgD"] = np.zeros((n, nbm, 2), f)
        A["dmg_broken"] = np.ones((n, nbm), f)
        A["dmg_w"] = np.ones((n, nbm), f)
        for k in ("dLtot", "TdLtot", "ddLtot", "TddLtot", "dnl", "dloc"):
            A[k] = np.zeros((n, 2), f)
        A["beta"] = np.zeros((n, 6, 3), f)
        A["alpha"] = np.zeros((n, 3), f)
        A["beta_eq"] = np.zeros((n, 3), f)
        for k in ("dlambda", "seq", "sm", "triax", "damage_visual"):
            A[k] = np.zeros(n, f)
        A["stress"] = np.zeros((n, 6), f)
        A["pl_flag"] = np.zeros(n, i32)
        A["Pin"] = np.zeros(3 * n, f)
        for k in ("Pex", "Pex_temp", "residual", "disp"):
            A[k] = np.zeros(d * n, f)
        A["dispBC"] = np.ones(d * n, i32)
        A["fixidx"] = np.ones(d * n, i32)
        A["reaction"] = np.zeros(d * n, f)
        A["dist"] = np.zeros((n, nbm), f)
        if cp is not None:
            # crystal plasticity: cp = dict(R=3x3, tau0, taus, h0, q, eta, p, dtime, maxloop)
            sch = np.zeros((64, 6), f)
            ns = lib().orc_slip_systems_fcc(ptr(np.ascontiguousarray(cp.get("R", np.eye(3)), f).reshape(9)),
                                             ptr(sch))
            A["schmid"] = np.ascontiguousarray(sch[:ns])
            A["cp_gy"] = np.full((n, ns, 3), cp["tau0"], f)
            A["cp_A_single"] = np.zeros((n, ns, 3), f)
            A["cp_A"] = np.zeros((n, 3), f)
            A["cp_dA"] = np.zeros(n, f)
            for k in ("cp_dgy", "cp_dA_single", "cp_RSS"):
                A[k] = np.zeros((n, ns), f)
            A["cp_Jact"] = np.zeros((n, ns), i32)
            A["cp_Cab"] = np.zeros((n, ns, ns), f)
            A["state_v"] = np.zeros(n, i32)
        self.A = A
        s = OrcModel()
        s.dim, s.lattice, s.n, s.nbmax = d, lattice, n, nbm
        s.nb1max, s.nb2max, s.connmax, s.ntype = lp.nb1max, lp.nb2max, self.connmax, Ce.shape[0]
        s.radius, s.cut1, s.cut2, s.volume = radius, lp.cut1, lp.cut2, lp.volume
        s.plmode, s.J2_H, s.J2_xi = plmode, J2_H, J2_xi
        s.damage_L, s.damage_threshold, s.damagec_A = damage_L, damage_threshold, damagec_A
        s.critical_bstrain, s.nbreak = critical_bstrain, nbreak
        s.J2_C, s.j2nl_jacobi = J2_C, j2nl_jacobi
        if cp is not None:
            s.nslip = A["schmid"].shape[0]
            s.cp_tau0, s.cp_taus, s.cp_h0, s.cp_q = cp["tau0"], cp["taus"], cp["h0"], cp["q"]
            s.cp_eta, s.cp_p, s.dtime, s.cp_maxloop = cp["eta"], cp["p"], cp["dtime"], cp["maxloop"]
        for k, v in A.items():
            setattr(s, k, ptr(v))
        s.nreaction = 0
        s.nK = 0
        self.s = s
        self.p = C.pointer(s)

    def __getattr__(self, k):
        A = self.__dict__.get("A")
        if A is not None and k in A:
            return A[k]
        raise AttributeError(k)

    # -- setup -----------------------------------------------------------------------
    def neighbors(self, cells=False):
        rc = (lib().orc_neighbor_search_cells if cells else lib().orc_neighbor_search)(self.p)
        assert rc == 0, "neighbour list overflow"

    def connectivity(self):
        nK = lib().orc_afem_connectivity(self.p)
        assert nK > 0, "conn overflow"
        self.A["K"] = np.zeros(nK, np.float64)
        self.A["JK"] = np.full(nK, -1, np.int32)
        self.A["IK"] = np.full(self.dim * self.n + 1, -1, np.int32)
        self.s.nK = nK
        self.s.K, self.s.JK, self.s.IK = ptr(self.A["K"]), ptr(self.A["JK"]), ptr(self.A["IK"])
        return nK

    def setup(self, cells=False):
        self.neighbors(cells)
        self.connectivity()
        lib().orc_kntv(self.p)
        lib().orc_computedL(self.p)

    def call(self, name, *args):
        return getattr(lib(), name)(self.p, *args)

    def compute_strain(self, strain=None):
        """computeStrain from the model's dL; strain [n][6] in and out (zeros if None)."""
        st = np.zeros((self.n, 6)) if strain is None else np.array(strain, dtype=np.float64, order="C")
        lib().orc_compute_strain(self.p, ptr(st))
        return st


def bc_tables(steps):
    """steps: list of (disp_list, force_list); disp entry (type, flag, step),
    force entry (type, (fx, fy, fz)).  Returns ctypes arrays for orc_run_steps."""
    maxbc = max(1, max(max(len(d), len(f)) for d, f in steps))
    ns = len(steps)
    dbc = (OrcDispBC * (ns * maxbc))()
    fbc = (OrcForceBC * (ns * maxbc))()
    nbd = np.zeros(ns, np.int32)
    nbf = np.zeros(ns, np.int32)
    for s, (dl, fl) in enumerate(steps):
        nbd[s], nbf[s] = len(dl), len(fl)
        for k, (t, flag, st) in enumerate(dl):
            dbc[s * maxbc + k].type, dbc[s * maxbc + k].flag, dbc[s * maxbc + k].step = t, flag.encode(), st
        for k, (t, st) in enumerate(fl):
            fbc[s * maxbc + k].type = t
            for c in range(3):
                fbc[s * maxbc + k].step[c] = st[c]
    return maxbc, nbd, dbc, nbf, fbc


def run_steps(model, steps, solver, out_axis="z", out_type=1):
    maxbc, nbd, dbc, nbf, fbc = bc_tables(steps)
    out = (OrcStepOut * len(steps))()
    rc = lib().orc_run_steps(model.p, C.byref(solver), len(steps), ptr(nbd), dbc, ptr(nbf), fbc, maxbc,
                             out_axis.encode(), out_type, out)
    return rc, list(out)


def set_output(bond=None, reaction=None, disp=None, axis="z", t1=1, step_offset=0):
    """Reference-format files orc_run_steps appends to (None: off)."""
    enc = lambda p: None if p is None else str(p).encode()
    lib().orc_set_output(enc(bond), enc(reaction), enc(disp), axis.encode(), t1, step_offset)


def adjoint_run(model, steps, solver, target, out_axis="z", out_type=1):
    """Discrete adjoint of a J2 load history with damage off: returns (rc, J, reactions,
    dJ/d modulus factor [n], dJ/d yield stress [n], dJ/dH) for J = sum (R_n - target_n)^2."""
    maxbc, nbd, dbc, nbf, fbc = bc_tables(steps)
    ns, n = len(steps), model.n
    J = np.zeros(1)
    reac, ge, gy, gh = np.zeros(ns), np.zeros(n), np.zeros(n), np.zeros(1)
    tgt = np.ascontiguousarray(target, np.float64)
    rc = lib().orc_adjoint_run(model.p, C.byref(solver), ns, ptr(nbd), dbc, ptr(nbf), fbc, maxbc,
                               out_axis.encode(), out_type, ptr(tgt), ptr(J), ptr(reac), ptr(ge), ptr(gy), ptr(gh))
    return rc, J[0], reac, ge, gy, gh[0]


def solver(rtol2=1e-8, atol2=1e-12, maxit=None, newton_tol=1e-4, newton_maxit=100, tangent=0, nthreads=1, n=0,
           fixed_diag=0, accel=0, skip_damage=0):
    s = OrcSolver()
    s.rtol2, s.atol2 = rtol2, atol2
    s.maxit = maxit if maxit is not None else max(1, n)
    s.newton_tol, s.newton_maxit, s.tangent, s.nthreads = newton_tol, newton_maxit, tangent, nthreads
    s.fixed_diag = fixed_diag
    s.accel = accel
    s.skip_damage = skip_damage
    return s


def upper_csr_dense(model):
    """Full dense matrix of the upper CSR tangent (small models only)."""
    nr = model.dim * model.n
    K = np.zeros((nr, nr))
    IK, JK, V = model.IK, model.JK, model.K
    for r in range(nr):
        for p in range(IK[r] - 1, IK[r + 1] - 1):
            c = JK[p] - 1
            K[r, c] += V[p]
            if c != r:
                K[c, r] += V[p]
    return K
