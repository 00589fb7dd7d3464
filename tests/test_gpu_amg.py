"""The AMG hierarchy built on the device (amg_setup.hip) against its definition.

The smoothed-aggregation preconditioner has no counterpart in the reference, so its
set-up is checked against a numpy restatement of the construction on the exported
hierarchy (lpmx_amg_export):
  - aggregates: greedy in index order over the bond graph (a root with all its
    neighbours free takes them), leftovers join the first aggregated neighbour, the
    rest form new aggregates; coarser levels use the aggregates' adjacency;
  - tentative prolongator T: the rigid-body modes about each aggregate's centroid,
    zero except on DOFs both displacement-constrained and fixed, orthonormalised per
    aggregate by modified Gram-Schmidt; the triangular factors are the next level's
    modes;
  - P = T - w D^-1 A T with w = 4 / (3 lambda), lambda from 20 power-iteration steps
    on D^-1 A from the fixed xorshift64 start;
  - R = P^T (bit for bit) and the next operator R A P.
numpy's sums run in another order than the device's, so values agree to round-off
(1e-10 relative), the patterns and level sizes exactly.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from lpm_harness import P

pytestmark = pytest.mark.gpu


def export(level, which):
    L = P.lib()
    dims = np.zeros(5, np.int32)
    P.check(L.lpmx_amg_export(level, which, 0, None, None, None, P.ptr(dims)), "export dims")
    nr, nc, br, bc, cnt = (int(v) for v in dims)
    rows, cols = np.zeros(cnt, np.int32), np.zeros(cnt, np.int32)
    vals = np.zeros(cnt * br * bc)
    P.check(L.lpmx_amg_export(level, which, cnt, P.ptr(rows), P.ptr(cols), P.ptr(vals), P.ptr(dims)), "export")
    r = (rows[:, None, None] * br + np.arange(br)[None, :, None]) + 0 * np.arange(bc)[None, None, :]
    c = (cols[:, None, None] * bc + np.arange(bc)[None, None, :]) + 0 * np.arange(br)[None, :, None]
    M = sp.csr_matrix((vals, (r.ravel(), c.ravel())), shape=(nr * br, nc * bc))
    M.sum_duplicates()
    return M, (nr, nc, br, bc)


def fine_tangent(n, d):
    L = P.lib()
    IK = np.zeros(d * n + 1, np.int32)
    cap = n * 64 * d * d
    JK, K = np.zeros(cap, np.int32), np.zeros(cap)
    P.check(L.lpmx_export_upper_csr(P.ptr(IK), P.ptr(JK), P.ptr(K)), "tangent export")
    nK = IK[-1] - 1
    rows = np.repeat(np.arange(d * n), np.diff(IK))
    U = sp.csr_matrix((K[:nK], (rows, JK[:nK] - 1)), shape=(d * n, d * n))
    return (U + sp.triu(U, 1).T).tocsr()


def xorshift_start(m):
    M = (1 << 64) - 1
    s = 0x9E3779B97F4A7C15
    out = np.empty(m)
    for i in range(m):
        s ^= (s << 13) & M
        s ^= s >> 7
        s ^= (s << 17) & M
        out[i] = (s >> 11) * (1.0 / 9007199254740992.0) - 0.5
    return out


def diag_block_inverses(A, d):
    nr = A.shape[0] // d
    D = np.zeros((nr, d, d))
    Ad = A.tocsr()
    for r in range(d):
        for c in range(d):
            D[:, r, c] = Ad[np.arange(nr) * d + r, np.arange(nr) * d + c].A1
    return np.linalg.inv(D)


def lambda_max(A, Dinv, d):
    x = xorshift_start(A.shape[0])
    lam = 0.0
    for _ in range(20):
        x = x / np.sqrt(x @ x)
        z = np.einsum("iab,ib->ia", Dinv, (A @ x).reshape(-1, d)).ravel()
        lam = np.sqrt(z @ z)
        x = z
    return lam


def aggregate(n, aptr, adj):
    agg = -np.ones(n, np.int64)
    na = 0
    for i in range(n):
        nb = adj[aptr[i]:aptr[i + 1]]
        if agg[i] >= 0 or (agg[nb] >= 0).any():
            continue
        agg[i] = na
        agg[nb] = na
        na += 1
    first = agg.copy()
    for i in range(n):
        if first[i] < 0:
            for j in adj[aptr[i]:aptr[i + 1]]:
                if first[j] >= 0:
                    agg[i] = first[j]
                    break
    for i in range(n):
        if agg[i] < 0:
            agg[i] = na
            for j in adj[aptr[i]:aptr[i + 1]]:
                if agg[j] < 0:
                    agg[j] = na
            na += 1
    return agg, na


def quotient(agg, na, aptr, adj):
    sets = [set() for _ in range(na)]
    for i in range(len(agg)):
        for j in adj[aptr[i]:aptr[i + 1]]:
            if agg[j] != agg[i]:
                sets[agg[i]].add(int(agg[j]))
    qadj = [sorted(s) for s in sets]
    qptr = np.concatenate([[0], np.cumsum([len(s) for s in qadj])]).astype(np.int64)
    return qptr, np.array([b for s in qadj for b in s], np.int64)


def tentative(agg, na, d, nm, B):
    nr = len(agg)
    T = np.zeros((nr * d, nm))
    Bc = np.zeros((na * nm, nm))
    order = np.argsort(agg, kind="stable")
    bounds = np.searchsorted(agg[order], np.arange(na + 1))
    for a in range(na):
        mem = order[bounds[a]:bounds[a + 1]]
        rows = (mem[:, None] * d + np.arange(d)).ravel()
        Q = B[rows].copy()
        ref = np.sqrt((Q * Q).sum(0)).max()
        R = np.zeros((nm, nm))
        for c in range(nm):
            for p in range(c):
                s = Q[:, p] @ Q[:, c]
                R[p, c] = s
                Q[:, c] -= s * Q[:, p]
            s = np.sqrt(Q[:, c] @ Q[:, c])
            if s > 1e-10 * ref and ref > 0:
                R[c, c] = s
                Q[:, c] /= s
            else:
                Q[:, c] = 0.0
        T[rows] = Q
        Bc[a * nm:(a + 1) * nm] = R
    r = np.repeat(np.arange(nr * d), nm)
    c = (np.repeat(agg, d)[:, None] * nm + np.arange(nm)).ravel()
    return sp.csr_matrix((T.ravel(), (r, c)), shape=(nr * d, na * nm)), Bc


def rigid_modes(x, agg, na, d, keep):
    cnt = np.bincount(agg, minlength=na).astype(np.float64)
    cx = np.stack([np.bincount(agg, weights=x[:, c], minlength=na) for c in range(3)], 1) / cnt[:, None]
    p = x - cx[agg]
    n = len(agg)
    if d == 3:
        B = np.zeros((n, 3, 6))
        B[:, 0, 0] = B[:, 1, 1] = B[:, 2, 2] = 1.0
        B[:, 0, 4], B[:, 0, 5] = p[:, 2], -p[:, 1]
        B[:, 1, 3], B[:, 1, 5] = -p[:, 2], p[:, 0]
        B[:, 2, 3], B[:, 2, 4] = p[:, 1], -p[:, 0]
    else:
        B = np.zeros((n, 2, 3))
        B[:, 0, 0] = B[:, 1, 1] = 1.0
        B[:, 0, 2], B[:, 1, 2] = -p[:, 1], p[:, 0]
    B *= keep.reshape(n, d)[:, :, None]
    return B.reshape(n * d, -1)


def assert_close(name, got, want, tol):
    diff = abs((got - want).tocsr().data).max(initial=0.0)
    scale = abs(want.tocsr().data).max()
    print(f"{name}: max diff {diff:.3e} of {scale:.3e}", flush=True)
    assert diff <= tol * scale, name


def build_constrained(case):
    L = P.lib()
    if case == "sc":
        n, x, t, c = P.build_sc_model(P.sc_box_for(20))  # 8,000 particles: three levels
        L.calcStiffness3DFiniteDifference(6)
        d = 3
        dbc = np.ones(d * n, np.int32)
        dbc.reshape(n, d)[t == 2, :] = 0
        dbc.reshape(n, d)[t == 3, :] = 0
        dbc.reshape(n, d)[t == 1, 2] = 0
    else:
        n, x, t, c = P.build_hex_model([0.0, 40.0, 0.0, 35.0, 0.0, 1.0], critical=1.0)  # two levels
        L.calcStiffness2DFiniteDifference(6)
        d = 2
        dbc = np.ones(d * n, np.int32)
        dbc.reshape(n, d)[t == 2, :] = 0
        dbc.reshape(n, d)[t == 1, 1] = 0
    P.write1d("dispBC_index", dbc)
    P.write1d("residual", np.random.default_rng(3).standard_normal(d * n) * dbc)
    (L.setDispBC_stiffnessUpdate3D if d == 3 else L.setDispBC_stiffnessUpdate2D)()
    x = P.field("XYZ", 3 * n).reshape(n, 3)  # the set-up's positions (the solve moves them)
    s = P.solver(precision=P.LPMX_FP64, rtol2=1e-16, atol2=0.0, maxit=10 * d * n, precond=1)
    L.lpm_ext_set_solver(C.byref(s))
    L.solverCG()  # the first preconditioned solve builds the hierarchy
    L.lpm_ext_set_solver(None)
    return n, d, x


@pytest.mark.parametrize("case,layout", [("sc", "default"), ("hex", "default"), ("sc", "sliced")])
def test_amg_hierarchy_matches_definition(case, layout, monkeypatch):
    # "sliced": every level's matrices in the sliced layout with their rows sorted by
    # length (at 1M only the prolongator is), which the export maps back to rows
    if layout == "sliced":
        monkeypatch.setenv("LPMX_AMG_SLICE_MIN", "1")
    n, d, x = build_constrained(case)
    nm = 6 if d == 3 else 3
    L = P.lib()
    lev, ts, dev = C.c_int(), C.c_double(), C.c_int64()
    P.check(L.lpmx_amg_info(C.byref(lev), C.byref(ts), C.byref(dev)), "amg info")
    print(f"{case}: {lev.value} levels, set-up {ts.value:.3f} s", flush=True)
    assert lev.value == (3 if case == "sc" else 2)
    # finest level: the model's tangent, its inverses, the bond graph
    A = fine_tangent(n, d)
    keep = (P.field("DISPBC_INDEX", d * n) != 0) & (P.field("FIX_INDEX", d * n) != 0)
    nbr = P.field("NEIGHBORS", n * P.gint("nneighbors").value).reshape(n, -1)
    nb0 = P.field("NB_INITIAL", n)
    rows = [np.unique([q for q in nbr[i, :nb0[i]] if q >= 0 and q != i]) for i in range(n)]
    aptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    adj = np.concatenate(rows).astype(np.int64)
    agg, na = aggregate(n, aptr, adj)
    B = rigid_modes(x, agg, na, d, keep)
    bd = d
    for level in range(lev.value - 1):
        Dinv = diag_block_inverses(A, bd)
        lam = lambda_max(A, Dinv, bd)
        T, Bc = tentative(agg, na, bd, nm, B)
        DA = sp.block_diag(list(Dinv), format="csr") @ A
        Pref = T - (4.0 / (3.0 * lam)) * (DA @ T)
        Pd, (nr, nc, br, bc) = export(level, 1)
        Rd, _ = export(level, 2)
        assert (nr, nc, br, bc) == (A.shape[0] // bd, na, bd, nm), (level, nr, nc, br, bc)
        assert_close(f"{case} level {level} P", Pd, Pref, 1e-10)
        assert (Rd != Pd.T.tocsr()).nnz == 0, "R is P^T bit for bit"
        Ac = (Rd @ A @ Pd).tocsr()
        if level + 1 < lev.value - 1:
            Ad, dims = export(level + 1, 0)
            assert dims[:2] == (na, na)
            assert_close(f"{case} level {level + 1} operator", Ad, Ac, 1e-12)
            A = Ad
        aptr, adj = quotient(agg, na, aptr, adj)
        bd = nm
        B = Bc
        if level + 1 < lev.value - 1:
            agg, na = aggregate(na, aptr, adj)
