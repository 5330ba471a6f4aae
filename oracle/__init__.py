"""ctypes view of oracle/liboracle.so, the FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker.  The product never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


if not os.path.exists(LIB_PATH):
    build()
lib = C.CDLL(LIB_PATH)

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


_sig("orc_rng_uniform", C.c_double, [C.c_uint64] * 3)
_sig("orc_rng_normal", C.c_double, [C.c_uint64] * 3)
_sig("orc_mesh_box", None, [C.c_int] * 3 + [C.c_double] * 6 + [dp, ip])
_sig("orc_field_inclusions", None, [C.c_int64, ip, dp, C.c_int, C.c_double, C.c_double, C.c_uint64, C.c_double,
                                    C.c_double, dp])
_sig("orc_field_lognormal", None, [C.c_int64, C.c_uint64, C.c_uint64, C.c_int64, C.c_double, C.c_double, dp])
_sig("orc_field_grf", None, [C.c_int64, C.c_int64, ip, dp, C.c_int, C.c_double, C.c_uint64, C.c_double, C.c_double,
                             dp, ip])
_sig("orc_tet_geometry", C.c_double, [dp, dp])
_sig("orc_isotropic_D_export", None, [C.c_double, C.c_double, dp])
_sig("orc_elem_elastic", C.c_double, [dp, C.c_double, C.c_double, C.c_int, dp])
_sig("orc_elem_scalar", C.c_double, [dp, C.c_double, dp])
_sig("orc_scalar_system", vp, [C.c_int, C.c_int64, dp, ip, dp, C.c_int, ip, dp])
_sig("orc_renumber", C.c_int, [C.c_int, C.c_int64, ip, ip, C.c_int, ip, ip])
_sig("orc_elastic_system", vp, [C.c_int, C.c_int64, dp, ip, ip, dp, C.c_double, C.c_int, ip, dp, C.c_int, ip, dp,
                                C.c_int, C.c_int])
_sig("orc_sys_n", C.c_int64, [vp])
_sig("orc_sys_nnz", C.c_int64, [vp])
_sig("orc_sys_missing", C.c_int, [vp])
_sig("orc_sys_csr", None, [vp, ip, ip, dp, dp, dp])
_sig("orc_sys_node_map", None, [vp, ip])
_sig("orc_sys_free", None, [vp])
_sig("orc_csr_mv", None, [C.c_int, ip, ip, dp, dp, dp])
_sig("orc_pcg", C.c_int, [C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int, C.c_int, dp])
_sig("orc_gmres", C.c_int, [C.c_int, ip, ip, dp, dp, dp, C.c_int, C.c_double, C.c_int, C.c_int, dp])
_sig("orc_dense_solve", C.c_int, [C.c_int, ip, ip, dp, dp, dp])
_sig("orc_bicgstab", C.c_int, [C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int, C.c_int, dp])
_sig("orc_refpath_solve", C.c_int, [C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp])
_sig("orc_bicg_diag", C.c_double, [C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int, dp, ip])
_sig("orc_msfem_cell", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, C.c_double,
                                 C.c_double, dp, dp, ip])
_sig("orc_msfem_cell_refpath", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, C.c_double,
                                         C.c_double, dp, dp, ip])
_sig("orc_bcg_stab2_csr", C.c_int, [C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int, dp])
_sig("orc_elem_seg", C.c_double, [dp, C.c_double, C.c_double, C.c_int, dp])
_sig("orc_elem_tri", C.c_double, [dp, C.c_double, C.c_double, C.c_int, dp])
_sig("orc_edge_1d", C.c_int, [C.c_int, dp, dp, C.c_double, C.c_int, dp, dp, dp])
_sig("orc_msfem_cell_reftrace", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, C.c_double,
                                          C.c_double, C.c_int, dp, dp, ip])
_sig("orc_mlp_loss_grad", C.c_double, [C.c_int, ip, C.POINTER(dp), C.POINTER(dp), C.c_int, dp, dp, C.POINTER(dp),
                                       C.POINTER(dp)])


def mesh_box(nx, ny=None, nz=None, origin=(0.0, 0.0, 0.0), h=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    h = (1.0 / nx, 1.0 / ny, 1.0 / nz) if h is None else h
    xyz = np.empty(((nx + 1) * (ny + 1) * (nz + 1), 3))
    tets = np.empty((6 * nx * ny * nz, 4), dtype=np.int32)
    lib.orc_mesh_box(nx, ny, nz, *origin, *h, _p(xyz, C.c_double), _p(tets, C.c_int))
    return xyz, tets


def field_inclusions(tets, xyz, count, rmin, rmax, seed, v_bg, v_in):
    out = np.empty(len(tets))
    lib.orc_field_inclusions(len(tets), _p(_i32(tets), C.c_int), _p(_f64(xyz), C.c_double), count, rmin, rmax, seed,
                             v_bg, v_in, _p(out, C.c_double))
    return out


def field_lognormal(nt, seed, stream, index0, mu, sigma):
    out = np.empty(nt)
    lib.orc_field_lognormal(nt, seed, stream, index0, mu, sigma, _p(out, C.c_double))
    return out


def field_grf(xyz, tets, modes, ell, seed, v_lo, v_hi):
    out = np.empty(len(tets))
    mat = np.empty(len(tets), dtype=np.int32)
    lib.orc_field_grf(len(xyz), len(tets), _p(_i32(tets), C.c_int), _p(_f64(xyz), C.c_double), modes, ell, seed, v_lo,
                      v_hi, _p(out, C.c_double), _p(mat, C.c_int))
    return out, mat


def tet_geometry(P):
    P = _f64(P).reshape(4, 3)
    a = np.empty(16)
    V = lib.orc_tet_geometry(_p(P, C.c_double), _p(a, C.c_double))
    return V, a.reshape(4, 4)


def isotropic_D(E, nu):
    D = np.empty(36)
    lib.orc_isotropic_D_export(E, nu, _p(D, C.c_double))
    return D.reshape(6, 6)


def elem_elastic(P, E, nu, convention=0):
    P = _f64(P).reshape(4, 3)
    K = np.empty(144)
    lib.orc_elem_elastic(_p(P, C.c_double), E, nu, convention, _p(K, C.c_double))
    return K.reshape(12, 12)


def elem_scalar(P, sigma):
    P = _f64(P).reshape(4, 3)
    K = np.empty(16)
    lib.orc_elem_scalar(_p(P, C.c_double), sigma, _p(K, C.c_double))
    return K.reshape(4, 4)


class System:
    """Assembled system (CSSD rows after BCs) held by the oracle."""

    def __init__(self, h):
        self.h = h
        self.n = lib.orc_sys_n(h)
        self.nnz = lib.orc_sys_nnz(h)
        self.missing = lib.orc_sys_missing(h)
        self.rowptr = np.empty(self.n + 1, dtype=np.int32)
        self.col = np.empty(self.nnz, dtype=np.int32)
        self.val = np.empty(self.nnz)
        self.rhs = np.empty(self.n)
        self.x0 = np.empty(self.n)
        lib.orc_sys_csr(h, _p(self.rowptr, C.c_int), _p(self.col, C.c_int), _p(self.val, C.c_double),
                        _p(self.rhs, C.c_double), _p(self.x0, C.c_double))

    def node_map(self, nn):
        m = np.empty(nn, dtype=np.int32)
        lib.orc_sys_node_map(self.h, _p(m, C.c_int))
        return m

    def __del__(self):
        try:
            lib.orc_sys_free(self.h)
        except Exception:
            pass

    def csr(self):
        return self.rowptr, self.col, self.val


def scalar_system(xyz, tets, sigma, fixed_nodes, fixed_vals):
    fixed_nodes = _i32(fixed_nodes)
    fixed_vals = _f64(fixed_vals)
    h = lib.orc_scalar_system(len(xyz), len(tets), _p(_f64(xyz), C.c_double), _p(_i32(tets), C.c_int),
                              _p(_f64(sigma), C.c_double), len(fixed_nodes), _p(fixed_nodes, C.c_int),
                              _p(fixed_vals, C.c_double))
    return System(h)


def renumber(nn, tets, material, fixed_nodes):
    m = np.empty(nn, dtype=np.int32)
    fixed_nodes = _i32(fixed_nodes)
    lib.orc_renumber(nn, len(tets), _p(_i32(tets), C.c_int),
                     _p(_i32(material), C.c_int) if material is not None else None, len(fixed_nodes),
                     _p(fixed_nodes, C.c_int), _p(m, C.c_int))
    return m


def elastic_system(xyz, tets, E, nu, fixed_nodes=(), fixed_vals=None, tris=None, traction=None, material=None,
                   convention=0, renumber=True):
    fixed_nodes = _i32(fixed_nodes)
    fixed_vals = _f64(fixed_vals if fixed_vals is not None else np.zeros((len(fixed_nodes), 3))).ravel()
    tris = _i32(tris if tris is not None else np.zeros((0, 3)))
    traction = _f64(traction if traction is not None else np.zeros((0, 3))).ravel()
    h = lib.orc_elastic_system(len(xyz), len(tets), _p(_f64(xyz), C.c_double), _p(_i32(tets), C.c_int),
                               _p(_i32(material), C.c_int) if material is not None else None,
                               _p(_f64(E), C.c_double), nu, len(fixed_nodes), _p(fixed_nodes, C.c_int),
                               _p(fixed_vals, C.c_double), len(tris), _p(tris, C.c_int), _p(traction, C.c_double),
                               convention, int(renumber))
    return System(h)


def csr_mv(rowptr, col, val, x):
    y = np.empty(len(rowptr) - 1)
    x = _f64(x)
    lib.orc_csr_mv(len(rowptr) - 1, _p(_i32(rowptr), C.c_int), _p(_i32(col), C.c_int), _p(_f64(val), C.c_double),
                   _p(x, C.c_double), _p(y, C.c_double))
    return y


def pcg(rowptr, col, val, b, x0=None, tol=1e-10, maxit=100000, precond=True):
    n = len(rowptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    rr = C.c_double()
    it = lib.orc_pcg(n, _p(_i32(rowptr), C.c_int), _p(_i32(col), C.c_int), _p(_f64(val), C.c_double),
                     _p(_f64(b), C.c_double), _p(x, C.c_double), tol, maxit, int(precond), C.byref(rr))
    return x, it, rr.value


def gmres(rowptr, col, val, b, m=30, x0=None, tol=1e-10, maxit=100000, precond=True):
    n = len(rowptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    rr = C.c_double()
    it = lib.orc_gmres(n, _p(_i32(rowptr), C.c_int), _p(_i32(col), C.c_int), _p(_f64(val), C.c_double),
                       _p(_f64(b), C.c_double), _p(x, C.c_double), m, tol, maxit, int(precond), C.byref(rr))
    return x, it, rr.value


def bicgstab(rowptr, col, val, b, x0=None, tol=1e-10, maxit=100000, precond=True):
    n = len(rowptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    tr = C.c_double()
    it = lib.orc_bicgstab(n, _p(_i32(rowptr), C.c_int), _p(_i32(col), C.c_int), _p(_f64(val), C.c_double),
                          _p(_f64(b), C.c_double), _p(x, C.c_double), tol, maxit, int(precond), C.byref(tr))
    return x, it, tr.value


def dense_solve(rowptr, col, val, b):
    n = len(rowptr) - 1
    x = np.zeros(n)
    code = lib.orc_dense_solve(n, _p(_i32(rowptr), C.c_int), _p(_i32(col), C.c_int), _p(_f64(val), C.c_double),
                               _p(_f64(b), C.c_double), _p(x, C.c_double))
    assert code == 0
    return x


def refpath_solve(rowptr, col, val, b, x0=None, tol=1e-8, maxit=100000, mode=0, threads=1):
    """The reference's CPU solver path (mode 0: SSOR-PCG, 1: MSG CG)."""
    n = len(rowptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    rr, sec = C.c_double(), C.c_double()
    it = lib.orc_refpath_solve(n, _p(_i32(rowptr), C.c_int), _p(_i32(col), C.c_int), _p(_f64(val), C.c_double),
                               _p(_f64(b), C.c_double), _p(x, C.c_double), tol, maxit, mode, threads, C.byref(rr),
                               C.byref(sec))
    return x, it, rr.value, sec.value


def bicg_diag(rowptr, col, val, b, diag, x0=None, E=1e-10, maxiter=100000):
    n = len(rowptr) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    it = C.c_int()
    res = lib.orc_bicg_diag(n, _p(_i32(rowptr), C.c_int), _p(_i32(col), C.c_int), _p(_f64(val), C.c_double),
                            _p(_f64(b), C.c_double), _p(x, C.c_double), E, maxiter, _p(_f64(diag), C.c_double),
                            C.byref(it))
    return x, it.value, res


def msfem_cell(n, origin, h, E, nu, tol=1e-12):
    N = 3 * (n + 1) ** 3
    basis = np.empty((24, N))
    Kc = np.empty((24, 24))
    it = C.c_int()
    lib.orc_msfem_cell(n, *origin, h, _p(_f64(E), C.c_double), nu, tol, _p(basis, C.c_double), _p(Kc, C.c_double),
                       C.byref(it))
    return basis, Kc, it.value


def bcg_stab2(rowptr, col, val, b, x0=None, E=1e-10, maxiter=100000):
    """The reference's unpreconditioned BCG_Stab2 (x, iterations, true relres)."""
    x = np.zeros(len(b)) if x0 is None else _f64(x0).copy()
    rr = C.c_double()
    it = lib.orc_bcg_stab2_csr(len(b), _p(_i32(rowptr), C.c_int), _p(_i32(col), C.c_int), _p(_f64(val), C.c_double),
                               _p(_f64(b), C.c_double), _p(x, C.c_double), E, maxiter, C.byref(rr))
    return x, it, rr.value


def msfem_cell_refpath(n, origin, h, E, nu, min_residual=1e-8):
    """Bases of one cell solved on the reference's CPU path (full system with
    eliminated boundary rows, component-major DOFs, SSOR-PCG continuation
    with BCG_Stab2 fallback, FEM.h:3699-3770); total iterations."""
    N = 3 * (n + 1) ** 3
    basis = np.empty((24, N))
    Kc = np.empty((24, 24))
    it = C.c_int()
    lib.orc_msfem_cell_refpath(n, *origin, h, _p(_f64(E), C.c_double), nu, min_residual, _p(basis, C.c_double),
                               _p(Kc, C.c_double), C.byref(it))
    return basis, Kc, it.value


def elem_seg(P, E, nu, convention=0):
    K = np.empty((6, 6))
    L = lib.orc_elem_seg(_p(_f64(P), C.c_double), E, nu, convention, _p(K, C.c_double))
    return K, L


def elem_tri(P, E, nu, convention=0):
    K = np.empty((9, 9))
    A = lib.orc_elem_tri(_p(_f64(P), C.c_double), E, nu, convention, _p(K, C.c_double))
    return K, A


def edge_1d(xyz, E, nu, u0, un, convention=0):
    xyz = _f64(xyz).reshape(-1, 3)
    U = np.empty(xyz.size)
    rc = lib.orc_edge_1d(len(xyz) - 1, _p(xyz, C.c_double), _p(_f64(E), C.c_double), nu, convention,
                         _p(_f64(u0), C.c_double), _p(_f64(un), C.c_double), _p(U, C.c_double))
    assert rc == 0
    return U.reshape(-1, 3)


def msfem_cell_reftrace(n, origin, h, E, nu, tol=1e-12, form=0):
    """Bases of one cell with the reference's edge -> face -> volume traces
    (basis 3c+d driven by e_d), and Kc for the per-direction (form 0) or
    component-wise (form 1) coarse functions."""
    N = 3 * (n + 1) ** 3
    basis = np.empty((24, N))
    Kc = np.empty((24, 24))
    it = C.c_int()
    rc = lib.orc_msfem_cell_reftrace(n, *origin, h, _p(_f64(E), C.c_double), nu, tol, form, _p(basis, C.c_double),
                                     _p(Kc, C.c_double), C.byref(it))
    assert rc == N, rc
    return basis, Kc, it.value


def mlp_loss_grad(Ws, bs, X, T, grads=True):
    L = len(Ws)
    dims = np.array([Ws[0].shape[1]] + [W.shape[0] for W in Ws], dtype=np.int32)
    Ws = [_f64(W) for W in Ws]
    bs = [_f64(b) for b in bs]
    X = _f64(X)
    T = _f64(T)
    Wp = (dp * L)(*[_p(W, C.c_double) for W in Ws])
    bp = (dp * L)(*[_p(b, C.c_double) for b in bs])
    if grads:
        gW = [np.empty_like(W) for W in Ws]
        gb = [np.empty_like(b) for b in bs]
        gWp = (dp * L)(*[_p(g, C.c_double) for g in gW])
        gbp = (dp * L)(*[_p(g, C.c_double) for g in gb])
    else:
        gW = gb = None
        gWp = gbp = None
    loss = lib.orc_mlp_loss_grad(L, _p(dims, C.c_int), Wp, bp, X.shape[0], _p(X, C.c_double), _p(T, C.c_double), gWp,
                                 gbp)
    return loss, gW, gb
