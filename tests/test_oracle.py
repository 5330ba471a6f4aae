"""Pins for the FP64 CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors for this path (its
own checks are unasserted prints), so each oracle function is pinned here by
what the mathematics fixes: closed forms, invariants, special cases that
reduce to a textbook routine, and brute force on tiny meshes.  Each check is
chosen so a plausible slip (dropped term, wrong sign or index, transposed
operand) fails it.
"""
import itertools
import os

import numpy as np
import pytest

import oracle as orc
from multiscalefemsolver3d_amd import problems

NU = 0.3


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def rand_tet(rng):
    while True:
        P = rng.normal(size=(4, 3))
        if abs(np.linalg.det(np.c_[np.ones(4), P])) > 0.3:
            return P


def numpy_B(g):
    """Engineering-strain B (6x3), Voigt xx,yy,zz,xy,yz,xz, written from the
    textbook definition (not from the oracle)."""
    gx, gy, gz = g
    return np.array([[gx, 0, 0], [0, gy, 0], [0, 0, gz], [gy, gx, 0], [0, gz, gy], [gz, 0, gx]])


def numpy_D(E, nu):
    lam, mu = lame(E, nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[range(3), range(3)] = lam + 2 * mu
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    return D


# --- geometry -----------------------------------------------------------------

def test_tet_alpha_is_inverse_of_vandermonde():
    rng = np.random.default_rng(0)
    for _ in range(20):
        P = rand_tet(rng)
        V, a = orc.tet_geometry(P)
        M = np.c_[np.ones(4), P]
        # phi_i(x_j) = delta_ij
        assert np.allclose(M @ a.T, np.eye(4), atol=1e-12)
        assert abs(V - abs(np.linalg.det(M)) / 6) < 1e-12 * max(1, V)
        assert np.allclose(a[:, 1:].sum(0), 0, atol=1e-12)   # sum grad phi = 0
        assert abs(a[:, 0].sum() - 1) < 1e-12               # sum phi = 1


def test_unit_tet_volume_and_gradients():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    V, a = orc.tet_geometry(P)
    assert V == pytest.approx(1 / 6, rel=1e-15)
    assert np.allclose(a[:, 1:], [[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]])


def test_isotropic_D_closed_form():
    E = 2.5e9
    D = orc.isotropic_D(E, NU)
    assert np.allclose(D, numpy_D(E, NU), rtol=1e-14)
    v = D @ np.array([1, 1, 1, 0, 0, 0.0])
    assert np.allclose(v, E / (1 - 2 * NU) * np.array([1, 1, 1, 0, 0, 0]), rtol=1e-14)


# --- element matrices ---------------------------------------------------------

def test_elastic_element_textbook_matches_numpy_BtDB():
    rng = np.random.default_rng(1)
    for _ in range(10):
        P = rand_tet(rng)
        E = float(rng.uniform(1e8, 1e10))
        K = orc.elem_elastic(P, E, NU, convention=1)
        V, a = orc.tet_geometry(P)
        B = np.hstack([numpy_B(a[i, 1:]) for i in range(4)])
        Kref = V * B.T @ numpy_D(E, NU) @ B
        assert np.allclose(K, Kref, rtol=1e-12, atol=1e-12 * abs(Kref).max())


def test_elastic_element_reference_is_blockwise_transpose():
    """FEM_Grid.h:2159-2191 forms B_J^T D B_I for block (I,J): every 3x3 block
    is the transpose of the textbook block."""
    rng = np.random.default_rng(2)
    P = rand_tet(rng)
    Kr = orc.elem_elastic(P, 1e9, NU, convention=0)
    Kt = orc.elem_elastic(P, 1e9, NU, convention=1)
    for I, J in itertools.product(range(4), range(4)):
        bt = Kt[3 * I:3 * I + 3, 3 * J:3 * J + 3]
        br = Kr[3 * I:3 * I + 3, 3 * J:3 * J + 3]
        assert np.allclose(br, bt.T, rtol=1e-13, atol=1e-13 * abs(Kt).max())
    assert np.allclose(Kr, Kr.T, atol=1e-12 * abs(Kr).max())
    # the two conventions differ for nu != 0.25 ...
    assert abs(Kr - Kt).max() > 1e-3 * abs(Kt).max()
    # ... and coincide when lambda == mu (nu = 1/4)
    Kr4 = orc.elem_elastic(P, 1e9, 0.25, 0)
    Kt4 = orc.elem_elastic(P, 1e9, 0.25, 1)
    assert np.allclose(Kr4, Kt4, rtol=1e-12, atol=1e-12 * abs(Kt4).max())


def rigid_modes(P):
    modes = []
    for d in range(3):
        u = np.zeros((4, 3))
        u[:, d] = 1
        modes.append(u.ravel())
    for w in np.eye(3):
        modes.append(np.cross(w, P).ravel())
    return np.array(modes)


def test_elastic_element_null_space():
    rng = np.random.default_rng(3)
    P = rand_tet(rng)
    Kt = orc.elem_elastic(P, 1e9, NU, 1)
    Kr = orc.elem_elastic(P, 1e9, NU, 0)
    R = rigid_modes(P)
    s = abs(Kt).max()
    assert abs(Kt @ R.T).max() < 1e-9 * s                 # all 6 rigid modes
    assert np.linalg.matrix_rank(Kt, tol=1e-9 * s) == 6
    assert abs(Kr @ R[:3].T).max() < 1e-9 * s             # translations
    assert np.linalg.eigvalsh(Kt).min() > -1e-9 * s       # PSD


def test_scalar_element():
    rng = np.random.default_rng(4)
    P = rand_tet(rng)
    K = orc.elem_scalar(P, 3.0)
    V, a = orc.tet_geometry(P)
    G = a[:, 1:]
    assert np.allclose(K, 3.0 * V * G @ G.T, rtol=1e-13)
    assert np.allclose(K.sum(1), 0, atol=1e-12 * abs(K).max())
    Pu = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    Ku = orc.elem_scalar(Pu, 1.0)
    assert np.allclose(Ku, np.array([[3, -1, -1, -1], [-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]]) / 6)


# --- mesh, portrait, assembly ---------------------------------------------------

def freudenthal_counts(n):
    nodes = (n + 1) ** 3
    # edges: 3 axis families, 3 face-diagonal families, 1 body diagonal family
    edges = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    return nodes, edges


def test_mesh_counts_and_volume():
    xyz, tets = orc.mesh_box(4)
    assert xyz.shape == (125, 3) and tets.shape == (384, 4)
    vols = [orc.tet_geometry(xyz[t])[0] for t in tets]
    assert np.allclose(vols, 1 / (6 * 64))
    assert sum(vols) == pytest.approx(1.0, rel=1e-13)


@pytest.mark.parametrize("n", [2, 3, 5])
def test_portrait_brute_force_and_counts(n):
    xyz, tets = orc.mesh_box(n)
    S = orc.scalar_system(xyz, tets, np.ones(len(tets)), [], [])
    pairs = set()
    for t in tets:
        for a in t:
            for b in t:
                pairs.add((a, b))
    rows = np.repeat(np.arange(S.n), np.diff(S.rowptr))
    assert set(zip(rows.tolist(), S.col.tolist())) == pairs
    assert all(np.all(np.diff(S.col[S.rowptr[i]:S.rowptr[i + 1]]) > 0) for i in range(S.n))
    nodes, edges = freudenthal_counts(n)
    assert S.nnz == nodes + 2 * edges


def test_elastic20_pattern_count_closed_form():
    """20^3 elastic cube: nnz = 9 (nodes + 2 edges) = 1,157,229."""
    nodes, edges = freudenthal_counts(20)
    assert nodes + 2 * edges == 128581
    assert 9 * (nodes + 2 * edges) == 1157229


def test_scalar_assembly_constant_in_null_space():
    xyz, tets = orc.mesh_box(4)
    sig = orc.field_inclusions(tets, xyz, 5, 0.1, 0.3, 2, 1.0, 100.0)
    S = orc.scalar_system(xyz, tets, sig, [], [])
    assert S.missing == 0
    y = orc.csr_mv(S.rowptr, S.col, S.val, np.ones(S.n))
    assert abs(y).max() < 1e-12 * abs(S.val).max()


def test_elastic_assembly_translations_in_null_space_and_block_order():
    xyz, tets = orc.mesh_box(3)
    E = orc.field_lognormal(len(tets), 1, 0, 0, np.log(1e9), 1.0)
    S = orc.elastic_system(xyz, tets, E, NU, renumber=False)
    assert S.missing == 0 and S.n == 3 * len(xyz)
    for d in range(3):
        u = np.zeros(S.n)
        u[d::3] = 1
        y = orc.csr_mv(S.rowptr, S.col, S.val, u)
        assert abs(y).max() < 1e-9 * abs(S.val).max()
    # Math.h:1963-2054 ordering: row 3i+a holds {3c+b} for every block column c
    # of node i, ascending - i.e. each scalar row is the 3-expansion of the
    # block row, identical for a = 0, 1, 2.
    Sb = orc.scalar_system(xyz, tets, np.ones(len(tets)), [], [])
    for i in range(len(xyz)):
        bc = Sb.col[Sb.rowptr[i]:Sb.rowptr[i + 1]]
        expect = (3 * bc[:, None] + np.arange(3)).ravel()
        for a in range(3):
            r = 3 * i + a
            assert np.array_equal(S.col[S.rowptr[r]:S.rowptr[r + 1]], expect)


def test_neumann_total_force():
    n = 4
    xyz, tets = orc.mesh_box(n)
    # top-face triangles: the two faces of each top square, split along (i,j)-(i+1,j+1)
    s = n + 1
    tris = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = [(i + di) + s * ((j + dj) + s * n) for di, dj in ((0, 0), (1, 0), (0, 1), (1, 1))]
            tris += [[a, b, d], [a, c, d]]
    tris = np.array(tris)
    bottom = np.arange(s * s)
    S = orc.elastic_system(xyz, tets, np.full(len(tets), 1e9), NU, bottom, np.zeros((len(bottom), 3)), tris,
                           np.tile([0, 0, -1e6], (len(tris), 1)))
    m = S.node_map(len(xyz))
    top_nodes = np.arange(s * s * n, s ** 3)
    fz = S.rhs[3 * m[top_nodes] + 2]
    assert fz.sum() == pytest.approx(-1e6, rel=1e-13)
    # corner nodes touch 1 or 2 triangles, edge nodes 3, interior 6: f = -1e6 * A/3 * count
    A = 0.5 / n ** 2
    counts = np.zeros(s * s)
    for t in tris:
        for v in t:
            counts[v - s * s * n] += 1
    assert np.allclose(fz, -1e6 * A / 3 * counts, rtol=1e-13)


# --- renumbering --------------------------------------------------------------

def test_renumbering_rule():
    """FEM.h:1990-2081: fixed vertices first in list order, then first
    appearance scanning material 0's tets, then material 1's."""
    xyz, tets = orc.mesh_box(2)
    mat = (np.arange(len(tets)) % 3 == 0).astype(np.int32)
    fixed = np.array([7, 3, 20])
    m = orc.renumber(len(xyz), tets, mat, fixed)
    assert list(m[fixed]) == [0, 1, 2]
    seen = list(fixed)
    for mm in (0, 1):
        for e in np.nonzero(mat == mm)[0]:
            for v in tets[e]:
                if v not in seen:
                    seen.append(v)
    expect = np.empty(len(xyz), dtype=np.int32)
    expect[np.array(seen)] = np.arange(len(seen))
    assert np.array_equal(m, expect)


# --- solvers ------------------------------------------------------------------

def small_elastic(n=3, conv=0):
    xyz, tets = orc.mesh_box(n)
    s = n + 1
    bottom = np.arange(s * s)
    tris = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = [(i + di) + s * ((j + dj) + s * n) for di, dj in ((0, 0), (1, 0), (0, 1), (1, 1))]
            tris += [[a, b, d], [a, c, d]]
    return orc.elastic_system(xyz, tets, np.full(len(tets), 1e9), NU, bottom, np.zeros((len(bottom), 3)),
                              np.array(tris), np.tile([0, 0, -1e6], (2 * n * n, 1)), convention=conv)


def test_pcg_cg_gmres_match_dense_solve():
    S = small_elastic(3)
    xd = orc.dense_solve(S.rowptr, S.col, S.val, S.rhs)
    for prec in (True, False):
        x, it, rr = orc.pcg(S.rowptr, S.col, S.val, S.rhs, tol=1e-12, precond=prec)
        assert rr < 1e-11 and np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    x, it, rr = orc.gmres(S.rowptr, S.col, S.val, S.rhs, m=30, tol=1e-12)
    assert rr < 1e-11 and np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    x, it, rr = orc.gmres(S.rowptr, S.col, S.val, S.rhs, m=5, tol=1e-12, precond=False)
    assert rr < 1e-11 and np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)


def test_bicgstab_matches_dense_solve_and_diagonal_special_case():
    rp, col, val, b = problems.convection_diffusion(12, 40.0)
    A = np.zeros((len(rp) - 1,) * 2)
    for i in range(len(rp) - 1):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    assert abs(A - A.T).max() > 1.0  # genuinely non-symmetric
    xd = np.linalg.solve(A, b)
    for prec in (True, False):
        x, it, tr = orc.bicgstab(rp, col, val, b, tol=1e-12, precond=prec)
        assert tr < 1e-11 and np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd), (prec, it, tr)
    # A = D: the right-preconditioned operator is the identity -> one step;
    # without the preconditioner (distinct entries) it needs more
    n = 40
    d = np.linspace(1.0, 50.0, n)
    rp1, col1 = np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32)
    x, it, tr = orc.bicgstab(rp1, col1, d, b[:n], tol=1e-14, precond=True)
    assert it == 1 and np.abs(x - b[:n] / d).max() < 1e-14
    x, it, tr = orc.bicgstab(rp1, col1, d, b[:n], tol=1e-14, precond=False)
    assert it > 1 and tr < 1e-13


def test_reference_operator_is_indefinite_under_free_surfaces():
    """Under the reference's block convention the clamped-bottom / free-side
    elastic operator has negative eigenvalues (nu=0.3), the textbook one does
    not.  CG still converges on it (checked above against dense LU)."""
    Sr, St = small_elastic(3, 0), small_elastic(3, 1)

    def dense(S):
        A = np.zeros((S.n, S.n))
        for i in range(S.n):
            A[i, S.col[S.rowptr[i]:S.rowptr[i + 1]]] = S.val[S.rowptr[i]:S.rowptr[i + 1]]
        return A
    assert np.linalg.eigvalsh(dense(Sr)).min() < 0
    assert np.linalg.eigvalsh(dense(St)).min() > 0


def test_patch_scalar_linear_solution_exact():
    xyz, tets = orc.mesh_box(4)
    bnd = np.nonzero((xyz == 0).any(1) | (xyz == 1).any(1))[0]
    g = 0.3 + 2 * xyz[bnd, 0] - xyz[bnd, 1] + 0.5 * xyz[bnd, 2]
    S = orc.scalar_system(xyz, tets, np.full(len(tets), 7.0), bnd, g)
    x, it, rr = orc.pcg(S.rowptr, S.col, S.val, S.rhs, tol=1e-13)
    exact = 0.3 + 2 * xyz[:, 0] - xyz[:, 1] + 0.5 * xyz[:, 2]
    assert abs(x - exact).max() < 1e-10


def test_patch_elastic_affine_solution_exact():
    """Affine displacement on the whole boundary, homogeneous E: P1 reproduces
    the affine field (the reference's own analytic check,
    MechanicalMain.cpp:2223-2300, and its homogenisation BCs, :243-290)."""
    xyz, tets = orc.mesh_box(3)
    G = np.array([[0.01, 0.002, 0], [0.001, -0.02, 0.003], [0, 0.004, 0.01]])
    bnd = np.nonzero((xyz == 0).any(1) | (xyz == 1).any(1))[0]
    for conv in (0, 1):
        S = orc.elastic_system(xyz, tets, np.full(len(tets), 1e9), NU, bnd, xyz[bnd] @ G.T, convention=conv)
        x, it, rr = orc.pcg(S.rowptr, S.col, S.val, S.rhs, tol=1e-13)
        m = S.node_map(len(xyz))
        u = x.reshape(-1, 3)[m]
        assert abs(u - xyz @ G.T).max() < 1e-11


def test_ssor_and_bicg_diag_restatements_reach_the_same_solution():
    xyz, tets = orc.mesh_box(4)
    sig = orc.field_inclusions(tets, xyz, 4, 0.1, 0.3, 2, 1.0, 100.0)
    fixed = np.nonzero((xyz[:, 2] == 0) | (xyz[:, 2] == 1))[0]
    S = orc.scalar_system(xyz, tets, sig, fixed, xyz[fixed, 2])
    xd = orc.dense_solve(S.rowptr, S.col, S.val, S.rhs)
    x, it, rr, _ = orc.refpath_solve(S.rowptr, S.col, S.val, S.rhs, S.rhs.copy(), tol=1e-12, mode=0)
    assert np.linalg.norm(x - xd) < 1e-9 * np.linalg.norm(xd)
    x, it, rr, _ = orc.refpath_solve(S.rowptr, S.col, S.val, S.rhs, tol=1e-12, mode=1)
    assert np.linalg.norm(x - xd) < 1e-9 * np.linalg.norm(xd)
    diag = np.array([S.val[S.rowptr[i]:S.rowptr[i + 1]][S.col[S.rowptr[i]:S.rowptr[i + 1]] == i][0]
                     for i in range(S.n)])
    x, it, res = orc.bicg_diag(S.rowptr, S.col, S.val, S.rhs, diag, E=1e-13)
    assert res < 1e-13  # BiCG_Diag returns the absolute preconditioned norm
    assert np.linalg.norm(x - xd) < 1e-9 * np.linalg.norm(xd)


# --- MsFEM cell ----------------------------------------------------------------

def test_msfem_cell_partition_of_unity_and_galerkin():
    n = 3
    E = orc.field_lognormal(6 * n ** 3, 1, 0, 0, np.log(1e9), 1.0)
    basis, Kc, it = orc.msfem_cell(n, (0, 0, 0), 0.5, E, NU, tol=1e-13)
    N = basis.shape[1]
    # sum over the 8 coarse nodes of basis (node, d) = e_d everywhere
    for d in range(3):
        s = basis[d::3].sum(0)
        e = np.zeros(N)
        e[d::3] = 1
        assert abs(s - e).max() < 1e-9
    assert np.allclose(Kc, Kc.T, atol=1e-9 * abs(Kc).max())
    for d in range(3):
        t = np.zeros(24)
        t[d::3] = 1
        assert abs(Kc @ t).max() < 1e-8 * abs(Kc).max()


def test_msfem_cell_homogeneous_affine_reproduction():
    """With homogeneous E, the sum of bases weighted by an affine field's
    nodal values reproduces the affine field only if the basis solves are
    right; on a cell whose boundary trace of a linear field is linear, the
    interior solution is that linear field."""
    n = 3
    E = np.full(6 * n ** 3, 1e9)
    basis, Kc, it = orc.msfem_cell(n, (0, 0, 0), 1.0, E, NU, tol=1e-13)
    xyz, _ = orc.mesh_box(n)
    G = np.array([[0.01, 0.002, 0], [0.001, -0.02, 0.003], [0, 0.004, 0.01]])
    corners = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], dtype=float)
    coef = (corners @ G.T).ravel()  # value at coarse node c, direction d -> index 3c+d
    u = coef @ basis
    assert abs(u.reshape(-1, 3) - xyz @ G.T).max() < 1e-10


# --- MsFEM reference-trace mode (edge -> face -> volume) ---------------------

def series_chain(xyz, E, nu, u0, un):
    """1-D chain of segments with stiffness blocks (E/L)((lam+mu) t t^T + mu I):
    the flux is the same in every segment, so u(t) = u0 + (sum_{s<t} S_s^-1) c
    with c = (sum_s S_s^-1)^-1 (un - u0)."""
    lam, mu = lame(1.0, nu)
    comp = []
    for s in range(len(E)):
        d = xyz[s + 1] - xyz[s]
        L = np.linalg.norm(d)
        t = d / L
        comp.append(np.linalg.inv((E[s] / L) * ((lam + mu) * np.outer(t, t) + mu * np.eye(3))))
    c = np.linalg.solve(sum(comp), np.asarray(un, float) - np.asarray(u0, float))
    U = [np.asarray(u0, float)]
    for S in comp:
        U.append(U[-1] + S @ c)
    return np.array(U)


@pytest.mark.parametrize("direction", [(1.0, 0, 0), (0.3, -0.5, 0.8)])
def test_edge_1d_matches_series_springs(direction):
    rng = np.random.default_rng(5)
    nseg = 7
    t = np.asarray(direction) / np.linalg.norm(direction)
    steps = rng.uniform(0.5, 1.5, nseg)  # unequal segment lengths
    xyz = np.vstack([np.zeros(3), np.cumsum(steps)[:, None] * t[None, :]]) + np.array([0.2, -0.1, 0.4])
    E = np.exp(rng.normal(size=nseg))
    u0, un = [1.0, -0.5, 0.25], [0.0, 0.3, -1.0]
    U = orc.edge_1d(xyz, E, NU, u0, un)
    assert abs(U - series_chain(xyz, E, NU, u0, un)).max() < 1e-13


def cst_plane_strain(P2, E, nu):
    """Plane-strain constant-strain triangle (textbook) for the in-plane
    components plus mu * P1 Laplacian for the normal component, 9x9 in
    node-major order with components (x, y, normal)."""
    lam, mu = lame(E, nu)
    M = np.c_[np.ones(3), P2]
    g = np.linalg.inv(M)[1:, :].T
    A = 0.5 * abs(np.linalg.det(M))
    Dp = np.array([[lam + 2 * mu, lam, 0], [lam, lam + 2 * mu, 0], [0, 0, mu]])
    K = np.zeros((9, 9))
    for I in range(3):
        for J in range(3):
            BI = np.array([[g[I, 0], 0], [0, g[I, 1]], [g[I, 1], g[I, 0]]])
            BJ = np.array([[g[J, 0], 0], [0, g[J, 1]], [g[J, 1], g[J, 0]]])
            K[3 * I:3 * I + 2, 3 * J:3 * J + 2] = A * BI.T @ Dp @ BJ
            K[3 * I + 2, 3 * J + 2] = A * mu * g[I] @ g[J]
    return K, A


def test_triangle_element_is_rotated_plane_strain():
    rng = np.random.default_rng(9)
    P2 = np.array([[0.0, 0.0], [0.7, 0.1], [0.2, 0.6]])
    Kp, A = cst_plane_strain(P2, 2.5, NU)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))  # rows: the plane's x, y axes and its normal
    P3 = np.c_[P2, np.zeros(3)] @ Q + np.array([0.3, -0.2, 0.5])
    K, area = orc.elem_tri(P3, 2.5, NU, convention=1)
    R = np.kron(np.eye(3), Q.T)
    assert area == pytest.approx(A, rel=1e-14)
    assert abs(K - R @ Kp @ R.T).max() < 1e-13 * abs(Kp).max()
    K0, _ = orc.elem_tri(P3, 2.5, NU, convention=0)
    for I, J in itertools.product(range(3), range(3)):
        assert abs(K0[3 * I:3 * I + 3, 3 * J:3 * J + 3] - K[3 * I:3 * I + 3, 3 * J:3 * J + 3].T).max() < 1e-13


def test_segment_element_closed_form():
    P = np.array([[0.1, 0.2, 0.3], [0.4, -0.1, 0.9]])
    K, L = orc.elem_seg(P, 3.0, NU, convention=0)
    d = P[1] - P[0]
    t = d / np.linalg.norm(d)
    lam, mu = lame(3.0, NU)
    S = ((lam + mu) * np.outer(t, t) + mu * np.eye(3)) / L
    assert L == pytest.approx(np.linalg.norm(d), rel=1e-15)
    assert abs(K - np.kron(np.array([[1, -1], [-1, 1]]), S)).max() < 1e-13 * abs(S).max()


def reftrace_cell(n, E):
    return orc.msfem_cell_reftrace(n, (0.0, 0.0, 0.0), 1.0, E, NU, tol=1e-13)


def test_reftrace_edges_are_harmonic_mean_profiles():
    """On coarse edge x (y = z = 0) the basis of corner 0 driven by e_d is
    e_d (1 - C(t)/C(n)), C(t) = sum_{s<t} 1/E_s, with E_s the modulus of the
    tet owning the boundary triangle on the face y = 0 that holds segment s
    (found here by brute force over the tets)."""
    n = 4
    E = orc.field_lognormal(6 * n ** 3, 1, 3, 0, 0.0, 1.0)
    basis, _, _ = reftrace_cell(n, E)
    xyz, tets = orc.mesh_box(n)
    s = n + 1
    Es = []
    for t in range(n):
        a, b = t, t + 1  # nodes (t,0,0), (t+1,0,0)
        owner = [k for k in range(len(tets)) if a in tets[k] and b in tets[k]
                 and sum(xyz[v][1] == 0 for v in tets[k]) == 3]
        assert len(owner) == 1
        Es.append(E[owner[0]])
    C = np.concatenate([[0.0], np.cumsum(1.0 / np.array(Es))])
    prof = 1 - C / C[-1]
    for d in range(3):
        phi = basis[d].reshape(-1, 3)
        e = np.zeros(3)
        e[d] = 1
        for t in range(s):
            assert abs(phi[t] - prof[t] * e).max() < 1e-13


def test_reftrace_partition_of_unity_and_coarse_invariants():
    n = 4
    E = orc.field_lognormal(6 * n ** 3, 1, 4, 0, np.log(1e9), 1.0)
    for form in (0, 1):
        basis, Kc, _ = orc.msfem_cell_reftrace(n, (0, 0, 0), 0.5, E, NU, tol=1e-13, form=form)
        N = basis.shape[1]
        for d in range(3):
            e = np.zeros(N)
            e[d::3] = 1
            assert abs(basis[d::3].sum(0) - e).max() < 1e-9, (form, d)
        assert np.allclose(Kc, Kc.T, atol=1e-9 * abs(Kc).max())
        for d in range(3):
            t = np.zeros(24)
            t[d::3] = 1
            assert abs(Kc @ t).max() < 1e-8 * abs(Kc).max()


def test_reftrace_face_traces_differ_from_hats_but_agree_for_homogeneous_affine():
    """Homogeneous material: every stage of the hierarchy reproduces affine
    data (the edge profile is linear, the face and volume problems have the
    affine field as their discrete solution), so the bases reproduce affine
    fields in both coarse forms; with a heterogeneous field the traces are
    not the hats."""
    n = 3
    xyz, _ = orc.mesh_box(n)
    G = np.array([[0.01, 0.002, 0], [0.001, -0.02, 0.003], [0, 0.004, 0.01]])
    corners = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], dtype=float)
    coef = (corners @ G.T).ravel()
    for form in (0, 1):
        basis, _, _ = orc.msfem_cell_reftrace(n, (0, 0, 0), 1.0, np.full(6 * n ** 3, 1e9), NU, tol=1e-13, form=form)
        u = coef @ basis
        assert abs(u.reshape(-1, 3) - xyz @ G.T).max() < 1e-10, form
    E = orc.field_lognormal(6 * n ** 3, 1, 5, 0, 0.0, 1.0)
    b_ref, _, _ = reftrace_cell(n, E)
    b_q1, _, _ = orc.msfem_cell(n, (0, 0, 0), 1.0, E, NU, tol=1e-13)
    assert abs(b_ref - b_q1).max() > 1e-3


# --- MLP oracle gradient --------------------------------------------------------

def test_mlp_oracle_gradient_central_differences():
    rng = np.random.default_rng(5)
    dims = [7, 5, 4, 3]
    Ws = [rng.normal(size=(dims[i + 1], dims[i])) * 0.5 for i in range(3)]
    bs = [rng.normal(size=dims[i + 1]) * 0.1 for i in range(3)]
    X = rng.normal(size=(6, 7))
    T = rng.normal(size=(6, 3))
    loss, gW, gb = orc.mlp_loss_grad(Ws, bs, X, T)
    h = 1e-6
    for l in range(3):
        for idx in [(0, 0), (1, 2), (dims[l + 1] - 1, dims[l] - 1)]:
            Wp = [W.copy() for W in Ws]
            Wm = [W.copy() for W in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            num = (orc.mlp_loss_grad(Wp, bs, X, T, False)[0] - orc.mlp_loss_grad(Wm, bs, X, T, False)[0]) / (2 * h)
            assert gW[l][idx] == pytest.approx(num, rel=1e-5, abs=1e-9)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][0] += h
        bm[l][0] -= h
        num = (orc.mlp_loss_grad(Ws, bp, X, T, False)[0] - orc.mlp_loss_grad(Ws, bm, X, T, False)[0]) / (2 * h)
        assert gb[l][0] == pytest.approx(num, rel=1e-5, abs=1e-9)


def test_bcg_stab2_restatement_matches_dense_solve():
    """BCG_Stab2 (unpreconditioned BiCGStab, CSSD_Matrix.h:634-729) reaches
    the dense solution of a non-symmetric system and of a diagonal one with
    distinct entries (with A = c I its first step has s = 0 and omega = 0/0,
    as in the reference, so that case is not used)."""
    rp, col, val, b = problems.convection_diffusion(10, 20.0)
    A = np.zeros((len(rp) - 1,) * 2)
    for i in range(len(rp) - 1):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    xd = np.linalg.solve(A, b)
    x, it, tr = orc.bcg_stab2(rp, col, val, b, E=1e-12)
    assert tr < 1e-11 and np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd), (it, tr)
    n = 30
    rp1, col1 = np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32)
    d = np.linspace(1.0, 9.0, n)
    x, it, tr = orc.bcg_stab2(rp1, col1, d, b[:n], E=1e-13)
    assert 1 <= it < n and tr < 1e-12 and np.abs(x - b[:n] / d).max() < 1e-11


def test_msfem_cell_on_the_reference_path_agrees_with_the_jacobi_path():
    """The CPU-baseline restatement of the reference's volume solves (full
    system, component-major DOFs, SSOR-PCG continuation with BCG_Stab2
    fallback) gives the same bases as the Jacobi-PCG cell solve to the level
    its stopping rule leaves (the best iterate of the step before the one
    that reaches the minimum residual), and the same coarse matrix."""
    n = 4
    E = orc.field_lognormal(6 * n ** 3, 1, 3, 0, np.log(1e9), 1.0)
    b0, K0, _ = orc.msfem_cell(n, (0, 0, 0), 0.25, E, NU, tol=1e-13)
    b1, K1, it = orc.msfem_cell_refpath(n, (0, 0, 0), 0.25, E, NU, min_residual=1e-10)
    assert it > 24
    assert np.linalg.norm(b1 - b0) <= 1e-6 * np.linalg.norm(b0)
    assert np.abs(K1 - K0).max() <= 1e-5 * np.abs(K0).max()


def test_reference_convention_two_phase_cube_defeats_bicg_diag_but_not_minres():
    """The 8^3 two-phase cube (E in {1e9, 2e11}, clamped bottom, free sides,
    reference element convention): the reference's own Mechanical solve,
    BiCG_Diag with diag = 1 in its continuation loop (needed = 10^-(ii+1)
    absolute, N/10 iterations per call, FEM.h:2399-2485), ends with a true
    relative residual above 1, while Jacobi MINRES (Elman-Silvester-Wathen
    Alg. 2.4, numpy) reaches 1e-10 within 2,000 steps.  The operator has 42
    negative eigenvalues and a Jacobi-scaled smallest |eigenvalue| of about
    2e-4 (textbook element: none negative)."""
    import importlib.util
    import scipy.sparse as sp
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("gpu_core", os.path.join(here, "test_gpu_core.py"))
    gc = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gc)
    p = problems.elastic_cube(8, field="grf")
    O = orc.elastic_system(p["xyz"], p["tets"], p["E"], p["nu"], p["fixed_nodes"], p["fixed_values"], p["tris"],
                           p["traction"], p["material"], convention=0)
    A = sp.csr_matrix((O.val, O.col, O.rowptr), shape=(O.n, O.n))
    b = O.rhs
    x, ii = O.x0.copy(), 1
    for _ in range(5):
        needed = 10.0 ** (-(ii + 1))
        x, it, res = orc.bicg_diag(O.rowptr, O.col, O.val, b, np.ones(O.n), x0=x, E=needed, maxiter=O.n // 10)
        if res < needed:
            ii += 1
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) > 1.0
    xm = gc.minres_esw(A, b, O.x0, 1.0 / A.diagonal(), 2000)
    assert np.linalg.norm(b - A @ xm) / np.linalg.norm(b) < 1e-10
    ev = np.linalg.eigvalsh((sp.diags(1 / np.sqrt(A.diagonal())) @ A @ sp.diags(1 / np.sqrt(A.diagonal()))).toarray())
    assert (ev < 0).sum() == 42 and np.abs(ev).min() < 5e-4
