"""GPU parity tests of the core path (SpMV, vector kernels, assembly, PCG,
mixed precision, GMRES) against the FP64 CPU oracle, through the C-ABI."""
import numpy as np
import pytest

import multiscalefemsolver3d_amd as mfs
import oracle as orc
from multiscalefemsolver3d_amd import problems

pytestmark = pytest.mark.gpu


def oracle_scalar(p):
    return orc.scalar_system(p["xyz"], p["tets"], p["sigma"], p["fixed_nodes"], p["fixed_values"])


def oracle_elastic(p, conv=0):
    return orc.elastic_system(p["xyz"], p["tets"], p["E"], p["nu"], p["fixed_nodes"], p["fixed_values"], p["tris"],
                              p["traction"], p["material"], convention=conv)


def assemble(ctx, p, conv=0):
    if p["problem"] == "scalar":
        S = mfs.Assembly(ctx, mfs.PROBLEM_SCALAR, p["xyz"], p["tets"], fixed_nodes=p["fixed_nodes"],
                         fixed_values=p["fixed_values"])
        S.run(p["sigma"])
    else:
        S = mfs.Assembly(ctx, mfs.PROBLEM_ELASTIC, p["xyz"], p["tets"], material=p["material"], renumber=True,
                         convention=conv, fixed_nodes=p["fixed_nodes"], fixed_values=p["fixed_values"],
                         tris=p["tris"], traction=p["traction"])
        S.run(p["E"], p["nu"])
    return S


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("kind,conv", [("scalar", 0), ("elastic", mfs.ELEM_REFERENCE), ("elastic", mfs.ELEM_TEXTBOOK)])
def test_fused_slice_assembly_matches_general_path(ctx, kind, conv, monkeypatch):
    """The per-slice assembly (tet records in LDS, one pass writing the FP64
    and FP32 values and the inverse diagonal, fixed DOFs in the same launch)
    gives the general path's bits (tet records in HBM, gather, elimination,
    derivation): matrix values, right-hand side, start vector, and a mixed
    PCG run (which reads the FP32 values and the FP32 inverse diagonal)
    with the same iterations and solution, for every chunk height (64, 32
    and 16 block rows per workgroup).  MFS_ASSEMBLY_SLICES=0 selects the
    general path, MFS_ASSEMBLY_CHUNK the chunk height."""
    if kind == "scalar":
        p = problems.laplace_inclusions(14, count=40, rmin=0.05, rmax=0.15)
    else:
        p = problems.elastic_cube(9, field="grf")
    out = []
    for flag, chunk in (("0", "64"), ("1", "64"), ("1", "32"), ("1", "16")):
        monkeypatch.setenv("MFS_ASSEMBLY_SLICES", flag)
        monkeypatch.setenv("MFS_ASSEMBLY_CHUNK", chunk)
        S = assemble(ctx, p, conv)
        rp, col, val = S.A.export_csr()
        x = mfs.Vec(ctx, S.A.n)
        x.copy_from(S.x0)
        # 200 iterations at most: identical iterates are the point, not convergence
        method = mfs.PCG_JACOBI if kind == "scalar" or conv == mfs.ELEM_TEXTBOOK else mfs.MINRES
        rep = mfs.Solver(S.A, precision=mfs.PREC_MIXED, tol=1e-9, maxit=200, method=method).run(S.b, x, check=False)
        out.append((rp, col, val, S.b.download(), S.x0.download(), x.download(), rep.iterations))
    rp0, c0, v0, b0, x00, xs0, it0 = out[0]
    for rp1, c1, v1, b1, x01, xs1, it1 in out[1:]:
        assert np.array_equal(rp1, rp0) and np.array_equal(c1, c0)
        assert np.array_equal(v1, v0), np.abs(v1 - v0).max()
        assert np.array_equal(b1, b0) and np.array_equal(x01, x00)
        assert it1 == it0 and np.array_equal(xs1, xs0)


# --- SpMV / BLAS -----------------------------------------------------------------

@pytest.mark.parametrize("kind", ["csr", "bsr3"])
def test_spmv_matches_oracle(ctx, kind):
    rng = np.random.default_rng(11)
    if kind == "csr":
        x, t = mfs.mesh_box(7)
        rp, col = mfs.portrait(len(x), t)
        val = rng.normal(size=len(col))
        A = mfs.Mat.csr(ctx, rp, col, val)
        srp, scol, sval = rp, col, val
    else:
        x, t = mfs.mesh_box(5)
        rp, col = mfs.portrait(len(x), t)
        val = rng.normal(size=(len(col), 9))
        A = mfs.Mat.bsr3(ctx, rp, col, val)
        srp, scol, sval = A.export_csr()
        # export is the MakeCopyMatrix expansion of the uploaded blocks
        k = 0
        for i in range(len(rp) - 1):
            for a in range(3):
                for j in range(rp[i], rp[i + 1]):
                    for b in range(3):
                        assert scol[k] == 3 * col[j] + b and sval[k] == val[j, 3 * a + b]
                        k += 1
    n = A.n
    xv = rng.normal(size=n)
    y_ref = orc.csr_mv(srp, scol, sval, xv)
    X = mfs.Vec(ctx, n, mfs.F64, xv)
    Y = mfs.Vec(ctx, n, mfs.F64)
    A.spmv(X, Y, mfs.F64)
    assert rel(Y.download(), y_ref) < 1e-14
    X32 = mfs.Vec(ctx, n, mfs.F32, xv)
    Y32 = mfs.Vec(ctx, n, mfs.F32)
    A.spmv(X32, Y32, mfs.F32)
    assert rel(Y32.download(), y_ref) < 2e-6


def test_dot_axpby(ctx):
    rng = np.random.default_rng(12)
    a, b = rng.normal(size=100001), rng.normal(size=100001)
    for prec, tol in ((mfs.F64, 1e-13), (mfs.F32, 1e-5)):
        X, Y = mfs.Vec(ctx, len(a), prec, a), mfs.Vec(ctx, len(a), prec, b)
        assert mfs.dot(X, Y) == pytest.approx(float(a @ b), rel=tol, abs=tol * np.linalg.norm(a) * np.linalg.norm(b))
        mfs.axpby(2.0, X, -0.5, Y)
        assert rel(Y.download(), 2 * a - 0.5 * b) < tol


def test_empty_and_tiny_inputs(ctx):
    # a 1-hex mesh (8 nodes) still assembles and solves
    p = problems.laplace_inclusions(1, count=0)
    S = assemble(ctx, p)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, tol=1e-12).run(S.b, x)
    assert rep.true_relres <= 1e-12
    assert np.allclose(x.download(), p["xyz"][:, 2])
    v = mfs.Vec(ctx, 0)
    assert v.download().size == 0


# --- assembly ---------------------------------------------------------------------

def test_scalar_assembly_matches_oracle(ctx):
    p = problems.laplace_inclusions(8, count=30, rmin=0.1, rmax=0.2)
    S = assemble(ctx, p)
    O = oracle_scalar(p)
    rp, col, val = S.A.export_csr()
    assert np.array_equal(rp, O.rowptr) and np.array_equal(col, O.col)
    scale = abs(O.val).max()
    assert abs(val - O.val).max() <= 1e-13 * scale
    assert abs(S.b.download() - O.rhs).max() <= 1e-13 * max(1.0, abs(O.rhs).max())
    assert np.array_equal(S.x0.download() != 0, O.x0 != 0)


@pytest.mark.parametrize("conv", [mfs.ELEM_REFERENCE, mfs.ELEM_TEXTBOOK])
def test_elastic_assembly_matches_oracle(ctx, conv):
    p = problems.elastic_cube(5, field="grf")
    S = assemble(ctx, p, conv)
    O = oracle_elastic(p, conv)
    assert np.array_equal(S.node_map(), O.node_map(len(p["xyz"])))
    rp, col, val = S.A.export_csr()
    assert np.array_equal(rp, O.rowptr) and np.array_equal(col, O.col)
    scale = abs(O.val).max()
    assert abs(val - O.val).max() <= 1e-13 * scale
    assert rel(S.b.download(), O.rhs) < 1e-13


# --- solvers ----------------------------------------------------------------------

def test_pcg_f64_matches_oracle_scalar(ctx):
    p = problems.laplace_inclusions(10, count=40, rmin=0.05, rmax=0.15)
    S = assemble(ctx, p)
    O = oracle_scalar(p)
    xo, _, rr = orc.pcg(O.rowptr, O.col, O.val, O.rhs, O.x0, tol=1e-13)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, tol=1e-11).run(S.b, x)
    assert rep.status == 0 and rep.true_relres <= 1e-11
    assert rel(x.download(), xo) < 1e-9


@pytest.mark.parametrize("method", [mfs.PCG_JACOBI, mfs.CG])
def test_mixed_refinement_reaches_fp64_tolerance(ctx, method):
    p = problems.laplace_inclusions(12)
    S = assemble(ctx, p)
    O = oracle_scalar(p)
    xo, _, _ = orc.pcg(O.rowptr, O.col, O.val, O.rhs, O.x0, tol=1e-13)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, method=method, precision=mfs.PREC_MIXED, tol=1e-10, inner_tol=1e-4).run(S.b, x)
    assert rep.true_relres <= 1e-10 and rep.outer_steps >= 2
    assert rel(x.download(), xo) < 1e-8


@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
def test_finish_sum_kernels_give_the_same_solve(ctx, monkeypatch, precision):
    """Large systems add the per-block partial sums once in a separate
    kernel instead of in every consumer block (MFS_FINISH_SUMS forces either
    path): both reach the tolerance with the same iterations and solution."""
    p = problems.laplace_inclusions(16)
    S = assemble(ctx, p)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("MFS_FINISH_SUMS", mode)
        x = mfs.Vec(ctx, S.A.n)
        x.copy_from(S.x0)
        rep = mfs.Solver(S.A, precision=precision, tol=1e-10).run(S.b, x)
        out[mode] = (rep, x.download())
        xs = [mfs.Vec(ctx, S.A.n) for _ in range(3)]
        for v in xs:
            v.copy_from(S.x0)
        reps = mfs.solve_multi(S.A, [S.b, S.b, S.b], xs, precision=precision, tol=1e-10)
        assert all(r.true_relres <= 1e-10 for r in reps)
        assert rel(xs[2].download(), out[mode][1]) < 1e-8
    r0, r1 = out["0"][0], out["1"][0]
    assert r0.true_relres <= 1e-10 and r1.true_relres <= 1e-10
    assert abs(r0.iterations - r1.iterations) <= 2
    assert rel(out["0"][1], out["1"][1]) < 1e-9


def test_mixed_cg_keeps_krylov_information(ctx):
    """The FP32 iteration is continued across FP64 residual replacements (the
    search direction is kept), so the mixed solve needs about as many
    iterations as FP64 PCG.  On this system (24^3, inner reduction 1e-3,
    tol 1e-9) FP64 PCG takes 381 iterations; restarting the FP32 iteration at
    each of the 3 replacements takes 485 (numpy emulation of the same
    algorithm), continuing it 427."""
    p = problems.laplace_inclusions(24)
    S = assemble(ctx, p)
    its = {}
    for prec in (mfs.PREC_F64, mfs.PREC_MIXED):
        x = mfs.Vec(ctx, S.A.n)
        x.copy_from(S.x0)
        rep = mfs.Solver(S.A, precision=prec, tol=1e-9, inner_tol=1e-3).run(S.b, x)
        assert rep.status == 0 and rep.true_relres <= 1e-9
        its[prec] = (rep.iterations, rep.outer_steps)
    assert its[mfs.PREC_MIXED][1] >= 3
    assert its[mfs.PREC_MIXED][0] <= 1.15 * its[mfs.PREC_F64][0] + 5, its


@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
def test_gmres_elastic_textbook_operator(ctx, precision):
    """GMRES(30) on the two-phase elastic system with the textbook (SPD)
    element convention.  (On the reference convention the clamped-bottom /
    free-side operator is indefinite and restarted GMRES(30) stagnates - the
    oracle's GMRES does too - while PCG converges.)"""
    p = problems.elastic_cube(5, field="grf")
    S = assemble(ctx, p, mfs.ELEM_TEXTBOOK)
    O = oracle_elastic(p, mfs.ELEM_TEXTBOOK)
    xd = orc.dense_solve(O.rowptr, O.col, O.val, O.rhs)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, method=mfs.GMRES, precision=precision, tol=1e-9, restart=30, maxit=20000).run(S.b, x)
    assert rep.true_relres <= 1e-9
    assert rel(x.download(), xd) < 1e-6
    # same number of Arnoldi steps as the oracle's GMRES(30) within a restart
    _, it_o, _ = orc.gmres(O.rowptr, O.col, O.val, O.rhs, m=30, x0=O.x0, tol=1e-9, maxit=20000)
    if precision == mfs.PREC_F64:
        assert abs(rep.iterations - it_o) <= 30


@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
def test_gmres_scalar_matches_oracle(ctx, precision):
    p = problems.laplace_inclusions(10, count=40, rmin=0.05, rmax=0.15)
    S = assemble(ctx, p)
    O = oracle_scalar(p)
    xo, _, _ = orc.pcg(O.rowptr, O.col, O.val, O.rhs, O.x0, tol=1e-13)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, method=mfs.GMRES, precision=precision, tol=1e-10, restart=30).run(S.b, x)
    assert rep.true_relres <= 1e-10
    assert rel(x.download(), xo) < 1e-8


def test_pcg_elastic_matches_dense(ctx):
    p = problems.elastic_cube(4)
    S = assemble(ctx, p)
    O = oracle_elastic(p)
    xd = orc.dense_solve(O.rowptr, O.col, O.val, O.rhs)
    for prec in (mfs.PREC_F64, mfs.PREC_MIXED):
        x = mfs.Vec(ctx, S.A.n)
        x.copy_from(S.x0)
        rep = mfs.Solver(S.A, precision=prec, tol=1e-10).run(S.b, x)
        assert rep.true_relres <= 1e-10
        assert rel(x.download(), xd) < 1e-7


def test_failure_paths(ctx):
    p = problems.laplace_inclusions(4, count=0)
    S = assemble(ctx, p)
    x = mfs.Vec(ctx, S.A.n)
    rep = mfs.Solver(S.A, tol=1e-12, maxit=1).run(S.b, x, check=False)
    assert rep.status == -3 and rep.iterations == 1
    bad = S.b.download()
    bad[3] = np.nan
    B = mfs.Vec(ctx, len(bad), mfs.F64, bad)
    with pytest.raises(mfs.MfsError) as e:
        mfs.Solver(S.A, tol=1e-8).run(B, x)
    assert e.value.code == -4


def test_solver_reuse_same_results(ctx):
    p = problems.laplace_inclusions(10)
    S = assemble(ctx, p)
    sol = mfs.Solver(S.A, precision=mfs.PREC_MIXED, tol=1e-8)
    outs = []
    for _ in range(3):
        x = mfs.Vec(ctx, S.A.n)
        x.copy_from(S.x0)
        sol.run(S.b, x)
        outs.append(x.download())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


@pytest.mark.slow
def test_laplace64_full_size_mixed_vs_fp64(ctx):
    """64^3 Laplace with 200 inclusions: mixed reaches the FP64 tolerance and
    agrees with the FP64 GPU PCG solution; the residual is re-checked in numpy
    from the exported matrix."""
    p = problems.laplace_inclusions(64)
    S = assemble(ctx, p)
    xs = {}
    for prec in (mfs.PREC_F64, mfs.PREC_MIXED):
        x = mfs.Vec(ctx, S.A.n)
        x.copy_from(S.x0)
        rep = mfs.Solver(S.A, precision=prec, tol=1e-8).run(S.b, x)
        assert rep.true_relres <= 1e-8
        xs[prec] = x.download()
    rp, col, val = S.A.export_csr()
    b = S.b.download()
    r = b - orc.csr_mv(rp, col, val, xs[mfs.PREC_MIXED])
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.0001e-8
    assert rel(xs[mfs.PREC_MIXED], xs[mfs.PREC_F64]) < 1e-6


@pytest.mark.parametrize("method,precision,tol,bound", [(mfs.PCG_JACOBI, mfs.PREC_F64, 1e-12, 1e-8),
                                                         (mfs.MINRES, mfs.PREC_F64, 1e-12, 1e-8),
                                                         (mfs.MINRES, mfs.PREC_MIXED, 1e-12, 1e-8)])
def test_elastic20_full_size_vs_oracle(ctx, method, precision, tol, bound):
    """20^3 elastic cube (27,783 DOFs; clamped bottom, pressure on top,
    reference element convention): the GPU solution matches the oracle's FP64
    CPU PCG solution on the identical seeded mesh (relative L2 error).  The
    operator is symmetric indefinite: FP64 PCG still converges, FP32 CG does
    not (the mixed path uses MINRES here)."""
    p = problems.elastic_cube(20)
    S = assemble(ctx, p)
    O = oracle_elastic(p)
    xo, _, _ = orc.pcg(O.rowptr, O.col, O.val, O.rhs, O.x0, tol=1e-13)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, method=method, precision=precision, tol=tol, maxit=200000).run(S.b, x)
    assert rep.status == 0 and rep.true_relres <= tol
    err = rel(x.download(), xo)
    print(f"20^3 elastic, method {method} precision {precision}: relative L2 error vs oracle {err:.2e}")
    assert err < bound


@pytest.mark.slow
def test_laplace64_vs_oracle_solution(ctx):
    """64^3 Laplace, 200 inclusions: the mixed-precision GPU solution matches
    the oracle's FP64 CPU PCG solution to a relative L2 error below 1e-8."""
    p = problems.laplace_inclusions(64)
    S = assemble(ctx, p)
    O = oracle_scalar(p)
    xo, _, _ = orc.pcg(O.rowptr, O.col, O.val, O.rhs, O.x0, tol=1e-12)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, precision=mfs.PREC_MIXED, tol=1e-11).run(S.b, x)
    assert rep.status == 0 and rep.true_relres <= 1e-11
    err = rel(x.download(), xo)
    print(f"64^3 Laplace, mixed: relative L2 error vs oracle {err:.2e}")
    assert err < 1e-8


@pytest.mark.parametrize("kind", ["scalar", "elastic"])
@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
@pytest.mark.parametrize("k", [1, 3, 5, 6, 8])
def test_multi_rhs_matches_single_solves(ctx, kind, precision, k):
    """k systems sharing one matrix, solved as one iteration stream (one
    SpMM per iteration), give the solutions of k separate solves; the padding
    systems of the interleaved layout (k = 1, 3 -> 4, 5 -> 6) stay inert, and
    the 6-system layout (24-byte FP32 elements, 8-byte loads) matches too."""
    if kind == "scalar":
        p = problems.laplace_inclusions(10, count=30, rmin=0.05, rmax=0.15)
        S = assemble(ctx, p)
    else:
        p = problems.elastic_cube(5, field="grf")
        S = assemble(ctx, p, mfs.ELEM_TEXTBOOK)
    n = S.A.n
    rng = np.random.default_rng(40 + k)
    b0 = S.b.download()
    bs = [mfs.Vec(ctx, n, mfs.F64, b0 * (1.0 + 0.5 * j) + rng.normal(size=n) * np.abs(b0).max() * 0.1)
          for j in range(k)]
    xs = []
    for j in range(k):
        x = mfs.Vec(ctx, n)
        x.copy_from(S.x0)
        xs.append(x)
    reps = mfs.solve_multi(S.A, bs, xs, precision=precision, tol=1e-10)
    rp, col, val = S.A.export_csr()
    for j in range(k):
        assert reps[j].status == 0 and reps[j].true_relres <= 1e-10, (j, reps[j])
        xm = xs[j].download()
        bj = bs[j].download()
        r = bj - orc.csr_mv(rp, col, val, xm)
        assert np.linalg.norm(r) / np.linalg.norm(bj) <= 1.01e-10
        xsgl = mfs.Vec(ctx, n)
        xsgl.copy_from(S.x0)
        ref = mfs.Solver(S.A, precision=mfs.PREC_F64, tol=1e-11).run(bs[j], xsgl, check=False)
        assert ref.true_relres <= 1e-10
        assert rel(xm, xsgl.download()) < 1e-7, j


@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
def test_minres_reference_elastic_operator(ctx, precision):
    """MINRES (Jacobi) on the indefinite reference-convention elastic system
    (two-phase E, clamped bottom, free sides) matches dense LU."""
    p = problems.elastic_cube(5, field="grf")
    S = assemble(ctx, p)
    O = oracle_elastic(p)
    xd = orc.dense_solve(O.rowptr, O.col, O.val, O.rhs)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, method=mfs.MINRES, precision=precision, tol=1e-9, maxit=20000).run(S.b, x)
    assert rep.true_relres <= 1e-9
    assert rel(x.download(), xd) < 1e-6


def test_minres_scalar_matches_pcg(ctx):
    p = problems.laplace_inclusions(8, count=20, rmin=0.1, rmax=0.2)
    S = assemble(ctx, p)
    O = oracle_scalar(p)
    xo, _, _ = orc.pcg(O.rowptr, O.col, O.val, O.rhs, O.x0, tol=1e-13)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, method=mfs.MINRES, tol=1e-10).run(S.b, x)
    assert rep.true_relres <= 1e-10 and rel(x.download(), xo) < 1e-8


@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
def test_bicgstab_nonsymmetric_matches_oracle(ctx, precision):
    """BiCGStab (right Jacobi) on an upwind convection-diffusion system:
    solution against the oracle's BiCGStab and numpy's dense solve; in FP64
    the iteration count follows the oracle's."""
    rp, col, val, b = problems.convection_diffusion(40, 60.0, seed=3)
    n = len(b)
    A = mfs.Mat.csr(ctx, rp, col, val)
    bv, x = mfs.Vec(ctx, n, mfs.F64, b), mfs.Vec(ctx, n, mfs.F64, np.zeros(n))
    rep = mfs.Solver(A, method=mfs.BICGSTAB, precision=precision, tol=1e-10).run(bv, x)
    assert rep.status == 0 and rep.true_relres <= 1e-10, rep
    xo, it_o, tr = orc.bicgstab(rp, col, val, b, tol=0.5e-10)
    assert tr <= 1e-10
    xg = x.download()
    assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)
    Ad = np.zeros((n, n))
    for i in range(n):
        Ad[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    xd = np.linalg.solve(Ad, b)
    assert np.linalg.norm(xg - xd) <= 1e-8 * np.linalg.norm(xd)
    if precision == mfs.PREC_F64:
        assert abs(rep.iterations - it_o) <= 0.15 * it_o + 3, (rep.iterations, it_o)


@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
def test_bicgstab_reduction_and_index_paths_agree(ctx, monkeypatch, precision):
    """The BiCGStab kernels sum their producer's partials in every block, or
    read totals that a separate kernel added once (MFS_FINISH_SUMS), and read
    16-bit or 32-bit column indices (MFS_NO_COL16 at matrix creation): all
    four paths reach the tolerance with the same iterations and solution."""
    rp, col, val, b = problems.convection_diffusion(48, 40.0, seed=5)
    n = len(b)
    out = {}
    for fin in ("0", "1"):
        for col16 in ("0", "1"):
            monkeypatch.setenv("MFS_FINISH_SUMS", fin)
            if col16 == "1":
                monkeypatch.setenv("MFS_NO_COL16", "1")
            else:
                monkeypatch.delenv("MFS_NO_COL16", raising=False)
            A = mfs.Mat.csr(ctx, rp, col, val)
            bv, x = mfs.Vec(ctx, n, mfs.F64, b), mfs.Vec(ctx, n, mfs.F64, np.zeros(n))
            rep = mfs.Solver(A, method=mfs.BICGSTAB, precision=precision, tol=1e-10).run(bv, x)
            assert rep.status == 0 and rep.true_relres <= 1e-10, (fin, col16, rep)
            out[fin, col16] = (rep.iterations, x.download())
    it0, x0 = out["0", "0"]
    for key, (it, xk) in out.items():
        assert abs(it - it0) <= 0.05 * it0 + 2, (key, it, it0)
        assert np.linalg.norm(xk - x0) <= 1e-8 * np.linalg.norm(x0), key


def test_bicgstab_reference_elastic_operator(ctx):
    """The reference element convention's two-phase elastic system (symmetric
    indefinite, clamped bottom, free sides): BiCGStab reaches dense LU's
    solution."""
    p = problems.elastic_cube(5, field="grf")
    S = assemble(ctx, p)
    O = oracle_elastic(p)
    xd = orc.dense_solve(O.rowptr, O.col, O.val, O.rhs)
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, method=mfs.BICGSTAB, tol=1e-10, maxit=50000).run(S.b, x)
    assert rep.status == 0 and rep.true_relres <= 1e-10, rep
    assert rel(x.download(), xd) < 1e-7


@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
def test_slice_order_paths_agree(ctx, monkeypatch, precision):
    """The SELL kernels take one contiguous range of slices per XCD (and the
    update kernels the matching rows) or keep the dispatch order
    (MFS_XCD_MAP=0/1, read at matrix creation): PCG and BiCGStab reach the
    tolerance in both orders with the same iterations (up to the rounding of
    the differently grouped partial sums) and solutions."""
    p = problems.laplace_inclusions(16)
    rp, col, val, b = problems.convection_diffusion(48, 40.0, seed=8)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("MFS_XCD_MAP", mode)
        S = assemble(ctx, p)
        x = mfs.Vec(ctx, S.A.n)
        x.copy_from(S.x0)
        rep = mfs.Solver(S.A, precision=precision, tol=1e-10).run(S.b, x)
        assert rep.status == 0 and rep.true_relres <= 1e-10, (mode, rep)
        A = mfs.Mat.csr(ctx, rp, col, val)
        bv, xb = mfs.Vec(ctx, len(b), mfs.F64, b), mfs.Vec(ctx, len(b), mfs.F64, np.zeros(len(b)))
        rb = mfs.Solver(A, method=mfs.BICGSTAB, precision=precision, tol=1e-10).run(bv, xb)
        assert rb.status == 0 and rb.true_relres <= 1e-10, (mode, rb)
        out[mode] = (rep.iterations, x.download(), rb.iterations, xb.download())
    (i0, x0, j0, y0), (i1, x1, j1, y1) = out["0"], out["1"]
    assert abs(i0 - i1) <= 0.02 * i0 + 2 and abs(j0 - j1) <= 0.05 * j0 + 3, (i0, i1, j0, j1)
    assert rel(x1, x0) < 1e-8 and rel(y1, y0) < 1e-8


@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
def test_gmres_graph_and_direct_launches_agree(ctx, monkeypatch, precision):
    """GMRES replays each Arnoldi step from a graph captured per step index,
    or launches the same kernels directly (MFS_NO_GRAPH=1): identical steps
    and solutions."""
    rp, col, val, b = problems.convection_diffusion(40, 60.0, seed=4)
    n = len(b)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("MFS_NO_GRAPH", mode)
        A = mfs.Mat.csr(ctx, rp, col, val)
        bv, x = mfs.Vec(ctx, n, mfs.F64, b), mfs.Vec(ctx, n, mfs.F64, np.zeros(n))
        rep = mfs.Solver(A, method=mfs.GMRES, precision=precision, tol=1e-9).run(bv, x)
        assert rep.status == 0 and rep.true_relres <= 1e-9, (mode, rep)
        out[mode] = (rep.iterations, x.download())
    assert out["0"][0] == out["1"][0]
    assert rel(out["0"][1], out["1"][1]) < 1e-12


@pytest.mark.parametrize("maxit", [1, 7, 13])
def test_gmres_and_bicgstab_stop_at_maxit(ctx, maxit):
    """An iteration cap that falls inside GMRES's 6-step check interval or
    BiCGStab's 8-iteration graph chunk is honoured exactly: the report gives
    the capped count and NOT_CONVERGED, and the residual is finite (GMRES's
    cannot grow)."""
    rp, col, val, b = problems.convection_diffusion(64, 100.0, seed=6)
    n = len(b)
    A = mfs.Mat.csr(ctx, rp, col, val)
    bv = mfs.Vec(ctx, n, mfs.F64, b)
    for method in (mfs.GMRES, mfs.BICGSTAB):
        x = mfs.Vec(ctx, n, mfs.F64, np.zeros(n))
        rep = mfs.Solver(A, method=method, precision=mfs.PREC_F64, tol=1e-12, maxit=maxit).run(bv, x, check=False)
        assert rep.iterations == maxit, (method, maxit, rep)
        assert rep.status == -3, (method, rep)  # MFS_ERR_NOT_CONVERGED
        assert 0 < rep.true_relres < (1.0 if method == mfs.GMRES else 10.0), (method, rep)


def minres_esw(A, b, x0, dinv, steps):
    """Jacobi-preconditioned MINRES, Elman, Silvester & Wathen, "Finite
    Elements and Fast Iterative Solvers", Algorithm 2.4, in numpy (FP64):
    the iterate after `steps` Lanczos steps from x0."""
    x = x0.copy()
    v_old = np.zeros_like(b)
    v = b - A @ x
    z = dinv * v
    g_old, g = 1.0, np.sqrt(z @ v)
    eta = g
    s_old = s = 0.0
    c_old = c = 1.0
    w_old = np.zeros_like(b)
    w = np.zeros_like(b)
    for _ in range(steps):
        z = z / g
        q = A @ z
        delta = q @ z
        v_new = q - (delta / g) * v - (g / g_old) * v_old
        z_new = dinv * v_new
        g_new = np.sqrt(z_new @ v_new)
        a0 = c * delta - c_old * s * g
        a1 = np.hypot(a0, g_new)
        a2 = s * delta + c_old * c * g
        a3 = s_old * g
        c_new, s_new = a0 / a1, g_new / a1
        w_new = (z - a3 * w_old - a2 * w) / a1
        x = x + c_new * eta * w_new
        eta = -s_new * eta
        v_old, v, z = v, v_new, z_new
        g_old, g = g, g_new
        w_old, w = w, w_new
        c_old, c, s_old, s = c, c_new, s, s_new
    return x


@pytest.mark.parametrize("maxit", [1, 7, 13])
def test_minres_stops_at_maxit_with_the_textbook_iterate(ctx, maxit):
    """An iteration cap inside MINRES's 8-step check interval is honoured
    exactly (capped count, NOT_CONVERGED, finite residual), and the FP64
    iterate after `maxit` steps is the one Algorithm 2.4 of Elman, Silvester
    & Wathen gives on the same indefinite system (reference-convention
    elastic operator, two-phase E, free sides): pins the device recurrence
    step by step."""
    import scipy.sparse as sp
    p = problems.elastic_cube(4, field="grf")
    S = assemble(ctx, p)
    O = oracle_elastic(p)
    A = sp.csr_matrix((O.val, O.col, O.rowptr), shape=(O.n, O.n))
    x = mfs.Vec(ctx, S.A.n)
    x.copy_from(S.x0)
    rep = mfs.Solver(S.A, method=mfs.MINRES, precision=mfs.PREC_F64, tol=1e-14, maxit=maxit).run(S.b, x, check=False)
    assert rep.iterations == maxit, rep
    assert rep.status == -3, rep  # MFS_ERR_NOT_CONVERGED
    assert np.isfinite(rep.true_relres) and rep.true_relres < 1.0, rep
    xe = minres_esw(A, O.rhs, O.x0, 1.0 / A.diagonal(), maxit)
    assert rel(x.download(), xe) < 1e-9, rel(x.download(), xe)
    r = O.rhs - A @ xe
    assert abs(np.linalg.norm(r) / np.linalg.norm(O.rhs) - rep.true_relres) < 1e-9
