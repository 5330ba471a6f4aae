"""mfs_solve_noalloc (no device allocation, fixed library scratch) against
mfs_solve through the C-ABI: two calls in a row, the second on a smaller
system, in FP64 and mixed precision; a system larger than the scratch is
rejected with an error."""
import ctypes as C

import numpy as np
import pytest

import multiscalefemsolver3d_amd as mfs
from multiscalefemsolver3d_amd import problems

pytestmark = pytest.mark.gpu

lib = mfs.lib
lib.mfs_solve_noalloc.restype = C.c_int
lib.mfs_solve_noalloc.argtypes = [C.c_void_p, C.POINTER(mfs.SolveDesc), C.c_void_p, C.c_void_p,
                                  C.POINTER(mfs.SolveReport)]


def solve_both(ctx, n, precision):
    p = problems.laplace_inclusions(n, count=30, rmin=0.05, rmax=0.15, seed=7)
    S = mfs.Assembly(ctx, mfs.PROBLEM_SCALAR, p["xyz"], p["tets"], fixed_nodes=p["fixed_nodes"],
                     fixed_values=p["fixed_values"])
    S.run(p["sigma"])
    d = mfs.solve_desc(precision=precision, tol=1e-10)
    xa = mfs.Vec(ctx, S.A.n)
    xa.copy_from(S.x0)
    ra = mfs.SolveReport()
    assert lib.mfs_solve(S.A.h, C.byref(d), S.b.h, xa.h, C.byref(ra)) == 0, lib.mfs_last_error()
    xb = mfs.Vec(ctx, S.A.n)
    xb.copy_from(S.x0)
    rb = mfs.SolveReport()
    assert lib.mfs_solve_noalloc(S.A.h, C.byref(d), S.b.h, xb.h, C.byref(rb)) == 0, lib.mfs_last_error()
    a, b = xa.download(), xb.download()
    S.close()
    return ra, rb, a, b


@pytest.mark.parametrize("precision", [mfs.PREC_F64, mfs.PREC_MIXED])
def test_noalloc_matches_solve_twice(ctx, precision):
    for n in (24, 10):  # the second call is smaller: stale scratch contents must not matter
        ra, rb, a, b = solve_both(ctx, n, precision)
        print(n, precision, ra.iterations, rb.iterations, ra.true_relres, rb.true_relres)
        assert rb.status == 0 and rb.true_relres <= 1e-10
        assert rb.iterations == ra.iterations and rb.outer_steps == ra.outer_steps
        assert np.linalg.norm(b - a) <= 1e-12 * np.linalg.norm(a)


def test_noalloc_rejects_a_system_larger_than_the_scratch(ctx):
    n = 1_200_000
    rp = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    A = mfs.Mat.csr(ctx, rp, col, np.full(n, 2.0))
    b = mfs.Vec(ctx, n, mfs.F64, np.ones(n))
    x = mfs.Vec(ctx, n, mfs.F64, np.zeros(n))
    d = mfs.solve_desc(precision=mfs.PREC_F64, tol=1e-10)
    r = mfs.SolveReport()
    assert lib.mfs_solve_noalloc(A.h, C.byref(d), b.h, x.h, C.byref(r)) == -1  # MFS_ERR_ARG
    assert b"bytes of scratch" in lib.mfs_last_error()
    # the rejected call left nothing behind: a small solve still works
    ra, rb, a, bb = solve_both(ctx, 6, mfs.PREC_F64)
    assert rb.status == 0 and np.linalg.norm(bb - a) <= 1e-12 * np.linalg.norm(a)
