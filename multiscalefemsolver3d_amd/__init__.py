"""Python view of libmfs.so, the MI355X (gfx950) FEM solve path.

The product is the C-ABI library declared in include/mfs.h; this module only
binds it with ctypes so tests and benchmarks can drive it from numpy.  Every
compute call goes to the HIP kernels in libmfs.so: if the library is missing
the import fails, there is no Python fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfs.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
        "or make -C multiscalefemsolver3d_amd)")
lib = C.CDLL(LIB_PATH)

# Objects still alive when the interpreter exits are not destroyed through
# the library: the process is ending and the device runtime may already be
# tearing down, so finalizers only release device memory before that point.
_EXITING = False


def _at_exit():
    global _EXITING
    _EXITING = True


atexit.register(_at_exit)

MFS_OK = 0
ERRORS = {-1: "ARG", -2: "OOM", -3: "NOT_CONVERGED", -4: "NONFINITE", -5: "HIP", -6: "UNSUPPORTED",
          -7: "NO_DEVICE"}
F64, F32 = 0, 1
PCG_JACOBI, CG, GMRES, MINRES, BICGSTAB = 0, 1, 2, 3, 4
PREC_F64, PREC_MIXED = 0, 1
PROBLEM_SCALAR, PROBLEM_ELASTIC = 0, 1
ELEM_REFERENCE, ELEM_TEXTBOOK = 0, 1
TRACE_Q1, TRACE_REFERENCE = 0, 1
BASIS_PER_DIRECTION, BASIS_COMPONENTWISE = 0, 1

c_i64p = C.POINTER(C.c_int64)
c_i32p = C.POINTER(C.c_int32)
c_dp = C.POINTER(C.c_double)
c_fp = C.POINTER(C.c_float)
vp = C.c_void_p


class MfsError(RuntimeError):
    def __init__(self, code, where):
        msg = lib.mfs_last_error().decode(errors="replace")
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


def _chk(code, where):
    if code != MFS_OK:
        raise MfsError(code, where)


class SolveDesc(C.Structure):
    _fields_ = [("method", C.c_int), ("precision", C.c_int), ("tol", C.c_double), ("maxit", C.c_int),
                ("restart", C.c_int), ("inner_tol", C.c_double), ("check_every", C.c_int)]


class SolveReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_steps", C.c_int), ("true_relres", C.c_double),
                ("recursive_relres", C.c_double), ("seconds", C.c_double), ("status", C.c_int)]


class FemDesc(C.Structure):
    _fields_ = [("problem", C.c_int), ("n_nodes", C.c_int64), ("n_tets", C.c_int64), ("xyz", c_dp),
                ("tets", c_i32p), ("material", c_i32p), ("renumber", C.c_int), ("convention", C.c_int),
                ("n_fixed", C.c_int64), ("fixed_nodes", c_i32p), ("fixed_values", c_dp), ("n_tris", C.c_int64),
                ("tris", c_i32p), ("traction", c_dp)]


class BatchDesc(C.Structure):
    _fields_ = [("n_cells", C.c_int), ("cell_n", c_i32p), ("cell_origin", c_dp), ("cell_h", C.c_double),
                ("E", c_dp), ("nu", C.c_double), ("convention", C.c_int), ("trace", C.c_int),
                ("coarse_basis", C.c_int), ("trace_tol", C.c_double)]


class BatchSolveDesc(C.Structure):
    _fields_ = [("precision", C.c_int), ("tol", C.c_double), ("maxit", C.c_int), ("inner_tol", C.c_double),
                ("use_graph", C.c_int), ("wg_size", C.c_int), ("use_guess", C.c_int), ("variant", C.c_int),
                ("fallback_wg_size", C.c_int), ("fallback_variant", C.c_int)]


class BatchReport(C.Structure):
    _fields_ = [("seconds", C.c_double), ("max_iterations", C.c_int), ("mean_iterations", C.c_double),
                ("max_true_relres", C.c_double), ("n_not_converged", C.c_int), ("n_fallback", C.c_int),
                ("n_precision_fallback", C.c_int), ("kernel_seconds", C.c_double)]


class MlpDesc(C.Structure):
    _fields_ = [("n_layers", C.c_int), ("dims", c_i32p), ("seed", C.c_uint64), ("lr", C.c_double),
                ("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double)]


def _sig(name, res, args):
    f = getattr(lib, name, None)
    if f is None:
        return None
    f.restype = res
    f.argtypes = args
    return f


_P = vp
_sig("mfs_last_error", C.c_char_p, [])
_sig("mfs_version", C.c_char_p, [])
_sig("mfs_ctx_create", C.c_int, [C.c_int, C.POINTER(vp)])
_sig("mfs_ctx_destroy", C.c_int, [vp])
_sig("mfs_ctx_sync", C.c_int, [vp])
_sig("mfs_ctx_info", C.c_int, [vp, C.c_char_p, C.c_int, C.POINTER(C.c_int)])
_sig("mfs_vec_create", C.c_int, [vp, C.c_int64, C.c_int, C.POINTER(vp)])
_sig("mfs_vec_destroy", C.c_int, [vp])
_sig("mfs_vec_size", C.c_int, [vp, c_i64p, C.POINTER(C.c_int)])
_sig("mfs_vec_upload", C.c_int, [vp, c_dp])
_sig("mfs_vec_download", C.c_int, [vp, c_dp])
_sig("mfs_vec_fill", C.c_int, [vp, C.c_double])
_sig("mfs_vec_copy", C.c_int, [vp, vp])
_sig("mfs_mat_create_csr", C.c_int, [vp, C.c_int64, c_i32p, c_i32p, c_dp, C.POINTER(vp)])
_sig("mfs_mat_create_bsr3", C.c_int, [vp, C.c_int64, c_i32p, c_i32p, c_dp, C.POINTER(vp)])
_sig("mfs_mat_update_values", C.c_int, [vp, c_dp])
_sig("mfs_mat_destroy", C.c_int, [vp])
_sig("mfs_mat_info", C.c_int, [vp, c_i64p, c_i64p, C.POINTER(C.c_int), c_i64p, c_i64p])
_sig("mfs_mat_export_csr", C.c_int, [vp, c_i32p, c_i32p, c_dp])
_sig("mfs_spmv", C.c_int, [vp, C.c_int, vp, vp])
_sig("mfs_dot", C.c_int, [vp, vp, c_dp])
_sig("mfs_axpby", C.c_int, [C.c_double, vp, C.c_double, vp])
_sig("mfs_spmv_time", C.c_int, [vp, C.c_int, vp, vp, C.c_int, c_dp])
_sig("mfs_solve_desc_default", None, [C.POINTER(SolveDesc)])
_sig("mfs_solver_create", C.c_int, [vp, C.POINTER(SolveDesc), C.POINTER(vp)])
_sig("mfs_solver_destroy", C.c_int, [vp])
_sig("mfs_solver_run", C.c_int, [vp, vp, vp, C.POINTER(SolveReport)])
_sig("mfs_solver_kernel_time", C.c_int, [vp, C.c_int, C.c_int, c_dp, c_dp])
_sig("mfs_solve_multi", C.c_int, [vp, C.POINTER(SolveDesc), C.c_int, C.POINTER(vp), C.POINTER(vp),
                                  C.POINTER(SolveReport)])


def solve_multi(A, bs, xs, check=True, **kw):
    """Solve A x_j = b_j for the k <= 8 FP64 vector pairs in one iteration
    stream (the matrix is read once per iteration for all of them); x_j hold
    the initial guesses.  Returns one Report per system."""
    k = len(bs)
    assert k == len(xs) and 1 <= k <= 8
    d = solve_desc(**kw)
    hb = (vp * k)(*[v.h for v in bs])
    hx = (vp * k)(*[v.h for v in xs])
    reps = (SolveReport * k)()
    code = lib.mfs_solve_multi(A.h, C.byref(d), k, hb, hx, reps)
    if check:
        _chk(code, "mfs_solve_multi")
    return [Report(r.iterations, r.outer_steps, r.true_relres, r.recursive_relres, r.seconds, r.status) for r in reps]
_sig("mfs_gemm_bf16_time", C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, c_dp, c_dp])


def gemm_bf16_time(ctx, M, N, K, reps=10):
    """(avg ms, TFLOP/s) of the BF16 MFMA GEMM kernel on an M x N x K product."""
    ms, tf = C.c_double(), C.c_double()
    _chk(lib.mfs_gemm_bf16_time(ctx.h, M, N, K, reps, C.byref(ms), C.byref(tf)), "mfs_gemm_bf16_time")
    return ms.value, tf.value
_sig("mfs_valu_fma_rate", C.c_int, [vp, C.c_int, c_dp])


def valu_fma_rate(ctx, dtype):
    """Measured VALU multiply-add rate in TFLOP/s for F32 (unpacked) or F64."""
    tf = C.c_double()
    _chk(lib.mfs_valu_fma_rate(ctx.h, dtype, C.byref(tf)), "mfs_valu_fma_rate")
    return tf.value
_sig("mfs_solve", C.c_int, [vp, C.POINTER(SolveDesc), vp, vp, C.POINTER(SolveReport)])
_sig("mfs_assemble_create", C.c_int, [vp, C.POINTER(FemDesc), C.POINTER(vp)])
_sig("mfs_assemble_run", C.c_int, [vp, c_dp, C.c_double])
_sig("mfs_assemble_destroy", C.c_int, [vp])
_sig("mfs_assemble_matrix", C.c_int, [vp, C.POINTER(vp)])
_sig("mfs_assemble_rhs", C.c_int, [vp, C.POINTER(vp)])
_sig("mfs_assemble_x0", C.c_int, [vp, C.POINTER(vp)])
_sig("mfs_assemble_node_map", C.c_int, [vp, c_i32p])
_sig("mfs_assemble_time", C.c_int, [vp, C.c_int, c_dp])
_sig("mfs_mesh_box_counts", C.c_int, [C.c_int, C.c_int, C.c_int, c_i64p, c_i64p])
_sig("mfs_mesh_box", C.c_int, [C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + [c_dp, c_i32p])
_sig("mfs_mesh_box_face_nodes", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, c_i32p, c_i64p])
_sig("mfs_mesh_box_face_tris", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, c_i32p, c_i64p])
_sig("mfs_field_inclusions", C.c_int, [C.c_int64, c_i32p, c_dp, C.c_int, C.c_double, C.c_double, C.c_uint64,
                                       C.c_double, C.c_double, c_dp])
_sig("mfs_field_lognormal", C.c_int, [C.c_int64, C.c_uint64, C.c_uint64, C.c_int64, C.c_double, C.c_double, c_dp])
_sig("mfs_field_grf", C.c_int, [C.c_int64, C.c_int64, c_i32p, c_dp, C.c_int, C.c_double, C.c_uint64, C.c_double,
                                C.c_double, c_dp, c_i32p])
_sig("mfs_renumber", C.c_int, [C.c_int64, C.c_int64, c_i32p, c_i32p, C.c_int64, c_i32p, c_i32p])
_sig("mfs_portrait", C.c_int, [C.c_int64, C.c_int64, c_i32p, c_i32p, c_i32p])
_sig("mfs_write_gmsh", C.c_int, [C.c_char_p, C.c_int64, c_dp, C.c_int64, c_i32p, c_i32p, C.c_int64, c_i32p,
                                 c_i32p])
_sig("mfs_batch_create", C.c_int, [vp, C.POINTER(BatchDesc), C.POINTER(vp)])
_sig("mfs_batch_destroy", C.c_int, [vp])
_sig("mfs_batch_cell_dofs", C.c_int, [vp, C.c_int, c_i64p, c_i64p])
_sig("mfs_batch_assemble", C.c_int, [vp])
_sig("mfs_batch_solve", C.c_int, [vp, C.POINTER(BatchSolveDesc), C.POINTER(BatchReport)])
_sig("mfs_batch_set_guess", C.c_int, [vp, c_dp])
_sig("mfs_batch_get_basis", C.c_int, [vp, C.c_int, c_dp])
_sig("mfs_batch_coarse_matrices", C.c_int, [vp, c_dp])
_sig("mfs_batch_iterations", C.c_int, [vp, c_i32p])
_sig("mfs_batch_status", C.c_int, [vp, c_i32p, c_dp, c_i32p])
_sig("mfs_batch_autotune", C.c_int, [vp, C.POINTER(BatchSolveDesc), C.POINTER(C.c_int), c_dp])
_sig("mfs_mlp_create", C.c_int, [vp, C.POINTER(MlpDesc), C.POINTER(vp)])
_sig("mfs_mlp_destroy", C.c_int, [vp])
_sig("mfs_mlp_get_params", C.c_int, [vp, C.c_int, c_fp, c_fp])
_sig("mfs_mlp_set_params", C.c_int, [vp, C.c_int, c_fp, c_fp])
_sig("mfs_mlp_forward", C.c_int, [vp, C.c_int, c_fp, c_fp])
_sig("mfs_mlp_loss_grad", C.c_int, [vp, C.c_int, c_fp, c_fp, c_dp, C.POINTER(c_fp), C.POINTER(c_fp)])
_sig("mfs_mlp_train", C.c_int, [vp, C.c_int64, c_fp, c_fp, C.c_int, C.c_int, c_dp])
_sig("mfs_mlp_time_step", C.c_int, [vp, C.c_int, C.c_int, c_dp, c_dp])
_sig("mfs_mlp_train_vec", C.c_int, [vp, C.c_int64, vp, vp, C.c_int, C.c_int, C.c_uint64, c_dp])
_sig("mfs_mlp_forward_vec", C.c_int, [vp, C.c_int64, vp, vp, C.c_int])
_sig("mfs_batch_features", C.c_int, [vp, C.c_double, C.c_double, vp])
_sig("mfs_batch_targets", C.c_int, [vp, vp])
_sig("mfs_batch_guess_from", C.c_int, [vp, vp])
_sig("mfs_msfem_coarse_system", C.c_int, [vp, c_dp, c_dp, C.c_int64, c_i32p, c_dp, C.POINTER(vp), C.POINTER(vp),
                                          C.POINTER(vp)])
_sig("mfs_msfem_top_traction", C.c_int, [vp, c_dp, c_dp])
_sig("mfs_msfem_downscale", C.c_int, [vp, c_dp, C.c_int, c_dp])
_sig("mfs_write_solver_params", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_double, C.c_int,
                                          vp, C.c_int, vp, C.c_int, vp, C.c_char_p])
_sig("mfs_read_gmsh", C.c_int, [C.c_char_p, c_i64p, c_dp, c_i64p, c_i32p, c_i32p, c_i64p, c_i32p, c_i32p, c_i32p])
_sig("mfs_msfem_error", C.c_int, [vp, c_dp, c_dp, c_dp])
_sig("mfs_batch_trace_status", C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double)])


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------
# Host utilities (no device needed)
# ---------------------------------------------------------------------------
def mesh_box(nx, ny=None, nz=None, origin=(0.0, 0.0, 0.0), h=None):
    """Structured Freudenthal tet mesh of a box: returns (xyz[nn,3], tets[nt,4])."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    if h is None:
        h = (1.0 / nx, 1.0 / ny, 1.0 / nz)
    nn, nt = C.c_int64(), C.c_int64()
    _chk(lib.mfs_mesh_box_counts(nx, ny, nz, C.byref(nn), C.byref(nt)), "mfs_mesh_box_counts")
    xyz = np.empty((nn.value, 3))
    tets = np.empty((nt.value, 4), dtype=np.int32)
    _chk(lib.mfs_mesh_box(nx, ny, nz, *origin, *h, ptr(xyz, C.c_double), ptr(tets, C.c_int32)), "mfs_mesh_box")
    return xyz, tets


def box_face_nodes(nx, ny, nz, face):
    cnt = C.c_int64()
    _chk(lib.mfs_mesh_box_face_nodes(nx, ny, nz, face, None, C.byref(cnt)), "face_nodes")
    out = np.empty(cnt.value, dtype=np.int32)
    _chk(lib.mfs_mesh_box_face_nodes(nx, ny, nz, face, ptr(out, C.c_int32), C.byref(cnt)), "face_nodes")
    return out


def box_face_tris(nx, ny, nz, face):
    cnt = C.c_int64()
    _chk(lib.mfs_mesh_box_face_tris(nx, ny, nz, face, None, C.byref(cnt)), "face_tris")
    out = np.empty((cnt.value, 3), dtype=np.int32)
    _chk(lib.mfs_mesh_box_face_tris(nx, ny, nz, face, ptr(out, C.c_int32), C.byref(cnt)), "face_tris")
    return out


def field_inclusions(tets, xyz, count, rmin, rmax, seed, v_bg, v_in):
    out = np.empty(len(tets))
    _chk(lib.mfs_field_inclusions(len(tets), ptr(i32(tets), C.c_int32), ptr(f64(xyz), C.c_double), count, rmin,
                                  rmax, seed, v_bg, v_in, ptr(out, C.c_double)), "field_inclusions")
    return out


def field_lognormal(nt, seed, stream, index0, mu, sigma):
    out = np.empty(nt)
    _chk(lib.mfs_field_lognormal(nt, seed, stream, index0, mu, sigma, ptr(out, C.c_double)), "field_lognormal")
    return out


def field_grf(xyz, tets, modes, ell, seed, v_lo, v_hi):
    out = np.empty(len(tets))
    mat = np.empty(len(tets), dtype=np.int32)
    _chk(lib.mfs_field_grf(len(xyz), len(tets), ptr(i32(tets), C.c_int32), ptr(f64(xyz), C.c_double), modes, ell,
                           seed, v_lo, v_hi, ptr(out, C.c_double), ptr(mat, C.c_int32)), "field_grf")
    return out, mat


def renumber(nn, tets, material, fixed_nodes):
    m = np.empty(nn, dtype=np.int32)
    fixed_nodes = i32(fixed_nodes)
    _chk(lib.mfs_renumber(nn, len(tets), ptr(i32(tets), C.c_int32),
                          ptr(i32(material), C.c_int32) if material is not None else None, len(fixed_nodes),
                          ptr(fixed_nodes, C.c_int32), ptr(m, C.c_int32)), "renumber")
    return m


def portrait(n_rows, tets):
    tets = i32(tets)
    rp = np.empty(n_rows + 1, dtype=np.int32)
    _chk(lib.mfs_portrait(n_rows, len(tets), ptr(tets, C.c_int32), ptr(rp, C.c_int32), None), "portrait")
    col = np.empty(int(rp[-1]), dtype=np.int32)
    _chk(lib.mfs_portrait(n_rows, len(tets), ptr(tets, C.c_int32), ptr(rp, C.c_int32), ptr(col, C.c_int32)),
         "portrait")
    return rp, col


def write_gmsh(path, xyz, tets, material=None, tris=None, tri_tags=None):
    tris = np.zeros((0, 3), np.int32) if tris is None else i32(tris)
    tri_tags = np.zeros(0, np.int32) if tri_tags is None else i32(tri_tags)
    _chk(lib.mfs_write_gmsh(path.encode(), len(xyz), ptr(f64(xyz), C.c_double), len(tets),
                            ptr(i32(tets), C.c_int32),
                            ptr(i32(material), C.c_int32) if material is not None else None, len(tris),
                            ptr(tris, C.c_int32), ptr(tri_tags, C.c_int32)), "write_gmsh")


class DirichletSpec(C.Structure):
    _fields_ = [("tag", C.c_int), ("fixed", C.c_int * 3), ("value", C.c_double * 3)]


class NeumannSpec(C.Structure):
    _fields_ = [("tag", C.c_int), ("value", C.c_double), ("direction", C.c_double * 3)]


class MaterialSpec(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double)]


def write_solver_params(path, mesh_path, result_dir, dirichlet=(), neumann=(), materials=((1e9, 0.3),),
                        threads=16, critical_residual=1e-8, mesh_format=10, log_flag=0):
    """Write the Mechanical driver's property file.  dirichlet: (tag,
    (vx|None, vy|None, vz|None)); neumann: (tag, value, (dx, dy, dz));
    materials: (E, nu) per domain."""
    D = (DirichletSpec * max(1, len(dirichlet)))()
    for i, (tag, vals) in enumerate(dirichlet):
        D[i].tag = tag
        for a, v in enumerate(vals):
            D[i].fixed[a] = 0 if v is None else 1
            D[i].value[a] = 0.0 if v is None else float(v)
    N = (NeumannSpec * max(1, len(neumann)))()
    for i, (tag, val, d) in enumerate(neumann):
        N[i].tag, N[i].value = tag, float(val)
        for a in range(3):
            N[i].direction[a] = float(d[a])
    M = (MaterialSpec * len(materials))()
    for i, (E, nu) in enumerate(materials):
        M[i].E, M[i].nu = float(E), float(nu)
    _chk(lib.mfs_write_solver_params(path.encode(), log_flag, threads, mesh_format, mesh_path.encode(),
                                     critical_residual, len(dirichlet), C.cast(D, vp), len(neumann),
                                     C.cast(N, vp), len(materials), C.cast(M, vp), result_dir.encode()),
         "mfs_write_solver_params")


def read_gmsh(path):
    """(xyz, tets, material, tris, tri_tags, tri_parent) of a Gmsh 2.2 file,
    0-based, read with SimpleGrid::ReadFromMSH's rules."""
    nn, nt, nf = C.c_int64(), C.c_int64(), C.c_int64()
    _chk(lib.mfs_read_gmsh(path.encode(), C.byref(nn), None, C.byref(nt), None, None, C.byref(nf), None, None,
                           None), "read_gmsh")
    xyz = np.empty((nn.value, 3))
    tets = np.empty((nt.value, 4), np.int32)
    mat = np.empty(nt.value, np.int32)
    tris = np.empty((nf.value, 3), np.int32)
    tags = np.empty(nf.value, np.int32)
    par = np.empty(nf.value, np.int32)
    _chk(lib.mfs_read_gmsh(path.encode(), C.byref(nn), ptr(xyz, C.c_double), C.byref(nt), ptr(tets, C.c_int32),
                           ptr(mat, C.c_int32), C.byref(nf), ptr(tris, C.c_int32), ptr(tags, C.c_int32),
                           ptr(par, C.c_int32)), "read_gmsh")
    return xyz, tets, mat, tris, tags, par


# ---------------------------------------------------------------------------
# Device objects
# ---------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        h = vp()
        _chk(lib.mfs_ctx_create(device, C.byref(h)), "mfs_ctx_create")
        self.h = h

    def info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int()
        _chk(lib.mfs_ctx_info(self.h, name, 256, C.byref(cu)), "mfs_ctx_info")
        return name.value.decode(), cu.value

    def sync(self):
        _chk(lib.mfs_ctx_sync(self.h), "mfs_ctx_sync")

    def close(self):
        if self.h:
            lib.mfs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


class Vec:
    def __init__(self, ctx, n, prec=F64, data=None, borrowed=None, owned=False):
        self.ctx = ctx
        if borrowed is not None:
            self.h, self.owned = borrowed, owned
        else:
            h = vp()
            _chk(lib.mfs_vec_create(ctx.h, n, prec, C.byref(h)), "mfs_vec_create")
            self.h, self.owned = h, True
        n_, p_ = C.c_int64(), C.c_int()
        _chk(lib.mfs_vec_size(self.h, C.byref(n_), C.byref(p_)), "mfs_vec_size")
        self.n, self.prec = n_.value, p_.value
        if data is not None:
            self.upload(data)

    def upload(self, a):
        a = f64(a).ravel()
        assert a.size == self.n
        _chk(lib.mfs_vec_upload(self.h, ptr(a, C.c_double)), "mfs_vec_upload")

    def download(self):
        out = np.empty(self.n)
        _chk(lib.mfs_vec_download(self.h, ptr(out, C.c_double)), "mfs_vec_download")
        return out

    def fill(self, v):
        _chk(lib.mfs_vec_fill(self.h, v), "mfs_vec_fill")

    def copy_from(self, other):
        _chk(lib.mfs_vec_copy(self.h, other.h), "mfs_vec_copy")

    def close(self):
        if self.h and self.owned:
            lib.mfs_vec_destroy(self.h)
        self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


class Mat:
    def __init__(self, ctx, h, owned=True):
        self.ctx, self.h, self.owned = ctx, h, owned
        n, nnz, nb, nnzb, fmt = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        _chk(lib.mfs_mat_info(h, C.byref(n), C.byref(nnz), C.byref(fmt), C.byref(nb), C.byref(nnzb)), "mat_info")
        self.n, self.nnz, self.format, self.nb, self.nnzb = n.value, nnz.value, fmt.value, nb.value, nnzb.value

    @classmethod
    def csr(cls, ctx, rowptr, col, val):
        h = vp()
        _chk(lib.mfs_mat_create_csr(ctx.h, len(rowptr) - 1, ptr(i32(rowptr), C.c_int32), ptr(i32(col), C.c_int32),
                                    ptr(f64(val), C.c_double), C.byref(h)), "mfs_mat_create_csr")
        return cls(ctx, h)

    @classmethod
    def bsr3(cls, ctx, browptr, bcol, val):
        h = vp()
        _chk(lib.mfs_mat_create_bsr3(ctx.h, len(browptr) - 1, ptr(i32(browptr), C.c_int32),
                                     ptr(i32(bcol), C.c_int32), ptr(f64(val).ravel(), C.c_double), C.byref(h)),
             "mfs_mat_create_bsr3")
        return cls(ctx, h)

    def update_values(self, val):
        _chk(lib.mfs_mat_update_values(self.h, ptr(f64(val).ravel(), C.c_double)), "mfs_mat_update_values")

    def export_csr(self):
        rp = np.empty(self.n + 1, dtype=np.int32)
        col = np.empty(self.nnz, dtype=np.int32)
        val = np.empty(self.nnz)
        _chk(lib.mfs_mat_export_csr(self.h, ptr(rp, C.c_int32), ptr(col, C.c_int32), ptr(val, C.c_double)),
             "mfs_mat_export_csr")
        return rp, col, val

    def spmv(self, x, y, prec=F64):
        _chk(lib.mfs_spmv(self.h, prec, x.h, y.h), "mfs_spmv")

    def spmv_time(self, x, y, reps, prec=F64):
        ms = C.c_double()
        _chk(lib.mfs_spmv_time(self.h, prec, x.h, y.h, reps, C.byref(ms)), "mfs_spmv_time")
        return ms.value

    def spmv_bytes(self, prec):
        sv = 8 if prec == F64 else 4
        return self.nnz * sv + self.nnzb * 4 + (self.nb + 1) * 4 + 2 * self.n * sv

    def close(self):
        if self.h and self.owned:
            lib.mfs_mat_destroy(self.h)
        self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


def dot(x, y):
    r = C.c_double()
    _chk(lib.mfs_dot(x.h, y.h, C.byref(r)), "mfs_dot")
    return r.value


def axpby(a, x, b, y):
    _chk(lib.mfs_axpby(a, x.h, b, y.h), "mfs_axpby")


def solve_desc(method=PCG_JACOBI, precision=PREC_F64, tol=1e-8, maxit=100000, restart=30, inner_tol=1e-4,
               check_every=0):
    d = SolveDesc()
    lib.mfs_solve_desc_default(C.byref(d))
    d.method, d.precision, d.tol, d.maxit = method, precision, tol, maxit
    d.restart, d.inner_tol, d.check_every = restart, inner_tol, check_every
    return d


@dataclass
class Report:
    iterations: int
    outer_steps: int
    true_relres: float
    recursive_relres: float
    seconds: float
    status: int


class Solver:
    def __init__(self, A, **kw):
        self.A = A
        self.desc = solve_desc(**kw)
        h = vp()
        _chk(lib.mfs_solver_create(A.h, C.byref(self.desc), C.byref(h)), "mfs_solver_create")
        self.h = h

    def run(self, b, x, check=True):
        rep = SolveReport()
        code = lib.mfs_solver_run(self.h, b.h, x.h, C.byref(rep))
        if check:
            _chk(code, "mfs_solver_run")
        return Report(rep.iterations, rep.outer_steps, rep.true_relres, rep.recursive_relres, rep.seconds,
                      rep.status)

    def kernel_time(self, kernel, reps=200):
        """(avg ms, algorithmic bytes) of one PCG iteration kernel: 0 = fused
        direction + SpMV, 1 = x/r/z update (inner precision)."""
        ms, by = C.c_double(), C.c_double()
        _chk(lib.mfs_solver_kernel_time(self.h, kernel, reps, C.byref(ms), C.byref(by)), "mfs_solver_kernel_time")
        return ms.value, by.value

    def close(self):
        if self.h:
            lib.mfs_solver_destroy(self.h)
        self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


class Assembly:
    """Device assembly of a scalar or elastic FEM system (mfs_assemble_*)."""

    def __init__(self, ctx, problem, xyz, tets, material=None, renumber=False, convention=ELEM_REFERENCE,
                 fixed_nodes=None, fixed_values=None, tris=None, traction=None):
        self.ctx = ctx
        self._keep = []
        k = self._keep

        def keep(a):
            k.append(a)
            return a

        xyz = keep(f64(xyz))
        tets = keep(i32(tets))
        d = FemDesc()
        d.problem, d.n_nodes, d.n_tets = problem, len(xyz), len(tets)
        d.xyz, d.tets = ptr(xyz, C.c_double), ptr(tets, C.c_int32)
        d.material = ptr(keep(i32(material)), C.c_int32) if material is not None else None
        d.renumber, d.convention = int(renumber), convention
        fixed_nodes = keep(i32(fixed_nodes if fixed_nodes is not None else []))
        fixed_values = keep(f64(fixed_values if fixed_values is not None else []).ravel())
        d.n_fixed, d.fixed_nodes, d.fixed_values = len(fixed_nodes), ptr(fixed_nodes, C.c_int32), ptr(
            fixed_values, C.c_double)
        tris = keep(i32(tris if tris is not None else np.zeros((0, 3))))
        traction = keep(f64(traction if traction is not None else np.zeros((0, 3))).ravel())
        d.n_tris, d.tris, d.traction = len(tris), ptr(tris, C.c_int32), ptr(traction, C.c_double)
        h = vp()
        _chk(lib.mfs_assemble_create(ctx.h, C.byref(d), C.byref(h)), "mfs_assemble_create")
        self.h = h
        self._keep = []
        mh, bh, xh = vp(), vp(), vp()
        _chk(lib.mfs_assemble_matrix(h, C.byref(mh)), "assemble_matrix")
        _chk(lib.mfs_assemble_rhs(h, C.byref(bh)), "assemble_rhs")
        _chk(lib.mfs_assemble_x0(h, C.byref(xh)), "assemble_x0")
        self.A = Mat(ctx, mh, owned=False)
        self.b = Vec(ctx, 0, borrowed=bh)
        self.x0 = Vec(ctx, 0, borrowed=xh)
        self.n_nodes = len(xyz)

    def run(self, coef, nu=0.3):
        coef = f64(coef)
        _chk(lib.mfs_assemble_run(self.h, ptr(coef, C.c_double), nu), "mfs_assemble_run")

    def node_map(self):
        m = np.empty(self.n_nodes, dtype=np.int32)
        _chk(lib.mfs_assemble_node_map(self.h, ptr(m, C.c_int32)), "node_map")
        return m

    def time(self, reps):
        ms = C.c_double()
        _chk(lib.mfs_assemble_time(self.h, reps, C.byref(ms)), "mfs_assemble_time")
        return ms.value

    def close(self):
        if self.h:
            lib.mfs_assemble_destroy(self.h)
        self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


def exported_symbols():
    """Names of the functions include/mfs.h declares (parsed from the header)."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "mfs.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\*?\s+\*?(mfs_[a-z0-9_]+)\(", text, re.M)))


# ---------------------------------------------------------------------------
# MsFEM batched local problems
# ---------------------------------------------------------------------------
@dataclass
class BatchResult:
    seconds: float
    max_iterations: int
    mean_iterations: float
    max_true_relres: float
    n_not_converged: int
    n_fallback: int
    n_precision_fallback: int
    status: int
    kernel_seconds: float = 0.0


def batch_solve_desc(precision=PREC_MIXED, tol=1e-8, maxit=20000, inner_tol=1e-4, use_graph=True, wg_size=0,
                     use_guess=False, variant=0, fallback_wg_size=0, fallback_variant=0):
    d = BatchSolveDesc()
    d.precision, d.tol, d.maxit, d.inner_tol = precision, tol, maxit, inner_tol
    d.use_graph, d.wg_size, d.use_guess, d.variant = int(use_graph), wg_size, int(use_guess), variant
    d.fallback_wg_size, d.fallback_variant = fallback_wg_size, fallback_variant
    return d


class Batch:
    """All local basis problems of a set of coarse cells (mfs_batch_*)."""

    def __init__(self, ctx, cell_n, cell_origin, cell_h, E, nu=0.3, convention=ELEM_REFERENCE, trace=TRACE_Q1,
                 coarse_basis=BASIS_PER_DIRECTION, trace_tol=0.0):
        self.ctx = ctx
        self.cell_n = i32(cell_n)
        self.cell_origin = f64(cell_origin).reshape(-1, 3)
        E = f64(E)
        assert E.size == int((6 * self.cell_n.astype(np.int64) ** 3).sum())
        d = BatchDesc()
        d.n_cells = len(self.cell_n)
        d.cell_n = ptr(self.cell_n, C.c_int32)
        d.cell_origin = ptr(self.cell_origin, C.c_double)
        d.cell_h, d.E, d.nu, d.convention = cell_h, ptr(E, C.c_double), nu, convention
        d.trace, d.coarse_basis, d.trace_tol = trace, coarse_basis, trace_tol
        h = vp()
        _chk(lib.mfs_batch_create(ctx.h, C.byref(d), C.byref(h)), "mfs_batch_create")
        self.h = h
        self.ncells = len(self.cell_n)

    def assemble(self):
        _chk(lib.mfs_batch_assemble(self.h), "mfs_batch_assemble")

    def solve(self, check=True, **kw):
        d = batch_solve_desc(**kw)
        rep = BatchReport()
        code = lib.mfs_batch_solve(self.h, C.byref(d), C.byref(rep))
        if check:
            _chk(code, "mfs_batch_solve")
        return BatchResult(rep.seconds, rep.max_iterations, rep.mean_iterations, rep.max_true_relres,
                           rep.n_not_converged, rep.n_fallback, rep.n_precision_fallback, code,
                           rep.kernel_seconds)

    def autotune(self, **kw):
        d = batch_solve_desc(**kw)
        best, ms = C.c_int(), C.c_double()
        _chk(lib.mfs_batch_autotune(self.h, C.byref(d), C.byref(best), C.byref(ms)), "mfs_batch_autotune")
        v = best.value  # workgroup * 100 + workgroups per CU * 10 + kernel variant
        return v // 100, (v // 10) % 10, v % 10, ms.value

    def n_interior(self, cell):
        ni, nt = C.c_int64(), C.c_int64()
        _chk(lib.mfs_batch_cell_dofs(self.h, cell, C.byref(ni), C.byref(nt)), "cell_dofs")
        return ni.value, nt.value

    def set_guess(self, guess):
        g = f64(guess).ravel()
        _chk(lib.mfs_batch_set_guess(self.h, ptr(g, C.c_double)), "mfs_batch_set_guess")

    def basis(self, cell):
        _, nt = self.n_interior(cell)
        out = np.empty((24, nt))
        _chk(lib.mfs_batch_get_basis(self.h, cell, ptr(out, C.c_double)), "mfs_batch_get_basis")
        return out

    def coarse_matrices(self):
        out = np.empty((self.ncells, 24, 24))
        _chk(lib.mfs_batch_coarse_matrices(self.h, ptr(out, C.c_double)), "mfs_batch_coarse_matrices")
        return out

    def features(self, X, mu, sigma):
        _chk(lib.mfs_batch_features(self.h, mu, sigma, X.h), "mfs_batch_features")

    def targets(self, T):
        _chk(lib.mfs_batch_targets(self.h, T.h), "mfs_batch_targets")

    def guess_from(self, G=None):
        _chk(lib.mfs_batch_guess_from(self.h, G.h if G is not None else None), "mfs_batch_guess_from")

    def coarse_system(self, Kc, load, fixed_nodes=(), fixed_values=None):
        Kc = f64(Kc)
        load = f64(load).ravel()
        fixed_nodes = i32(fixed_nodes)
        fixed_values = f64(fixed_values if fixed_values is not None else np.zeros((len(fixed_nodes), 3))).ravel()
        A, b, x = vp(), vp(), vp()
        _chk(lib.mfs_msfem_coarse_system(self.h, ptr(Kc, C.c_double), ptr(load, C.c_double), len(fixed_nodes),
                                         ptr(fixed_nodes, C.c_int32), ptr(fixed_values, C.c_double), C.byref(A),
                                         C.byref(b), C.byref(x)), "mfs_msfem_coarse_system")
        return Mat(self.ctx, A), Vec(self.ctx, 0, borrowed=b, owned=True), Vec(self.ctx, 0, borrowed=x, owned=True)

    def top_traction(self, t, n_coarse_nodes):
        t = f64(t)
        out = np.empty(3 * n_coarse_nodes)
        _chk(lib.mfs_msfem_top_traction(self.h, ptr(t, C.c_double), ptr(out, C.c_double)), "top_traction")
        return out

    def downscale(self, u_coarse, cell):
        u_coarse = f64(u_coarse).ravel()
        _, nt = self.n_interior(cell)
        out = np.empty(nt)
        _chk(lib.mfs_msfem_downscale(self.h, ptr(u_coarse, C.c_double), cell, ptr(out, C.c_double)), "downscale")
        return out

    def msfem_error(self, u_coarse, u_fem):
        """Per-component relative L2 error of the downscaled MsFEM field
        against a fine FEM field on the whole box's fine grid."""
        u_coarse = f64(u_coarse).ravel()
        u_fem = f64(u_fem).ravel()
        out = np.empty(3)
        _chk(lib.mfs_msfem_error(self.h, ptr(u_coarse, C.c_double), ptr(u_fem, C.c_double), ptr(out, C.c_double)),
             "mfs_msfem_error")
        return out

    def trace_status(self):
        """Reference-trace mode: (largest face-solve relative residual, face
        problems above the trace tolerance, device ms of the edge + face
        launches in the last assemble)."""
        it, bad, ms = C.c_double(), C.c_int(), C.c_double()
        _chk(lib.mfs_batch_trace_status(self.h, C.byref(it), C.byref(bad), C.byref(ms)), "mfs_batch_trace_status")
        return it.value, bad.value, ms.value

    def status(self):
        it = np.empty((self.ncells, 24), dtype=np.int32)
        rr = np.empty((self.ncells, 24))
        fl = np.empty((self.ncells, 24), dtype=np.int32)
        _chk(lib.mfs_batch_status(self.h, ptr(it, C.c_int32), ptr(rr, C.c_double), ptr(fl, C.c_int32)), "status")
        return it, rr, fl

    def iterations(self):
        out = np.empty((self.ncells, 24), dtype=np.int32)
        _chk(lib.mfs_batch_iterations(self.h, ptr(out, C.c_int32)), "mfs_batch_iterations")
        return out

    def close(self):
        if self.h:
            lib.mfs_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Learned initial guess network
# ---------------------------------------------------------------------------
class MLP:
    """ReLU MLP on BF16 MFMA GEMMs with FP32 accumulation (mfs_mlp_*)."""

    def __init__(self, ctx, dims, seed=1, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx = ctx
        self.dims = [int(d) for d in dims]
        dd = np.array(self.dims, dtype=np.int32)
        d = MlpDesc()
        d.n_layers, d.dims, d.seed, d.lr = len(dims) - 1, ptr(dd, C.c_int32), seed, lr
        d.beta1, d.beta2, d.eps = beta1, beta2, eps
        h = vp()
        _chk(lib.mfs_mlp_create(ctx.h, C.byref(d), C.byref(h)), "mfs_mlp_create")
        self.h = h
        self.L = len(dims) - 1

    def get_params(self):
        out = []
        for l in range(self.L):
            W = np.empty((self.dims[l + 1], self.dims[l]), np.float32)
            b = np.empty(self.dims[l + 1], np.float32)
            _chk(lib.mfs_mlp_get_params(self.h, l, ptr(W, C.c_float), ptr(b, C.c_float)), "get_params")
            out.append((W, b))
        return out

    def set_params(self, params):
        for l, (W, b) in enumerate(params):
            W = np.ascontiguousarray(W, np.float32)
            b = np.ascontiguousarray(b, np.float32)
            _chk(lib.mfs_mlp_set_params(self.h, l, ptr(W, C.c_float), ptr(b, C.c_float)), "set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.empty((X.shape[0], self.dims[-1]), np.float32)
        _chk(lib.mfs_mlp_forward(self.h, X.shape[0], ptr(X, C.c_float), ptr(Y, C.c_float)), "mfs_mlp_forward")
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        gW = [np.empty((self.dims[l + 1], self.dims[l]), np.float32) for l in range(self.L)]
        gb = [np.empty(self.dims[l + 1], np.float32) for l in range(self.L)]
        gWp = (c_fp * self.L)(*[ptr(g, C.c_float) for g in gW])
        gbp = (c_fp * self.L)(*[ptr(g, C.c_float) for g in gb])
        loss = C.c_double()
        _chk(lib.mfs_mlp_loss_grad(self.h, X.shape[0], ptr(X, C.c_float), ptr(T, C.c_float), C.byref(loss), gWp,
                                   gbp), "mfs_mlp_loss_grad")
        return loss.value, gW, gb

    def train(self, X, T, batch, epochs):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = C.c_double()
        _chk(lib.mfs_mlp_train(self.h, X.shape[0], ptr(X, C.c_float), ptr(T, C.c_float), batch, epochs,
                               C.byref(loss)), "mfs_mlp_train")
        return loss.value

    def train_vec(self, n, X, T, batch, epochs, seed=7):
        loss = C.c_double()
        _chk(lib.mfs_mlp_train_vec(self.h, n, X.h, T.h, batch, epochs, seed, C.byref(loss)), "mfs_mlp_train_vec")
        return loss.value

    def forward_vec(self, n, X, Y, batch=256):
        _chk(lib.mfs_mlp_forward_vec(self.h, n, X.h, Y.h, batch), "mfs_mlp_forward_vec")

    def time_step(self, batch, reps=20):
        ms, tf = C.c_double(), C.c_double()
        _chk(lib.mfs_mlp_time_step(self.h, batch, reps, C.byref(ms), C.byref(tf)), "mfs_mlp_time_step")
        return ms.value, tf.value

    def close(self):
        if self.h:
            lib.mfs_mlp_destroy(self.h)
        self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)
