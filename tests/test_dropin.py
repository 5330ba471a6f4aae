"""The C++ drop-in CSSD_Matrix (include/mfs_cssd.hpp) used the way the
reference's drivers use Library/CSSD_Matrix.h, for the scalar
CSSD_Matrix<double, double> and the 3x3 block CSSD_Matrix<Tensor2Rank3D,
Point<double>> (BSR3 on the device)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from multiscalefemsolver3d_amd import problems

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")
EXE = os.path.join(HERE, "test_cssd_dropin")
METHODS = ["MSG", "MSG_PreconditioningSSOR", "BiCG_Diag", "GMRES", "BiCG_Stab"]


def _build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def test_dropin_host_methods():
    """Host methods of both instantiations: assembly by SummPartOfMatrix,
    GetValue/SetValue with the reference's bool semantics, SummMatrix,
    MakeDivideOnSqrtFormSumm, the SSOR factor and its application against
    the dense formula, RandomMatrix, SolveMaxEigenvalue, Minimization_Problem
    against the normal equations."""
    _build()
    out = subprocess.run([EXE, "cpu"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "OK" in out.stdout, out.stdout + out.stderr


def block_system(n=4):
    """Clamped n^3 elastic cube under top pressure (reference element
    convention, the Mechanical driver's renumbering), assembled by the oracle
    and regrouped into 3x3 blocks: the CSSD_Matrix<Tensor2Rank3D,
    Point<double>> the drivers hold before math::MakeCopyMatrix_A_into_B."""
    p = problems.elastic_cube(n)
    O = orc.elastic_system(p["xyz"], p["tets"], p["E"], p["nu"], p["fixed_nodes"], p["fixed_values"], p["tris"],
                           p["traction"], p["material"])
    nb = O.n // 3
    rp, col, val = O.rowptr, O.col, O.val
    brp = [0]
    bcol, bval = [], []
    for i in range(nb):
        rows = [dict(zip(col[rp[3 * i + a]:rp[3 * i + a + 1]], val[rp[3 * i + a]:rp[3 * i + a + 1]])) for a in range(3)]
        blocks = sorted({c // 3 for c in rows[0]})
        for j in blocks:
            bcol.append(j)
            bval.append([rows[a].get(3 * j + b, 0.0) for a in range(3) for b in range(3)])
        brp.append(len(bcol))
    return O, nb, np.array(brp, np.int32), np.array(bcol, np.int32), np.array(bval).ravel()


def test_block_regrouping_is_the_scalar_expansion():
    """The 3x3 grouping used by the block test expands back to the oracle's
    scalar CSR in MakeCopyMatrix_A_into_B's order (no value dropped)."""
    O, nb, brp, bcol, bval = block_system(3)
    assert brp[-1] * 9 == O.nnz
    x = np.random.default_rng(0).normal(size=O.n)
    y = np.zeros(O.n)
    B = bval.reshape(-1, 3, 3)
    for i in range(nb):
        for k in range(brp[i], brp[i + 1]):
            y[3 * i:3 * i + 3] += B[k] @ x[3 * bcol[k]:3 * bcol[k] + 3]
    assert np.allclose(y, orc.csr_mv(O.rowptr, O.col, O.val, x), rtol=0, atol=1e-12 * abs(O.val).max())


@pytest.mark.gpu
def test_dropin_solvers_on_gpu():
    assert os.path.exists(EXE), "build tests/cpp first (__graft_entry__.build does)"
    out = subprocess.run([EXE, "gpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "OK" in out.stdout, out.stdout + out.stderr


@pytest.mark.gpu
def test_dropin_block_system_vs_oracle(tmp_path):
    """CSSD_Matrix<3x3 block, 3-vector> filled member by member as the
    drivers do, solved on the GPU (BSR3) by MSG, MSG_PreconditioningSSOR,
    BiCG_Diag, GMRES and BiCG_Stab to 1e-12: every solution within 1e-8
    (relative L2) of the oracle's FP64 PCG solution."""
    assert os.path.exists(EXE), "build tests/cpp first (__graft_entry__.build does)"
    O, nb, brp, bcol, bval = block_system(4)
    fin, fout = tmp_path / "sys.bin", tmp_path / "x.bin"
    with open(fin, "wb") as f:
        np.array([nb], np.int64).tofile(f)
        brp.tofile(f)
        bcol.tofile(f)
        bval.astype(np.float64).tofile(f)
        O.rhs.astype(np.float64).tofile(f)
        O.x0.astype(np.float64).tofile(f)
    out = subprocess.run([EXE, "bsr3", str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0 and "OK" in out.stdout, out.stdout + out.stderr
    xs = np.fromfile(fout, dtype=np.float64).reshape(len(METHODS), 3 * nb)
    xo, _, _ = orc.pcg(O.rowptr, O.col, O.val, O.rhs, O.x0, tol=1e-14)
    for name, x in zip(METHODS, xs):
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        assert err < 1e-8, (name, err)
