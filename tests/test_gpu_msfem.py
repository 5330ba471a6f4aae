"""GPU tests of the batched MsFEM local problems against the oracle's
per-cell FP64 solves (orc_msfem_cell) and the method's invariants."""
import numpy as np
import pytest

import multiscalefemsolver3d_amd as mfs
import oracle as orc
from multiscalefemsolver3d_amd import problems

pytestmark = pytest.mark.gpu


def make_batch(ctx, p):
    B = mfs.Batch(ctx, p["cell_n"], p["origin"], p["h"], p["E"], p["nu"])
    B.assemble()
    return B


@pytest.fixture(scope="module")
def small(ctx):
    p = problems.msfem_cells(2, n_lo=3, n_hi=5)
    return p, make_batch(ctx, p)


@pytest.mark.parametrize("precision", [mfs.PREC_MIXED, mfs.PREC_F64])
def test_bases_match_oracle(ctx, small, precision):
    p, B = small
    rep = B.solve(precision=precision, tol=1e-10, use_graph=False)
    assert rep.n_not_converged == 0 and rep.max_true_relres <= 1e-10
    Kc = B.coarse_matrices()
    for c in range(B.ncells):
        n = int(p["cell_n"][c])
        basis_o, Kc_o, _ = orc.msfem_cell(n, p["origin"][c], p["h"], problems.cell_E(p, c), p["nu"], tol=1e-13)
        basis = B.basis(c)
        err = np.linalg.norm(basis - basis_o) / np.linalg.norm(basis_o)
        assert err < 1e-8, (c, err)
        assert np.abs(Kc[c] - Kc_o).max() < 1e-7 * np.abs(Kc_o).max()


def test_partition_of_unity_and_coarse_invariants(ctx, small):
    p, B = small
    B.solve(precision=mfs.PREC_MIXED, tol=1e-9)
    Kc = B.coarse_matrices()
    for c in range(B.ncells):
        basis = B.basis(c)
        N = basis.shape[1]
        for d in range(3):
            e = np.zeros(N)
            e[d::3] = 1
            assert np.abs(basis[d::3].sum(0) - e).max() < 1e-7
        K = Kc[c]
        assert np.abs(K - K.T).max() < 1e-7 * np.abs(K).max()
        for d in range(3):
            t = np.zeros(24)
            t[d::3] = 1
            assert np.abs(K @ t).max() < 1e-6 * np.abs(K).max()


def test_graph_and_direct_launch_agree(ctx, small):
    p, B = small
    B.solve(precision=mfs.PREC_MIXED, tol=1e-9, use_graph=False)
    a = [B.basis(c) for c in range(B.ncells)]
    it_a = B.iterations()
    for _ in range(3):  # replays with eager work in between
        B.solve(precision=mfs.PREC_MIXED, tol=1e-9, use_graph=True)
        b = [B.basis(c) for c in range(B.ncells)]
        assert all(np.array_equal(x, y) for x, y in zip(a, b))
        assert np.array_equal(it_a, B.iterations())


def test_guess_exact_and_bad_guess_fallback(ctx, small):
    p, B = small
    B.solve(precision=mfs.PREC_MIXED, tol=1e-9)
    it0 = B.iterations()
    # interior values in the [cell][interior node][3][24] layout
    guess = []
    for c in range(B.ncells):
        n = int(p["cell_n"][c])
        s, m = n + 1, n - 1
        basis = B.basis(c).reshape(24, -1, 3)
        ids = [(i + s * (j + s * k)) for k in range(1, n) for j in range(1, n) for i in range(1, n)]
        guess.append(np.transpose(basis[:, ids, :], (1, 2, 0)).ravel())
    guess = np.concatenate(guess)
    B.set_guess(guess)
    rep = B.solve(precision=mfs.PREC_MIXED, tol=1e-9, use_guess=True)
    assert rep.n_not_converged == 0 and rep.max_iterations <= 2 and rep.n_fallback == 0
    B.set_guess(guess * 1e6)
    rep = B.solve(precision=mfs.PREC_MIXED, tol=1e-9, use_guess=True)
    assert rep.n_not_converged == 0 and rep.n_fallback == B.ncells * 24
    assert np.array_equal(B.iterations(), it0)


@pytest.mark.slow
def test_config2_full_batch(ctx):
    """16^3 coarse cells, n ~ U{6..10}, lognormal E: 98,304 local problems."""
    p = problems.msfem_cells(16)
    B = make_batch(ctx, p)
    rep = B.solve(precision=mfs.PREC_MIXED, tol=1e-8)
    print("config2 mixed:", rep)
    assert rep.n_not_converged == 0 and rep.max_true_relres <= 1e-8
    # spot-check one cell of each size against the oracle
    for n in np.unique(p["cell_n"])[:2]:
        c = int(np.nonzero(p["cell_n"] == n)[0][0])
        basis_o, _, _ = orc.msfem_cell(int(n), p["origin"][c], p["h"], problems.cell_E(p, c), p["nu"], tol=1e-12)
        err = np.linalg.norm(B.basis(c) - basis_o) / np.linalg.norm(basis_o)
        assert err < 1e-6, (c, n, err)
