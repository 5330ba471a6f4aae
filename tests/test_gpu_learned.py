"""The learned initial guess end to end on a small configuration: the MsFEM
solver generates training data, the BF16-MFMA MLP is trained on it, and on
held-out cells the learned guess must converge to the same tolerance as the
zero guess while saving iterations over both the zero and the hat guess."""
import pytest

import multiscalefemsolver3d_amd as mfs
from multiscalefemsolver3d_amd import learned

pytestmark = pytest.mark.gpu


def test_learned_guess_small(ctx):
    out = learned.run(ctx, n=4, n_train=2048, n_test=128, hidden=128, layers=3, epochs=60, batch=128, lr=2e-3,
                      tol=1e-8)
    for k in ("zero_guess", "hat_guess", "learned_guess"):
        assert out[k]["not_converged"] == 0 and out[k]["max_true_relres"] <= 1e-8, (k, out[k])
    assert out["correction_rel_error"]["learned_guess"] < 0.8, out
    assert out["iterations_saved_vs_zero"]["learned"] > 0.05, out
    assert out["iterations_saved_learned_vs_hat"] > 0.0, out


def q1_hat_values(n):
    """The Q1 coarse basis of an n-cell at its interior nodes, laid out as
    the batch's guesses: entry (v*3 + a)*24 + q of interior node v = i + m*(j
    + m*k) (m = n - 1, fine node (i+1, j+1, k+1)/n), basis q = 3*c + d of
    corner c (bit 0: x = 1, bit 1: y = 1, bit 2: z = 1) and direction d.
    The corner's hat is prod(1 - |x - corner|), written independently of the
    library's form."""
    import numpy as np
    m = n - 1
    g = (np.arange(m) + 1.0) / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")  # node v = i + m*(j + m*k)
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    out = np.zeros((m ** 3, 3, 24))
    for c in range(8):
        corner = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1], dtype=float)
        hat = np.prod(1.0 - np.abs(pts - corner), axis=1)
        for d in range(3):
            out[:, d, 3 * c + d] = hat
    return out.ravel()


def test_guess_and_targets_use_the_q1_hat(ctx):
    """targets(T) = x - hat and guess_from(G) = hat + G against an
    independent restatement of the hat functions, and their round trip."""
    import numpy as np
    n, count = 5, 3
    cn, org, h, E = learned.cells(count, n, 7)
    B = mfs.Batch(ctx, cn, org, h, E)
    hat = np.tile(q1_hat_values(n), count)
    rng = np.random.default_rng(3)
    x = rng.normal(size=hat.size)
    B.set_guess(x)
    T = mfs.Vec(ctx, hat.size, mfs.F32)
    B.targets(T)
    np.testing.assert_allclose(T.download(), (x - hat).astype(np.float32), rtol=1e-6, atol=1e-7)
    # guess_from(None): x = hat exactly, so the targets are exactly zero
    B.guess_from(None)
    B.targets(T)
    assert not np.any(T.download())
    # guess_from(G) then targets: (hat + G) - hat gives G back bit for bit
    G = rng.normal(size=hat.size).astype(np.float32)
    Gv = mfs.Vec(ctx, hat.size, mfs.F32, G)
    B.guess_from(Gv)
    B.targets(T)
    np.testing.assert_array_equal(T.download(), G)
    B.close()
