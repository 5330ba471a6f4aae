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
