"""Learned initial guess for the MsFEM local problems.

Pipeline (all data stays in device memory):
  1. training cells with seeded lognormal moduli are solved by the batched
     solver (the MsFEM method generates its own training data);
  2. features = normalised log-moduli per fine tet, targets = basis values at
     the interior nodes minus the trilinear hat interpolant;
  3. a ReLU MLP (BF16 MFMA GEMMs, FP32 accumulation, Adam) is trained on them;
  4. on held-out cells the prediction (hat + network output) seeds the
     batched CG; the zero guess stays the fallback (a guess worse than zero is
     rejected per problem) and every solve still runs to the same tolerance.

`train` does steps 1-3 once; `LearnedSolver.solve` is the per-batch hot path
(features, network inference, guess assembly and the mixed batched solve);
`run` chains both and compares against the zero and hat guesses.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import (MLP, PREC_F64, PREC_MIXED, Batch, F32, Vec, field_lognormal)

MU_LOGE = math.log(1e9)
SIGMA_LOGE = 1.0
INFER_BATCH = 8192  # samples per inference chunk (large GEMMs for the MFMA units)


def cells(count, n, seed, h=1.0 / 32):
    """`count` cells of n^3 hexes on a line (the local problems are
    translation invariant), lognormal E per tet (stream 1000 + cell)."""
    cell_n = np.full(count, n, np.int32)
    origin = np.zeros((count, 3))
    origin[:, 0] = np.arange(count) * h
    E = np.concatenate([field_lognormal(6 * n ** 3, seed, 1000 + c, 0, MU_LOGE, SIGMA_LOGE) for c in range(count)])
    return cell_n, origin, h, E


@dataclass
class Trained:
    net: MLP
    n: int
    shape: dict  # precision -> batch-solve launch shape (autotuned or default)
    info: dict


def train(ctx, n=8, n_train=20000, hidden=256, layers=4, epochs=30, batch=256, lr=1e-3, tol=1e-8, seed_train=4,
          log=print, tune=False, shape=None):
    """Solve `n_train` cells with the batched mixed solver, train the MLP on
    (features -> basis corrections) and return it with the launch shapes
    (autotuned when `tune`, else `shape` or the library's defaults)."""
    nt = 6 * n ** 3
    nout = 3 * (n - 1) ** 3 * 24
    info = {"n": n, "n_train": n_train, "input": nt, "output": nout, "hidden": hidden, "layers": layers,
            "epochs": epochs, "batch": batch, "lr": lr}
    cn, org, h, E = cells(n_train, n, seed_train)
    Btr = Batch(ctx, cn, org, h, E)
    Btr.assemble()
    shape = dict(shape) if shape else {PREC_MIXED: {}, PREC_F64: {}}
    if tune:
        # launch shapes tuned on the training cells (same subdivision as the
        # test cells) and reused for the test batch
        for prec in (PREC_MIXED, PREC_F64):
            wg, _, var, _ = Btr.autotune(precision=prec, tol=tol)
            shape[prec] = {"wg_size": wg, "variant": var + 1}
        info["autotuned_shape"] = {"mixed": shape[PREC_MIXED], "fp64": shape[PREC_F64]}
    rep = Btr.solve(precision=PREC_MIXED, tol=tol, **mixed_shape(shape))
    info["train_solve_ms"] = round(rep.seconds * 1e3, 1)
    X = Vec(ctx, n_train * nt, F32)
    T = Vec(ctx, n_train * nout, F32)
    Btr.features(X, MU_LOGE, SIGMA_LOGE)
    Btr.targets(T)
    Btr.close()
    log(f"[learned] training data: {n_train} cells solved in {rep.seconds * 1e3:.0f} ms")
    dims = [nt] + [hidden] * (layers - 1) + [nout]
    net = MLP(ctx, dims, seed=11, lr=lr)
    t1 = time.time()
    loss = net.train_vec(n_train, X, T, batch, epochs, seed=3)
    info["train_seconds"] = round(time.time() - t1, 2)
    info["final_train_loss"] = loss
    ms, tf = net.time_step(batch, 10)
    info["mlp_step_ms"] = round(ms, 3)
    info["mlp_tflops"] = round(tf, 2)
    info["mlp_mfma_frac_bf16_peak"] = round(tf / 2500.0, 4)
    log(f"[learned] trained {epochs} epochs in {info['train_seconds']} s, loss {loss:.3e}, "
        f"fwd+bwd {ms:.2f} ms/batch = {tf:.1f} TFLOP/s")
    del X, T
    return Trained(net, n, shape, info)


def mixed_shape(shape):
    """Launch shape of a mixed batched solve: the mixed shape for the FP32
    launch and the FP64 shape for the FP64 pass over its failures."""
    out = dict(shape[PREC_MIXED])
    f64 = shape.get(PREC_F64) or {}
    if f64.get("wg_size"):
        out["fallback_wg_size"] = f64["wg_size"]
    if f64.get("variant"):
        out["fallback_variant"] = f64["variant"]
    return out


class LearnedSolver:
    """The per-batch hot path of the learned guess on one batch of cells:
    features of the moduli, network inference in chunks of INFER_BATCH,
    guess = hat + prediction, mixed batched solve from that guess.  Work
    vectors are allocated once per batch size."""

    def __init__(self, ctx, model: Trained, B: Batch, ncells: int, tol=1e-8):
        self.ctx, self.m, self.B, self.ncells, self.tol = ctx, model, B, ncells, tol
        n = model.n
        self.nt, self.nout = 6 * n ** 3, 3 * (n - 1) ** 3 * 24
        self.X = Vec(ctx, ncells * self.nt, F32)
        self.Y = Vec(ctx, ncells * self.nout, F32)

    def infer(self):
        self.B.features(self.X, MU_LOGE, SIGMA_LOGE)
        self.m.net.forward_vec(self.ncells, self.X, self.Y, min(INFER_BATCH, self.ncells))
        self.B.guess_from(self.Y)

    def solve(self):
        """One step: inference + guess + batched solve (returns the batch
        report; the caller times the whole step)."""
        self.infer()
        return self.B.solve(precision=PREC_MIXED, tol=self.tol, use_guess=True, **mixed_shape(self.m.shape))

    def close(self):
        del self.X, self.Y


def run(ctx, n=8, n_train=20000, n_test=4096, hidden=256, layers=4, epochs=30, batch=256, lr=1e-3, tol=1e-8,
        seed_train=4, seed_test=5, log=print, tune=False, model: Trained | None = None):
    """Train (unless `model` is given), then solve `n_test` held-out cells
    from the zero, the hat and the learned guess (mixed) and from the zero
    guess in plain FP64; report iterations, guess quality and times."""
    t0 = time.time()
    if model is None:
        model = train(ctx, n=n, n_train=n_train, hidden=hidden, layers=layers, epochs=epochs, batch=batch, lr=lr,
                      tol=tol, seed_train=seed_train, log=log, tune=tune)
    out = dict(model.info)
    out["n_test"] = n_test
    shape = model.shape
    cn, org, h, E = cells(n_test, n, seed_test)
    B = Batch(ctx, cn, org, h, E)
    B.assemble()
    res = {}
    # the plain FP64 batched solve from the zero guess (the comparison point)
    res["fp64_zero"] = B.solve(precision=PREC_F64, tol=tol, **shape[PREC_F64])
    r0 = B.solve(precision=PREC_MIXED, tol=tol, **mixed_shape(shape))
    it0, rr0, _ = B.status()
    res["zero"] = r0
    # reference solution for guess-quality figures
    Tref = Vec(ctx, n_test * model.info["output"], F32)
    B.targets(Tref)
    corr = Tref.download()
    del Tref
    B.guess_from(None)
    r_hat = B.solve(precision=PREC_MIXED, tol=tol, use_guess=True, **mixed_shape(shape))
    it_hat, _, _ = B.status()
    res["hat"] = r_hat
    L = LearnedSolver(ctx, model, B, n_test, tol)
    ctx.sync()
    t1 = time.perf_counter()
    L.infer()
    ctx.sync()
    infer_s = time.perf_counter() - t1
    pred = L.Y.download()
    r_l = B.solve(precision=PREC_MIXED, tol=tol, use_guess=True, **mixed_shape(shape))
    it_l, rr_l, fl_l = B.status()
    res["learned"] = r_l
    L.close()
    # guess quality: relative error of the interior correction (basis - hat)
    learned_err = float(np.linalg.norm(pred - corr) / max(np.linalg.norm(corr), 1e-300))
    for name, r in res.items():
        out[f"{name}_guess"] = {"ms": round(r.seconds * 1e3, 1), "mean_iterations": round(r.mean_iterations, 2),
                                "max_iterations": r.max_iterations, "max_true_relres": r.max_true_relres,
                                "not_converged": r.n_not_converged, "rejected_guesses": r.n_fallback,
                                "fp32_to_fp64_fallbacks": r.n_precision_fallback}
    out["inference_ms"] = round(infer_s * 1e3, 2)
    out["correction_rel_error"] = {"hat_guess": 1.0, "learned_guess": round(learned_err, 4)}
    out["iterations_saved_vs_zero"] = {
        "hat": round(1 - float(it_hat.mean()) / float(it0.mean()), 4),
        "learned": round(1 - float(it_l.mean()) / float(it0.mean()), 4)}
    out["iterations_saved_learned_vs_hat"] = round(1 - float(it_l.mean()) / float(it_hat.mean()), 4)
    out["learned_all_converged"] = bool(r_l.n_not_converged == 0 and r_l.max_true_relres <= tol)
    out["speedup_mixed_learned_vs_fp64_zero"] = round(res["fp64_zero"].seconds / (r_l.seconds + infer_s), 3)
    out["speedup_mixed_zero_vs_fp64_zero"] = round(res["fp64_zero"].seconds / r0.seconds, 3)
    out["seconds_total"] = round(time.time() - t0, 1)
    B.close()
    log(f"[learned] {out}")
    return out
