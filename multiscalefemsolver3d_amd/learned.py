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
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import (MLP, PREC_F64, PREC_MIXED, Batch, F32, Vec, field_lognormal)

MU_LOGE = math.log(1e9)
SIGMA_LOGE = 1.0


def cells(count, n, seed, h=1.0 / 32):
    """`count` cells of n^3 hexes on a line (the local problems are
    translation invariant), lognormal E per tet (stream 1000 + cell)."""
    cell_n = np.full(count, n, np.int32)
    origin = np.zeros((count, 3))
    origin[:, 0] = np.arange(count) * h
    E = np.concatenate([field_lognormal(6 * n ** 3, seed, 1000 + c, 0, MU_LOGE, SIGMA_LOGE) for c in range(count)])
    return cell_n, origin, h, E


def run(ctx, n=8, n_train=20000, n_test=4096, hidden=256, layers=4, epochs=30, batch=256, lr=1e-3, tol=1e-8,
        seed_train=4, seed_test=5, log=print, tune=False):
    t0 = time.time()
    nt = 6 * n ** 3
    nout = 3 * (n - 1) ** 3 * 24
    out = {"n": n, "n_train": n_train, "n_test": n_test, "input": nt, "output": nout, "hidden": hidden,
           "layers": layers, "epochs": epochs}
    # 1. training data from the solver itself
    cn, org, h, E = cells(n_train, n, seed_train)
    Btr = Batch(ctx, cn, org, h, E)
    Btr.assemble()
    shape = {PREC_MIXED: {}, PREC_F64: {}}
    if tune:
        # launch shapes tuned on the training cells (same subdivision as the
        # test cells) and reused for the test batch
        for prec in (PREC_MIXED, PREC_F64):
            wg, _, var, _ = Btr.autotune(precision=prec, tol=tol)
            shape[prec] = {"wg_size": wg, "variant": var + 1}
        out["autotuned_shape"] = {"mixed": shape[PREC_MIXED], "fp64": shape[PREC_F64]}
    rep = Btr.solve(precision=PREC_MIXED, tol=tol, **shape[PREC_MIXED])
    out["train_solve_ms"] = round(rep.seconds * 1e3, 1)
    X = Vec(ctx, n_train * nt, F32)
    T = Vec(ctx, n_train * nout, F32)
    Btr.features(X, MU_LOGE, SIGMA_LOGE)
    Btr.targets(T)
    Btr.close()
    log(f"[learned] training data: {n_train} cells solved in {rep.seconds * 1e3:.0f} ms")
    # 2. train
    dims = [nt] + [hidden] * (layers - 1) + [nout]
    net = MLP(ctx, dims, seed=11, lr=lr)
    t1 = time.time()
    loss = net.train_vec(n_train, X, T, batch, epochs, seed=3)
    out["train_seconds"] = round(time.time() - t1, 2)
    out["final_train_loss"] = loss
    ms, tf = net.time_step(batch, 10)
    out["mlp_step_ms"] = round(ms, 3)
    out["mlp_tflops"] = round(tf, 2)
    out["mlp_mfma_frac_bf16_peak"] = round(tf / 2500.0, 4)
    log(f"[learned] trained {epochs} epochs in {out['train_seconds']} s, loss {loss:.3e}, "
        f"fwd+bwd {ms:.2f} ms/batch = {tf:.1f} TFLOP/s")
    del X, T
    # 3. held-out cells: zero guess vs hat guess vs learned guess
    cn, org, h, E = cells(n_test, n, seed_test)
    B = Batch(ctx, cn, org, h, E)
    B.assemble()
    res = {}
    # the plain FP64 batched solve from the zero guess (the comparison point)
    res["fp64_zero"] = B.solve(precision=PREC_F64, tol=tol, **shape[PREC_F64])
    r0 = B.solve(precision=PREC_MIXED, tol=tol, **shape[PREC_MIXED])
    it0, rr0, _ = B.status()
    res["zero"] = r0
    # reference solution for guess-quality figures
    Tref = Vec(ctx, n_test * nout, F32)
    B.targets(Tref)
    corr = Tref.download()
    B.guess_from(None)
    r_hat = B.solve(precision=PREC_MIXED, tol=tol, use_guess=True, **shape[PREC_MIXED])
    it_hat, _, _ = B.status()
    res["hat"] = r_hat
    Xt = Vec(ctx, n_test * nt, F32)
    B.features(Xt, MU_LOGE, SIGMA_LOGE)
    Y = Vec(ctx, n_test * nout, F32)
    net.forward_vec(n_test, Xt, Y, batch)
    pred = Y.download()
    B.guess_from(Y)
    r_l = B.solve(precision=PREC_MIXED, tol=tol, use_guess=True, **shape[PREC_MIXED])
    it_l, rr_l, fl_l = B.status()
    res["learned"] = r_l
    # guess quality: relative error of the interior correction (basis - hat)
    # and of the full interior basis
    hat_err = float(np.linalg.norm(corr) / max(np.linalg.norm(corr), 1e-300))
    learned_err = float(np.linalg.norm(pred - corr) / max(np.linalg.norm(corr), 1e-300))
    for name, r in res.items():
        out[f"{name}_guess"] = {"ms": round(r.seconds * 1e3, 1), "mean_iterations": round(r.mean_iterations, 2),
                                "max_iterations": r.max_iterations, "max_true_relres": r.max_true_relres,
                                "not_converged": r.n_not_converged, "rejected_guesses": r.n_fallback,
                                "fp32_to_fp64_fallbacks": r.n_precision_fallback}
    out["correction_rel_error"] = {"hat_guess": round(hat_err, 4), "learned_guess": round(learned_err, 4)}
    out["iterations_saved_vs_zero"] = {
        "hat": round(1 - float(it_hat.mean()) / float(it0.mean()), 4),
        "learned": round(1 - float(it_l.mean()) / float(it0.mean()), 4)}
    out["iterations_saved_learned_vs_hat"] = round(1 - float(it_l.mean()) / float(it_hat.mean()), 4)
    out["speedup_mixed_learned_vs_fp64_zero"] = round(res["fp64_zero"].seconds / r_l.seconds, 3)
    out["speedup_mixed_zero_vs_fp64_zero"] = round(res["fp64_zero"].seconds / r0.seconds, 3)
    out["seconds_total"] = round(time.time() - t0, 1)
    B.close()
    log(f"[learned] {out}")
    return out
