#!/usr/bin/env python3
"""Benchmark of the MI355X FEM solve path.

Headline workload: the learned-guess MsFEM elasticity solve on a 32^3 coarse
grid (32,768 coarse cells, each a fine 8^3 x 6-tet sub-grid with lognormal
Young's modulus, 24 local basis problems per cell = 786,432 local problems).
A "step" is the per-batch hot path with the cell matrices resident in HBM:
features of the moduli, inference of the 4-layer width-256 MLP (BF16 MFMA,
FP32 accumulation) for all cells, guess assembly, and the mixed-precision
batched persistent PCG (FP32 inner iterations, FP64 refinement) from that
guess until every problem reaches a true FP64 relative residual of 1e-8.
The network is trained once beforehand on 20,000 cells the batched solver
solves itself (reported, not part of a step).  The comparison point is the
plain FP64 batched PCG from the zero guess on the same cells.  The CPU
baseline is the oracle's FP64 restatement of the reference's per-cell path
(CSSD assembly, 24 PCG solves per cell), sampled and extrapolated.

Sub-objects: 64^3 Laplace with inclusions (mixed vs FP64 Jacobi-PCG, the
fused PCG kernel's roofline), six right-hand sides, non-symmetric BiCGStab,
MsFEM on 16^3 cells with n ~ U{6..10} (Q1 and reference traces) and the
128^3 elastic cube (SpMV at HBM scale, assembly, solvers).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-1e-8 solve & SpMV HBM GB/s: mixed+batched+learned vs FP64 GPU CG"
HBM_PEAK_GBS = 8000.0
# Bytes per launch of k_cg_dir_spmv<float,1> (16-bit column offsets) on the
# 64^3 Laplace system between the L2 caches and the memory fabric, from
# rocprofv3 --pmc FETCH_SIZE (13,300.2 KiB, x2.000 from a 512 MiB copy) and
# WRITE_SIZE (2,179.2 KiB, x1.000), medians over 7,752 / 7,731 dispatches
TRAFFIC_K1 = 29470310
# HBM bytes of one k_batch_solve<float,1,true> launch on the headline
# workload (learned guess, launch shape 384 / variant 3): rocprofv3 --pmc
# FETCH_SIZE 57,946,409.9 KiB x2.000 + WRITE_SIZE 14,526,056.7 KiB x1.000
# (the copy calibration above).  The register-resident kernel keeps r, q, d
# and p on chip, so this is 0.058 of the algorithmic bytes.
TRAFFIC_BATCH = 133548929536


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def batch_bytes(iters, n, sv=4):
    """Algorithmic bytes of one batched-solve launch on cells of n^3 hexes:
    per cell-iteration one stencil SpMV (15 x 9 matrix values per interior
    node, operand and result vectors) with the matrix read once for all 24
    right-hand sides, plus the 11 vector passes of a PCG iteration per
    right-hand side; a cell iterates as long as its slowest problem."""
    import numpy as np
    nint = (n - 1) ** 3
    N = 3 * nint
    it = np.asarray(iters, dtype=np.int64).reshape(-1, 24)
    spmv = 135 * nint * sv + 2 * N * sv
    return int(it.max(axis=1).sum() * spmv + it.sum() * 11 * N * sv)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=32768, help="headline: coarse cells (32^3)")
    ap.add_argument("--train-cells", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--shape", default="", help="headline batch-solve launch shape WG,VARIANT for both precisions, "
                                               "or WG,VARIANT,WG64,VARIANT64, instead of the autotuner (counter runs)")
    ap.add_argument("--n", type=int, default=64, help="Laplace leg: mesh hexes per edge")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work per leg")
    ap.add_argument("--no-laplace", action="store_true", help="skip the 64^3 Laplace measurement")
    ap.add_argument("--no-msfem", action="store_true", help="skip the 16^3-cell MsFEM measurement")
    ap.add_argument("--no-large", action="store_true", help="skip the 128^3 elastic measurement")
    ap.add_argument("--no-bicgstab", action="store_true", help="skip the non-symmetric BiCGStab measurement")
    ap.add_argument("--headline-only", action="store_true", help="only the headline workload")
    ap.add_argument("--large-only", action="store_true",
                    help="only the 128^3 elastic leg (assembly, SpMV, solvers; --large-spmv-only skips the solvers)")
    ap.add_argument("--gemm-only", action="store_true",
                    help="only the network's BF16 MFMA GEMM (8192^3) and one inference pass of the headline "
                         "network over 32,768 random inputs (counter runs)")
    ap.add_argument("--large-n", type=int, default=128)
    ap.add_argument("--large-spmv-only", action="store_true", help="128^3: assembly and SpMV only, no solves")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 512 MiB device copy (reference kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    if args.headline_only:
        args.no_laplace = args.no_msfem = args.no_large = args.no_bicgstab = True

    import multiscalefemsolver3d_amd as mfs
    from multiscalefemsolver3d_amd import problems

    ctx = mfs.Context(0)
    if args.gemm_only or args.large_only:
        res = bench_gemm(mfs, ctx) if args.gemm_only else bench_large(mfs, problems, ctx, args.large_n,
                                                                        not args.large_spmv_only)
        print(json.dumps(res), flush=True)
        ctx.close()
        return
    head = bench_headline(mfs, ctx, args)
    lap = None if args.no_laplace else bench_laplace(mfs, problems, ctx, args)
    nonsym = None if args.no_bicgstab else bench_bicgstab(mfs, problems, ctx, cpu=not args.no_cpu)
    msfem = None if args.no_msfem else bench_msfem(mfs, problems, ctx, cpu=not args.no_cpu)
    large = None if args.no_large else bench_large(mfs, problems, ctx, args.large_n, not args.large_spmv_only)
    if args.copy_calibration:
        # one plain copy of 512 MiB: FETCH_SIZE and WRITE_SIZE of this kernel
        # should each read 524,288 KiB; their ratio to that scales the counters
        big = 64 * 1024 * 1024
        src = mfs.Vec(ctx, big, mfs.F64)
        dst = mfs.Vec(ctx, big, mfs.F64)
        src.fill(1.0)
        ctx.sync()
        dst.copy_from(src)
        ctx.sync()
        log(f"[bench] copy calibration: {big * 8} bytes copied")
        del src, dst

    out = {
        "metric": METRIC,
        "value": head["ms_per_step"],
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer (bf16 mfma inference)",
        "data": "synthetic (seeded lognormal moduli; the network is trained on cells the solver generates)",
        "config": head["config"],
        # FP64 GPU CG on the same problems: the batched register-resident
        # kernel (same design as the mixed one, the ratio quoted), and plain
        # FP64 PCG through mfs_solve one problem at a time (sampled)
        "fp64_gpu_cg_ms": head["fp64_batched_zero_ms"],
        "speedup_vs_fp64_gpu_cg": round(head["fp64_batched_zero_ms"] / head["ms_per_step"], 3),
        "fp64_plain_gpu_cg_one_at_a_time_ms": head["detail"]["fp64_plain_cg_one_problem_at_a_time"]["ms_extrapolated"],
        "speedup_vs_fp64_plain_gpu_cg": head["detail"]["fp64_plain_cg_one_problem_at_a_time"]["speedup_learned_vs_plain"],
        "headline": head["detail"],
        "roofline": head["roofline"],
        "cpu_baseline": head["cpu_baseline"],
        "laplace64_inclusions": lap,
        "bicgstab_nonsymmetric": nonsym,
        "msfem_16cubed_cells": msfem,
        "elastic_large": large,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_headline(mfs, ctx, args):
    """Learned-guess MsFEM solve of 32^3 cells (n = 8): timed steps of
    inference + guess + mixed batched solve; FP64 zero-guess batched PCG and
    mixed zero-guess beside it; roofline of the persistent batch kernel."""
    import numpy as np
    from multiscalefemsolver3d_amd import learned
    n, ncells = 8, args.cells
    t0 = time.time()
    shape = None
    if args.shape:
        v = [int(x) for x in args.shape.split(",")]
        wg, var = v[0], v[1]
        wg64, var64 = (v[2], v[3]) if len(v) == 4 else (wg, var)
        shape = {mfs.PREC_MIXED: {"wg_size": wg, "variant": var}, mfs.PREC_F64: {"wg_size": wg64, "variant": var64}}
    model = learned.train(ctx, n=n, n_train=args.train_cells, hidden=256, layers=4, epochs=args.epochs, batch=256,
                          lr=1e-3, tol=1e-8, log=log, tune=shape is None, shape=shape)
    train_wall = time.time() - t0
    cn, org, h, E = learned.cells(ncells, n, 5)
    B = mfs.Batch(ctx, cn, org, h, E)
    t0 = time.perf_counter()
    B.assemble()
    ctx.sync()
    asm_s = time.perf_counter() - t0
    L = learned.LearnedSolver(ctx, model, B, ncells, tol=1e-8)
    for _ in range(args.warmup):
        L.solve()
    ctx.sync()
    t_start = time.perf_counter()
    reps = [L.solve() for _ in range(args.steps)]
    ctx.sync()
    elapsed = time.perf_counter() - t_start
    ms = 1000.0 * elapsed / args.steps
    it_l, rr_l, fl_l = B.status()
    ok = all(r.n_not_converged == 0 and r.max_true_relres <= 1e-8 for r in reps)
    # inference alone (features + MLP + guess assembly)
    ctx.sync()
    t1 = time.perf_counter()
    for _ in range(3):
        L.infer()
    ctx.sync()
    infer_ms = 1000.0 * (time.perf_counter() - t1) / 3
    kern_ms = 1000.0 * float(np.mean([r.kernel_seconds for r in reps]))
    solve_ms = 1000.0 * float(np.mean([r.seconds for r in reps]))
    nbytes = batch_bytes(it_l, n)
    gbs = nbytes / (kern_ms * 1e-3) / 1e9
    log(f"[bench] headline: {ms:.1f} ms per step (solve {solve_ms:.1f} ms device, persistent kernel "
        f"{kern_ms:.1f} ms = {gbs:.0f} GB/s algorithmic, inference {infer_ms:.1f} ms), mean iterations "
        f"{it_l.mean():.1f}, all converged {ok}")
    # comparison points on the same cells: plain FP64 batched PCG and mixed
    # batched PCG, both from the zero guess (wall clock like the steps)
    def timed(**kw):
        best = None
        for _ in range(3):
            ctx.sync()
            t = time.perf_counter()
            r = B.solve(tol=1e-8, **kw)
            ctx.sync()
            w = time.perf_counter() - t
            if best is None or w < best[0]:
                best = (w, r)
        return best
    w64, r64 = timed(precision=mfs.PREC_F64, **model.shape[mfs.PREC_F64])
    it64, _, _ = B.status()
    w0, r0 = timed(precision=mfs.PREC_MIXED, **learned.mixed_shape(model.shape))
    it0, _, _ = B.status()
    ms64 = 1000.0 * w64
    log(f"[bench] FP64 batched zero guess {ms64:.1f} ms (mean it {it64.mean():.1f}); mixed zero guess "
        f"{1000 * w0:.1f} ms (mean it {it0.mean():.1f}); learned {ms:.1f} ms -> {ms64 / ms:.2f}x")
    # plain FP64 GPU CG (mfs_solve, Jacobi-PCG) on one local problem at a time,
    # sampled on two cells of the batch and extrapolated to all problems
    nu = 0.3  # the batch's Poisson ratio (mfs.Batch default)
    sec = float(np.mean([fp64_plain_cg_seconds(mfs, ctx, n, tuple(org[c]), h, E[6 * n ** 3 * c:6 * n ** 3 * (c + 1)], nu)
                         for c in (0, 1)]))
    plain = {"ms_per_problem": round(sec * 1e3, 3), "sample": "2 cells x 24 problems",
             "ms_extrapolated": round(sec * 1e3 * 24 * ncells, 1),
             "speedup_learned_vs_plain": round(sec * 1e3 * 24 * ncells / ms, 1)}
    detail = {
        "problems": 24 * ncells, "cell_n": n, "interior_dofs_per_cell": 3 * (n - 1) ** 3,
        "all_converged_to_1e-8": ok, "max_true_relres": max(r.max_true_relres for r in reps),
        "mean_iterations": {"learned_mixed": round(float(it_l.mean()), 2), "zero_mixed": round(float(it0.mean()), 2),
                            "zero_fp64": round(float(it64.mean()), 2)},
        "max_iterations": {"learned_mixed": int(it_l.max()), "zero_mixed": int(it0.max()),
                           "zero_fp64": int(it64.max())},
        "iterations_saved_learned_vs_zero_mixed": round(1 - float(it_l.mean()) / float(it0.mean()), 4),
        "step_breakdown_ms": {"inference_features_guess": round(infer_ms, 2), "batched_solve_device": round(solve_ms, 2),
                              "persistent_kernel": round(kern_ms, 2)},
        "fp32_to_fp64_fallbacks": reps[-1].n_precision_fallback, "rejected_guesses": reps[-1].n_fallback,
        "mixed_zero_guess_ms": round(1000 * w0, 2),
        "speedup_vs_mixed_zero_guess": round(1000 * w0 / ms, 3),
        "fp64_plain_cg_one_problem_at_a_time": plain,
        "training": dict(model.info, wall_s=round(train_wall, 1)),
        "cell_assembly_s": round(asm_s, 3),
        "timing": "wall clock of K steps with the context synchronised on both sides",
    }
    roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_BATCH,
            "kernel": "k_batch_solve<float,R,REG> (persistent batched PCG, FP32 inner / FP64 refinement)",
            "bytes_per_launch": nbytes, "avg_ms": round(kern_ms, 3),
            "measured_hbm_gbs": round(TRAFFIC_BATCH / (kern_ms * 1e6), 1),
            "note": "algorithmic bytes: per cell-iteration the 15-point stencil matrix once for all 24 "
                    "right-hand sides plus 11 vector passes per right-hand side and iteration; the "
                    "register-resident kernel keeps the vectors on chip (registers, LDS, L2), so the "
                    "algorithmic rate is an effective one that can exceed the HBM peak, while the "
                    "measured traffic (FETCH_SIZE/WRITE_SIZE) is far below it: the kernel is bound by "
                    "FP32 VALU issue and latency, not by HBM"}
    # the GEMM of the network alone at a large size (MFMA utilisation)
    gms, gtf = mfs.gemm_bf16_time(ctx, 8192, 8192, 8192, reps=5)
    detail["gemm_bf16_kernel_8192cubed"] = {"ms": round(gms, 3), "tflops": round(gtf, 1),
                                            "frac_dense_bf16_peak": round(gtf / 2500.0, 4)}
    # the VALU multiply-add rate the solver kernels are compiled for (plain
    # FP32, no packed instructions) beside FP64: the arithmetic ceiling of
    # the FP32-inner / FP64 speedup
    detail["valu_fma_tflops"] = {"f32_unpacked": round(mfs.valu_fma_rate(ctx, mfs.F32), 1),
                                 "f64": round(mfs.valu_fma_rate(ctx, mfs.F64), 1),
                                 "published_peaks": {"f32_vector": 157.3, "f64_vector": 78.6}}
    cpu = None
    if not args.no_cpu:
        cpu = msfem_cpu_baseline_uniform(learned, n, ncells, args.cpu_budget)
    L.close()
    B.close()
    config = {"workload": f"msfem_learned_{round(ncells ** (1 / 3))}^3_cells_n{n}", "coarse_cells": ncells,
              "fine_hexes_per_cell": n ** 3, "local_problems": 24 * ncells, "tol": 1e-8,
              "network": "MLP 3072-256-256-256-24696, BF16 MFMA / FP32 accumulation",
              "training_pairs": args.train_cells,
              "solver": "batched persistent Jacobi-PCG, FP32 inner + FP64 refinement, learned initial guess",
              "parallelism": "single GPU"}
    return {"ms_per_step": round(ms, 3), "fp64_batched_zero_ms": round(ms64, 3), "detail": detail,
            "roofline": roof, "cpu_baseline": cpu, "config": config}


def bench_gemm(mfs, ctx):
    """The MLP's BF16 MFMA GEMM kernel alone at 8192^3, then one inference
    pass of the headline network (3072-256-256-256-24696) over 32,768
    random inputs in chunks of 8,192."""
    import numpy as np
    ms, tf = mfs.gemm_bf16_time(ctx, 8192, 8192, 8192, reps=5)
    net = mfs.MLP(ctx, [3072, 256, 256, 256, 24696], seed=11, lr=1e-3)
    ns = 32768
    X = mfs.Vec(ctx, ns * 3072, mfs.F32, np.random.default_rng(0).normal(size=ns * 3072).astype(np.float32))
    Y = mfs.Vec(ctx, ns * 24696, mfs.F32)
    net.forward_vec(ns, X, Y, 8192)
    ctx.sync()
    t0 = time.perf_counter()
    net.forward_vec(ns, X, Y, 8192)
    ctx.sync()
    inf_ms = (time.perf_counter() - t0) * 1e3
    flops = 2.0 * ns * (3072 * 256 + 256 * 256 * 2 + 256 * 24696)
    del X, Y, net
    return {"gemm_8192cubed_ms": round(ms, 3), "gemm_tflops": round(tf, 1),
            "frac_dense_bf16_peak": round(tf / 2500.0, 4), "inference_ms": round(inf_ms, 3),
            "inference_tflops": round(flops / (inf_ms * 1e-3) / 1e12, 1),
            "vendor_library_8192cubed": vendor_gemm_bf16(8192)}


def vendor_gemm_bf16(n, reps=5):
    """The vendor library's BF16 GEMM (torch.mm, which calls hipBLASLt) on
    the same n^3 product, row-major contiguous operands, for comparison
    with the kernel above.  Timed like it: HIP events on torch's stream
    around reps back-to-back launches after three warm-up launches.  It runs
    in a child process that loads only torch, so that torch's HIP runtime
    and this library's do not share a process."""
    import subprocess
    code = (
        "import json, torch\n"
        f"n, reps = {int(n)}, {int(reps)}\n"
        "g = torch.Generator(device='cuda').manual_seed(0)\n"
        "a = torch.randn(n, n, device='cuda', dtype=torch.bfloat16, generator=g)\n"
        "b = torch.randn(n, n, device='cuda', dtype=torch.bfloat16, generator=g)\n"
        "c = torch.empty(n, n, device='cuda', dtype=torch.bfloat16)\n"
        "for _ in range(3): torch.mm(a, b, out=c)\n"
        "torch.cuda.synchronize()\n"
        "e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)\n"
        "e0.record()\n"
        "for _ in range(reps): torch.mm(a, b, out=c)\n"
        "e1.record(); e1.synchronize()\n"
        "print(json.dumps({'ms': e0.elapsed_time(e1) / reps, 'torch': torch.__version__}))\n")
    try:
        out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    except (OSError, subprocess.TimeoutExpired) as e:
        log(f"[bench] vendor GEMM not run: {e}")
        return None
    if out.returncode != 0:
        log(f"[bench] vendor GEMM failed ({out.returncode}): {out.stderr.strip()[-400:]}")
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    tf = 2.0 * n ** 3 / (r["ms"] * 1e-3) / 1e12
    return {"library": "torch.mm (hipBLASLt), BF16 output", "torch": r["torch"],
            "ms": round(r["ms"], 3), "tflops": round(tf, 1), "frac_dense_bf16_peak": round(tf / 2500.0, 4)}


def msfem_cpu_baseline_uniform(learned, n, ncells, budget):
    """The oracle's FP64 restatement of the reference's per-cell path on
    test cells of the headline batch (CSSD assembly; per basis the full
    cell system with eliminated boundary rows in component-major order,
    solved by the SSOR-PCG continuation with BCG_Stab2 fallback of
    FEM.h:3699-3770 down to 1e-8; the Galerkin product), one cell per host
    thread as the reference's OpenMP loop over cells runs them; as many
    cells as fit the time budget, extrapolated to the whole batch."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    import oracle as orc
    threads = min(os.cpu_count() or 1, 16)
    cn, org, h, E = learned.cells(threads, n, 5)
    nt = 6 * n ** 3

    def one(c):
        t0 = time.perf_counter()
        orc.msfem_cell_refpath(n, tuple(org[c]), h, E[c * nt:(c + 1) * nt], 0.3, min_residual=1e-8)
        return time.perf_counter() - t0

    t0 = time.perf_counter()
    rounds = 0
    per = []
    with ThreadPoolExecutor(max_workers=threads) as ex:
        while rounds == 0 or (time.perf_counter() - t0 < budget and rounds < 8):
            per += list(ex.map(one, range(threads)))
            rounds += 1
    wall = time.perf_counter() - t0
    cells_done = rounds * threads
    est_ms = wall / cells_done * ncells * 1e3
    return {"value": round(est_ms, 1), "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"{cells_done} cells (n={n}; 24 local problems each on the reference's path: SSOR-PCG "
                      f"continuation with BCG_Stab2 fallback to 1e-8, oracle restatement) on {threads} threads in "
                      f"{wall:.1f}s, mean {np.mean(per):.3f}s per cell under that load; extrapolated to {ncells} "
                      f"cells"}


def bench_laplace(mfs, problems, ctx, args):
    """64^3 Laplace with 200 inclusions: mixed vs FP64 Jacobi-PCG
    to 1e-8, the fused PCG kernel's roofline, plain SpMV, six right-hand
    sides, the reference CPU solver path."""
    import numpy as np
    t0 = time.time()
    p = problems.laplace_inclusions(args.n)
    ctx.sync()
    S = mfs.Assembly(ctx, mfs.PROBLEM_SCALAR, p["xyz"], p["tets"], fixed_nodes=p["fixed_nodes"],
                     fixed_values=p["fixed_values"])
    setup_s = time.time() - t0
    S.run(p["sigma"])
    asm_ms = S.time(10)
    A = S.A
    mixed = mfs.Solver(A, precision=mfs.PREC_MIXED, tol=1e-8, inner_tol=1e-4)
    fp64 = mfs.Solver(A, precision=mfs.PREC_F64, tol=1e-8)
    x = mfs.Vec(ctx, A.n)

    def run(solver):
        x.copy_from(S.x0)
        return solver.run(S.b, x)

    for _ in range(2):
        run(mixed)
        run(fp64)
    reps = 10
    ctx.sync()
    t = time.perf_counter()
    for _ in range(reps):
        r_m = run(mixed)
    ctx.sync()
    ms_mixed = 1000.0 * (time.perf_counter() - t) / reps
    x_mixed = x.download()
    ctx.sync()
    t = time.perf_counter()
    for _ in range(5):
        r_d = run(fp64)
    ctx.sync()
    ms_fp64 = 1000.0 * (time.perf_counter() - t) / 5
    x_fp64 = x.download()
    agree = float(np.linalg.norm(x_mixed - x_fp64) / np.linalg.norm(x_fp64))
    log(f"[bench] laplace: mixed {ms_mixed:.2f} ms ({r_m.iterations} inner it, {r_m.outer_steps} outer), "
        f"fp64 {ms_fp64:.2f} ms ({r_d.iterations} it); |x_mixed-x_fp64|/|x_fp64| = {agree:.2e}")
    k1_ms, k1_bytes = mixed.kernel_time(0, 200)
    k2_ms, k2_bytes = mixed.kernel_time(1, 200)
    k1_gbs = k1_bytes / (k1_ms * 1e-3) / 1e9
    k2_gbs = k2_bytes / (k2_ms * 1e-3) / 1e9
    X32, Y32 = mfs.Vec(ctx, A.n, mfs.F32, x_mixed), mfs.Vec(ctx, A.n, mfs.F32)
    spmv_ms = A.spmv_time(X32, Y32, 200, mfs.F32)
    spmv_gbs = A.spmv_bytes(mfs.F32) / (spmv_ms * 1e-3) / 1e9
    X64, Y64 = mfs.Vec(ctx, A.n, mfs.F64, x_mixed), mfs.Vec(ctx, A.n, mfs.F64)
    spmv64_ms = A.spmv_time(X64, Y64, 200, mfs.F64)
    spmv64_gbs = A.spmv_bytes(mfs.F64) / (spmv64_ms * 1e-3) / 1e9
    multi6 = bench_multi6(mfs, ctx, S, A, mixed)
    cpu = None
    if not args.no_cpu:
        cpu = laplace_cpu_baseline(A, S, r_d.iterations, args.cpu_budget)
    out = {"nodes": len(p["xyz"]), "tets": len(p["tets"]), "dofs": A.n, "nnz": A.nnz, "tol": 1e-8,
           "host_setup_s": round(setup_s, 2), "assembly_ms": round(asm_ms, 4),
           "mixed_ms": round(ms_mixed, 3), "fp64_ms": round(ms_fp64, 3),
           "speedup_mixed_vs_fp64": round(ms_fp64 / ms_mixed, 3),
           "iterations": {"mixed_inner": r_m.iterations, "mixed_outer": r_m.outer_steps, "fp64": r_d.iterations},
           "true_relres": {"mixed": r_m.true_relres, "fp64": r_d.true_relres},
           "mixed_vs_fp64_solution_rel_diff": agree,
           "k_cg_dir_spmv_f32": {"avg_us": round(k1_ms * 1e3, 3), "bytes_per_launch": k1_bytes,
                                 "gbs": round(k1_gbs, 1), "frac_hbm": round(k1_gbs / HBM_PEAK_GBS, 4),
                                 "traffic": TRAFFIC_K1,
                                 "note": "working set (35 MB) is resident in the 256 MB Infinity Cache: "
                                         "effective bandwidth"},
           "k_cg_update_f32": {"avg_us": round(k2_ms * 1e3, 3), "bytes_per_launch": k2_bytes,
                               "gbs": round(k2_gbs, 1)},
           "spmv_f32_us": round(spmv_ms * 1e3, 3), "spmv_f32_gbs": round(spmv_gbs, 1),
           "spmv_f64_gbs": round(spmv64_gbs, 1), "multi_rhs_6": multi6, "cpu_baseline": cpu}
    S.close()
    return out


def bench_multi6(mfs, ctx, S, A, mixed):
    """Six right-hand sides with one matrix (the homogenisation driver's load
    cases): one iteration stream with one SpMM per iteration vs six solves."""
    import numpy as np
    rng = np.random.default_rng(6)
    b0 = S.b.download()
    bs6 = [mfs.Vec(ctx, A.n, mfs.F64, b0 * (1.0 + 0.2 * j) + rng.normal(size=A.n) * np.abs(b0).max() * 0.05)
           for j in range(6)]
    xs6 = [mfs.Vec(ctx, A.n) for _ in range(6)]

    def reset6():
        for v in xs6:
            v.copy_from(S.x0)

    reset6()
    mfs.solve_multi(A, bs6, xs6, precision=mfs.PREC_MIXED, tol=1e-8, inner_tol=1e-4)  # warm-up
    reset6()
    ctx.sync()
    t6 = time.perf_counter()
    reps6 = mfs.solve_multi(A, bs6, xs6, precision=mfs.PREC_MIXED, tol=1e-8, inner_tol=1e-4)
    ctx.sync()
    joint_ms = 1000.0 * (time.perf_counter() - t6)
    reset6()
    ctx.sync()
    t6 = time.perf_counter()
    for j in range(6):
        mixed.run(bs6[j], xs6[j])
    ctx.sync()
    seq_ms = 1000.0 * (time.perf_counter() - t6)
    log(f"[bench] 6 right-hand sides: joint {joint_ms:.1f} ms vs sequential {seq_ms:.1f} ms")
    return {"systems": 6, "joint_ms": round(joint_ms, 2), "sequential_ms": round(seq_ms, 2),
            "speedup": round(seq_ms / joint_ms, 3), "max_true_relres": max(r.true_relres for r in reps6),
            "iterations": [r.iterations for r in reps6], "precision": "f32 inner / f64 outer"}


def laplace_cpu_baseline(A, S, gpu_it, budget):
    """The reference's CPU solver path restated in the oracle
    (MSG_PreconditioningSSOR on CSSD rows, OpenMP dynamic SpMV, serial dots
    and sweeps) on the same system: to 1e-8 if it fits the budget, else a
    timed prefix projected with the SSOR-PCG's own convergence rate."""
    import numpy as np
    import oracle as orc
    rp, col, val = A.export_csr()
    b = S.b.download()
    x0 = S.x0.download()
    threads = min(os.cpu_count() or 1, 16)
    _, it_p, _, sec_p = orc.refpath_solve(rp, col, val, b, x0, tol=1e-8, maxit=6, mode=0, threads=threads)
    per_it = sec_p / max(it_p, 1)
    budget_it = int(max(10, budget / max(per_it, 1e-6)))
    xc, it_c, rr_c, sec_c = orc.refpath_solve(rp, col, val, b, x0, tol=1e-8, maxit=budget_it, mode=0,
                                              threads=threads)
    if rr_c <= 1e-8:
        value = sec_c * 1000.0
        sample = f"full SSOR-PCG solve to 1e-8 ({it_c} iterations, true relres {rr_c:.2e})"
    else:
        # the SSOR-PCG's own rate over the timed iterations: relres falls
        # like rr_c ** (k / it_c); iterations to 1e-8 at that rate
        rate_it = it_c * np.log(1e-8) / np.log(rr_c) if 0 < rr_c < 1 else float("nan")
        value = sec_c / it_c * rate_it * 1000.0
        sample = (f"{it_c} SSOR-PCG iterations timed ({sec_c:.1f}s, true relres {rr_c:.2e}), projected to "
                  f"{rate_it:.0f} iterations to 1e-8 with their own average convergence rate (the FP64 GPU "
                  f"Jacobi-PCG needs {gpu_it})")
    log(f"[bench] laplace cpu baseline {value:.0f} ms ({sample})")
    return {"value": round(value, 2), "unit": "ms", "cores": threads, "kind": "port",
            "sample": "reference CPU solver path (MSG_PreconditioningSSOR, CSSD rows, OpenMP dynamic SpMV, "
                      "serial dots/sweeps) on the same system: " + sample}


def bench_bicgstab(mfs, problems, ctx, cpu=True):
    """Non-symmetric solve (the BiCG_Stab surface): upwind convection-
    diffusion on a 512^2 grid (262,144 unknowns, Peclet 200) to 1e-8 with
    right-Jacobi BiCGStab, FP64 and mixed, against GMRES(30) mixed; the
    oracle's BiCGStab on the host threads as the CPU reference (a bounded
    sample of iterations, projected)."""
    import numpy as np
    rp, col, val, b = problems.convection_diffusion(512, 200.0, seed=9)
    n = len(b)
    A = mfs.Mat.csr(ctx, rp, col, val)
    bv = mfs.Vec(ctx, n, mfs.F64, b)
    res = {"unknowns": n, "nnz": int(len(val)), "tol": 1e-8}
    for name, method, prec in (("bicgstab_fp64", mfs.BICGSTAB, mfs.PREC_F64),
                               ("bicgstab_mixed", mfs.BICGSTAB, mfs.PREC_MIXED),
                               ("gmres30_mixed", mfs.GMRES, mfs.PREC_MIXED)):
        S = mfs.Solver(A, method=method, precision=prec, tol=1e-8, maxit=20000)
        x = mfs.Vec(ctx, n, mfs.F64, np.zeros(n))
        S.run(bv, x, check=False)  # warm (graph capture, first launches)
        x.upload(np.zeros(n))
        r = S.run(bv, x, check=False)
        res[name] = {"ms": round(r.seconds * 1e3, 2), "iterations": r.iterations, "outer": r.outer_steps,
                     "true_relres": r.true_relres, "status": r.status}
    if cpu:
        import oracle as orc
        t0 = time.perf_counter()
        _, it_c, _ = orc.bicgstab(rp, col, val, b, tol=1e-8, maxit=200)
        sec = time.perf_counter() - t0
        it_g = res["bicgstab_fp64"]["iterations"]
        res["cpu_oracle_bicgstab"] = {"ms_projected": round(sec / max(it_c, 1) * it_g * 1e3, 1),
                                      "sample": f"{it_c} iterations timed ({sec:.2f}s, 1 thread), projected to "
                                                f"{it_g} (the FP64 GPU count)"}
    log(f"[bench] non-symmetric (BiCGStab): {res}")
    return res


def bench_msfem(mfs, problems, ctx, cpu=True):
    """MsFEM local problems: 16^3 coarse cells, n ~ U{6..10}, lognormal E (sigma 1):
    98,304 local problems to 1e-8.  Mixed batched persistent launch (hipGraph)
    vs the same launch in plain FP64 vs one-problem-at-a-time FP64 Jacobi-PCG
    through mfs_solve (sampled cells, extrapolated per problem)."""
    import numpy as np
    p = problems.msfem_cells(16)
    B = mfs.Batch(ctx, p["cell_n"], p["origin"], p["h"], p["E"], p["nu"])
    B.assemble()
    res = {"autotuned_shape": {}}
    for name, prec in (("mixed", mfs.PREC_MIXED), ("fp64", mfs.PREC_F64)):
        wg, per_cu, var, ms = B.autotune(precision=prec, tol=1e-8)
        res["autotuned_shape"][name] = {"workgroup": wg, "workgroups_per_cu": per_cu,
                                        "rows_per_thread_variant": var, "tuning_best_ms": round(ms, 2)}
    log(f"[bench] msfem autotuned shapes: {res['autotuned_shape']}")
    for name, prec in (("mixed_batched", mfs.PREC_MIXED), ("fp64_batched", mfs.PREC_F64)):
        reps = [B.solve(precision=prec, tol=1e-8, use_graph=True) for _ in range(3)]
        r = min(reps, key=lambda x: x.seconds)
        res[name] = {"ms": round(r.seconds * 1e3, 2), "timing": "device time of the graph launch (HIP events)",
                     "mean_iterations": round(r.mean_iterations, 1),
                     "max_iterations": r.max_iterations, "max_true_relres": r.max_true_relres,
                     "not_converged": r.n_not_converged, "fp32_to_fp64_fallbacks": r.n_precision_fallback}
    # plain FP64 GPU CG, one local problem at a time (sample of cells)
    sample = [int(np.nonzero(p["cell_n"] == n)[0][0]) for n in (6, 8, 10)]
    per_problem = [(int(p["cell_n"][c]), fp64_plain_cg_seconds(mfs, ctx, int(p["cell_n"][c]), tuple(p["origin"][c]),
                                                                p["h"], problems.cell_E(p, c), p["nu"]))
                   for c in sample]
    # extrapolate with the batch's cell-size histogram
    counts = {int(n): int((p["cell_n"] == n).sum()) for n in np.unique(p["cell_n"])}
    t_of = dict(per_problem)
    ns = sorted(t_of)
    est = sum(24 * cnt * float(np.interp(n, ns, [t_of[k] for k in ns])) for n, cnt in counts.items())
    res["fp64_one_at_a_time_ms_extrapolated"] = round(est * 1e3, 1)
    res["fp64_one_at_a_time_sample"] = {f"n={n}": round(t * 1e3, 3) for n, t in per_problem}
    res["speedup_mixed_batched_vs_fp64_one_at_a_time"] = round(est * 1e3 / res["mixed_batched"]["ms"], 1)
    res["speedup_mixed_batched_vs_fp64_batched"] = round(res["fp64_batched"]["ms"] / res["mixed_batched"]["ms"], 3)
    res["problems"] = 24 * len(p["cell_n"])
    B.coarse_matrices()  # warm
    t0 = time.perf_counter()
    Kc = B.coarse_matrices()
    res["coarse_matrices_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    res["coarse_matrices_note"] = "wall clock of mfs_batch_coarse_matrices incl. the 24x24 per cell download"
    del Kc
    B.close()
    res["reference_trace"] = bench_msfem_reference_trace(mfs, ctx, p)
    if cpu:
        res["cpu_baseline"] = msfem_cpu_baseline(p, counts)
    log(f"[bench] msfem: {res}")
    return res


def fp64_plain_cg_seconds(mfs, ctx, n, origin, h, E, nu):
    """Mean device time of one local problem of a cell (n^3 hexes, per-tet
    modulus E) solved on its own through mfs_solve: FP64 Jacobi-PCG to 1e-8,
    the 24 hat-function boundary conditions one after the other."""
    import numpy as np
    xyz, tets = mfs.mesh_box(n, origin=origin, h=(h / n,) * 3)
    s = n + 1
    ijk = np.array([[i, j, k] for k in range(s) for j in range(s) for i in range(s)])
    bnd = np.nonzero(((ijk == 0) | (ijk == n)).any(1))[0]
    X = ijk / n
    times = []
    for q in range(24):
        cc, d = q // 3, q % 3
        hat = ((X[:, 0] if cc & 1 else 1 - X[:, 0]) * (X[:, 1] if cc & 2 else 1 - X[:, 1]) *
               (X[:, 2] if cc & 4 else 1 - X[:, 2]))
        g = np.zeros((len(bnd), 3))
        g[:, d] = hat[bnd]
        S = mfs.Assembly(ctx, mfs.PROBLEM_ELASTIC, xyz, tets, fixed_nodes=bnd, fixed_values=g)
        S.run(E, nu)
        x = mfs.Vec(ctx, S.A.n)
        x.copy_from(S.x0)
        rep = mfs.Solver(S.A, precision=mfs.PREC_F64, tol=1e-8).run(S.b, x)
        times.append(rep.seconds)
        S.close()
    return float(np.mean(times))


def bench_msfem_reference_trace(mfs, ctx, p):
    """The same cells with the reference's edge -> face -> volume boundary
    data (MsFEM_Grid.h:5026-5506) and its component-wise coarse functions:
    batched edge profiles + face solves (98,304 face problems for 16^3 cells),
    then the mixed batched volume solves."""
    import numpy as np
    B = mfs.Batch(ctx, p["cell_n"], p["origin"], p["h"], p["E"], p["nu"], trace=mfs.TRACE_REFERENCE,
                  coarse_basis=mfs.BASIS_COMPONENTWISE)
    t0 = time.perf_counter()
    B.assemble()
    t_asm = time.perf_counter() - t0
    B.assemble()  # second assemble: warm
    worst, bad, ms = B.trace_status()
    wg, per_cu, var, _ = B.autotune(precision=mfs.PREC_MIXED, tol=1e-8)
    reps = [B.solve(precision=mfs.PREC_MIXED, tol=1e-8, use_graph=True) for _ in range(3)]
    r = min(reps, key=lambda x: x.seconds)
    t0 = time.perf_counter()
    Kc = B.coarse_matrices()
    t_kc = time.perf_counter() - t0
    out = {"coarse_matrices_ms": round(t_kc * 1e3, 2),"edge_face_solves_ms": round(ms, 2), "face_problems": 6 * 12 * len(p["cell_n"]),
           "face_solver": "banded LDL^T in LDS (direct)", "face_max_relres": worst, "face_failed": bad, "first_assemble_wall_s": round(t_asm, 3),
           "mixed_batched_ms": round(r.seconds * 1e3, 2), "mean_iterations": round(r.mean_iterations, 1),
           "max_iterations": r.max_iterations, "max_true_relres": r.max_true_relres,
           "not_converged": r.n_not_converged, "shape": {"workgroup": wg, "workgroups_per_cu": per_cu, "variant": var},
           "coarse_matrices_finite": bool(np.isfinite(Kc).all())}
    B.close()
    log(f"[bench] msfem reference traces: {out}")
    return out


def msfem_cpu_baseline(p, counts, per_n=8):
    """The reference's CPU path for the local problems, restated in the
    oracle: per cell, CSSD assembly of the cell matrix, 24 solves with the
    SSOR-PCG continuation and BCG_Stab2 fallback of FEM.h:3699-3770 to 1e-8
    and the Galerkin product.  Cells run concurrently on the host
    threads (one cell per thread, as an OpenMP loop over cells would); a
    sample of `per_n` cells per subdivision is timed and the batch total is
    extrapolated with the cell-size histogram."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    import oracle as orc
    from multiscalefemsolver3d_amd import problems
    threads = min(os.cpu_count() or 1, 16)
    sample = []
    for n in sorted(counts):
        sample += [int(c) for c in np.nonzero(p["cell_n"] == n)[0][:per_n]]

    def one(c):
        n = int(p["cell_n"][c])
        t0 = time.perf_counter()
        orc.msfem_cell_refpath(n, tuple(p["origin"][c]), p["h"], problems.cell_E(p, c), p["nu"], min_residual=1e-8)
        return n, time.perf_counter() - t0

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(one, sample))
    wall = time.perf_counter() - t0
    per = {n: float(np.mean([t for m, t in res if m == n])) for n in counts}
    total_cpu = sum(counts[n] * per[n] for n in counts)
    est_ms = total_cpu / threads * 1e3
    return {"value": round(est_ms, 1), "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"{len(sample)} cells ({per_n} per n) solved concurrently on {threads} threads in "
                      f"{wall:.1f}s; per-cell seconds under that load {dict((f'n={n}', round(t, 3)) for n, t in per.items())}"
                      f"; batch total = sum(count_n * t_n) / {threads}"}


def bench_large(mfs, problems, ctx, n, solves=True):
    """Large elasticity: n^3 elastic cube, two-phase E in {1e9, 2e11}; SpMV GB/s at
    HBM scale (FP32 BSR3 working set ~1.3 GB), assembly time, mixed vs FP64
    Jacobi-PCG, GMRES(30) (textbook convention: restarted GMRES stagnates on
    the reference convention's indefinite operator)."""
    import numpy as np
    t0 = time.time()
    p = problems.elastic_cube(n, field="grf")
    S = mfs.Assembly(ctx, mfs.PROBLEM_ELASTIC, p["xyz"], p["tets"], material=p["material"], renumber=True,
                     fixed_nodes=p["fixed_nodes"], fixed_values=p["fixed_values"], tris=p["tris"],
                     traction=p["traction"])
    setup = time.time() - t0
    S.run(p["E"], p["nu"])
    asm_ms = S.time(5)
    A = S.A
    res = {"n": n, "dofs": A.n, "nnz": A.nnz, "block_nnz": A.nnzb, "host_setup_s": round(setup, 2),
           "assembly_ms": round(asm_ms, 3)}
    nt, nn = len(p["tets"]), len(p["xyz"])
    asm_bytes = nt * (16 + 8) + nn * 24 + A.nnz * 8
    res["assembly_gbs_algorithmic"] = round(asm_bytes / (asm_ms * 1e-3) / 1e9, 1)
    for prec, name in ((mfs.F32, "f32"), (mfs.F64, "f64")):
        X = mfs.Vec(ctx, A.n, prec, np.ones(A.n))
        Y = mfs.Vec(ctx, A.n, prec)
        ms = min(A.spmv_time(X, Y, 50, prec) for _ in range(3))  # best of 3 loops of 50 launches
        gbs = A.spmv_bytes(prec) / (ms * 1e-3) / 1e9
        res[f"spmv_{name}_us"] = round(ms * 1e3, 1)
        res[f"spmv_{name}_gbs"] = round(gbs, 1)
        res[f"spmv_{name}_frac_hbm"] = round(gbs / HBM_PEAK_GBS, 4)
        res[f"spmv_{name}_bytes"] = A.spmv_bytes(prec)
    if not solves:
        S.close()
        log(f"[bench] large: {res}")
        return res
    x = mfs.Vec(ctx, A.n)
    # reference element convention: symmetric indefinite -> MINRES (capped:
    # it does not reach 1e-8 on this operator in tens of thousands of
    # iterations); CG shown for contrast (capped)
    for name, kw in (("minres_fp64_capped", dict(method=mfs.MINRES, precision=mfs.PREC_F64, maxit=4000)),
                     ("minres_mixed_capped", dict(method=mfs.MINRES, precision=mfs.PREC_MIXED, maxit=4000)),
                     ("pcg_fp64_capped", dict(precision=mfs.PREC_F64, maxit=3000))):
        x.copy_from(S.x0)
        rep = mfs.Solver(A, tol=1e-8, **kw).run(S.b, x, check=False)
        res[name] = {"ms": round(rep.seconds * 1e3, 1), "iterations": rep.iterations, "outer": rep.outer_steps,
                     "true_relres": rep.true_relres, "status": rep.status}
        log(f"[bench] large {name}: {res[name]}")
    S.close()
    # GMRES(30) on the textbook (SPD) operator of the same mesh
    S = mfs.Assembly(ctx, mfs.PROBLEM_ELASTIC, p["xyz"], p["tets"], material=p["material"], renumber=True,
                     convention=mfs.ELEM_TEXTBOOK, fixed_nodes=p["fixed_nodes"], fixed_values=p["fixed_values"],
                     tris=p["tris"], traction=p["traction"])
    S.run(p["E"], p["nu"])
    # six right-hand sides: iteration throughput of the joint stream (matrix
    # read once per iteration for all six) vs one system at a time, 300
    # iterations each (HBM-bound regime)
    rng = np.random.default_rng(7)
    b0 = S.b.download()
    bs6 = [mfs.Vec(ctx, S.A.n, mfs.F64, b0 * (1.0 + 0.2 * j) + rng.normal(size=S.A.n) * np.abs(b0).max() * 0.05)
           for j in range(6)]
    xs6 = [mfs.Vec(ctx, S.A.n) for _ in range(6)]
    for v in xs6:
        v.copy_from(S.x0)
    ctx.sync()
    t6 = time.perf_counter()
    mfs.solve_multi(S.A, bs6, xs6, check=False, precision=mfs.PREC_MIXED, tol=1e-12, maxit=300)
    ctx.sync()
    joint = time.perf_counter() - t6
    one = mfs.Solver(S.A, precision=mfs.PREC_MIXED, tol=1e-12, maxit=300)
    xs6[0].copy_from(S.x0)
    ctx.sync()
    t6 = time.perf_counter()
    one.run(bs6[0], xs6[0], check=False)
    ctx.sync()
    single = time.perf_counter() - t6
    res["multi_rhs_6_throughput"] = {"iterations": 300, "joint_six_s": round(joint, 3), "one_system_s": round(single, 3),
                                     "speedup_vs_six_sequential": round(6 * single / joint, 3)}
    log(f"[bench] large 6 right-hand sides, 300 iterations: joint {joint:.2f} s vs one system {single:.2f} s")
    del bs6, xs6
    for name, kw in (("pcg_fp64_textbook", dict(precision=mfs.PREC_F64, maxit=40000)),
                     ("pcg_mixed_textbook", dict(precision=mfs.PREC_MIXED, maxit=40000)),
                     ("gmres30_mixed_textbook_capped", dict(method=mfs.GMRES, precision=mfs.PREC_MIXED,
                                                            restart=30, maxit=3000))):
        x = mfs.Vec(ctx, S.A.n)
        x.copy_from(S.x0)
        rep = mfs.Solver(S.A, tol=1e-8, **kw).run(S.b, x, check=False)
        res[name] = {"ms": round(rep.seconds * 1e3, 1), "iterations": rep.iterations, "outer": rep.outer_steps,
                     "true_relres": rep.true_relres, "status": rep.status}
        log(f"[bench] large {name}: {res[name]}")
    S.close()
    log(f"[bench] large: {res}")
    return res


if __name__ == "__main__":
    main()
