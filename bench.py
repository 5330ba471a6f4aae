#!/usr/bin/env python3
"""Benchmark of the MI355X FEM solve path.

Workload (default): electrostatics on a 64^3 structured tet mesh (1,572,864
tets, 274,625 nodes), permittivity 1 with 200 seeded spherical inclusions of
radius U(0.02,0.06) and permittivity 100, potential 0 at z=0 and 1 at z=1.
A "step" is one complete solve to a true FP64 relative residual of 1e-8 with
the mixed-precision Jacobi-PCG (FP32 inner iterations inside FP64 iterative
refinement), starting from the Dirichlet-lifted zero guess, with the matrix
and right-hand side resident in HBM.  The plain FP64 GPU Jacobi-PCG on the
same system is timed beside it (the judge), and the reference's CPU solver
path (CSSD rows, OpenMP SpMV, serial SSOR sweeps and dot products, restated
in the oracle) is timed on the host cores.

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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="mesh hexes per edge")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=25.0, help="seconds of CPU baseline work")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import multiscalefemsolver3d_amd as mfs
    from multiscalefemsolver3d_amd import problems

    t0 = time.time()
    p = problems.laplace_inclusions(args.n)
    log(f"[bench] mesh+field {time.time() - t0:.1f}s: nodes {len(p['xyz'])} tets {len(p['tets'])}")
    ctx = mfs.Context(0)
    t0 = time.time()
    S = mfs.Assembly(ctx, mfs.PROBLEM_SCALAR, p["xyz"], p["tets"], fixed_nodes=p["fixed_nodes"],
                     fixed_values=p["fixed_values"])
    setup_s = time.time() - t0
    S.run(p["sigma"])
    asm_ms = S.time(10)
    A = S.A
    log(f"[bench] host setup {setup_s:.2f}s, GPU assembly {asm_ms:.3f} ms, N={A.n} nnz={A.nnz}")

    mixed = mfs.Solver(A, precision=mfs.PREC_MIXED, tol=1e-8, inner_tol=1e-4)
    fp64 = mfs.Solver(A, precision=mfs.PREC_F64, tol=1e-8)
    x = mfs.Vec(ctx, A.n)

    def run(solver):
        x.copy_from(S.x0)
        return solver.run(S.b, x)

    for _ in range(args.warmup):
        run(mixed)
        run(fp64)
    ctx.sync()
    t_start = time.perf_counter()
    reps = []
    for _ in range(args.steps):
        reps.append(run(mixed))
    ctx.sync()
    elapsed = time.perf_counter() - t_start
    ms_mixed = 1000.0 * elapsed / args.steps
    x_mixed = x.download()
    reps64 = []
    ctx.sync()
    t64 = time.perf_counter()
    for _ in range(max(3, args.steps // 2)):
        reps64.append(run(fp64))
    ctx.sync()
    ms_fp64 = 1000.0 * (time.perf_counter() - t64) / len(reps64)
    x_fp64 = x.download()
    r_m, r_d = reps[-1], reps64[-1]
    agree = float(np.linalg.norm(x_mixed - x_fp64) / np.linalg.norm(x_fp64))
    log(f"[bench] mixed {ms_mixed:.2f} ms ({r_m.iterations} inner it, {r_m.outer_steps} outer, "
        f"relres {r_m.true_relres:.2e}); fp64 {ms_fp64:.2f} ms ({r_d.iterations} it, relres {r_d.true_relres:.2e}); "
        f"|x_mixed-x_fp64|/|x_fp64| = {agree:.2e}")

    # Dominant kernel: the FP32 SELL SpMV of the inner iterations, timed with
    # HIP events on the library stream.
    X32, Y32 = mfs.Vec(ctx, A.n, mfs.F32, x_mixed), mfs.Vec(ctx, A.n, mfs.F32)
    spmv_ms = A.spmv_time(X32, Y32, 200, mfs.F32)
    spmv_bytes = A.spmv_bytes(mfs.F32)
    X64, Y64 = mfs.Vec(ctx, A.n, mfs.F64, x_mixed), mfs.Vec(ctx, A.n, mfs.F64)
    spmv64_ms = A.spmv_time(X64, Y64, 200, mfs.F64)
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    spmv64_gbs = A.spmv_bytes(mfs.F64) / (spmv64_ms * 1e-3) / 1e9
    log(f"[bench] SpMV f32 {spmv_ms * 1e3:.1f} us = {spmv_gbs:.0f} GB/s; f64 {spmv64_ms * 1e3:.1f} us = "
        f"{spmv64_gbs:.0f} GB/s (algorithmic bytes; working set is cache-resident)")

    cpu = None
    if not args.no_cpu:
        import oracle as orc
        rp, col, val = A.export_csr()
        b = S.b.download()
        x0 = S.x0.download()
        threads = min(os.cpu_count() or 1, 16)
        # probe the per-iteration cost, then run as many iterations as fit the budget
        _, it_p, _, sec_p = orc.refpath_solve(rp, col, val, b, x0, tol=1e-8, maxit=6, mode=0, threads=threads)
        per_it = sec_p / max(it_p, 1)
        budget_it = int(max(10, args.cpu_budget / max(per_it, 1e-6)))
        xc, it_c, rr_c, sec_c = orc.refpath_solve(rp, col, val, b, x0, tol=1e-8, maxit=budget_it, mode=0,
                                                  threads=threads)
        done = rr_c <= 1e-8
        if done:
            value = sec_c * 1000.0
            sample = f"full SSOR-PCG solve to 1e-8 ({it_c} iterations, true relres {rr_c:.2e})"
        else:
            # project with the FP64 GPU PCG iteration count (labelled projection)
            value = sec_c / it_c * r_d.iterations * 1000.0
            sample = (f"{it_c} SSOR-PCG iterations timed ({sec_c:.1f}s), projected to {r_d.iterations} "
                      f"iterations (the FP64 GPU PCG count)")
        cpu = {"value": round(value, 2), "unit": "ms", "cores": threads, "kind": "port",
               "sample": "reference CPU solver path (MSG_PreconditioningSSOR, CSSD rows, OpenMP dynamic SpMV, "
                         "serial dots/sweeps) on the same system: " + sample}
        log(f"[bench] cpu baseline {value:.0f} ms ({sample})")

    out = {
        "metric": METRIC,
        "value": round(ms_mixed, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mixed, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic (seeded structured mesh and inclusion field)",
        "config": {"workload": f"laplace{args.n}^3_inclusions200", "nodes": len(p["xyz"]), "tets": len(p["tets"]),
                   "dofs": A.n, "nnz": A.nnz, "tol": 1e-8, "solver": "Jacobi-PCG, FP32 inner + FP64 refinement",
                   "parallelism": "single GPU"},
        "fp64_gpu_cg_ms": round(ms_fp64, 3),
        "speedup_vs_fp64_gpu_cg": round(ms_fp64 / ms_mixed, 3),
        "iterations": {"mixed_inner": r_m.iterations, "mixed_outer": r_m.outer_steps, "fp64": r_d.iterations},
        "true_relres": {"mixed": r_m.true_relres, "fp64": r_d.true_relres},
        "mixed_vs_fp64_solution_rel_diff": agree,
        "assembly_ms": round(asm_ms, 4),
        "spmv_f64_gbs": round(spmv64_gbs, 1),
        "roofline": {"bound": "hbm", "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_sell_spmv<float,1> (FP32 SELL-64 SpMV)",
                     "bytes_per_launch": spmv_bytes, "avg_us": round(spmv_ms * 1e3, 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
