#!/usr/bin/env python3
"""Benchmark: solve time to ||r||/||b|| <= 1e-10 of the 3D linear-elastic block
(C3D8 B-bar, 64^3 elements, 811,200 DOF, clamped at x=0, pressure 1e6 Pa on
x=1), multifrontal LDL^T through the C ABI on one MI355X.

One step = numeric factorization from values already in HBM + triangular
solves + FP64 iterative refinement until the relative residual is <= 1e-10,
with the right-hand side already in HBM.  Two paths are timed on the same
input: mixed (FP32 Schur updates on MFMA with FP32 accumulation, FP64 pivot
blocks, FP64 refinement) and plain FP64.  ``value`` is the mixed path's time
per step; the FP64 time, the speedup, the error of the mixed solution against
the FP64 one, the roofline of the dominant kernel (the Schur-update GEMM,
timed per launch with HIP events on the solver's stream) and the CPU baseline
(the C oracle on the host cores, a bounded sample extrapolated by flop count)
ride along on the same JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

METRIC = "solve time to ‖r‖/‖b‖≤1e-10, mixed vs FP64 GPU; front TFLOP/s & %peak"
PEAK = {"fp32_mfma_tflops": 157.3, "fp64_mfma_tflops": 78.6, "hbm_gbs": 8000.0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_path(ffi, hip, p, precision, d_b, steps, warmup, tol):
    import numpy as np
    t0 = time.perf_counter()
    h, err, _ = ffi.Handle.factor(p, precision=precision, max_refine=30, refine_tol=tol)
    if h is None:
        raise RuntimeError(f"{precision} factorization failed: err={err} {ffi.last_error()}")
    log(f"{precision}: build+first factor {time.perf_counter() - t0:.1f}s  {h.stats()}")
    x = hip.DeviceArray(d_b.count)

    def step():
        if h.refactor(None) != -1:
            raise RuntimeError("refactor failed")
        x.copy_from(d_b)
        h.solve_device(x.ptr, p.rhs.shape[1])

    for i in range(warmup):
        step()
        log(f"{precision}: warmup {i} done")
    hip.synchronize()
    times, stats = [], []
    for i in range(steps):
        hip.synchronize()
        t0 = time.perf_counter()
        step()
        hip.synchronize()
        times.append(time.perf_counter() - t0)
        stats.append(h.stats())
        log(f"{precision}: step {i} {times[-1] * 1e3:.1f} ms  factor {stats[-1]['factor_ms']:.1f} ms  "
            f"solve {stats[-1]['solve_ms']:.1f} ms  refine steps {stats[-1]['refine_steps']:.0f}  "
            f"rel {stats[-1]['rel_residual']:.2e}")
    sol = x.download().reshape(p.rhs.shape, order="F")
    a = p.to_scipy()
    rel = np.linalg.norm(p.rhs - a @ sol, axis=0) / np.linalg.norm(p.rhs, axis=0)
    x.free()
    h.free()
    return {"time_s": float(np.median(times)), "times_s": times, "stats": stats, "solution": sol,
            "true_rel_residual": float(rel.max())}


def cpu_baseline(size, work_full, nrhs):
    """The C oracle (FP64 restatement of the reference algorithms, all host
    threads) on a size^3 block; extrapolated to the full workload by flops."""
    import numpy as np
    import oracle
    from rustyfem_amd import meshes
    oracle.build()
    q = meshes.hex_block(size, nrhs=nrhs)
    t0 = time.perf_counter()
    h, err, _ = oracle.Handle.factor(q)
    t1 = time.perf_counter()
    x = h.solve(q.rhs)
    t2 = time.perf_counter()
    work = float(h.info()[0])
    rate = work / (t1 - t0)
    est = work_full / rate + (t2 - t1) * (work_full / work) ** (2.0 / 3.0)
    rel = float(np.max(np.linalg.norm(q.rhs - q.to_scipy() @ x, axis=0) / np.linalg.norm(q.rhs, axis=0)))
    return {"value": est, "unit": "s", "cores": int(oracle.threads()), "kind": "port",
            "sample": (f"C oracle FP64 multifrontal LDL^T of the {size}^3 block (n={q.n}, {work:.3e} flops): "
                       f"factor {t1 - t0:.2f} s ({rate / 1e9:.1f} GFLOP/s), solve {t2 - t1:.3f} s, rel. residual "
                       f"{rel:.1e}; value = full-workload flops / sample rate (+ solve scaled by nnz(L))")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=64, help="elements per cube edge (64 = the 811,200-DOF block)")
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=32, help="cube edge of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB copied device-to-device twice before the run (a kernel of known traffic for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the multifrontal factorization of one problem is not sharded "
              "across GPUs in this build (see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("RUSTYFEM_KERNEL_TIMING", "1")
    import numpy as np
    from rustyfem_amd import ffi, hip, meshes
    ffi.lib()
    if args.copy_calibration > 0:
        cnt = args.copy_calibration * (1 << 20) // 8
        src, dst = hip.DeviceArray(cnt), hip.DeviceArray(cnt)
        for _ in range(2):
            dst.copy_from(src)
        hip.synchronize()
        src.free()
        dst.free()
        log(f"copy calibration: 2 x {args.copy_calibration} MiB device-to-device")
    t0 = time.perf_counter()
    p = meshes.hex_block(args.size, nrhs=args.nrhs)
    log(f"problem {p.name}: n={p.n} nnz={p.nnz} nrhs={args.nrhs} built in {time.perf_counter() - t0:.1f}s")
    d_b = hip.DeviceArray.from_host(np.asfortranarray(p.rhs).reshape(-1, order="F"))
    tol = 1e-10
    mixed = run_path(ffi, hip, p, "mixed", d_b, args.steps, args.warmup, tol)
    fp64 = None if args.skip_fp64 else run_path(ffi, hip, p, "fp64", d_b, args.steps, args.warmup, tol)
    st = mixed["stats"][-1]
    gemm_ms, gemm_n, gemm_flops = st["gemm_ms"], st["gemm_launches"], st["gemm_flops"]
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else None
    roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK["fp32_mfma_tflops"], "unit": "TFLOP/s",
            "frac": (achieved / PEAK["fp32_mfma_tflops"]) if achieved else None, "traffic": None,
            "kernel": "k_gemm<float> (Schur update, v_mfma_f32_16x16x4f32)",
            "launches": gemm_n, "avg_launch_ms": gemm_ms / gemm_n if gemm_n else None,
            "algorithmic_flops_per_launch": gemm_flops / gemm_n if gemm_n else None}
    factor_tflops = st["flops"] / (st["factor_ms"] * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": mixed["time_s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["time_s"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"3D C3D8 B-bar block {args.size}^3 elements (n={p.n} DOF), E=210e9 nu=0.3, "
                               f"x=0 clamped, pressure 1e6 Pa on x=1; multifrontal LDL^T, ND leaf 16; "
                               f"{args.nrhs} RHS; step = refactor + solve + FP64 refinement to 1e-10",
                   "n": p.n, "nrhs": args.nrhs, "factor_flops": st["flops"], "fronts": st["fronts"],
                   "levels": st["levels"]},
        "mixed": {"time_s": mixed["time_s"], "factor_ms": st["factor_ms"], "solve_ms": st["solve_ms"],
                  "refine_steps": st["refine_steps"], "rel_residual": st["rel_residual"],
                  "true_rel_residual": mixed["true_rel_residual"], "factor_tflops": factor_tflops,
                  "factor_frac_fp32_peak": factor_tflops / PEAK["fp32_mfma_tflops"]},
        "roofline": roof,
    }
    if fp64 is not None:
        s6 = fp64["stats"][-1]
        f6 = s6["flops"] / (s6["factor_ms"] * 1e-3) / 1e12
        err = float(np.max(np.abs(mixed["solution"] - fp64["solution"])) / np.max(np.abs(fp64["solution"])))
        out["fp64"] = {"time_s": fp64["time_s"], "factor_ms": s6["factor_ms"], "solve_ms": s6["solve_ms"],
                       "refine_steps": s6["refine_steps"], "rel_residual": s6["rel_residual"],
                       "true_rel_residual": fp64["true_rel_residual"], "factor_tflops": f6,
                       "factor_frac_fp64_peak": f6 / PEAK["fp64_mfma_tflops"],
                       "gemm_tflops": (s6["gemm_flops"] / (s6["gemm_ms"] * 1e-3) / 1e12) if s6["gemm_ms"] else None}
        out["speedup_mixed_vs_fp64"] = fp64["time_s"] / mixed["time_s"]
        out["mixed_vs_fp64_displacement_rel_error"] = err
    if args.cpu_sample > 0:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_sample, st["flops"], args.nrhs)
        except Exception as ex:  # the baseline must not hide the GPU result
            out["cpu_baseline"] = {"value": None, "error": str(ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
