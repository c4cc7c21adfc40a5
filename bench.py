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
ride along on the same JSON line, with both arms' error against the oracle's
FP64 solution on that sample.

The same line carries the other configurations as sub-objects, each with its
own roofline, CPU baseline and error against the FP64 oracle: ``c2_16rhs``
(the same block with 16 right-hand sides, both arms), ``c3_batch`` (512
plates factored and solved together, 8 load cases each), ``c5_learned``
(the C3D4 block with the learned-smoother FCG against plain AMG-PCG) and
``f1_assembly`` (the block's stiffness assembled on the GPU, bit for bit
against the oracle).  ``--no-extras`` leaves the last three out.
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


def run_path(ffi, hip, p, precision, d_b, steps, warmup, tol, d_b16=None, rhs16=None):
    """Steps of one arm (refactor from HBM + solve + refinement to tol) with
    one right-hand side, then, on the same handle, with the 16 of rhs16."""
    import numpy as np
    t0 = time.perf_counter()
    h, err, _ = ffi.Handle.factor(p, precision=precision, max_refine=30, refine_tol=tol)
    if h is None:
        raise RuntimeError(f"{precision} factorization failed: err={err} {ffi.last_error()}")
    log(f"{precision}: build+first factor {time.perf_counter() - t0:.1f}s  {h.stats()}")

    def leg(db, nrhs, label):
        x = hip.DeviceArray(db.count)

        def step():
            if h.refactor(None) != -1:
                raise RuntimeError("refactor failed")
            x.copy_from(db)
            h.solve_device(x.ptr, nrhs)

        for i in range(warmup):
            step()
        hip.synchronize()
        times, stats = [], []
        for i in range(steps):
            hip.synchronize()
            t0 = time.perf_counter()
            step()
            hip.synchronize()
            times.append(time.perf_counter() - t0)
            stats.append(h.stats())
            log(f"{precision}{label}: step {i} {times[-1] * 1e3:.1f} ms  factor {stats[-1]['factor_ms']:.1f} ms  "
                f"solve {stats[-1]['solve_ms']:.1f} ms  refine steps {stats[-1]['refine_steps']:.0f}  "
                f"rel {stats[-1]['rel_residual']:.2e}")
        sol = x.download().reshape((p.n, nrhs), order="F")
        x.free()
        return times, stats, sol

    times, stats, sol = leg(d_b, 1, "")
    a = p.to_scipy()
    b1 = p.rhs[:, :1]
    rel = np.linalg.norm(b1 - a @ sol, axis=0) / np.linalg.norm(b1, axis=0)
    out = {"time_s": float(np.median(times)), "times_s": times, "stats": stats, "solution": sol,
           "true_rel_residual": float(rel.max()), "factor_entries": float(h.info()[1])}
    if d_b16 is not None:
        t16, s16, x16 = leg(d_b16, rhs16.shape[1], " 16 RHS")
        rel16 = np.linalg.norm(rhs16 - a @ x16, axis=0) / np.linalg.norm(rhs16, axis=0)
        out["rhs16"] = {"time_s": float(np.median(t16)), "stats": s16[-1], "solution": x16,
                        "true_rel_residual": float(rel16.max())}
    h.free()
    return out


def cpu_baseline(size, work_full, nrhs):
    """The C oracle (FP64 restatement of the reference algorithms, all host
    threads) on a size^3 block, extrapolated to the full workload by flops;
    and, on the same sample, both GPU arms against the oracle's FP64 solution
    (max-norm relative displacement error over the right-hand sides)."""
    import numpy as np
    import oracle
    from rustyfem_amd import ffi, meshes
    oracle.build()
    q = meshes.hex_block(size, nrhs=nrhs)
    t0 = time.perf_counter()
    h, err, _ = oracle.Handle.factor(q)
    t1 = time.perf_counter()
    x = h.solve(q.rhs)
    t2 = time.perf_counter()
    work = float(h.info()[0])
    rate = work / (t1 - t0)
    est = work_full / rate + (t2 - t1) / nrhs * (work_full / work) ** (2.0 / 3.0)
    rel = float(np.max(np.linalg.norm(q.rhs - q.to_scipy() @ x, axis=0) / np.linalg.norm(q.rhs, axis=0)))
    base = {"value": est, "unit": "s", "cores": int(oracle.threads()), "kind": "port",
            "sample": (f"C oracle FP64 multifrontal LDL^T of the {size}^3 block (n={q.n}, {work:.3e} flops): "
                       f"factor {t1 - t0:.2f} s ({rate / 1e9:.1f} GFLOP/s), {nrhs}-RHS solve {t2 - t1:.3f} s, "
                       f"rel. residual {rel:.1e}; value = full-workload flops / sample rate (+ one RHS's solve "
                       f"scaled by nnz(L)); the oracle at the full size would take minutes")}
    errs = {"sample": f"the {size}^3 block (n={q.n}), {nrhs} RHS; reference: the oracle's FP64 solution"}
    for prec in ("mixed", "fp64"):
        g, e, _ = ffi.Handle.factor(q, precision=prec, max_refine=30, refine_tol=1e-10)
        xg = g.solve(q.rhs)
        errs[prec] = float(np.max(np.abs(xg - x)) / np.max(np.abs(x)))
        g.free()
    return base, errs


def measured_traffic():
    """HBM bytes per launch of the Schur-update GEMM from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), if present."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_k_gemm_float_hex64.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def lower_of(p):
    import numpy as np
    import scipy.sparse as sp
    a = p.to_scipy().tocoo()
    m = a.row >= a.col
    low = sp.csc_matrix((a.data[m], (a.row[m], a.col[m])), shape=a.shape)
    low.sort_indices()
    return low.indptr.astype(np.int64), low.indices.astype(np.int64), low.data.copy()


def amg_cpu_baseline(p, n_full, tol, label):
    """The oracle's AMG-PCG (the reference's CPU algorithm: smoothed aggregation,
    Gauss-Seidel V-cycle, src/solver/iterative*.rs) on a smaller problem of the
    same family, setup + solve timed on the host threads and scaled by DOF
    count to the full problem (AMG iterations grow slowly with n, so the
    scaled figure is a lower bound)."""
    import oracle
    from rustyfem_amd import meshes
    _, coords, ndof, _, _ = meshes.node_graph(p)
    cp, ri, v = lower_of(p)
    b = p.rhs[:, 0].copy()
    t0 = time.perf_counter()
    x, rep = oracle.solve_lower(p.n, cp, ri, v, b, "amg", coords, ndof, tol=tol)
    dt = time.perf_counter() - t0
    log(f"cpu baseline: oracle AMG-PCG on {label} (n={p.n}) {dt:.2f}s, {rep['iterations']} iterations")
    return {"value": dt * n_full / p.n, "unit": "s", "cores": int(oracle.threads()), "kind": "port",
            "sample": f"{label} (n={p.n}): oracle AMG-PCG setup + solve to {tol:g} in {dt:.2f} s, "
                      f"{rep['iterations']} iterations, scaled by DOF count to n={n_full}",
            "sample_s": dt, "sample_iterations": rep["iterations"], "sample_converged": rep["converged"]}


def run_amg(args):
    """The heterogeneous-stiffness block (C3D8 B-bar, lognormal E with log-std 1.5,
    correlation length 0.1, contrast clipped to 1e3, clamped base, body load):
    AMG preconditioned Krylov to ||r|| <= 1e-8 ||b||, FP32 V-cycle inside FP64 FCG
    vs all-FP64 PCG.  One step = one solve from x0 = 0 with the preconditioner
    already set up and b in HBM."""
    import numpy as np
    from rustyfem_amd import ffi, hip, meshes
    ffi.lib()
    t0 = time.perf_counter()
    p = meshes.hetero_block(args.size, seed=4)
    _, coords, ndof, _, _ = meshes.node_graph(p)
    cp, ri, v = lower_of(p)
    log(f"problem hetero {args.size}^3: n={p.n} built in {time.perf_counter() - t0:.1f}s")
    b = np.ascontiguousarray(p.rhs[:, 0])
    d_b = hip.DeviceArray.from_host(b)
    tol = 1e-8
    res = {}
    for name, kw in (("fp32_fcg", {"krylov": "fcg", "vcycle": "fp32"}), ("fp64_pcg", {"krylov": "pcg", "vcycle": "fp64"})):
        t0 = time.perf_counter()
        s = ffi.Pcg(p.n, cp, ri, v, coords=coords, ndof=ndof, **kw)
        st = s.stats()
        log(f"{name}: setup {time.perf_counter() - t0:.1f}s levels {st['sizes']} complexity {st['complexity']:.3f}")
        d_x = hip.DeviceArray(p.n)
        for i in range(args.warmup):
            s.solve_device(d_b.ptr, d_x.ptr, tol=tol)
        times = []
        for i in range(args.steps):
            hip.synchronize()
            t1 = time.perf_counter()
            it, rel, conv = s.solve_device(d_b.ptr, d_x.ptr, tol=tol)
            hip.synchronize()
            times.append(time.perf_counter() - t1)
            log(f"{name}: step {i} {times[-1] * 1e3:.1f} ms iterations {it} rel {rel:.2e} converged {conv}")
        x = d_x.download()
        true_rel = float(np.linalg.norm(b - p.to_scipy() @ x) / np.linalg.norm(b))
        st = s.stats()
        res[name] = {"time_s": float(np.median(times)), "iterations": it, "converged": bool(conv),
                     "rel_residual": rel, "true_rel_residual": true_rel, "setup_s": st["setup_ms"] * 1e-3,
                     "levels": st["sizes"], "complexity": st["complexity"], "spmv_ms": st["spmv_ms"],
                     "spmv_bytes": st["spmv_bytes"], "solution": x}
        d_x.free()
        del s
    a, f = res["fp32_fcg"], res["fp64_pcg"]
    achieved = a["spmv_bytes"] / (a["spmv_ms"] * 1e-3) / 1e9
    err = float(np.max(np.abs(a["solution"] - f["solution"])) / np.max(np.abs(f["solution"])))
    for r in res.values():
        r.pop("solution")
    out = {"metric": "AMG-Krylov solve time to ||r||/||b|| <= 1e-8, FP32 V-cycle in FP64 FCG vs all-FP64 PCG",
           "value": a["time_s"], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": a["time_s"] * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32+fp64", "data": "synthetic",
           "config": {"workload": f"heterogeneous C3D8 B-bar block {args.size}^3 elements (n={p.n}), lognormal E "
                                  f"(log-std 1.5, corr. length 0.1, contrast 1e3), clamped base, body load 1e4 N/m^3",
                      "n": p.n},
           "fp32_fcg": a, "fp64_pcg": f, "speedup_fp32_vs_fp64": f["time_s"] / a["time_s"],
           "fp32_vs_fp64_displacement_rel_error": err,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                        "frac": achieved / PEAK["hbm_gbs"], "traffic": None,
                        "kernel": "k_spmv<double,3,G> (BSR 3x3 fine-matrix product, FP64)"}}
    if args.cpu_sample_iter > 0:
        q = meshes.hetero_block(args.cpu_sample_iter, seed=4)
        out["cpu_baseline"] = amg_cpu_baseline(q, p.n, tol, f"the heterogeneous block at {args.cpu_sample_iter}^3")
    print(json.dumps(out), flush=True)


def learned_result(args):
    """C3D4 block (Kuhn split of tet_size^3 cubes, perturbed interior vertices,
    lognormal stiffness per cube with a log-std drawn from U[0.5, 2], seed 42,
    clamped base, body load), to ||r|| <= 1e-8 ||b||, three solvers on the same
    AMG hierarchy: plain AMG-PCG (FP64 V-cycle, Chebyshev smoothing; the
    baseline and the fallback), AMG-FCG with the FP32 V-cycle and Chebyshev
    smoothing (the method without the network), and the same FCG whose finest
    smoother is the learned one (parameters from
    rustyfem_amd/data/smoother_c3d4.npz, one sweep).  One step = one solve
    from x0 = 0 with the preconditioner set up and b in HBM."""
    import numpy as np
    from rustyfem_amd import ffi, hip, meshes
    path = args.smoother or os.path.join(os.path.dirname(os.path.abspath(__file__)), "rustyfem_amd", "data",
                                         "smoother_c3d4.npz")
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].copy()
        meta = json.loads(str(z["meta"]))
    ffi.lib()
    t0 = time.perf_counter()
    p = meshes.tet_block(args.tet_size, seed=42)
    _, coords, ndof, _, _ = meshes.node_graph(p)
    cp, ri, v = lower_of(p)
    log(f"problem tet {args.tet_size}^3 cubes: n={p.n} log-std {p.meta['log_std']:.3f} built in "
        f"{time.perf_counter() - t0:.1f}s")
    b = np.ascontiguousarray(p.rhs[:, 0])
    d_b = hip.DeviceArray.from_host(b)
    tol = 1e-8
    a_mat = p.to_scipy()
    res = {}

    def timed(name, s):
        st = s.stats()
        d_x = hip.DeviceArray(p.n)
        for _ in range(args.warmup):
            s.solve_device(d_b.ptr, d_x.ptr, tol=tol)
        times = []
        for i in range(args.steps):
            hip.synchronize()
            t1 = time.perf_counter()
            it, rel, conv = s.solve_device(d_b.ptr, d_x.ptr, tol=tol)
            hip.synchronize()
            times.append(time.perf_counter() - t1)
            log(f"{name}: step {i} {times[-1] * 1e3:.1f} ms iterations {it} rel {rel:.2e} converged {conv}")
        x = d_x.download()
        d_x.free()
        st = s.stats()
        res[name] = {"time_s": float(np.median(times)), "iterations": it, "converged": bool(conv),
                     "rel_residual": rel, "true_rel_residual": float(np.linalg.norm(b - a_mat @ x) / np.linalg.norm(b)),
                     "setup_s": st["setup_ms"] * 1e-3, "levels": st["sizes"], "spmv_ms": st["spmv_ms"],
                     "spmv_bytes": st["spmv_bytes"], "fallback": st["fallback"], "solution": x}

    t0 = time.perf_counter()
    s = ffi.Pcg(p.n, cp, ri, v, coords=coords, ndof=ndof, krylov="pcg", vcycle="fp64")
    log(f"AMG setup (FP64 V-cycle) {time.perf_counter() - t0:.1f}s levels {s.stats()['sizes']}")
    timed("amg_pcg_fp64", s)
    del s
    t0 = time.perf_counter()
    s = ffi.Pcg(p.n, cp, ri, v, coords=coords, ndof=ndof, krylov="fcg", vcycle="fp32")
    log(f"AMG setup (FP32 V-cycle) {time.perf_counter() - t0:.1f}s")
    timed("fcg_fp32_chebyshev", s)
    t0 = time.perf_counter()
    s.set_learned_smoother(params, sweeps=1)
    log(f"learned smoother M computed in {time.perf_counter() - t0:.2f}s")
    timed("learned_fcg_fp32", s)
    del s
    a, f, c = res["learned_fcg_fp32"], res["amg_pcg_fp64"], res["fcg_fp32_chebyshev"]
    achieved = a["spmv_bytes"] / (a["spmv_ms"] * 1e-3) / 1e9
    err = float(np.max(np.abs(a["solution"] - f["solution"])) / np.max(np.abs(f["solution"])))
    for r in res.values():
        r.pop("solution")
    out = {"metric": "AMG-Krylov solve time to ||r||/||b|| <= 1e-8, learned-smoother FCG vs plain AMG-PCG",
           "value": a["time_s"], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": a["time_s"] * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp64+fp32", "data": "synthetic",
           "config": {"workload": f"C3D4 block, {args.tet_size}^3 cubes x 6 tetrahedra (n={p.n}), perturbed interior "
                                  f"vertices, lognormal E (log-std {p.meta['log_std']:.3f}), clamped base, body load",
                      "n": p.n, "smoother_training": meta},
           "learned_fcg_fp32": a, "fcg_fp32_chebyshev": c, "amg_pcg_fp64": f,
           "speedup_vs_plain_amg_pcg": f["time_s"] / a["time_s"],
           "gain_of_network_over_chebyshev": c["time_s"] / a["time_s"],
           "learned_vs_plain_displacement_rel_error": err,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                        "frac": achieved / PEAK["hbm_gbs"], "traffic": None,
                        "kernel": "k_spmv<double,3,G> (BSR 3x3 fine-matrix product, FP64)"}}
    if args.cpu_sample_iter > 0:
        q = meshes.tet_block(args.cpu_sample_iter, seed=42)
        out["cpu_baseline"] = amg_cpu_baseline(q, p.n, tol, f"the C3D4 block at {args.cpu_sample_iter}^3 cubes")
        out["error_vs_fp64_oracle"] = learned_oracle_error(meshes.tet_block(args.oracle_sample_iter, seed=42),
                                                            params, tol)
    return out


def run_learned(args):
    print(json.dumps(learned_result(args)), flush=True)


def learned_oracle_error(q, params, tol):
    """The learned-smoother FCG (FP32 V-cycle) and the plain FP64 AMG-PCG on the
    GPU against the oracle's FP64 multifrontal solution of the same sample
    problem (max-norm relative displacement error)."""
    import numpy as np
    import oracle
    from rustyfem_amd import ffi, meshes
    _, coords, ndof, _, _ = meshes.node_graph(q)
    cp, ri, v = lower_of(q)
    b = np.ascontiguousarray(q.rhs[:, 0])
    t0 = time.perf_counter()
    h, err, _ = oracle.Handle.factor(q)
    xo = h.solve(b)
    log(f"oracle direct solve of the sample (n={q.n}) {time.perf_counter() - t0:.1f}s")
    res = {"sample": f"C3D4 block (n={q.n}); reference: the oracle's FP64 multifrontal LDL^T solution"}
    for name, kw, learned in (("learned_fcg_fp32", {"krylov": "fcg", "vcycle": "fp32"}, True),
                              ("amg_pcg_fp64", {"krylov": "pcg", "vcycle": "fp64"}, False)):
        s = ffi.Pcg(q.n, cp, ri, v, coords=coords, ndof=ndof, **kw)
        if learned:
            s.set_learned_smoother(params, sweeps=1)
        x, it, rel, conv = s.solve(b, tol=tol)
        res[name] = float(np.max(np.abs(x - xo)) / np.max(np.abs(xo)))
        del s
    return res


def run_ic0(args):
    """IC(0)-preconditioned CG (src/solver/iterative/csr.rs:219-380) on the C3D8
    block (the direct workload's problem at --size elements per edge) to
    ||r|| <= 1e-10 ||b||: factorization and level-scheduled solves on the GPU,
    against the oracle's serial IC(0)-PCG on the host cores.  One step = one
    solve from x0 = 0 with the factor in HBM."""
    import numpy as np
    from rustyfem_amd import ffi, hip, meshes
    ffi.lib()
    t0 = time.perf_counter()
    p = meshes.hex_block(args.size, nrhs=1)
    cp, ri, v = lower_of(p)
    log(f"problem hex {args.size}^3: n={p.n} built in {time.perf_counter() - t0:.1f}s")
    b = np.ascontiguousarray(p.rhs[:, 0])
    d_b = hip.DeviceArray.from_host(b)
    d_x = hip.DeviceArray(p.n)
    tol = 1e-10
    t0 = time.perf_counter()
    s = ffi.Pcg(p.n, cp, ri, v, kind="ic0")
    setup_s = time.perf_counter() - t0
    st = s.stats()
    log(f"IC(0) setup {setup_s:.2f}s shift {st['ic0_shift']} levels fwd {st['ic0_forward_levels']} "
        f"bwd {st['ic0_backward_levels']}")
    for _ in range(args.warmup):
        s.solve_device(d_b.ptr, d_x.ptr, tol=tol)
    times = []
    for i in range(args.steps):
        hip.synchronize()
        t1 = time.perf_counter()
        it, rel, conv = s.solve_device(d_b.ptr, d_x.ptr, tol=tol)
        hip.synchronize()
        times.append(time.perf_counter() - t1)
        log(f"ic0 pcg: step {i} {times[-1] * 1e3:.1f} ms iterations {it} rel {rel:.2e} converged {conv}")
    x = d_x.download()
    true_rel = float(np.linalg.norm(b - p.to_scipy() @ x) / np.linalg.norm(b))
    # one preconditioner application (y = L^-1 r, z = L^-T D^-1 y), device to device
    d_z = hip.DeviceArray(p.n)
    s.precondition_device(d_b.ptr, d_z.ptr)
    hip.synchronize()
    t1 = time.perf_counter()
    reps = 20
    for _ in range(reps):
        s.precondition_device(d_b.ptr, d_z.ptr)
    hip.synchronize()
    apply_ms = (time.perf_counter() - t1) / reps * 1e3
    nnz_l = (len(ri) - p.n)
    apply_bytes = 2 * nnz_l * 12 + 5 * p.n * 8
    out = {"metric": "IC(0)-PCG solve time to ||r||/||b|| <= 1e-10", "value": float(np.median(times)), "unit": "s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": float(np.median(times)) * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
           "config": {"workload": f"C3D8 block {args.size}^3 (n={p.n}), IC(0) on the natural order", "n": p.n},
           "iterations": it, "rel_residual": rel, "true_rel_residual": true_rel, "setup_s": setup_s,
           "ic0_shift": st["ic0_shift"], "forward_levels": st["ic0_forward_levels"],
           "backward_levels": st["ic0_backward_levels"], "apply_ms": apply_ms,
           "roofline": {"bound": "latency", "achieved": apply_bytes / (apply_ms * 1e-3) / 1e9, "peak": PEAK["hbm_gbs"],
                        "unit": "GB/s", "frac": apply_bytes / (apply_ms * 1e-3) / 1e9 / PEAK["hbm_gbs"],
                        "traffic": None,
                        "kernel": "k_ic0_fwd / k_ic0_bwd (one launch per level; bytes = 2 nnz(L) 12 + 5 n 8 per application)"}}
    d_z.free()
    d_x.free()
    d_b.free()
    if args.cpu_sample > 0:
        import oracle
        t1 = time.perf_counter()
        xo, rep = oracle.solve_lower(p.n, cp, ri, v, b, "ic0", tol=tol)
        cpu_s = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": cpu_s, "unit": "s", "cores": oracle.threads(), "kind": "port",
                               "sample": f"the same problem (n={p.n}), setup + solve, iterations {rep['iterations']}",
                               "iterations": rep["iterations"]}
        out["gpu_vs_oracle_solution_rel_error"] = float(np.max(np.abs(x - xo)) / np.max(np.abs(xo)))
    print(json.dumps(out), flush=True)


def batch_result(args):
    """512 independent CPE4 plates (nx ~ U{20..200}, ny = nx/10, E ~ U[70e9, 210e9],
    nu ~ U[0.25, 0.35], 8 edge-traction load cases each, seed 3), factored and
    solved together.  One step = refactor of every plate from HBM + the 8-RHS
    solve of every plate + FP64 refinement check to 1e-10."""
    import numpy as np
    from rustyfem_amd import ffi, hip, meshes
    ffi.lib()
    t0 = time.perf_counter()
    plates = meshes.plate_batch(count=args.batch_count, seed=3, nrhs=8)
    ntot = sum(p.n for p in plates)
    log(f"{len(plates)} plates, {ntot} DOF, built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    b = ffi.Batch(plates, max_refine=30, refine_tol=1e-10)
    if not b.ok:
        raise RuntimeError(f"batch factorization failed: {ffi.last_error()} {b.err}")
    info = b.info()
    log(f"analysis + first factorization {time.perf_counter() - t0:.1f}s  {info}")
    allb = np.concatenate([p.rhs for p in plates], axis=0)
    d_b = hip.DeviceArray.from_host(np.asfortranarray(allb).reshape(-1, order="F"))
    d_x = hip.DeviceArray(d_b.count)

    def step():
        rc, _ = b.refactor(None)
        if rc != 0:
            raise RuntimeError("batch refactor failed")
        d_x.copy_from(d_b)
        b.solve_device(d_x.ptr, 8)

    for _ in range(args.warmup):
        step()
    times = []
    for i in range(args.steps):
        hip.synchronize()
        t1 = time.perf_counter()
        step()
        hip.synchronize()
        times.append(time.perf_counter() - t1)
        st = b.stats()
        log(f"step {i} {times[-1] * 1e3:.1f} ms factor {st['factor_ms']:.1f} solve {st['solve_ms']:.1f} "
            f"refine steps {st['refine_steps']:.0f} rel {st['rel_residual']:.2e}")
    x = d_x.download().reshape(allb.shape, order="F")
    worst = 0.0
    off = 0
    for p in plates:
        xi = x[off:off + p.n]
        off += p.n
        worst = max(worst, float(np.max(np.linalg.norm(p.rhs - p.to_scipy() @ xi, axis=0) /
                                        np.linalg.norm(p.rhs, axis=0))))
    t = float(np.median(times))
    out = {"metric": "batched factor + 8-RHS solve of 512 plates, FP64 refined per plate to 1e-10 or the FP64 floor",
           "value": t, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
           "config": {"workload": f"{len(plates)} CPE4 plates, nx~U{{20..200}}, ny=nx/10, 8 load cases, seed 3; "
                                  f"one forest, level-batched launches; {ntot} DOF, {info['fronts']} fronts, "
                                  f"{info['levels']} levels", "n": ntot, "work": info["work"]},
           "problems_per_s": len(plates) / t, "true_rel_residual_max": worst, "stats": b.stats()}
    st = out["stats"]
    if st.get("factor_ms", 0) > 0:
        ach = info["work"] / (st["factor_ms"] * 1e-3) / 1e12
        out["roofline"] = {"bound": "mfma", "achieved": ach, "peak": PEAK["fp64_mfma_tflops"], "unit": "TFLOP/s",
                           "frac": ach / PEAK["fp64_mfma_tflops"], "traffic": None,
                           "kernel": "whole batched FP64 factorization (flops of every plate's fronts / factor "
                                     "time; small fronts, latency-bound)"}
    if args.cpu_sample > 0:
        # the C oracle on every plate of the batch, one after another (each
        # factorization uses the host threads): the baseline at the workload
        # itself, and every plate's FP64 oracle solution
        import oracle
        oracle.build()
        t1 = time.perf_counter()
        worst_o, err = 0.0, 0.0
        sols = []
        for p in plates:
            h, _, _ = oracle.Handle.factor(p)
            sols.append(h.solve(p.rhs))
        dt = time.perf_counter() - t1
        off = 0
        for p, xo in zip(plates, sols):
            xi = x[off:off + p.n]
            off += p.n
            worst_o = max(worst_o, float(np.max(np.linalg.norm(p.rhs - p.to_scipy() @ xo, axis=0) /
                                                np.linalg.norm(p.rhs, axis=0))))
            err = max(err, float(np.max(np.abs(xi - xo)) / np.max(np.abs(xo))))
        out["oracle_rel_residual_max"] = worst_o
        out["error_vs_fp64_oracle"] = err
        out["cpu_baseline"] = {"value": dt, "unit": "s", "cores": int(oracle.threads()), "kind": "port",
                               "sample": f"the whole batch: C oracle FP64 factor + 8-RHS solve of all "
                                         f"{len(plates)} plates in sequence ({dt:.2f} s)"}
    return out


def run_batch(args):
    print(json.dumps(batch_result(args)), flush=True)


def indef_result(args):
    """The reference's indefinite benchmark (examples/solver_bench.rs:87-157):
    the 7-point Laplacian on a g^3 grid shifted by sigma midway between two
    eigenvalues near the 10 % quantile, factored by the indefinite LDL^T
    (1x1/2x2 pivots, delayed pivots, Bunch-Kaufman at the roots).  One step =
    the numeric factorization from the host values (the pattern's analysis is
    done once, as the reference times indef::factor after sparse::analyse).
    Checks: the inertia against the analytic count (51,205 at 80^3,
    docs/REFERENCE.md:419-422) and the solve residual."""
    import numpy as np
    from rustyfem_amd import ffi, meshes
    ffi.lib()
    g = args.indef_size
    t0 = time.perf_counter()
    sigma, below = meshes.laplacian_sigma(g)
    p = meshes.laplacian3d(g, sigma)
    log(f"shifted Laplacian {g}^3: n={p.n} sigma={sigma:.6f} analytic negatives {below}, "
        f"built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    h, err, ine = ffi.Handle.factor_indef(p)
    if h is None:
        raise RuntimeError(f"indefinite factorization failed: err={err} {ffi.last_error()}")
    first = time.perf_counter() - t0
    times = []
    for i in range(max(1, min(args.steps, 3))):
        t0 = time.perf_counter()
        if h.refactor(p.data) != -1:
            raise RuntimeError("indefinite refactorization failed")
        times.append(time.perf_counter() - t0)
        log(f"indef {g}^3: factorization {times[-1]:.2f}s")
    ine2 = h.inertia()
    x = h.solve(p.rhs)
    rel = float(np.linalg.norm(p.rhs[:, 0] - p.to_scipy() @ x[:, 0]) / np.linalg.norm(p.rhs[:, 0]))
    st = h.factor_stats()
    t = float(np.median(times))
    out = {"value": t, "unit": "s", "ms_per_step": t * 1e3, "dtype": "fp64",
           "config": {"workload": f"shifted 3D Laplacian {g}^3 (n={p.n}), sigma={sigma:.6f} at the 10 % eigenvalue "
                                  f"quantile; indefinite LDL^T with delayed pivots, ND leaf 16",
                      "n": p.n, "sigma": sigma},
           "inertia": [int(v) for v in ine2], "analytic_negatives": below,
           "inertia_matches_analytic": bool(int(ine2[0]) == below and int(ine2[1]) == 0),
           "first_factor_with_analysis_s": first, "true_rel_residual": rel,
           "delayed_pivots": int(st[3]), "two_by_two_pivots": int(st[4]), "factor_entries": int(st[5]),
           "reference_published_s": 44.0 if g == 80 else None,
           "reference_published": "docs/REFERENCE.md:419-422 (CPU, 512,000 unknowns, 51,205 negative pivots)"}
    return out


def run_indef(args):
    print(json.dumps(indef_result(args)), flush=True)


def assembly_result(args):
    """GPU element assembly of the benchmark block (C3D8 B-bar, size^3
    elements, x = 0 clamped, E = 210e9, nu = 0.3 per element): one step =
    the element-matrix kernel + the gather of every stored entry into device
    memory, materials given on the host.  The pattern and contribution lists
    are built once (setup_s).  Checked bit for bit against the oracle's
    sequential restatement of the reference's element loop and scatter, which
    is also the CPU baseline (one core; the reference's own scatter is
    serial, src/fem/assembly.rs:378-398).  Also timed: the refactorization
    fed straight from the assembly (rustyfem_assembly_refactor) against the
    host round trip (values to the host, rustyfem_refactor)."""
    import numpy as np
    import oracle
    from rustyfem_amd import ffi, hip, meshes
    ffi.lib()
    n = args.size
    kk, jj, ii = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    coords = np.stack([ii.ravel() / n, jj.ravel() / n, kk.ravel() / n], axis=1)
    free = (ii > 0).ravel()
    eq = meshes.equation_numbers(free, 3)
    ndof = 3 * int(free.sum())
    conn = meshes.grid_elements((n, n, n))
    ne = conn.shape[0]
    young, nu = np.full(ne, 210e9), np.full(ne, 0.3)
    t0 = time.perf_counter()
    a = ffi.Assembly("C3D8", conn, coords, eq, ndof)
    setup = time.perf_counter() - t0
    out_d = hip.DeviceArray(a.nnz)
    steps = max(args.steps, 3)
    for _ in range(max(args.warmup, 1)):
        a.values(young, nu, out=out_d.ptr)
    hip.synchronize()
    ke_ms, ga_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        a.values(young, nu, out=out_d.ptr)
        st = a.stats()
        ke_ms.append(st["element_ms"])
        ga_ms.append(st["gather_ms"])
    hip.synchronize()
    step_s = (time.perf_counter() - t0) / steps
    gpu = out_d.download()
    out_d.free()
    t0 = time.perf_counter()
    ref = oracle.assemble("C3D8", conn, coords, eq, ndof, young, nu, a.indptr, a.indices)
    cpu_s = time.perf_counter() - t0
    mismatches = int(np.count_nonzero(gpu != ref))
    # algorithmic work: the B^T (D B) products of the 8 points, 2 flops per
    # term over the structurally nonzero rows (3 normal rows of 24 columns,
    # 3 shear rows of 16), and the gather's bytes (lists, element entries,
    # output)
    flops_elem = 2 * 8 * (3 * 24 * 24 + 3 * 16 * 16)
    ncon = st["contributions"]
    gbytes = 4 * (a.nnz + 1) + 4 * ncon + 8 * ncon + 8 * a.nnz
    kms, gms = float(np.median(ke_ms)), float(np.median(ga_ms))
    el_tflops = flops_elem * ne / (kms * 1e-3) / 1e12
    ga_gbs = gbytes / (gms * 1e-3) / 1e9
    res = {"value": step_s, "unit": "s", "ms_per_step": step_s * 1e3, "dtype": "fp64",
           "config": {"workload": f"C3D8 B-bar stiffness assembly of the {n}^3 block (n={ndof} DOF, {ne} elements, "
                                  f"nnz={a.nnz} both triangles); step = element kernel + gather into HBM",
                      "elements": ne, "n": ndof, "nnz": a.nnz, "contributions": ncon},
           "setup_s": setup, "element_kernel_ms": kms, "gather_kernel_ms": gms,
           "roofline": {"bound": "valu", "achieved": el_tflops, "peak": 78.6, "unit": "TFLOP/s",
                        "frac": el_tflops / 78.6, "traffic": None,
                        "kernel": f"k_elem<C3D8> (FP64 vector ALU, no FMA contraction): {flops_elem} flops per "
                                  f"element (the B^T D B terms)"},
           "gather_roofline": {"bound": "hbm", "achieved": ga_gbs, "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                               "frac": ga_gbs / PEAK["hbm_gbs"], "traffic": None,
                               "kernel": f"k_gather_sum: {gbytes} bytes per launch (offsets, contribution "
                                         f"indices, element entries, output)"},
           "cpu_baseline": {"value": cpu_s, "unit": "s", "cores": 1, "kind": "port",
                            "sample": f"the whole workload: the C oracle's sequential element loop and scatter "
                                      f"({ne} elements)"},
           "mismatches_vs_oracle": mismatches}
    # assembly straight into the factorization against the host round trip
    prob = meshes.Problem("asm", ndof, coords.shape[0], np.repeat(np.nonzero(free)[0], 3).astype(np.int64), coords,
                          a.indptr, a.indices, gpu, np.zeros((ndof, 1)))
    h, err, _ = ffi.Handle.factor(prob, precision="mixed")
    if h is not None and err == -1:
        tr, th = [], []
        for _ in range(2):
            hip.synchronize()
            t0 = time.perf_counter()
            a.refactor(h, young, nu)
            hip.synchronize()
            tr.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            h.refactor(a.values(young, nu))
            hip.synchronize()
            th.append(time.perf_counter() - t0)
        res["refactor_from_assembly_s"] = min(tr)
        res["refactor_host_round_trip_s"] = min(th)
        h.free()
    log(f"assembly {n}^3: step {step_s * 1e3:.2f} ms (element {kms:.3f} ms, gather {gms:.3f} ms), "
        f"oracle {cpu_s:.2f} s, mismatches {mismatches}")
    return res


def run_assembly(args):
    print(json.dumps(assembly_result(args)), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=64, help="elements per cube edge (64 = the 811,200-DOF block)")
    ap.add_argument("--cpu-sample", type=int, default=32, help="cube edge of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--cpu-sample-iter", type=int, default=48,
                    help="cube edge of the CPU-baseline sample of the AMG workloads (0 = skip)")
    ap.add_argument("--oracle-sample-iter", type=int, default=24,
                    help="cubes per edge of the C3D4 sample solved by the oracle's direct solver (learned workload)")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="direct workload only: skip the batch and learned-smoother sub-benchmarks")
    ap.add_argument("--indef-size", type=int, default=80, help="grid edge of the shifted Laplacian (indef workload)")
    ap.add_argument("--workload", choices=["direct", "amg", "batch", "learned", "ic0", "indef", "assembly"],
                    default="direct",
                    help="direct: the multifrontal block (default); amg: the heterogeneous block with AMG-Krylov; "
                         "batch: 512 plates factored and solved together; learned: the C3D4 block with the "
                         "learned-smoother FCG")
    ap.add_argument("--tet-size", type=int, default=96, help="cubes per edge of the C3D4 block (learned workload)")
    ap.add_argument("--smoother", default=None, help="learned-smoother parameters (.npz; default rustyfem_amd/data)")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB copied device-to-device twice before the run (a kernel of known traffic for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the multifrontal factorization of one problem is not sharded "
              "across GPUs in this build (see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "amg":
        run_amg(args)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "ic0":
        return run_ic0(args)
    if args.workload == "learned":
        run_learned(args)
        return
    if args.workload == "indef":
        run_indef(args)
        return
    if args.workload == "assembly":
        run_assembly(args)
        return
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
    p = meshes.hex_block(args.size, nrhs=16)
    log(f"problem {p.name}: n={p.n} nnz={p.nnz} 16 RHS built in {time.perf_counter() - t0:.1f}s")
    d_b = hip.DeviceArray.from_host(np.ascontiguousarray(p.rhs[:, 0]))
    d_b16 = hip.DeviceArray.from_host(np.asfortranarray(p.rhs).reshape(-1, order="F"))
    tol = 1e-10
    mixed = run_path(ffi, hip, p, "mixed", d_b, args.steps, args.warmup, tol, d_b16, p.rhs)
    fp64 = None if args.skip_fp64 else run_path(ffi, hip, p, "fp64", d_b, args.steps, args.warmup, tol, d_b16,
                                                  p.rhs)
    st = mixed["stats"][-1]
    gemm_ms, gemm_n, gemm_flops = st["gemm_ms"], st["gemm_launches"], st["gemm_flops"]
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else None
    roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK["fp32_mfma_tflops"], "unit": "TFLOP/s",
            "frac": (achieved / PEAK["fp32_mfma_tflops"]) if achieved else None, "traffic": measured_traffic(),
            "kernel": "k_gemm<float> / k_gemm_agpr<float> / k_gemm_ilv<float> / k_gemm_k64<float> / k_gemm_q "
                      "(Schur update, v_mfma_f32_16x16x4f32; kernel autotuned per launch)",
            "autotuned_agpr_launches": st.get("agpr_launches"),
            "autotuned_double_width_k_launches": st.get("ilv_launches"),
            "autotuned_16byte_launches": st.get("q_launches"),
            "autotune_ms": st.get("tune_ms"),
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
                               f"1 RHS; step = refactor + solve + FP64 refinement to 1e-10",
                   "n": p.n, "nrhs": 1, "factor_flops": st["flops"], "fronts": st["fronts"],
                   "levels": st["levels"]},
        "mixed": {"time_s": mixed["time_s"], "factor_ms": st["factor_ms"], "solve_ms": st["solve_ms"],
                  "refine_steps": st["refine_steps"], "rel_residual": st["rel_residual"],
                  "refine_history": st["refine_history"],
                  "true_rel_residual": mixed["true_rel_residual"], "factor_tflops": factor_tflops,
                  "factor_frac_fp32_peak": factor_tflops / PEAK["fp32_mfma_tflops"]},
        "roofline": roof,
    }
    # the 16-RHS steps of both arms: refactor + 16-RHS solve + refinement; the
    # solve's roofline counts one read of the factor per triangular sweep
    m16 = mixed["rhs16"]
    sub16 = {"workload": "the same block and factorization with 16 RHS (pressure + 15 seeded random nodal "
                         "loads); step = refactor + 16-RHS solve + FP64 refinement to 1e-10",
             "mixed": {"time_s": m16["time_s"], "factor_ms": m16["stats"]["factor_ms"],
                       "solve_ms": m16["stats"]["solve_ms"], "refine_steps": m16["stats"]["refine_steps"],
                       "true_rel_residual": m16["true_rel_residual"]}}
    nsolve = m16["stats"]["refine_steps"] + 1
    sbytes = 2.0 * mixed["factor_entries"] * 4 * nsolve
    ach16 = sbytes / (m16["stats"]["solve_ms"] * 1e-3) / 1e9
    sub16["roofline"] = {"bound": "hbm", "achieved": ach16, "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                         "frac": ach16 / PEAK["hbm_gbs"], "traffic": None,
                         "kernel": "mixed 16-RHS triangular solves with refinement: 2 nnz(L) x 4 bytes per solve "
                                   "(the FP32 factor read forward and backward), %d solves" % nsolve}
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
        f16 = fp64["rhs16"]
        sub16["fp64"] = {"time_s": f16["time_s"], "factor_ms": f16["stats"]["factor_ms"],
                         "solve_ms": f16["stats"]["solve_ms"], "true_rel_residual": f16["true_rel_residual"]}
        sub16["speedup_mixed_vs_fp64"] = f16["time_s"] / m16["time_s"]
        sub16["mixed_vs_fp64_displacement_rel_error"] = float(
            np.max(np.abs(m16["solution"] - f16["solution"])) / np.max(np.abs(f16["solution"])))
    out["c2_16rhs"] = sub16
    d_b.free()
    d_b16.free()
    if args.cpu_sample > 0:
        try:
            base, errs = cpu_baseline(args.cpu_sample, st["flops"], 16)
            out["cpu_baseline"] = base
            out["error_vs_fp64_oracle"] = errs
            sub16["cpu_baseline"] = dict(base)
        except Exception as ex:  # the baseline must not hide the GPU result
            out["cpu_baseline"] = {"value": None, "error": str(ex)}
    if not args.no_extras:
        # the other configurations, each with its own roofline, CPU baseline
        # and error against the FP64 oracle; a failure is recorded, not raised
        for key, fn in (("c3_batch", batch_result), ("c5_learned", learned_result),
                        ("f1_assembly", assembly_result)):
            t0 = time.perf_counter()
            try:
                r = fn(args)
                for k in ("metric", "n_gpus", "steps", "warmup", "higher_is_better", "scaling", "vs_baseline",
                          "data"):
                    r.pop(k, None)
                out[key] = r
            except Exception as ex:
                out[key] = {"error": f"{type(ex).__name__}: {ex}"}
            log(f"{key} done in {time.perf_counter() - t0:.1f}s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
