#!/usr/bin/env python3
"""MLFMA throughput on one MI355X.

Workload (default): PEC sphere of radius 4 lambda, CFIE with alpha = 0.5,
geodesic mesh with 300,000 RWG functions (edge ~ lambda/21), leaf box 0.3 lambda,
6-level tree (levels 0..5), single right-hand side, FP64 (plain method).

A step is one full MLFMA matvec y = A x (near-field product, aggregation,
interpolation, translation, anterpolation, disaggregation) with x and y
resident in HBM.  `value` is matvecs per second.  The same run also reports the
GMRES(30) solve of the theta = 0 plane-wave excitation to 1e-3 true residual,
the per-kernel-group times measured with hipEvents on the operator's stream
over the timed steps, the roofline of the dominant group, and the CPU
baseline (the OpenMP FP64 restatement of the reference matvec on the same
operator structure, timed on this host).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "solve s to 1e-3 residual, learned-precond mixed vs plain FP64; MLFMA matvec/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(op, group):
    """HBM bytes one launch of a kernel group must move at minimum (FP64 complex = 16 B).

    near:         every stored near-field entry once + x gather + y write
    disaggregate: receiving patterns once + finest Bm + y read/modify/write
    aggregate:    radiation patterns once + x + finest Sm write
    """
    c = 16
    n = op.n
    K0 = op.k0
    cur = 2 if op.params.integral_equation == 2 else 1
    waves = cur
    if group == "near":
        return op.near_entries * c + 2 * n * c
    if group == "disaggregate":
        return waves * (op.n_rwg * 3 * K0 * c + op.n_leaves * cur * 3 * K0 * c) + 2 * n * c
    if group == "aggregate":
        return op.n_rwg * 3 * K0 * c + op.n_leaves * cur * 3 * K0 * c + n * c
    return None


def cpu_baseline(mesh, args):
    """Oracle matvec (C++/OpenMP FP64, reference parity mode) on this host."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import ctypes
    import numpy as np
    import pyoracle
    lib = pyoracle.library()
    lib.orc_threads.restype = ctypes.c_int
    threads = lib.orc_threads()
    t0 = time.perf_counter()
    orc = pyoracle.Oracle(mesh, ie=1, alpha=0.5, mode=4)
    build_s = time.perf_counter() - t0
    x = np.ones(orc.n, dtype=np.complex128)
    orc.apply(x)   # first touch
    times = []
    t_start = time.perf_counter()
    while len(times) < args.cpu_matvecs or (time.perf_counter() - t_start) < 10.0 and len(times) < 8:
        t = time.perf_counter()
        orc.apply(x)
        times.append(time.perf_counter() - t)
    del orc
    per = sum(times) / len(times)
    log(f"cpu baseline: {len(times)} matvecs, {per:.3f} s each, {threads} threads (structure build {build_s:.1f} s)")
    return {"value": 1.0 / per, "unit": "matvec/s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} full matvecs of the same {mesh.n_edges_interior}-RWG CFIE operator "
                      f"(same octree, grids, translation tables and near-field sparsity; pattern and near-field "
                      f"values synthetic, which does not change CPU matvec time)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frequency", type=int, default=100, help="geodesic frequency n (30 n^2 RWG)")
    ap.add_argument("--radius", type=float, default=4.0)
    ap.add_argument("--alpha", type=float, default=0.5)
    ap.add_argument("--cpu-matvecs", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--batch", type=int, default=32, help="right-hand sides of the batched-matvec measurement")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the whole operator on a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not launched)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from mlfma_pmchwt_amd import api, geometry

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)

    t0 = time.perf_counter()
    mesh = geometry.geodesic_sphere(args.frequency, args.radius)
    log(f"mesh: {mesh.n_faces} triangles in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    op = api.Operator(mesh, integral_equation=api.CFIE, cfie_alpha=args.alpha, far_precision=api.FP32,
                      near_precision=api.FP32, max_rhs=args.batch, verbose=1)
    setup_s = time.perf_counter() - t0
    op.k0 = 2 * op.grid(1, op.max_level)[0] ** 2
    log(f"setup {setup_s:.2f} s (FP64 and FP32 tables): N={op.n_rwg} levels={op.max_level + 1} "
        f"leaves={op.n_leaves} near={op.near_entries} device={op.device_bytes / 1e9:.2f} GB")

    gen = torch.Generator(device="cuda")
    gen.manual_seed(12345)
    x = torch.randn(op.n, dtype=torch.complex128, device="cuda", generator=gen)
    ys = {}

    def timed(prec):
        op.set_apply_precision(prec)
        y = torch.empty_like(x)
        for _ in range(args.warmup):
            op.apply_device(x.data_ptr(), y.data_ptr())
        torch.cuda.synchronize()
        op.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            op.apply_device(x.data_ptr(), y.data_ptr())
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        prof = op.profile()
        op.set_profiling(False)
        ys[prec] = y
        return elapsed, prof

    def solve_with(prec):
        op.set_apply_precision(prec)
        b = op.excitation(0.0, 0.0, 0.0)
        xs, st = op.solve(b, restart=30, tol_rel=1e-3)
        return {"solver": "GMRES(30)", "tol_rel": 1e-3, "iterations": st["iterations"], "matvecs": st["matvecs"],
                "fp64_matvecs": st["fp64_matvecs"], "converged": st["converged"],
                "true_rel_residual": st["rel_residual"], "seconds": round(st["seconds"], 4)}

    elapsed, prof = timed(api.FP64)
    ms_step = 1e3 * elapsed / args.steps
    log(f"FP64: {args.steps} matvecs in {elapsed:.3f} s -> {ms_step:.3f} ms/matvec")
    elapsed32, prof32 = timed(api.FP32)
    ms32 = 1e3 * elapsed32 / args.steps
    err32 = float(torch.linalg.vector_norm(ys[api.FP32] - ys[api.FP64]) / torch.linalg.vector_norm(ys[api.FP64]))
    log(f"FP32 mode (far field and near storage): {ms32:.3f} ms/matvec, rel. difference to FP64 {err32:.2e}")

    batched = None
    if args.batch > 1:
        Xb = torch.randn(args.batch, op.n, dtype=torch.complex128, device="cuda", generator=gen)
        Yb = torch.empty_like(Xb)
        batched = {"rhs": args.batch}
        for prec, name in ((api.FP64, "fp64"), (api.FP32, "fp32_mode")):
            op.set_apply_precision(prec)
            op.apply_device(Xb.data_ptr(), Yb.data_ptr(), args.batch)
            torch.cuda.synchronize()
            reps = max(1, args.steps // 4)
            t0 = time.perf_counter()
            for _ in range(reps):
                op.apply_device(Xb.data_ptr(), Yb.data_ptr(), args.batch)
            torch.cuda.synchronize()
            ms_b = 1e3 * (time.perf_counter() - t0) / reps
            batched[name] = {"ms_per_batch": round(ms_b, 3), "rhs_matvec_per_s": round(1e3 * args.batch / ms_b, 2)}
            log(f"batched {args.batch} RHS {name}: {ms_b:.2f} ms per batch -> "
                f"{1e3 * args.batch / ms_b:.1f} RHS-matvecs/s")
        del Xb, Yb

    groups = prof["groups"]
    per_launch = {g: v["ms"] / max(v["launches"], 1) for g, v in groups.items()}
    dom = max(("near", "disaggregate", "aggregate"), key=lambda g: groups[g]["ms"])
    nbytes = algorithmic_bytes(op, dom)
    achieved = nbytes / (per_launch[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": nbytes, "ms_per_launch": round(per_launch[dom], 4)}
    stage_ms = {g: round(v["ms"] / max(prof["matvecs"], 1), 4) for g, v in groups.items()}
    stage_ms32 = {g: round(v["ms"] / max(prof32["matvecs"], 1), 4) for g, v in prof32["groups"].items()}

    solve = solve32 = None
    if not args.no_solve:
        solve = solve_with(api.FP64)
        log(f"solve FP64: {solve}")
        solve32 = solve_with(api.FP32)
        log(f"solve FP32 mode + FP64 residual: {solve32}")
    fp32 = {"mode": "far field FP32 storage+accumulation, near entries FP32 storage (FP64 accumulation)",
            "value": round(args.steps / elapsed32, 3), "unit": "matvec/s", "ms_per_step": round(ms32, 4),
            "rel_diff_vs_fp64": err32, "stage_ms_per_matvec": stage_ms32, "solve": solve32}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, args)

    out = {
        "metric": METRIC,
        "value": round(args.steps / elapsed, 3),
        "unit": "matvec/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"PEC sphere r={args.radius} lambda, CFIE alpha={args.alpha}, geodesic mesh "
                               f"{op.n_rwg} RWG, leaf 0.3 lambda, {op.max_level + 1}-level MLFMA, 1 RHS, FP64",
                   "n_rwg": op.n_rwg, "levels": op.max_level + 1, "near_entries": op.near_entries},
        "setup_s": round(setup_s, 3),
        "stage_ms_per_matvec": stage_ms,
        "roofline": roofline,
        "solve": solve,
        "fp32_mode": fp32,
        "batched": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
