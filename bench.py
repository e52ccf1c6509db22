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


HEADLINE_RESTART = 50   # GMRES restart of every path in the headline (the reference drivers' mr = 50)


def headline(args):
    """End-to-end solve of 64 incidences (32 theta x pol V/H, phi = 0) on a dielectric sphere r = 2 lambda,
    eps_r = 2.5 - 0.1j, PMCHWT, geodesic mesh of 151,230 RWG (302,460 unknowns, edge ~ lambda / 35), 5-level
    tree, GMRES(50) (the reference drivers' restart, Code/RCS.h:215-219) to a true FP64 residual of 1e-3 per
    incidence.  (The reference's multilevel product loses
    accuracy as the coarsest boxes grow: on an r = 2.84 lambda sphere, 6-level tree, it is 25% off the dense
    Galerkin product and no solver reaches 1e-3 on it, so the 302k unknowns come from a finer mesh of the
    r = 2 lambda sphere.)

    mixed: FP32 far field and FP32 near storage, all 64 right-hand sides in one lockstep batched GMRES (flexible,
    restarts from the true FP64 residual), the learned polynomial preconditioner with its near products in BF16
    on the matrix cores (network trained here on a family of smaller spheres; the training time is reported
    apart).  FP64 sequential: the FP64 operator with the same preconditioner coefficients (FP64 near products), one
    incidence at a time, solved for `--plain-sample` incidences spread over theta; its 64-incidence time is
    setup + 64 x the mean (stated in the output).  The reference's own method (FP64, unpreconditioned GMRES(50)) is
    run on two of them with an iteration cap: on this sphere it stalls near 1e-2.  Backscatter RCS of the mixed
    solutions is compared with FP64 solutions converged to 1e-6."""
    import numpy as np
    from mlfma_pmchwt_amd import api, geometry, learned

    thetas = np.arange(32) * 180.0 / 31.0
    inc = [(pol, th, 0.0) for pol in (0.0, 90.0) for th in thetas]
    R = len(inc)
    eps = (2.5, -0.1)
    t0 = time.perf_counter()
    mesh = geometry.geodesic_sphere(71, 2.0)
    log(f"mesh {mesh.n_faces} triangles ({time.perf_counter() - t0:.1f} s)")

    # ---- learned preconditioner: training family (offline cost)
    t0 = time.perf_counter()
    rng = np.random.default_rng(2024)
    feats, Gs, hs, cs = [], [], [], []
    for radius in (0.5, 0.75, 1.0, 1.5):
        for density in (25, 35):   # edge ~ lambda / 21 and lambda / 29
            for _ in range(3):
                e = (float(rng.uniform(1.5, 6.0)), -float(rng.uniform(0.0, 0.3)))
                m = geometry.geodesic_sphere(int(round(density * radius)), radius)
                o = api.Operator(m, integral_equation=api.PMCHWT, eps_r_re=e[0], eps_r_im=e[1])
                f, G, h, c = learned.collect(o, o.excitation(0.0, float(rng.uniform(0, 180)), 0.0), steps=30)
                feats.append(f), Gs.append(G), hs.append(h), cs.append(c)
                o.close()
    data_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    lp = learned.LearnedPreconditioner(seed=7)
    hist = lp.fit(np.array(feats), np.array(Gs), np.array(hs), np.array(cs), steps=2000, lr=2e-3)
    train_s = time.perf_counter() - t0
    log(f"training: {len(feats)} problems, data {data_s:.1f} s, fit {train_s:.1f} s, loss {hist[0]:.3f} -> {hist[-1]:.3f}")

    # ---- mixed + batched + learned, end to end
    t_all = time.perf_counter()
    near_prec = api.BF16 if args.near == "bf16x3" else api.FP32
    op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=eps[0], eps_r_im=eps[1], far_precision=api.FP32,
                      near_precision=near_prec, precond_precision=api.BF16, max_rhs=R, verbose=1)
    setup_mixed = time.perf_counter() - t_all
    B = np.stack([op.excitation(*a) for a in inc], axis=1)
    t0 = time.perf_counter()
    alpha = lp.apply_to(op)
    predict_s = time.perf_counter() - t0
    op.set_apply_precision(api.FP32)
    t0 = time.perf_counter()
    tune = op.autotune(R, api.FP32)
    tune_s = time.perf_counter() - t0
    op.set_graphs(True)
    log(f"autotune {tune_s:.2f} s: {tune}")
    X, st = op.solve_batch(B, tol_rel=1e-3, restart=HEADLINE_RESTART, itr_max=3000, precond=api.PRECOND_POLY)
    mixed_total = time.perf_counter() - t_all
    its = [s["iterations"] for s in st]
    conv = all(s["converged"] for s in st)
    log(f"mixed: setup {setup_mixed:.2f} s, predict {predict_s:.2f} s, solve {st[0]['seconds']:.2f} s, "
        f"total {mixed_total:.2f} s; iterations {min(its)}..{max(its)}, converged {conv}, "
        f"fallbacks {st[0]['precond_fallbacks']}, max residual {max(s['rel_residual'] for s in st):.2e}")
    rcs_mixed = [op.rcs(X[:, j], inc[j][1], 0.0, pol=inc[j][0])[0, 0] for j in range(R)]
    n = op.n
    op.close()

    # ---- plain FP64 paths, one incidence at a time, on a sample of incidences
    t0 = time.perf_counter()
    opp = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=eps[0], eps_r_im=eps[1])
    setup_plain = time.perf_counter() - t0
    opp.set_precond_poly(alpha)
    sample = [int(round(k * (R // 2 - 1) / max(args.plain_sample - 1, 1))) for k in range(args.plain_sample)]
    ref_method, fp64_learned = [], []
    for k, j in enumerate(sample):
        b = opp.excitation(*inc[j])
        if k < 2:   # the reference's method: FP64, unpreconditioned GMRES(50), capped
            x, s = opp.solve(b, tol_rel=1e-3, restart=HEADLINE_RESTART, itr_max=args.plain_cap)
            ref_method.append({"incidence": inc[j], "iterations": s["iterations"], "seconds": round(s["seconds"], 3),
                               "converged": s["converged"], "rel_residual": s["rel_residual"]})
            log(f"plain FP64 unpreconditioned {inc[j]}: {s['iterations']} iterations, {s['seconds']:.2f} s, "
                f"converged {s['converged']}, residual {s['rel_residual']:.2e}")
        x, s = opp.solve(b, tol_rel=1e-3, restart=HEADLINE_RESTART, itr_max=3000, precond=api.PRECOND_POLY)
        # RCS reference: the same incidence solved to 1e-6 in FP64
        xr, sr = opp.solve(b, tol_rel=1e-6, restart=HEADLINE_RESTART, itr_max=6000, precond=api.PRECOND_POLY)
        rref = opp.rcs(xr, inc[j][1], 0.0, pol=inc[j][0])[0, 0]
        fp64_learned.append({"incidence": inc[j], "iterations": s["iterations"], "seconds": round(s["seconds"], 3),
                             "converged": s["converged"], "rel_residual": s["rel_residual"],
                             "rcs_fp64_1e-6_dbsm": round(rref, 4),
                             "rcs_diff_mixed_db": round(abs(rcs_mixed[j] - rref), 4),
                             "rcs_diff_fp64_1e-3_db": round(abs(opp.rcs(x, inc[j][1], 0.0, pol=inc[j][0])[0, 0] - rref), 4)})
        log(f"FP64 learned sequential {inc[j]}: {s['iterations']} iterations, {s['seconds']:.2f} s; backscatter "
            f"RCS mixed vs FP64 (1e-6) {abs(rcs_mixed[j] - rref):.3f} dB")
    opp.close()
    per = float(np.mean([p["seconds"] for p in fp64_learned]))
    seq_total = setup_plain + per * R
    out = {
        "metric": METRIC, "workload": "PMCHWT sphere r=2 lambda eps_r=2.5-0.1j, 151,230 RWG (302,460 unknowns), "
                                      "5-level tree, 64 incidences (32 theta x V/H), GMRES(50) to 1e-3",
        "n_unknowns": n, "rhs": R,
        "mixed": {"total_s": round(mixed_total, 3), "setup_s": round(setup_mixed, 3), "predict_s": round(predict_s, 3),
                  "autotune_s": round(tune_s, 3), "tuned": tune, "graphs": True,
                  "preconditioner_near_products": "bf16 matrix cores, fp32 accumulation",
                  "operator_near_products": ("split-pair bf16 matrix cores (hi + lo, three partial products), "
                                             "fp32 accumulation" if args.near == "bf16x3"
                                             else "fp32 matrix cores"),
                  "solve_s": round(st[0]["seconds"], 3), "iterations_min": min(its), "iterations_max": max(its),
                  "converged": conv, "precond_fallbacks": st[0]["precond_fallbacks"],
                  "max_true_rel_residual": max(s["rel_residual"] for s in st),
                  "alpha": [[round(a.real, 4), round(a.imag, 4)] for a in alpha]},
        "plain_fp64_unpreconditioned": {"sampled": ref_method, "iteration_cap": args.plain_cap,
                                        "note": "the reference's method; does not reach 1e-3 within the cap, so "
                                                "its time to 1e-3 exceeds the capped time (lower-bound speedup)"},
        "fp64_learned_sequential": {"setup_s": round(setup_plain, 3), "sampled": fp64_learned,
                                    "mean_s_per_incidence": round(per, 3), "total_s_estimate": round(seq_total, 2),
                                    "estimate": f"setup + {R} x mean of {len(fp64_learned)} measured incidences"},
        "speedup_vs_fp64_learned_sequential": round(seq_total / mixed_total, 2),
        "speedup_vs_plain_fp64_lower_bound": round((setup_plain + R * float(np.mean([p["seconds"] for p in ref_method])))
                                                   / mixed_total, 2) if ref_method and not any(p["converged"] for p in ref_method) else None,
        "max_rcs_diff_mixed_vs_fp64_db": max(f["rcs_diff_mixed_db"] for f in fp64_learned),
        "training": {"problems": len(feats), "data_s": round(data_s, 2), "fit_s": round(train_s, 2),
                     "loss_first": float(hist[0]), "loss_last": float(hist[-1])},
    }
    print(json.dumps(out))


def multi_object(args):
    """Several objects of uneven size in one batch run (the reference's batch
    configuration format): eight PMCHWT spheres of 3k-35k RWG, 6 incidences x 2
    polarisations each, GMRES(50) to 1e-4 (the reference drivers' settings) in
    FP32 mode, first one object after another, then `--concurrency` objects at
    once on that many host threads.  The results must not depend on it."""
    import tempfile
    import numpy as np
    from mlfma_pmchwt_amd import api, geometry
    objs = [(10, 0.3), (12, 0.35), (14, 0.4), (16, 0.45), (18, 0.5), (20, 0.55), (22, 0.6), (24, 0.65)]
    tmp = tempfile.mkdtemp(prefix="mlfma_objects_")
    paths = []
    for k, (n, r) in enumerate(objs):
        p = os.path.join(tmp, f"object{k}.nas")
        geometry.write_nas(p, geometry.geodesic_sphere(n, r))
        paths.append(p)
    cfg = (f"nas_files = {'; '.join(paths)}\nfrequencies = 299792458\nincident_theta_range = 0:150:30\n"
           "material_list = 4.0,-0.1\nsca_theta = 0:180\nsca_phi = 0:0\nstep = 2\nintegral_equation = 2\n")
    sp = api.SolveParams(api.GMRES, 2000, 50, 1e-4, 0.0, 0)
    runs = {}
    for conc in (1, args.concurrency):
        b = api.Batch(text=cfg)
        t0 = time.perf_counter()
        rc = b.run(sp, concurrency=conc, far_precision=api.FP32, near_precision=api.FP32)
        dt = time.perf_counter() - t0
        res = [[b.result(i, pol) for pol in (0, 1)] for i in range(b.n_tasks)]
        b.close()
        its = [r[2]["iterations"] for t in res for r in t]
        log(f"concurrency {conc}: {dt:.2f} s, status {rc}, iterations {min(its)}..{max(its)}")
        runs[conc] = (dt, res)
    (t1, r1), (tk, rk) = runs[1], runs[args.concurrency]
    same_its = all(a[2]["iterations"] == c[2]["iterations"] for ta, tc in zip(r1, rk) for a, c in zip(ta, tc))
    max_db = max(float(np.max(np.abs(a[1] - c[1]))) for ta, tc in zip(r1, rk) for a, c in zip(ta, tc))
    print(json.dumps({
        "metric": "multi-object batch wall s (8 PMCHWT spheres, 12 runs each)", "value": round(tk, 3), "unit": "s",
        "n_gpus": 1, "higher_is_better": False, "data": "synthetic", "dtype": "f32",
        "config": {"workload": "8 PMCHWT spheres eps_r=4-0.1j, geodesic n=10..24 (3,000-17,280 RWG), 6 theta x 2 "
                               "polarisations each, GMRES(50) to 1e-4, FP32 mode",
                   "objects": [{"geodesic_n": n, "radius_lambda": r, "n_rwg": 30 * n * n} for n, r in objs]},
        "sequential_s": round(t1, 3), "concurrency": args.concurrency, "concurrent_s": round(tk, 3),
        "speedup": round(t1 / tk, 3), "same_iterations": same_its, "max_rcs_diff_db": max_db}))


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
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed region, copy 1 GiB device to device once (counter calibration run)")
    ap.add_argument("--concurrency", type=int, default=4, help="multi_object: objects processed at once")
    ap.add_argument("--workload", default="pec_cfie", choices=["pec_cfie", "pmchwt_batch", "multi_object"],
                    help="pec_cfie: matvec/s on the 300k-RWG CFIE sphere (default); pmchwt_batch: end-to-end solve "
                         "of 64 incidences on the 300k-unknown PMCHWT sphere, mixed + batched + learned vs plain FP64; "
                         "multi_object: eight objects of uneven size in one batch run, sequential vs concurrent")
    ap.add_argument("--near", default="bf16x3", choices=["bf16x3", "fp32"],
                    help="pmchwt_batch: the mixed path's batched near products (split-pair BF16 or FP32)")
    ap.add_argument("--plain-sample", type=int, default=4, help="pmchwt_batch: incidences solved one at a time in FP64")
    ap.add_argument("--plain-cap", type=int, default=3000, help="pmchwt_batch: iteration cap of the unpreconditioned solves")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the whole operator runs on a single MI355X "
              "(sharding over several GPUs is not launched)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from mlfma_pmchwt_amd import api, geometry

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.workload == "pmchwt_batch":
        headline(args)
        return
    if args.workload == "multi_object":
        multi_object(args)
        return

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

    if args.calibrate_copy:   # a plain copy kernel of known size for the HBM-counter calibration
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.5, out=dst)   # elementwise copy kernel: 1 GiB read, 1 GiB written
        torch.cuda.synchronize()
        log("calibration copy kernel: 1 GiB read, 1 GiB written")
        del src, dst

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
    traffic_file = ROOT / "profiles" / f"{dom}_traffic_pec_cfie_300k_rwg.json"
    if traffic_file.exists() and args.frequency == 100 and args.radius == 4.0:
        measured = json.loads(traffic_file.read_text())
        roofline["traffic"] = measured["traffic_bytes_per_launch"]
        roofline["traffic_source"] = f"profiles/{traffic_file.name} (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, " \
                                     f"scaled by a 1 GiB copy)"
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
