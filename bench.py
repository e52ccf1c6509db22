#!/usr/bin/env python3
"""MLFMA solve time and matvec throughput on one MI355X.

Default workload (the metric's configuration): dielectric ellipsoid with semi-axes (3, 2, 1)
lambda, eps_r drawn from U[2, 8] by default_rng(1000) (5.128), PMCHWT, geodesic mesh of 151,230
RWG (302,460 unknowns), 6-level tree, 64 incidences (32 theta x V/H polarisation, phi = 0), every
incidence solved to a true FP64 relative residual of 1e-3 with GMRES(50).

* mixed (the value): FP32 far field, split-pair BF16 near field and disaggregation on the matrix
  cores, all 64 right-hand sides in one lockstep batched flexible GMRES with recorded matvecs,
  preconditioned by the learned near-field polynomial (coefficients predicted by the coefficient
  network, trained on the GPU beforehand on 200 seeded smaller members of the C5 family; BF16
  matrix-core near products); every restart judged by the FP64 operator.  value = end-to-end
  seconds: operator set-up, excitations, features and prediction, launch autotuning and the solve
  (the training is reported apart).
* the inner near-field solve (GMRES steps on the stored near field per application) on the same
  operator and right-hand sides, as the non-learned alternative.
* plain FP64: the reference's method on the GPU (FP64 operator, unpreconditioned GMRES(50), one
  incidence at a time), timed on a sample of incidences with an iteration cap; it does not reach
  1e-3 within the cap, so its 64-incidence time is a lower bound, stated as such.
* FP64 with the same (learned) preconditioner, one incidence at a time (sampled, extrapolated and
  labelled), whose solutions also give the RCS reference of the mixed solutions.
* a step: one 64-right-hand-side matvec of the mixed operator (K timed after W warm-up steps).
* roofline: the dominant kernel group of a profiled 64-RHS matvec (hipEvents on the operator's
  stream); cpu_baseline: the oracle's FP64 matvec on the same operator structure on this host.
* extra key pec_cfie_matvec: FP64 matvec/s of the PEC CFIE sphere (r = 4 lambda, 300,000 RWG).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload c5|pec_cfie|multi_object]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "solve s to 1e-3 residual, learned-precond mixed vs plain FP64; MLFMA matvec/s"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFS = 157.3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------- host description
def host_cpu():
    """lscpu model, physical cores usable by this process."""
    model, cores, sockets, tpc = "unknown", None, 1, 1
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k.strip() == "Model name":
                model = v
            elif k.strip() == "Core(s) per socket":
                cores = int(v)
            elif k.strip() == "Socket(s)":
                sockets = int(v)
            elif k.strip() == "Thread(s) per core":
                tpc = int(v)
    except Exception:
        pass
    allowed = len(os.sched_getaffinity(0))
    physical = (cores * sockets) if cores else allowed
    return {"model": model, "physical_cores": physical, "threads_per_core": tpc,
            "cpus_allowed": allowed, "threads_used": max(1, min(physical, allowed // max(tpc, 1)))}


# ---------------------------------------------------------------- roofline of a kernel group
def translation_flops(op, R, ncur):
    """FLOPs of translation + anterpolation for R right-hand sides: 8 per complex MAC, 3
    components, ncur currents; per level every distant-relative pair x K samples, and the
    9-tap anterpolation of every child sample from its parent."""
    f = 0.0
    nw = 2 if op.params.integral_equation == 2 else 1
    for wn in range(1, nw + 1):
        for lvl in range(2, op.max_level + 1):
            L = op.grid(wn, lvl)[0]
            K = 2 * L * L
            pairs = op.level_list(lvl, 1)[1].size
            f += 8.0 * 3 * ncur * R * pairs * K
            if lvl > 2:
                f += 8.0 * 3 * ncur * R * 9 * op.level(lvl)[0].size * K
    return f


def far_flops(op, R, group):
    """Algorithmic FLOPs (8 per complex multiply-add) of one R-right-hand-side matvec's
    aggregation or disaggregation: every RWG's 3K pattern samples against every current
    (aggregation: J and M into the leaf workspaces; disaggregation: the two outputs), per
    wavenumber."""
    nw = 2 if op.params.integral_equation == 2 else 1
    cur = 2 if op.params.integral_equation == 2 else 1
    f = 0.0
    for wn in range(1, nw + 1):
        L = op.grid(wn, op.max_level)[0]
        f += 8.0 * op.n_rwg * 3 * (2 * L * L) * R * cur
    return f


def group_roofline(op, R, dom, groups):
    """Roofline of the dominant kernel group of the profiled matvec (per launch, from the
    group's hipEvent time on the operator's stream)."""
    g = groups[dom]
    launches = max(g["launches"], 1)
    per_launch_ms = g["ms"] / launches
    if dom == "translate_anterpolate":
        kern = "k_m2l_cl + k_l2l_sep (sibling-cluster translation, then separable anterpolation; FP32 vector ALUs)"
        work, unit, peak, bound = translation_flops(op, R, 2) / launches, "TFLOP/s", FP32_PEAK_TFS, "mfma"
        note = ("compute roofline at the FP32 dense peak; the kernel does its FP32 FMAs on the vector ALUs, where "
                "packed FP32 instructions are disabled in this build (half that rate); counters: the vector-memory "
                "address unit is busy about 60% of the kernel's time, VALU issue about 38%")
    elif dom == "aggregate" and op.params.near_precision == 2:
        kern = "k_aggregate_bf3 (split-pair BF16 matrix cores, all right-hand sides)"
        work, unit, peak, bound = far_flops(op, R, dom) / launches, "TFLOP/s", 2500.0, "mfma"
        note = "BF16 MFMA dense peak; the split-pair product issues three BF16 products per complex multiply-add"
    elif dom == "aggregate":
        kern = "k_aggregate_all (FP32 matrix cores, all right-hand sides)"
        work, unit, peak, bound = far_flops(op, R, dom) / launches, "TFLOP/s", FP32_PEAK_TFS, "mfma"
        note = "FP32 MFMA dense peak"
    elif dom == "disaggregate":
        kern = "k_disagg_bf3 (split-pair BF16 matrix cores, all right-hand sides) + k_beff_all"
        work, unit, peak, bound = far_flops(op, R, dom) / launches, "TFLOP/s", 2500.0, "mfma"
        note = "BF16 MFMA dense peak; the split-pair product issues three BF16 products per complex multiply-add"
    elif dom == "near":
        kern = "k_near_bf16 (split-pair BF16 near field, all right-hand sides)"
        work, unit, peak, bound = (op.near_entries * 8.0 + 2 * op.n * R * 8 + 2 * op.n * R * 16) / launches, \
            "GB/s", HBM_PEAK_GBS, "hbm"
        note = "BF16 hi/lo record per stored entry (8 B) + split X (8 B) + FP64 Y (16 B)"
    else:
        return {"kernel_group": dom, "ms_per_matvec": round(g["ms"], 3)}
    scale = 1e12 if unit == "TFLOP/s" else 1e9
    achieved = work / (per_launch_ms * 1e-3) / scale
    return {"bound": bound, "kernel": kern, "achieved": round(achieved, 2), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "traffic": None, "work_per_launch": work,
            "ms_per_matvec": round(g["ms"], 3), "launches_per_matvec": g["launches"],
            "ms_per_launch": round(per_launch_ms, 4), "note": note}


def matvec_steps(op, args, R):
    """W warm-up and K timed R-right-hand-side mixed matvecs on device buffers (recorded
    graph, the solver's form), then one profiled matvec for the stage groups (hipEvents on
    the operator's stream) and the dominant group's roofline."""
    import torch
    gen = torch.Generator(device="cuda")
    gen.manual_seed(12345)
    Xd = torch.randn(R, op.n, dtype=torch.complex128, device="cuda", generator=gen)
    Yd = torch.empty_like(Xd)
    op.set_graphs(True)
    for _ in range(args.warmup):
        op.apply_device(Xd.data_ptr(), Yd.data_ptr(), R)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.apply_device(Xd.data_ptr(), Yd.data_ptr(), R)
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - t0) / args.steps
    log(f"step: {R}-RHS mixed matvec {1e3 * step_s:.2f} ms ({R / step_s:.0f} RHS-matvecs/s)")
    op.set_graphs(False)
    op.set_profiling(True)
    op.apply_device(Xd.data_ptr(), Yd.data_ptr(), R)
    torch.cuda.synchronize()
    prof = op.profile()
    op.set_profiling(False)
    groups = prof["groups"]
    dom = max(groups, key=lambda g: groups[g]["ms"])
    stage_ms = {g: round(v["ms"], 3) for g, v in groups.items()}
    log(f"profiled matvec stage groups (ms): {stage_ms}")
    return step_s, stage_ms, group_roofline(op, R, dom, groups)


def c5_matvec(args):
    """The C5 operator's 64-right-hand-side mixed matvec alone (set-up, autotune, K steps),
    for kernel profiling; --with-precond adds two batched applications of a degree-6
    near-field polynomial (the preconditioner's BF16 products)."""
    import numpy as np
    import torch
    from mlfma_pmchwt_amd import api, configs
    eps = configs.c5_eps(args.seed)
    R = 64
    op = api.Operator(configs.c5_mesh(args.frequency_c5), integral_equation=api.PMCHWT, eps_r_re=eps,
                      far_precision=api.FP32, near_precision=api.BF16, precond_precision=api.BF16, max_rhs=R)
    op.set_apply_precision(api.FP32)
    tune = op.autotune(R, api.FP32)
    step_s, stage_ms, roofline = matvec_steps(op, args, R)
    if args.with_precond:
        op.set_precond_poly(np.array([1.0, 0.5, 0.25, 0.1, 0.05, 0.02, 0.01], dtype=complex))
        Xh = np.random.default_rng(2).standard_normal((op.n, R)) + 0j
        for _ in range(2):
            op.precond_apply(Xh)
    print(json.dumps({"metric": "RHS-matvecs/s (64-RHS mixed matvec, C5 operator)", "value": round(R / step_s, 1),
                      "unit": "RHS-matvec/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(1e3 * step_s, 3), "higher_is_better": True, "data": "synthetic",
                      "dtype": "f32+bf16", "config": {"workload": f"C5 operator, {op.n} unknowns, 64 RHS",
                                                      "tuned": tune},
                      "stage_ms_per_64rhs_matvec_profiled": stage_ms, "roofline": roofline}))
    op.close()


# ---------------------------------------------------------------- CPU baseline (child process)
def cpu_baseline_child(argv):
    """Runs in its own process (OpenMP placement set before the runtime starts): the oracle's
    FP64 matvec (C++/OpenMP restatement of the reference) on the C5 operator structure, with
    synthetic near-field and pattern values (matvec time does not depend on the values)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import pyoracle
    from mlfma_pmchwt_amd import configs
    freq, eps, reps = int(argv[0]), float(argv[1]), int(argv[2])
    lib = pyoracle.library()
    t0 = time.perf_counter()
    orc = pyoracle.Oracle(configs.c5_mesh(freq), ie=2, eps=(eps, 0.0), mode=4)
    build = time.perf_counter() - t0
    x = np.ones(orc.n, dtype=np.complex128)
    orc.apply(x)
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        orc.apply(x)
        times.append(time.perf_counter() - t)
    print(json.dumps({"matvec_s": sum(times) / len(times), "matvecs": reps, "build_s": build, "n": orc.n,
                      "threads": int(lib.orc_threads())}))


def cpu_baseline(freq, eps, reps, cpu):
    env = dict(os.environ, OMP_NUM_THREADS=str(cpu["threads_used"]), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-baseline-child", str(freq), str(eps),
                        str(reps)], capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


# ---------------------------------------------------------------- learned preconditioner
def train_learned(args):
    """Coefficient network trained on the GPU on the seeded C5 family: the (3, 2, 1) lambda
    ellipsoid scaled by args.train_scale on a geodesic mesh of frequency args.train_freq (0.25 and
    18 mesh it as densely as the full problem's frequency 71: edges of 0.046 / 0.031 / 0.015
    lambda along the three axes), eps_r from default_rng(seed) for seeds 0..args.train_runs-1, one
    seeded incidence each; features of each problem and least-squares statistics of its own
    GMRES Krylov pairs."""
    import numpy as np
    from mlfma_pmchwt_amd import api, configs, learned
    mesh = configs.c5_mesh(args.train_freq, scale=args.train_scale)
    rng = np.random.default_rng(5)
    feats, Gs, hs, cs = [], [], [], []
    t0 = time.perf_counter()
    for s in range(args.train_runs):
        op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=configs.c5_eps(s))
        inc = (float(rng.integers(2)) * 90.0, float(rng.uniform(0.0, 180.0)), 0.0)
        f, G, h, c = learned.collect(op, op.excitation(*inc), steps=30)
        feats.append(f), Gs.append(G), hs.append(h), cs.append(c)
        op.close()
    data_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    lp = learned.LearnedPreconditioner(seed=7)
    hist = lp.fit(np.array(feats), np.array(Gs), np.array(hs), np.array(cs), steps=2000, lr=2e-3)
    fit_s = time.perf_counter() - t0
    log(f"learned: {args.train_runs} problems ({mesh.n_edges_interior} RWG each) in {data_s:.1f} s, fit {fit_s:.1f} s, "
        f"loss {hist[0]:.3f} -> {hist[-1]:.3f}")
    return lp, {"problems": args.train_runs, "rwg_per_problem": mesh.n_edges_interior,
                "family": f"ellipsoid (3,2,1) x {args.train_scale} lambda, geodesic {args.train_freq}, eps_r U[2,8] "
                          f"seeds 0..{args.train_runs - 1}",
                "data_s": round(data_s, 2), "fit_s": round(fit_s, 2), "loss_first": float(hist[0]),
                "loss_last": float(hist[-1])}


# ---------------------------------------------------------------- C5 headline
def c5_headline(args):
    import numpy as np
    import torch
    from mlfma_pmchwt_amd import api, configs

    eps = configs.c5_eps(args.seed)
    inc = configs.incidences_64()
    R = len(inc)
    t0 = time.perf_counter()
    mesh = configs.c5_mesh(args.frequency_c5)
    log(f"C5 mesh: {mesh.n_faces} triangles ({time.perf_counter() - t0:.1f} s), eps_r {eps:.4f}")

    learned_net, learned_info = train_learned(args)

    # ---- mixed precision + batched + learned preconditioner, end to end
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=eps, far_precision=api.FP32,
                      near_precision=api.BF16, precond_precision=api.BF16, max_rhs=R, verbose=1)
    setup_s = time.perf_counter() - t_all
    B = op.excitation_batch([a[0] for a in inc], [a[1] for a in inc], [a[2] for a in inc])
    t0 = time.perf_counter()
    alpha = learned_net.apply_to(op)
    predict_s = time.perf_counter() - t0
    op.set_apply_precision(api.FP32)
    t0 = time.perf_counter()
    tune = op.autotune(R, api.FP32)
    tune_s = time.perf_counter() - t0
    op.set_graphs(True)
    X, st = op.solve_batch(B, tol_rel=1e-3, restart=50, itr_max=3000, precond=api.PRECOND_POLY)
    mixed_total = time.perf_counter() - t_all
    its = [s["iterations"] for s in st]
    conv = all(s["converged"] for s in st)
    log(f"mixed + learned: setup {setup_s:.2f} s, predict {predict_s:.2f} s, autotune {tune_s:.2f} s, solve "
        f"{st[0]['seconds']:.2f} s, total {mixed_total:.2f} s; iterations {min(its)}..{max(its)}, converged {conv}, "
        f"fallbacks {st[0]['precond_fallbacks']}")
    rcs_mixed = np.array([op.rcs(X[:, j], inc[j][1], 0.0, pol=inc[j][0])[0, 0] for j in range(R)])

    # ---- the non-learned alternative on the same operator: inner near-field solve
    op.set_precond_inner(args.inner)
    Xi, si = op.solve_batch(B, tol_rel=1e-3, restart=50, itr_max=3000, precond=api.PRECOND_INNER)
    iits = [s["iterations"] for s in si]
    inner = {"preconditioner": f"inner near-field solve, {args.inner} GMRES steps per application (BF16 matrix-core "
                               f"near products)", "solve_s": round(si[0]["seconds"], 3),
             "iterations_min": min(iits), "iterations_max": max(iits),
             "converged": all(s["converged"] for s in si), "max_true_rel_residual": max(s["rel_residual"] for s in si),
             "total_s_with_same_setup": round(setup_s + tune_s + si[0]["seconds"], 3)}
    log(f"mixed + inner near-field solve: {inner}")
    Xi = None

    # ---- steps: K timed 64-RHS mixed matvecs (device buffers, recorded graph) and a profiled one
    step_s, stage_ms, roofline = matvec_steps(op, args, R)
    torch.cuda.empty_cache()

    n = op.n
    op.close()

    # ---- FP64 paths on a plain FP64 operator, one incidence at a time
    t0 = time.perf_counter()
    opp = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=eps)
    setup_plain = time.perf_counter() - t0
    k = args.plain_sample
    sample = sorted({int(round(i * (R - 1) / max(k - 1, 1))) for i in range(k)})
    plain, pre64 = [], []
    for j in sample:
        b = opp.excitation(*inc[j])
        x, s = opp.solve(b, tol_rel=1e-3, restart=50, itr_max=args.plain_cap)
        plain.append({"incidence": list(inc[j]), "iterations": s["iterations"], "seconds": round(s["seconds"], 3),
                      "converged": s["converged"], "true_rel_residual": s["rel_residual"]})
        log(f"plain FP64 {inc[j]}: {s['iterations']} iterations, {s['seconds']:.2f} s, converged {s['converged']}, "
            f"residual {s['rel_residual']:.2e}")
    opp.set_precond_poly(alpha)
    for j in sample[:args.fp64_pre_sample]:
        b = opp.excitation(*inc[j])
        x, s = opp.solve(b, tol_rel=1e-3, restart=50, itr_max=3000, precond=api.PRECOND_POLY)
        r64 = opp.rcs(x, inc[j][1], 0.0, pol=inc[j][0])[0, 0]
        pre64.append({"incidence": list(inc[j]), "iterations": s["iterations"], "seconds": round(s["seconds"], 3),
                      "converged": s["converged"], "true_rel_residual": s["rel_residual"],
                      "rcs_fp64_dbsm": round(float(r64), 4),
                      "rcs_diff_mixed_db": round(float(abs(rcs_mixed[j] - r64)), 4)})
        log(f"FP64 + learned polynomial {inc[j]}: {s['iterations']} iterations, {s['seconds']:.2f} s; "
            f"backscatter RCS mixed vs FP64 {abs(rcs_mixed[j] - r64):.4f} dB")
    opp.close()

    plain_per = float(np.mean([p["seconds"] for p in plain]))
    plain_any_conv = any(p["converged"] for p in plain)
    plain_total = setup_plain + R * plain_per
    pre_per = float(np.mean([p["seconds"] for p in pre64])) if pre64 else None
    pre_total = setup_plain + R * pre_per if pre64 else None

    cpu = None
    if not args.no_cpu:
        info = host_cpu()
        cb = cpu_baseline(args.frequency_c5, eps, args.cpu_matvecs, info)
        log(f"cpu baseline: {cb}")
        cpu = {"value": round(1.0 / cb["matvec_s"], 4), "unit": "matvec/s", "cores": cb["threads"], "kind": "port",
               "sample": f"{cb['matvecs']} FP64 matvecs of the same {cb['n']}-unknown C5 PMCHWT operator (same octree, "
                         f"grids, translation tables and near-field sparsity; pattern and near-field values synthetic, "
                         f"which does not change CPU matvec time), OMP_PROC_BIND=close OMP_PLACES=cores",
               "cpu_model": info["model"], "physical_cores": info["physical_cores"],
               "threads_per_core": info["threads_per_core"], "matvec_s": round(cb["matvec_s"], 3),
               "gpu_fp32_64rhs_matvec_speedup_per_rhs": round(cb["matvec_s"] / (step_s / R), 1)}

    pec = None if args.no_pec else pec_cfie_matvec(args)

    out = {
        "metric": METRIC,
        "value": round(mixed_total, 3),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * step_s, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16 (FP32 far field, split-pair BF16 near field, BF16 preconditioner; FP64 residuals)",
        "data": "synthetic",
        "config": {"workload": f"C5: dielectric ellipsoid (3,2,1) lambda, eps_r={eps:.4f} (U[2,8], default_rng({args.seed})), "
                               f"PMCHWT, geodesic {args.frequency_c5} mesh {n // 2} RWG ({n} unknowns), 6-level MLFMA, "
                               f"64 incidences (32 theta x V/H, phi 0), GMRES(50) to a true FP64 residual of 1e-3",
                   "n_unknowns": n, "rhs": R, "step": "one 64-right-hand-side mixed-precision matvec"},
        "rhs_matvec_per_s": round(R / step_s, 1),
        "mixed": {"total_s": round(mixed_total, 3), "setup_s": round(setup_s, 3), "predict_s": round(predict_s, 3),
                  "autotune_s": round(tune_s, 3),
                  "solve_s": round(st[0]["seconds"], 3), "iterations_min": min(its), "iterations_max": max(its),
                  "converged": conv, "max_true_rel_residual": max(s["rel_residual"] for s in st),
                  "fp64_matvecs_per_rhs": st[0]["fp64_matvecs"], "precond_fallbacks": st[0]["precond_fallbacks"],
                  "preconditioner": "learned near-field polynomial (network prediction, BF16 matrix-core near "
                                    "products)", "alpha": [[round(a.real, 4), round(a.imag, 4)] for a in alpha],
                  "tuned": tune},
        "mixed_inner_near_solve": inner,
        "plain_fp64": {"method": "FP64 operator, unpreconditioned GMRES(50), one incidence at a time (the reference's "
                                 "method)", "setup_s": round(setup_plain, 3), "iteration_cap": args.plain_cap,
                       "sampled": plain, "mean_s_per_incidence_at_cap": round(plain_per, 3),
                       "total_s_64": round(plain_total, 2),
                       "total_is": ("extrapolated lower bound: setup + 64 x the capped per-incidence time; no sampled "
                                    "incidence reached 1e-3" if not plain_any_conv else
                                    "extrapolated: setup + 64 x the sampled mean")},
        "fp64_same_preconditioner_sequential": {"method": "FP64 operator with the same learned polynomial (FP64 near "
                                                          "products), one incidence at a time",
                                                "sampled": pre64, "mean_s_per_incidence": pre_per and round(pre_per, 3),
                                                "total_s_64": pre_total and round(pre_total, 2),
                                                "total_is": "extrapolated: setup + 64 x the sampled mean"},
        "speedup_vs_plain_fp64": round(plain_total / mixed_total, 2),
        "speedup_vs_plain_fp64_is": "lower bound" if not plain_any_conv else "extrapolated",
        "speedup_vs_fp64_same_preconditioner": pre_total and round(pre_total / mixed_total, 2),
        "max_rcs_diff_mixed_vs_fp64_db": max((p["rcs_diff_mixed_db"] for p in pre64), default=None),
        "learned_training": learned_info,
        "stage_ms_per_64rhs_matvec_profiled": stage_ms,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pec_cfie_matvec": pec,
    }
    print(json.dumps(out))


# ---------------------------------------------------------------- C2: FP64 matvec/s (extra key)
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
    if group == "near":
        return op.near_entries * c + 2 * n * c
    if group == "disaggregate":
        return cur * (op.n_rwg * 3 * K0 * c + op.n_leaves * cur * 3 * K0 * c) + 2 * n * c
    if group == "aggregate":
        return op.n_rwg * 3 * K0 * c + op.n_leaves * cur * 3 * K0 * c + n * c
    return None


def pec_cfie_matvec(args):
    """FP64 single-vector matvec/s of the PEC CFIE sphere (r = 4 lambda, alpha = 0.5, 300,000 RWG,
    6-level tree), with per-group times and the near-field kernel's HBM roofline."""
    import torch
    from mlfma_pmchwt_amd import api, configs
    mesh = configs.c2_mesh(100)
    op = api.Operator(mesh, integral_equation=api.CFIE, cfie_alpha=configs.C2_ALPHA)
    op.k0 = 2 * op.grid(1, op.max_level)[0] ** 2
    gen = torch.Generator(device="cuda")
    gen.manual_seed(12345)
    x = torch.randn(op.n, dtype=torch.complex128, device="cuda", generator=gen)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        op.apply_device(x.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()
    op.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.apply_device(x.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    prof = op.profile()
    op.set_profiling(False)
    g = prof["groups"]
    near_ms = g["near"]["ms"] / max(g["near"]["launches"], 1)
    nbytes = algorithmic_bytes(op, "near")
    achieved = nbytes / (near_ms * 1e-3) / 1e9
    out = {"workload": f"PEC sphere r=4 lambda, CFIE alpha=0.5, {op.n_rwg} RWG, {op.max_level + 1}-level MLFMA, 1 RHS, FP64",
           "value": round(args.steps / el, 2), "unit": "matvec/s", "ms_per_matvec": round(1e3 * el / args.steps, 4),
           "stage_ms_per_matvec": {k: round(v["ms"] / max(prof["matvecs"], 1), 4) for k, v in g.items()},
           "near_roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": nbytes,
                             "ms_per_launch": round(near_ms, 4)}}
    log(f"PEC CFIE 300k: {out['value']} matvec/s, near {achieved:.0f} GB/s")
    op.close()
    return out


def multi_object(args):
    """C4: the 16 seeded objects of uneven size (ellipsoids, cubes, cylinders, eps_r U[1.5, 6],
    mesh edge lambda/10, 2.6k-29k RWG), 8 incidences each, PMCHWT, FP32 far field with
    split-pair BF16 near field, GMRES(50) to a true FP64 residual of 1e-3.  Three ways over the
    same operators: one object after another (batched solve per object), `--concurrency` host
    threads taking objects from a queue (each object on its own stream), and all 16 in one
    lockstep solve whose every Arnoldi step is ONE recorded graph with a branch per object
    (mlfma_solve_multi).  The solutions must agree."""
    import queue
    import threading
    import numpy as np
    from mlfma_pmchwt_amd import api, configs
    objs = configs.c4_objects()
    inc = configs.c4_incidences()
    t0 = time.perf_counter()
    ops, Bs = [], []
    for o in objs:
        op = api.Operator(o["mesh"], integral_equation=api.PMCHWT, eps_r_re=o["eps_r"], far_precision=api.FP32,
                          near_precision=api.BF16, max_rhs=len(inc))
        op.set_apply_precision(api.FP32)
        op.set_graphs(True)
        ops.append(op)
        Bs.append(op.excitation_batch([a[0] for a in inc], [a[1] for a in inc], [a[2] for a in inc]))
    setup_s = time.perf_counter() - t0
    log(f"C4: {len(ops)} operators ({min(op.n for op in ops)}-{max(op.n for op in ops)} unknowns, "
        f"{min(op.max_level for op in ops) + 1}-{max(op.max_level for op in ops) + 1} levels) in {setup_s:.1f} s")
    kw = dict(restart=50, tol_rel=1e-3, itr_max=5000)
    # warm-up: record every object's graphs once
    for op, B in zip(ops, Bs):
        op.solve_batch(B, itr_max=2, restart=50, tol_rel=1e-3)
    t0 = time.perf_counter()
    seq = [op.solve_batch(B, **kw) for op, B in zip(ops, Bs)]
    seq_s = time.perf_counter() - t0
    log(f"one after another: {seq_s:.2f} s")
    q = queue.Queue()
    for k in range(len(ops)):
        q.put(k)
    thr = [None] * len(ops)

    def worker():
        while True:
            try:
                k = q.get_nowait()
            except queue.Empty:
                return
            thr[k] = ops[k].solve_batch(Bs[k], **kw)
    t0 = time.perf_counter()
    pool = [threading.Thread(target=worker) for _ in range(args.concurrency)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    thr_s = time.perf_counter() - t0
    log(f"{args.concurrency} host threads: {thr_s:.2f} s")
    api.solve_multi(ops, Bs, itr_max=2, **{k: v for k, v in kw.items() if k != "itr_max"})
    t0 = time.perf_counter()
    Xm, stm = api.solve_multi(ops, Bs, **kw)
    multi_s = time.perf_counter() - t0
    log(f"one lockstep solve (a graph per Arnoldi step over all objects): {multi_s:.2f} s")
    same_its = all([s["iterations"] for s in a[1]] == [s["iterations"] for s in b]
                   for a, b in zip(seq, stm))
    max_diff = max(float(np.linalg.norm(xm - a[0]) / np.linalg.norm(a[0])) for xm, a in zip(Xm, seq))
    thr_diff = max(float(np.linalg.norm(t[0] - a[0]) / np.linalg.norm(a[0])) for t, a in zip(thr, seq))
    conv = all(s["converged"] for st in stm for s in st)
    its = [max(s["iterations"] for s in st) for st in stm]
    print(json.dumps({
        "metric": "C4 solve wall s (16 uneven objects x 8 incidences, to 1e-3)", "value": round(multi_s, 3),
        "unit": "s", "n_gpus": 1, "higher_is_better": False, "data": "synthetic", "dtype": "f32+bf16",
        "config": {"workload": "C4: 16 seeded objects (ellipsoids, cubes, cylinders; default_rng(4)), eps_r U[1.5, 6], "
                               "edge lambda/10, 8 incidences each, PMCHWT, GMRES(50) to 1e-3, FP32 far field, split-pair "
                               "BF16 near field",
                   "objects": [{"kind": o["kind"], "eps_r": round(o["eps_r"], 3), "unknowns": op.n,
                                "levels": op.max_level + 1, "max_iterations": it,
                                "sequential_solve_s": round(sq[1][0]["seconds"], 3)}
                               for o, op, it, sq in zip(objs, ops, its, seq)]},
        "setup_s": round(setup_s, 2), "sequential_s": round(seq_s, 3), "threads": args.concurrency,
        "threads_s": round(thr_s, 3), "lockstep_graph_s": round(multi_s, 3),
        "speedup_vs_sequential": round(seq_s / multi_s, 2), "speedup_vs_threads": round(thr_s / multi_s, 2),
        "all_converged": conv, "same_iterations_as_sequential": same_its,
        "max_solution_diff_lockstep_vs_sequential": max_diff, "max_solution_diff_threads_vs_sequential": thr_diff}))
    for op in ops:
        op.close()


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-baseline-child":
        cpu_baseline_child(sys.argv[2:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed 64-RHS mixed matvecs (and C2 FP64 matvecs)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c5", choices=["c5", "c5_matvec", "pec_cfie", "multi_object"])
    ap.add_argument("--copy-reference", action="store_true", help="a plain 1 GiB device read + write first")
    ap.add_argument("--with-precond", action="store_true", help="c5_matvec: two batched polynomial applications")
    ap.add_argument("--seed", type=int, default=1000, help="C5: eps_r = U[2,8] from default_rng(seed)")
    ap.add_argument("--frequency-c5", type=int, default=71, help="C5 geodesic frequency (30 n^2 RWG)")
    ap.add_argument("--inner", type=int, default=10, help="GMRES steps of the inner near-field solve")
    ap.add_argument("--plain-sample", type=int, default=2, help="incidences solved by the plain FP64 method")
    ap.add_argument("--plain-cap", type=int, default=1000, help="iteration cap of the plain FP64 solves")
    ap.add_argument("--fp64-pre-sample", type=int, default=2, help="incidences solved in FP64 with the preconditioner")
    ap.add_argument("--train-runs", type=int, default=200)
    ap.add_argument("--train-scale", type=float, default=0.25)
    ap.add_argument("--train-freq", type=int, default=18)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pec", action="store_true")
    ap.add_argument("--cpu-matvecs", type=int, default=2)
    ap.add_argument("--concurrency", type=int, default=4, help="multi_object: objects processed at once")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the whole operator runs on a single MI355X "
              "(sharding over several GPUs is not launched)", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.copy_reference:
        # a plain 1 GiB device read + write (the fill before it only writes): the
        # reference that scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run
        a = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.add(a, 1.0, out=b)
        torch.cuda.synchronize()
        a = b = None
    if args.workload == "multi_object":
        multi_object(args)
    elif args.workload == "c5_matvec":
        c5_matvec(args)
    elif args.workload == "pec_cfie":
        print(json.dumps({"metric": "MLFMA matvec/s", **pec_cfie_matvec(args)}))
    else:
        c5_headline(args)


if __name__ == "__main__":
    main()
