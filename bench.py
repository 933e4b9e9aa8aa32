#!/usr/bin/env python3
"""Benchmark: matrix-free Neo-Hookean tangent vmult on one MI355X.

Workload (default): the precision / large-size configuration -- unit cube,
128^3 hexes, Q2 elements (3^3 Gauss points), 50,923,779 DoFs, nu = 0.45 with
200 seeded ball inclusions (r = 0.03, 10x stiffness), tensor4 caching, FP64,
x = 0 clamped and u_x prescribed at x = 1, linearised about the seeded smooth
field u_i = 0.02 sin(pi k_i . x).  One step = one operator application
y = A x over the whole mesh (the reference's vmult, timed the same way as its
n_times = 10 loop, include/mf_elasticity.h:2354-2448), with x uniform in
[0,1) and zero on constrained DoFs.  Inputs are resident in HBM before the
timed region.

Prints one JSON line (stdout) with the throughput in GDoF/s, the HBM roofline
of the fused cell kernel (algorithmic bytes of the reference's memory model,
doc/large_strain.tex:739-753, divided by the kernel's event-timed duration)
and the CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "large-strain-matrix-free_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "Q2-3D tangent vmult GDoF/s; Newton-solve s, FP32-GMG vs FP64, same tol"


def progress(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(dim, caching, precision, n_qp, n_nodes):
    """memory model of the reference (doc/large_strain.tex:739-753):
    vector access per node + physics and metric terms per quadrature point"""
    es = 8 if precision == "fp64" else 4
    nv = dim * (dim + 1) // 2
    if caching == "scalar":
        per_qp = 1 + 2 * dim * dim + 1
        per_node = 3 * dim
    elif caching == "tensor2":
        per_qp = 2 + nv + dim * dim + 1
        per_node = 2 * dim
    else:
        per_qp = nv + nv * nv + dim * dim + 1
        per_node = 2 * dim
    return es * (per_qp * n_qp + per_node * n_nodes)


def cpu_baseline(cfg_id, caching, n_sample, seconds):
    """FP64 oracle (OpenMP, sum factorization) on a smaller cube of the same
    configuration; GDoF/s of its vmult."""
    import numpy as np
    import lsmf
    import oracle

    mesh, mu, lam = lsmf.build_config_mesh(cfg_id, n=n_sample)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, lsmf.workload.SEED + cfg_id)
    cons = mesh.constrained_mask({0: list(range(mesh.dim)), 1: [0]})
    op = oracle.OracleOperator(mesh.dim, mesh.degree, mesh.cell_nodes, mesh.coords, cons, caching=caching)
    op.set_material(mu, lam)
    op.set_linearization(u)
    op.cache()
    x = np.random.default_rng(0).random(mesh.n_dofs)
    x[cons == 1] = 0.0
    op.vmult(x)
    n, t0 = 0, time.perf_counter()
    while True:
        op.vmult(x)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 1000:
            break
    return {
        "value": mesh.n_dofs * n / el / 1e9,
        "unit": "GDoF/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": (f"{n_sample}^3-cell Q{mesh.degree} {caching} FP64 vmult of the same configuration "
                   f"({mesh.n_dofs} DoF), {n} vmults in {el:.1f} s, oracle C++ (OpenMP, sum factorization)"),
    }


def newton_solve(cfg_id, n_override, mu, lam, precision, tol, coarse="direct", verbose=False):
    """Load-stepping Newton with FP64 CG preconditioned by a GMG V-cycle whose
    level operators (tangents, Chebyshev smoother, transfer) run in
    `precision`; u_x prescribed at x = 1 as an increment per load step, x = 0
    clamped (include/mf_elasticity.h:2932-3007 semantics)."""
    import numpy as np
    import lsmf

    cfg = lsmf.CONFIGS[cfg_id]
    n = cfg["n"] if n_override is None else n_override
    dim = cfg["dim"]
    n_levels = int(round(np.log2(n // 2))) + 1
    steps = cfg["load_steps"]
    o = lsmf.solver_options(dim=dim, degree=cfg["degree"], n_q_1d=cfg["degree"] + 1, n_cells_coarse=[2] * dim,
                            n_levels=n_levels, caching=cfg["caching"], mg_caching=cfg["caching"],
                            mg_precision=precision, smoother_degree=cfg.get("cheb_degree", 5), tol_lin=tol,
                            tol_f=tol, dt=1.0 / steps, t_end=1.0 + 1e-9, verbose=1 if verbose else 0,
                            coarse_solver=coarse, use_graph=1,
                            dirichlet={0: (list(range(dim)), [0.0] * dim), 1: ([0], [cfg["increment"]])})
    t0 = time.perf_counter()
    S = lsmf.Solver(o, mu, lam)
    t_create = time.perf_counter() - t0
    stats = S.run()
    u = S.solution()
    return {"create_s": t_create, "stats": stats, "u": u, "n_levels": n_levels}


def time_ms(fn, reps, torch):
    """average ms of fn() over reps launches on the current stream (HIP events)"""
    stream = torch.cuda.current_stream()
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def variant_sweep(cfg_id, reps):
    """vmult throughput of every caching variant (and formulation 0) on one
    configuration, FP64 and FP32"""
    import torch
    import lsmf
    mesh, mu, lam = lsmf.build_config_mesh(cfg_id)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, lsmf.workload.SEED + cfg_id)
    cons = mesh.constrained_mask({0: list(range(mesh.dim)), 1: [0]})
    dev = torch.device("cuda:0")
    ud = torch.tensor(u, device=dev)
    out = {"workload": f"config {cfg_id}: {mesh.n_cells_dir[0]}^{mesh.dim} Q{mesh.degree}, {mesh.n_dofs} DoF",
           "gdofs": {}, "cache_bytes": {}}
    cases = [(c, p, 1) for c in ("tensor4", "tensor2", "scalar", "tensor4_ns", "scalar_referential")
             for p in ("fp64", "fp32")] + [("scalar", "fp64", 0), ("tensor4", "fp64", 0)]
    for caching, prec, form in cases:
        op = lsmf.NeoHookOperator(mesh, precision=prec, caching=caching, constrained=cons, formulation=form)
        op.set_material(mu, lam)
        op.set_linearization(ud)
        op.cache()
        dt = torch.float64 if prec == "fp64" else torch.float32
        x = torch.rand(mesh.n_dofs, dtype=dt, device=dev)
        y = torch.empty_like(x)
        ms = time_ms(lambda: op.vmult(y, x), reps, torch)
        key = f"{caching}_{prec}" + ("_formulation0" if form == 0 else "")
        out["gdofs"][key] = round(mesh.n_dofs / ms / 1e6, 3)
        out["cache_bytes"][key] = op.memory_consumption()
        del op, x, y
    torch.cuda.empty_cache()
    return out


def adjoint_trade(cfg_id, reps, caching="tensor4"):
    """residual VJP that recomputes the kinematics (no stored state) against
    the cache a cached backward would keep (the tangent cache of `caching`
    plus b - I and ln J per point): time and bytes"""
    import torch
    import lsmf
    mesh, mu, lam = lsmf.build_config_mesh(cfg_id)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.01, lsmf.workload.SEED + cfg_id)
    dev = torch.device("cuda:0")
    ud = torch.tensor(u, device=dev)
    del u
    op = lsmf.NeoHookOperator(mesh, precision="fp64", caching=caching)
    op.set_material(mu, lam)
    op.set_linearization(ud)
    w = torch.rand(mesh.n_dofs, dtype=torch.float64, device=dev)
    du, r = torch.empty_like(w), torch.empty_like(w)
    dmu = torch.empty(mesh.n_cells, dtype=torch.float64, device=dev)
    dla = torch.empty_like(dmu)
    t_res = time_ms(lambda: op.residual(r), reps, torch)
    t_vjp = time_ms(lambda: op.residual_vjp(w, du, dmu, dla), reps, torch)
    t_cache = time_ms(lambda: op.cache(), max(1, reps // 4), torch)
    t_vmult = time_ms(lambda: op.vmult(r, w, mask="plain"), reps, torch)
    n_qp = mesh.n_cells * (mesh.degree + 1) ** mesh.dim
    nv = mesh.dim * (mesh.dim + 1) // 2
    n_fields = {"tensor4": mesh.dim ** 2 + nv + nv * (nv + 1) // 2, "tensor2": mesh.dim ** 2 + nv + 2}[caching]
    out = {"workload": f"config {cfg_id}: {mesh.n_cells_dir[0]}^{mesh.dim} Q{mesh.degree} log-normal mu/lambda, "
                       f"{mesh.n_dofs} DoF, FP64",
           "residual_ms": round(t_res, 4), "vjp_recompute_ms": round(t_vjp, 4),
           "cached_caching": caching,
           "cache_build_ms": round(t_cache, 4), "cached_vmult_ms": round(t_vmult, 4),
           "vjp_state_bytes": 0, "cache_bytes": n_fields * 8 * n_qp,
           "cached_backward_bytes": (n_fields + nv + 1) * 8 * n_qp,
           "note": f"VJP w.r.t. u, mu, lambda recomputes F, J, ln J per point; the cached alternative "
                   f"({caching} cache + per-point b - I and ln J) would hold {n_fields} + {nv + 1} FP64 values "
                   f"per point and apply the cached tangent for the u part"}
    del op
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--n", type=int, default=None, help="override cells per edge")
    ap.add_argument("--caching", default=None)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-n", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="skip the FP32-GMG vs FP64-GMG Newton solves")
    ap.add_argument("--solve-tol", type=float, default=1e-8)
    ap.add_argument("--coarse", default="direct", choices=["direct", "chebyshev"],
                    help="coarse-level solver of the V-cycle (chebyshev = the reference's choice)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the caching-variant sweep (32^3 Q2) and the adjoint time/memory trade (48^3 Q2)")
    ap.add_argument("--solve-only", default=None, choices=["fp32", "fp64"],
                    help="run only the Newton solve with this level precision and print its statistics "
                         "as one JSON line (no vmult timing; for kernel traces of the solve)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after timing, run 3 plain device copies of the source vector (known bytes) so "
                         "HBM counters (FETCH_SIZE / WRITE_SIZE) can be scaled")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on exactly one GPU: the operator is measured on one MI355X "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import lsmf

    cfg = lsmf.CONFIGS[args.config]
    caching = args.caching or cfg["caching"]
    progress(f"building config {args.config} mesh")
    mesh, mu, lam = lsmf.build_config_mesh(args.config, n=args.n)
    if args.solve_only:
        if "load_steps" not in cfg:
            print(f"config {args.config} has no Newton solve", file=sys.stderr)
            sys.exit(2)
        progress(f"Newton solve with {args.solve_only} GMG levels (tol {args.solve_tol:g})")
        r = newton_solve(args.config, args.n, mu, lam, args.solve_only, args.solve_tol, args.coarse)
        st = r["stats"]
        print(json.dumps({"config": args.config, "level_precision": args.solve_only, "mg_levels": r["n_levels"],
                          "solve_s": round(sum(x["time_total_s"] for x in st), 3),
                          "newton_iterations": [x["newton_iterations"] for x in st],
                          "cg_iterations": [x["linear_iterations"] for x in st]}))
        return
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, lsmf.workload.SEED + args.config)
    cons = mesh.constrained_mask({0: list(range(mesh.dim)), 1: [0]})
    dev = torch.device("cuda:0")
    progress(f"creating operator: {mesh.n_cells} cells, {mesh.n_dofs} DoF, {caching} {args.precision}")
    op = lsmf.NeoHookOperator(mesh, precision=args.precision, caching=caching, constrained=cons)
    op.set_material(mu, lam)
    ud = torch.tensor(u, device=dev)
    del u
    op.set_linearization(ud)
    op.cache()
    dt = torch.float64 if args.precision == "fp64" else torch.float32
    g = torch.Generator(device=dev)
    g.manual_seed(2354)
    x = torch.rand(mesh.n_dofs, dtype=dt, device=dev, generator=g)
    x[torch.tensor(cons, device=dev).bool()] = 0
    y = torch.empty_like(x)
    progress("warmup")
    for _ in range(args.warmup):
        op.vmult(y, x)
    torch.cuda.synchronize()

    progress(f"timing {args.steps} vmults")
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    op.profile_enable(True)
    step_ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    e0.record(stream)
    step_ev[0].record(stream)
    for k in range(args.steps):
        op.vmult(y, x)
        step_ev[k + 1].record(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    total_ms = e0.elapsed_time(e1)
    step_ms = [step_ev[k].elapsed_time(step_ev[k + 1]) for k in range(args.steps)]
    kern_ms, n_launch = op.profile_read()
    op.profile_enable(False)
    ms_per_step = total_ms / args.steps
    if args.calibrate_copy:
        progress(f"copy calibration: 3 x {x.numel() * x.element_size()} bytes read + written")
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()

    nq = mesh.degree + 1
    n_qp = mesh.n_cells * nq ** mesh.dim
    alg_bytes = algorithmic_bytes(mesh.dim, caching, args.precision, n_qp, mesh.n_nodes)
    launches_per_vmult = n_launch / args.steps
    bytes_per_launch = alg_bytes / launches_per_vmult
    avg_launch_s = kern_ms / n_launch / 1e3
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_config{args.config}_{caching}_{args.precision}.json")
    if args.n is None and os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")

    del op, ud, x, y
    torch.cuda.empty_cache()

    solve = None
    if not args.no_solve and "load_steps" in cfg:
        res = {}
        for prec in ("fp32", "fp64"):
            progress(f"Newton solve with {prec} GMG levels (tol {args.solve_tol:g})")
            res[prec] = newton_solve(args.config, args.n, mu, lam, prec, args.solve_tol, args.coarse)
            st = res[prec]["stats"]
            progress(f"  {prec}: {sum(s['time_total_s'] for s in st):.2f} s, Newton {[s['newton_iterations'] for s in st]}"
                     f", CG {[s['linear_iterations'] for s in st]}")
        t32 = sum(s["time_total_s"] for s in res["fp32"]["stats"])
        t64 = sum(s["time_total_s"] for s in res["fp64"]["stats"])
        u32, u64 = res["fp32"]["u"], res["fp64"]["u"]
        solve = {
            "fp32_gmg_s": round(t32, 3), "fp64_gmg_s": round(t64, 3), "speedup_fp32_vs_fp64": round(t64 / t32, 3),
            "tol": args.solve_tol, "load_steps": len(res["fp32"]["stats"]), "mg_levels": res["fp32"]["n_levels"],
            "coarse_solver": args.coarse,
            "smoother": f"Chebyshev degree {cfg.get('cheb_degree', 5)}, range 20, 30 eigen-CG iterations",
            "newton_iterations": {p: [s["newton_iterations"] for s in res[p]["stats"]] for p in res},
            "cg_iterations": {p: [s["linear_iterations"] for s in res[p]["stats"]] for p in res},
            "final_rel_residual": {p: [s["residual_final"] / s["residual_0"] for s in res[p]["stats"]] for p in res},
            "setup_s": {p: round(sum(s["time_setup_s"] for s in res[p]["stats"]), 3) for p in res},
            "setup_caches_diagonals_s": {p: round(sum(s["time_setup_cache_s"] for s in res[p]["stats"]), 3)
                                         for p in res},
            "linear_s": {p: round(sum(s["time_linear_s"] for s in res[p]["stats"]), 3) for p in res},
            "rel_diff_u_fp32_vs_fp64": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
            "timed": "Newton loop of all load steps (residual, caches, level setup, CG); hierarchy creation excluded",
        }
        del res

    variants = adjoint = None
    if not args.no_extras and args.n is None:
        progress("caching-variant sweep on config 1 (32^3 Q2)")
        variants = variant_sweep(1, 20)
        progress(f"  GDoF/s {variants['gdofs']}")
        if args.config == 4:
            progress("adjoint time/memory trade on config 4 (64^3 Q4, tensor2)")
            adjoint = adjoint_trade(4, 10, "tensor2")
        else:
            progress("adjoint time/memory trade on config 3 (48^3 Q2)")
            adjoint = adjoint_trade(3, 20)
        progress(f"  {adjoint}")

    cpu = None
    if not args.no_cpu:
        progress(f"CPU baseline: oracle on {args.cpu_n}^3 sample for ~{args.cpu_seconds:.0f} s")
        cpu = cpu_baseline(args.config, caching, args.cpu_n, args.cpu_seconds)

    workload = (f"config {args.config}: {mesh.n_cells_dir[0]}^{mesh.dim} Q{mesh.degree} {caching} "
                f"{args.precision} vmult, {mesh.n_dofs} DoF")
    line = {
        "metric": METRIC,
        "value": round(mesh.n_dofs / (ms_per_step / 1e3) / 1e9, 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "ms_per_step_min": round(min(step_ms), 4) if step_ms else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": workload, "dim": mesh.dim, "degree": mesh.degree, "n_q_1d": nq,
                   "cells": mesh.n_cells, "dofs": mesh.n_dofs, "qp": n_qp, "caching": caching,
                   "precision": args.precision, "nu": cfg["nu"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_vmult (fused gather/SF/tangent/integrate/scatter, one launch per colour)",
                     "kernel_avg_ms": round(avg_launch_s * 1e3, 5), "launches_per_step": launches_per_vmult,
                     "algorithmic_bytes_per_launch": int(bytes_per_launch)},
        "cpu_baseline": cpu,
        "newton_solve": solve,
        "caching_variants": variants,
        "adjoint": adjoint,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
