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
and a CPU baseline timed on the host.  Beside the headline it reports
  * the other caching variants and precisions on the same 128^3 mesh, each
    against its roofline (scalar also against the FP64 VALU peak),
  * renumbered meshes: the cells layout on the lexicographic numbering, and
    the cells layout and the tiles the operator recovers from the cells'
    face neighbours on a Cuthill-McKee numbering (the reference's DoF
    renumbering) and on a random one,
  * the FP32-V-cycle against the all-FP64 Newton solve (3 load steps, CG to
    1e-8) on the 128^3 Q2 and the 64^3 Q4 meshes, and the 32^3 Q2 5-step
    solve,
  * the 64^3 Q4 tensor2 configuration's vmult and roofline, its residual VJP
    and the misfit gradient through its Newton solve,
  * the 48^3 Q2 differentiation workload: residual and its VJP, and the
    gradient of the x = 1 face misfit through a Newton solve,
  * hanging-node constraints on a half-refined cube (vmult and diagonal with
    and without the constraints).
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "large-strain-matrix-free_amd"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "Q2-3D tangent vmult GDoF/s; Newton-solve s, FP32-GMG vs FP64, same tol"
KERNEL_SOURCE = os.path.join(ROOT, "large-strain-matrix-free_amd", "csrc", "lsmf_kernels.hpp")


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


def kernel_source_sha():
    return hashlib.sha256(open(KERNEL_SOURCE, "rb").read()).hexdigest()[:16]


def measured_traffic(cfg_id, caching, precision):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this kernel, if it was measured on the current
    kernel source (else None: the figure would be stale)"""
    f = os.path.join(ROOT, "profiles", f"traffic_config{cfg_id}_{caching}_{precision}.json")
    if not os.path.exists(f):
        return None, "not measured"
    d = json.load(open(f))
    if d.get("kernel_source_sha") != kernel_source_sha():
        return None, f"stale: {os.path.basename(f)} was measured on another kernel source"
    return d.get("hbm_bytes_per_launch"), os.path.relpath(f, ROOT)


def physical_cores():
    """physical cores of the host (lscpu: unique core,socket pairs)"""
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=30).stdout
        pairs = {l for l in out.splitlines() if l and not l.startswith("#")}
        if pairs:
            return len(pairs)
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count() or 1


def cpu_baseline(cfg_id, caching, n_sample, seconds):
    """the reference's MatrixFree algorithm on the host: cell batches as SIMD
    lanes (VectorizedArray), sum factorization, the same caches, OpenMP over
    the colour-ordered batches (oracle/cpu_baseline.cpp, a CPU restatement
    checked against the oracle), on a smaller cube of the same
    configuration; GDoF/s of its vmult, mean of the timed products"""
    import numpy as np
    import lsmf
    from oracle import cpu_baseline as cb

    mesh, mu, lam = lsmf.build_config_mesh(cfg_id, n=n_sample)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, lsmf.workload.SEED + cfg_id)
    cons = mesh.constrained_mask({0: list(range(mesh.dim)), 1: [0]})
    cores = physical_cores()
    op = cb.BatchOperator(mesh, cons, caching=caching, threads=cores)
    op.set_material(mu, lam)
    op.set_linearization(u)
    op.cache()
    x = np.random.default_rng(0).random(mesh.n_dofs)
    x[cons == 1] = 0.0
    y = np.empty_like(x)
    op.vmult(y, x)
    n, t0 = 0, time.perf_counter()
    while True:
        op.vmult(y, x)
        n += 1
        el = time.perf_counter() - t0
        if (el >= seconds and n >= 10) or n >= 1000:
            break
    return {
        "value": mesh.n_dofs * n / el / 1e9,
        "unit": "GDoF/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{n_sample}^3-cell Q{mesh.degree} {caching} FP64 vmult of the same configuration "
                   f"({mesh.n_dofs} DoF; GDoF/s is size-independent once the working set leaves the caches), "
                   f"mean of {n} vmults in {el:.1f} s after one warm-up; reference-algorithm CPU restatement: "
                   f"{op.lanes}-wide cell batches (VectorizedArray), sum factorization, {caching} cache, "
                   f"OpenMP on {cores} physical cores (lscpu)"),
        "published_context": "reference on 20 Cascade Lake cores: tensor4 0.182, tensor2 0.327 GDoF/s (Q2, "
                             "Calculations/CSL_Munich/holes_3d_p2q3r3_MF_CG_gmg_*/timings.txt:51)",
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


def solve_pair(cfg_id, n_override, mu, lam, tol, coarse):
    """FP32-V-cycle against all-FP64 Newton solve of one configuration"""
    import numpy as np
    import lsmf
    cfg = lsmf.CONFIGS[cfg_id]
    res = {}
    for prec in ("fp32", "fp64"):
        progress(f"config {cfg_id} Newton solve with {prec} GMG levels (tol {tol:g})")
        res[prec] = newton_solve(cfg_id, n_override, mu, lam, prec, tol, coarse)
        st = res[prec]["stats"]
        progress(f"  {prec}: {sum(s['time_total_s'] for s in st):.2f} s, Newton {[s['newton_iterations'] for s in st]}"
                 f", CG {[s['linear_iterations'] for s in st]}")
    t32 = sum(s["time_total_s"] for s in res["fp32"]["stats"])
    t64 = sum(s["time_total_s"] for s in res["fp64"]["stats"])
    u32, u64 = res["fp32"]["u"], res["fp64"]["u"]
    return {
        "fp32_gmg_s": round(t32, 3), "fp64_gmg_s": round(t64, 3), "speedup_fp32_vs_fp64": round(t64 / t32, 3),
        "tol": tol, "load_steps": len(res["fp32"]["stats"]), "mg_levels": res["fp32"]["n_levels"],
        "coarse_solver": coarse,
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


def timed_vmults(op, x, y, steps, warmup, torch):
    """steps products y = A x between events, each product also bracketed
    (mean and fastest), and the fused cell kernels' event-timed launches"""
    for _ in range(warmup):
        op.vmult(y, x)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    op.profile_enable(True)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    ev[0].record(stream)
    for k in range(steps):
        op.vmult(y, x)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = [ev[k].elapsed_time(ev[k + 1]) for k in range(steps)]
    kern_ms, n_launch = op.profile_read()
    op.profile_enable(False)
    return ev[0].elapsed_time(ev[steps]) / steps, min(step_ms), kern_ms, n_launch


def vmult_record(mesh, mu, lam, ud, cons, caching, precision, steps, warmup, layout="auto", cfg_id=None,
                 flops_per_dof=None, node_constraints=None):
    """one operator's vmult throughput and roofline on `mesh`"""
    import torch
    import lsmf
    dev = ud.device
    op = lsmf.NeoHookOperator(mesh, precision=precision, caching=caching, constrained=cons, layout=layout,
                              use_mesh_colouring=layout == "auto" and getattr(mesh, "cell_colour", None) is not None,
                              node_constraints=node_constraints)
    op.set_material(mu, lam)
    op.set_linearization(ud)
    op.cache()
    dt = torch.float64 if precision == "fp64" else torch.float32
    g = torch.Generator(device=dev)
    g.manual_seed(2354)
    x = torch.rand(mesh.n_dofs, dtype=dt, device=dev, generator=g)
    x[torch.tensor(cons, device=dev).bool()] = 0
    y = torch.empty_like(x)
    ms, ms_min, kern_ms, n_launch = timed_vmults(op, x, y, steps, warmup, torch)
    nq = mesh.degree + 1
    n_qp = mesh.n_cells * nq ** mesh.dim
    alg = algorithmic_bytes(mesh.dim, caching, precision, n_qp, mesh.n_nodes)
    launches = n_launch / steps
    avg_s = kern_ms / n_launch / 1e3
    achieved = alg / launches / avg_s / 1e9
    rec = {"caching": caching, "precision": precision, "layout": layout, "ms": round(ms, 4),
           "ms_min": round(ms_min, 4), "gdofs": round(mesh.n_dofs / ms / 1e6, 3),
           "kernel_avg_ms": round(avg_s * 1e3, 5), "launches_per_vmult": launches,
           "algorithmic_bytes_per_vmult": int(alg),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4)},
           "cache_bytes": op.memory_consumption()}
    if cfg_id is not None:
        tr, src = measured_traffic(cfg_id, caching, precision)
        rec["roofline"]["traffic"] = tr
        rec["roofline"]["traffic_source"] = src
    if flops_per_dof is not None:
        tf = flops_per_dof * mesh.n_dofs / (kern_ms / steps / 1e3) / 1e12
        rec["compute_roofline"] = {"bound": "fp64 valu", "achieved": round(tf, 2), "peak": FP64_PEAK_TFLOPS,
                                   "unit": "TFLOP/s", "frac": round(tf / FP64_PEAK_TFLOPS, 4),
                                   "flops_per_dof": flops_per_dof}
    del op, x, y
    torch.cuda.empty_cache()
    return rec


def hanging_record(caching, precision, n=64):
    """vmult throughput with hanging-node constraints: the unit cube with
    its half x > 1/2 refined once (lsmf.half_refined_box, Q2, n^3 coarse-cell
    resolution), x = 0 clamped, against the same cells without the
    constraints (the cells layout either way: the mesh is not one grid)"""
    import numpy as np
    import torch
    import lsmf
    mesh, nc = lsmf.half_refined_box(2, n)
    x = mesh.coords
    cons = np.zeros(mesh.n_dofs, np.uint8)
    left = np.where(np.abs(x[:, 0]) < 1e-12)[0]
    for c in range(3):
        cons[left * 3 + c] = 1
    rng = np.random.default_rng(lsmf.workload.SEED)
    mu, lam = 1.0 + rng.random(mesh.n_cells), 1.5 + rng.random(mesh.n_cells)
    # a displacement continuous across the interface: hanging values from their masters
    u = (0.02 * np.sin(np.pi * (x @ np.arange(1, 4)))[:, None] * np.ones(3))
    nodes, ptr, masters, w = nc
    for k, h in enumerate(nodes):
        u[h] = w[ptr[k]:ptr[k + 1]] @ u[masters[ptr[k]:ptr[k + 1]]]
    ud = torch.tensor(u.reshape(-1), device="cuda:0")
    out = {"workload": f"unit cube, half x > 1/2 refined once: {mesh.n_cells} Q2 cells, {mesh.n_dofs} DoF, "
                       f"{len(nodes)} hanging nodes (3 or 9 masters each on the coarse faces), {caching} {precision}, "
                       f"cells layout"}
    out["with_constraints"] = vmult_record(mesh, mu, lam, ud, cons, caching, precision, 10, 2, layout="cells",
                                           node_constraints=nc)
    out["without_constraints"] = vmult_record(mesh, mu, lam, ud, cons, caching, precision, 10, 2, layout="cells")
    out["overhead"] = round(out["with_constraints"]["ms"] / out["without_constraints"]["ms"] - 1.0, 4)
    # compute_diagonal (closed form; cells with constrained nodes per coupled node)
    for key, ncs in (("diagonal_ms_with_constraints", nc), ("diagonal_ms_without_constraints", None)):
        op = lsmf.NeoHookOperator(mesh, precision=precision, caching=caching, constrained=cons, layout="cells",
                                  use_mesh_colouring=False, node_constraints=ncs)
        op.set_material(mu, lam)
        op.set_linearization(ud)
        op.cache()
        out[key] = round(time_ms(op.compute_diagonal, 5, torch), 4)
        del op
        torch.cuda.empty_cache()
    del ud
    torch.cuda.empty_cache()
    return out


def permuted_mesh(mesh, seed):
    """the same mesh with its nodes and its cells in a seeded random order
    (no lexicographic numbering left)"""
    import numpy as np
    import lsmf
    rng = np.random.default_rng(seed)
    perm = rng.permutation(mesh.n_nodes).astype(np.int32)
    order = rng.permutation(mesh.n_cells)
    return lsmf.renumber_mesh(mesh, perm, order), perm, order


def variant_sweep(cfg_id, reps):
    """vmult throughput of every caching variant (and formulation 0) on the
    32^3 configuration, FP64 and FP32 (0.8 M DoF: launch-bound and largely
    in the Infinity Cache, so no roofline is claimed)"""
    import torch
    import lsmf
    mesh, mu, lam = lsmf.build_config_mesh(cfg_id)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, lsmf.workload.SEED + cfg_id)
    cons = mesh.constrained_mask({0: list(range(mesh.dim)), 1: [0]})
    dev = torch.device("cuda:0")
    ud = torch.tensor(u, device=dev)
    out = {"workload": f"config {cfg_id}: {mesh.n_cells_dir[0]}^{mesh.dim} Q{mesh.degree}, {mesh.n_dofs} DoF",
           "note": "launch-bound and cache-resident: no roofline", "gdofs": {}}
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


def misfit_gradient_timing(cfg_id=3):
    """the differentiation workload at full size: the x = 1 face
    displacement misfit J against the solution for a second seeded
    log-normal field; dJ/dmu, dJ/dlambda through the Newton solve (one
    adjoint GMG-CG solve on the converged tangent and one residual VJP)"""
    import numpy as np
    import torch
    import lsmf
    from lsmf.autograd import misfit_gradient
    cfg = lsmf.CONFIGS[cfg_id]
    n = cfg["n"]
    mesh = lsmf.Mesh(3, cfg["degree"], [n] * 3)
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    nc = 3 if n % 3 == 0 else 2  # coarse cells per edge: 48 = 3 * 2^4, 64 = 2 * 2^5
    n_levels = int(round(np.log2(n // nc))) + 1
    o = lsmf.solver_options(dim=3, degree=cfg["degree"], n_q_1d=cfg["degree"] + 1, n_cells_coarse=[nc] * 3,
                            n_levels=n_levels, caching=cfg["caching"], mg_caching=cfg["caching"], mg_precision="fp32",
                            coarse_solver="direct", tol_lin=1e-10, tol_f=1e-10, tol_u=1e-10, dt=1.0, t_end=1.0,
                            dirichlet={0: ([0, 1, 2], [0.0] * 3), 1: ([0], [0.01])})
    mu_t, lam_t = lsmf.lognormal_material(mesh.n_cells, 1.0, cfg["nu"], cfg["sigma"], lsmf.workload.SEED + 31)
    St = lsmf.Solver(o, mu_t, lam_t)
    St.run()
    target = St.solution()
    del St
    mu, lam = lsmf.lognormal_material(mesh.n_cells, 1.0, cfg["nu"], cfg["sigma"], lsmf.workload.SEED + cfg_id)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S = lsmf.Solver(o, mu, lam)
    stats = S.run()
    u = S.solution()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    J, gmu, gla = misfit_gradient(S, mesh, u, mu, lam, target, mesh.boundary_nodes(1), cons, cg_tol=1e-10)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {"workload": f"{n}^3 Q{cfg['degree']} ({mesh.n_dofs} DoF), {cfg['caching']}, log-normal mu/lambda sigma "
                        f"{cfg['sigma']}, u_x = 0.01 at x = 1 in one load step, FP32-level GMG, CG to 1e-10",
            "forward_solve_s": round(t1 - t0, 3), "gradient_s": round(t2 - t1, 3),
            "newton_iterations": [s["newton_iterations"] for s in stats],
            "J": J, "grad_norm_mu": float(np.linalg.norm(gmu)), "grad_norm_lambda": float(np.linalg.norm(gla)),
            "timed": "forward: solver creation + Newton; gradient: adjoint CG on the converged tangent + VJP"}


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
    ap.add_argument("--cpu-n", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="skip the FP32-GMG vs FP64-GMG Newton solves")
    ap.add_argument("--solve-tol", type=float, default=1e-8)
    ap.add_argument("--coarse", default="direct", choices=["direct", "chebyshev"],
                    help="coarse-level solver of the V-cycle (chebyshev = the reference's choice)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip everything beside the headline vmult (and the solve unless --no-solve)")
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
    progress(f"timing {args.steps} vmults")
    ms_per_step, ms_min, kern_ms, n_launch = timed_vmults(op, x, y, args.steps, args.warmup, torch)
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
    traffic, traffic_src = (None, "not measured")
    if args.n is None:
        traffic, traffic_src = measured_traffic(args.config, caching, args.precision)
    del op, x, y
    torch.cuda.empty_cache()

    extras = not args.no_extras and args.n is None
    variants_128 = cells = None
    if extras:
        variants_128 = {"workload": f"config {args.config}: {mesh.n_cells_dir[0]}^{mesh.dim} Q{mesh.degree}, "
                                    f"{mesh.n_dofs} DoF, tiled layout", "records": []}
        for c2, p2 in (("tensor2", "fp64"), ("tensor4", "fp32"), ("tensor2", "fp32"), ("scalar", "fp64")):
            if (c2, p2) == (caching, args.precision):
                continue
            progress(f"variant {c2} {p2} on the config {args.config} mesh")
            flops = SCALAR_FLOPS_PER_DOF.get((mesh.dim, mesh.degree)) if c2 == "scalar" else None
            rec = vmult_record(mesh, mu, lam, ud, cons, c2, p2, 10, 2, cfg_id=args.config, flops_per_dof=flops)
            progress(f"  {rec['ms']} ms, frac {rec['roofline']['frac']}")
            variants_128["records"].append(rec)
        progress("renumbered meshes: the cells layout and tiles recovered from the face topology")
        cells = {"workload": f"config {args.config} mesh, {caching} {args.precision}; cells layout = one cell per "
                             f"lane group, greedy cell colouring; auto = tiles recovered from the cells' face "
                             f"neighbours whatever the numbering", "records": []}
        rec = vmult_record(mesh, mu, lam, ud, cons, caching, args.precision, 10, 2, layout="cells")
        rec["numbering"] = "lexicographic"
        cells["records"].append(rec)
        pm, perm, order = permuted_mesh(mesh, 7)
        pcons = np.empty_like(cons)
        pcons.reshape(-1, mesh.dim)[perm] = cons.reshape(-1, mesh.dim)
        upd = torch.empty_like(ud)
        upd.view(-1, mesh.dim)[torch.tensor(perm, device=dev, dtype=torch.int64)] = ud.view(-1, mesh.dim)
        for layout, steps in (("cells", 5), ("auto", 10)):
            rec = vmult_record(pm, mu[order], lam[order], upd, pcons, caching, args.precision, steps, 1,
                               layout=layout)
            rec["numbering"] = "random permutation of the nodes and of the cells (seed 7)"
            cells["records"].append(rec)
        del pm, perm, pcons, upd
        if mesh.dim == 3:
            pm, perm, order = lsmf.cuthill_mckee_numbering(mesh)
            pcons = np.empty_like(cons)
            pcons.reshape(-1, mesh.dim)[perm] = cons.reshape(-1, mesh.dim)
            upd = torch.empty_like(ud)
            upd.view(-1, mesh.dim)[torch.tensor(perm, device=dev, dtype=torch.int64)] = ud.view(-1, mesh.dim)
            for layout in ("cells", "auto"):
                rec = vmult_record(pm, mu[order], lam[order], upd, pcons, caching, args.precision, 10, 2,
                                   layout=layout)
                rec["numbering"] = ("Cuthill-McKee nodes (reverse Cuthill-McKee of the grid-line graph, read "
                                    "backwards), Morton-ordered cells")
                cells["records"].append(rec)
            del pm, perm, pcons, upd
        progress(f"  {[r['ms'] for r in cells['records']]} ms")
    del ud
    torch.cuda.empty_cache()

    solve = None
    if not args.no_solve and "load_steps" in cfg:
        solve = solve_pair(args.config, args.n, mu, lam, args.solve_tol, args.coarse)

    config4 = variants = adjoint = gradient = hanging = None
    if extras:
        if args.config != 4:
            progress("config 4: 64^3 Q4 tensor2")
            m4, mu4, lam4 = lsmf.build_config_mesh(4)
            u4, _ = lsmf.smooth_displacement(m4.coords, 0.02, lsmf.workload.SEED + 4)
            cons4 = m4.constrained_mask({0: [0, 1, 2], 1: [0]})
            ud4 = torch.tensor(u4, device=dev)
            del u4
            config4 = {"workload": f"config 4: 64^3 Q4 (5^3 Gauss points), {m4.n_dofs} DoF, nu 0.45, log-normal "
                                   f"mu/lambda sigma 0.3",
                       "vmult": vmult_record(m4, mu4, lam4, ud4, cons4, "tensor2", "fp64", 20, 3, cfg_id=4),
                       "vmult_fp32": vmult_record(m4, mu4, lam4, ud4, cons4, "tensor2", "fp32", 10, 2, cfg_id=4)}
            progress(f"  vmult {config4['vmult']['ms']} ms, frac {config4['vmult']['roofline']['frac']}")
            del ud4
            torch.cuda.empty_cache()
            if not args.no_solve:
                config4["newton_solve"] = solve_pair(4, None, mu4, lam4, args.solve_tol, args.coarse)
            progress("config 4 adjoint time/memory trade (tensor2)")
            config4["adjoint"] = adjoint_trade(4, 10, "tensor2")
            progress("config 4: the misfit gradient through a Newton solve (64^3 Q4)")
            config4["misfit_gradient"] = misfit_gradient_timing(4)
            progress(f"  {config4['misfit_gradient']}")
            del m4, mu4, lam4, cons4
        progress("hanging-node constraints: half-refined cube, Q2 tensor4 FP64")
        hanging = hanging_record("tensor4", "fp64")
        progress(f"  {hanging['with_constraints']['ms']} ms with, {hanging['without_constraints']['ms']} without")
        progress("config 3: residual VJP trade and the misfit gradient through a Newton solve (48^3 Q2)")
        adjoint = adjoint_trade(3, 20)
        gradient = misfit_gradient_timing(3)
        progress(f"  gradient {gradient}")
        progress("caching-variant sweep on config 1 (32^3 Q2)")
        variants = variant_sweep(1, 20)
        if not args.no_solve:
            progress("config 1: 5-load-step Newton solve (32^3 Q2, tensor4)")
            m1, mu1, lam1 = lsmf.build_config_mesh(1)
            variants["newton_solve"] = solve_pair(1, None, mu1, lam1, args.solve_tol, args.coarse)
            del m1

    cpu = None
    if not args.no_cpu:
        progress(f"CPU baseline: {args.cpu_n}^3 sample for ~{args.cpu_seconds:.0f} s")
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
        "ms_per_step_min": round(ms_min, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": workload, "dim": mesh.dim, "degree": mesh.degree, "n_q_1d": nq,
                   "cells": mesh.n_cells, "dofs": mesh.n_dofs, "qp": n_qp, "caching": caching,
                   "precision": args.precision, "nu": cfg["nu"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_vmult (fused gather/SF/tangent/integrate/scatter, one launch per colour)",
                     "kernel_avg_ms": round(avg_launch_s * 1e3, 5), "launches_per_step": launches_per_vmult,
                     "algorithmic_bytes_per_launch": int(bytes_per_launch)},
        "cpu_baseline": cpu,
        "newton_solve": solve,
        "variants_same_mesh": variants_128,
        "cells_layout": cells,
        "config4": config4,
        "hanging_nodes": hanging,
        "config3_adjoint": adjoint,
        "config3_misfit_gradient": gradient,
        "caching_variants_32": variants,
    }
    print(json.dumps(line), flush=True)


# FP64 floating-point operations per DoF of one scalar-variant vmult
# (the kernel's FP64 VALU adds, multiplies and fused multiply-adds counted
# by rocprofv3 SQ_INSTS_VALU_{ADD,MUL,FMA}_F64 (FMA = 2), divided by the DoFs):
# profiles/flops_scalar_fp64.json
SCALAR_FLOPS_PER_DOF = {}
_sf = os.path.join(ROOT, "profiles", "flops_scalar_fp64.json")
if os.path.exists(_sf):
    for k, v in json.load(open(_sf)).get("flops_per_dof", {}).items():
        d, p = k.split("_")
        SCALAR_FLOPS_PER_DOF[(int(d), int(p))] = v


if __name__ == "__main__":
    main()
