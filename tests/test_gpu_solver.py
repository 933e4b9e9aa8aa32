"""GPU solver stack: multigrid transfer, V-cycle, PCG and the Newton driver.

  * transfer: prolongation reproduces Q_p fields exactly, restriction is its
    transpose, interpolating a prolongated field returns it (R P = I)
  * the V-cycle is a symmetric positive definite preconditioner
  * GPU Newton with GMG on the reference's tests/cook_mf_gmg_form_1.prm
    (2x2 coarse mesh refined 4 times, Q1, GMG) reproduces the blessed output
    (Newton counts, corner displacement to 6 digits), with FP32 and FP64
    level operators
  * the 2D unit-square configuration (Q2, 5 levels, Chebyshev degree 4, 5
    prescribed-displacement steps) agrees with the oracle's Newton solve
"""
import json
import os

import numpy as np
import pytest

import lsmf
import oracle

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_outputs.json")))


def log(msg):
    print(msg, flush=True)


def tensor_poly(coords, p):
    """vector field with components in Q_p (degree <= p per direction)"""
    x = coords
    dim = x.shape[1]
    f = np.zeros_like(x)
    for c in range(dim):
        v = np.ones(len(x))
        for d in range(dim):
            v *= (0.3 + c + d) + 0.7 * x[:, d] ** p - 0.2 * x[:, d]
        f[:, c] = v
    return f.reshape(-1)


@pytest.mark.parametrize("dim,p,nc", [(2, 1, (2, 3)), (2, 2, (2, 3)), (3, 2, (2, 3, 2)), (3, 4, (2, 3, 2)),
                                      (3, 2, (350, 1, 1))])
def test_transfer_exactness(dim, p, nc, device):
    """prolongation reproduces Q_p fields, interpolation inverts it and
    restriction is its transpose, on every level; the 1400x4x4 mesh's
    x-lines (2,801 nodes, 67 KB in FP64) exceed what the x pass stages in
    LDS and take its global-memory form"""
    import torch
    log(f"[transfer] dim={dim} p={p} nc={nc}")
    nc = list(nc)
    o = lsmf.solver_options(dim=dim, degree=p, n_q_1d=p + 1, n_cells_coarse=nc, n_levels=3, caching="tensor4",
                            mg_caching="tensor4", mg_precision="fp64")
    mesh = lsmf.Mesh(dim, p, [k * 4 for k in nc])
    S = lsmf.Solver(o, np.ones(mesh.n_cells), np.ones(mesh.n_cells))
    for level in (1, 2):
        fine = lsmf.Mesh(dim, p, [k << level for k in nc])
        coarse = lsmf.Mesh(dim, p, [k << (level - 1) for k in nc])
        uc = torch.tensor(tensor_poly(coarse.coords, p), device=device)
        uf = torch.zeros(fine.n_dofs, dtype=torch.float64, device=device)
        S.transfer(level, 0, uf, uc)
        assert np.allclose(uf.cpu().numpy(), tensor_poly(fine.coords, p), atol=1e-12)
        # interpolation of the prolongated field returns the coarse values
        back = torch.zeros(coarse.n_dofs, dtype=torch.float64, device=device)
        S.transfer(level, 2, back, uf)
        assert np.allclose(back.cpu().numpy(), uc.cpu().numpy(), atol=1e-12)
        # restriction = transpose of prolongation
        a = torch.rand(coarse.n_dofs, dtype=torch.float64, device=device)
        b = torch.rand(fine.n_dofs, dtype=torch.float64, device=device)
        Pa = torch.zeros_like(b)
        S.transfer(level, 0, Pa, a)
        PTb = torch.zeros_like(a)
        S.transfer(level, 1, PTb, b)
        assert abs(float(Pa @ b) - float(a @ PTb)) < 1e-12 * abs(float(Pa @ b))


def small_problem(dim, p, nc, n_levels, prec, caching="tensor4", seed=3):
    n = [k << (n_levels - 1) for k in nc]
    mesh = lsmf.Mesh(dim, p, n)
    mu, lam, _ = lsmf.inclusion_material(mesh.cell_centres(), dim, 1.0, 0.3, 4, 0.2, 10.0, seed)
    o = lsmf.solver_options(dim=dim, degree=p, n_q_1d=p + 1, n_cells_coarse=nc, n_levels=n_levels, caching=caching,
                            mg_caching=caching, mg_precision=prec,
                            dirichlet={0: (list(range(dim)), [0.0] * dim), 1: ([0], [0.01])})
    return mesh, mu, lam, o


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_vcycle_is_spd_and_cg_converges(prec, device):
    import torch
    log(f"[vcycle] {prec}")
    mesh, mu, lam, o = small_problem(3, 2, [2, 2, 2], 3, prec)
    S = lsmf.Solver(o, mu, lam)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.002, 1)
    S.set_state(u)
    for l in range(3):
        lmax, lmin, deg = S.level_info(l)
        assert lmax > lmin > 0 and deg >= 1
        log(f"  level {l}: lambda_max {lmax:.4f} lambda_min {lmin:.4e} degree {deg}")
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    mask = torch.tensor(1 - cons, dtype=torch.float64, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(1)
    r = torch.rand(mesh.n_dofs, dtype=torch.float64, device=device, generator=g) * mask
    w = torch.rand(mesh.n_dofs, dtype=torch.float64, device=device, generator=g) * mask
    Mr, Mw = torch.zeros_like(r), torch.zeros_like(r)
    S.precondition(Mr, r)
    S.precondition(Mw, w)
    tol = 1e-10 if prec == "fp64" else 1e-5
    assert abs(float(w @ Mr) - float(r @ Mw)) < tol * abs(float(w @ Mr))
    assert float(r @ Mr) > 0
    assert float((Mr * (1 - mask)).abs().max()) == 0.0
    x = torch.zeros_like(r)
    its, res = S.cg(x, r, 1e-8 * float(r.norm()))
    log(f"  PCG its {its} residual {res:.3e}")
    assert its < 30


GMG_CASES = [(2, 2, [2, 2], 3, "tensor4", "chebyshev"), (3, 2, [2, 2, 2], 2, "tensor4", "chebyshev"),
             (3, 4, [2, 2, 2], 2, "tensor2", "chebyshev"), (3, 4, [1, 1, 1], 3, "tensor2", "direct")]


@pytest.mark.parametrize("dim,p,nc,L,caching,coarse", GMG_CASES,
                         ids=[f"{d}d-Q{p}-L{L}-{c}-{cs}" for d, p, _, L, c, cs in GMG_CASES])
def test_vcycle_and_pcg_match_oracle_gmg(dim, p, nc, L, caching, coarse, device):
    """GPU V-cycle (FP64 levels) = the oracle's restated V-cycle; PCG
    iteration counts within +-1 (FP64 levels) and +-2 (FP32 levels).  The
    Q4 tensor2 cases are the 64^3 Q4 configuration's multigrid (its
    prolongation / restriction bands, level tangents and eigenvalue
    estimates) on small hierarchies"""
    import torch
    from oracle.gmg import GMG
    log(f"[gmg parity] dim={dim} Q{p} levels={L} {caching} coarse={coarse}")
    meshes = [lsmf.Mesh(dim, p, [k << l for k in nc]) for l in range(L)]
    cons = [m.constrained_mask({0: list(range(dim)), 1: [0]}) for m in meshes]
    fm = meshes[-1]
    mu, lam, _ = lsmf.inclusion_material(fm.cell_centres(), dim, 1.0, 0.3, 4, 0.2, 10.0, 3)
    u, _ = lsmf.smooth_displacement(fm.coords, 0.002, 1)
    ref = GMG(meshes, cons, mu, lam, caching=caching, coarse_solver=coarse)
    ref.setup(u)
    r = np.random.default_rng(1).random(fm.n_dofs)
    r[cons[-1] == 1] = 0
    z_ref = ref.precondition(r)
    _, its_ref, _ = ref.pcg(ref.levels[-1].op, r, 1e-8 * np.linalg.norm(r))
    for prec, z_tol, it_tol in [("fp64", 1e-8, 1), ("fp32", 1e-3, 2)]:
        o = lsmf.solver_options(dim=dim, degree=p, n_q_1d=p + 1, n_cells_coarse=nc, n_levels=L, caching=caching,
                                mg_caching=caching, mg_precision=prec, coarse_solver=coarse,
                                dirichlet={0: (list(range(dim)), [0.0] * dim), 1: ([0], [0.0])})
        S = lsmf.Solver(o, mu, lam)
        S.set_state(u)
        for l in range(1 if coarse == "direct" else 0, L):
            lmax, lmin, deg = S.level_info(l)
            assert abs(lmax - ref.levels[l].lmax_est) < 1e-6 * lmax, (prec, l)
            if prec == "fp64":
                assert deg == ref.levels[l].degree
        rd = torch.tensor(r, device=device)
        z = torch.zeros_like(rd)
        S.precondition(z, rd)
        err = float(np.abs(z.cpu().numpy() - z_ref).max() / np.abs(z_ref).max())
        log(f"  {prec}: V-cycle rel diff {err:.2e}")
        assert err < z_tol
        x = torch.zeros_like(rd)
        its, _ = S.cg(x, rd, 1e-8 * float(np.linalg.norm(r)))
        log(f"  {prec}: PCG its {its} oracle {its_ref}")
        assert abs(its - its_ref) <= it_tol


def test_direct_coarse_solve_and_graph_replays(device):
    """exact FP64 coarse solve (assembled coarse matrix) matches the oracle's
    V-cycle with a dense coarse solve; the captured V-cycle graph gives the
    eager result on every replay, with eager work in between"""
    import torch
    from oracle.gmg import GMG
    dim, nc, L = 3, [2, 2, 2], 3
    meshes = [lsmf.Mesh(dim, 2, [k << l for k in nc]) for l in range(L)]
    cons = [m.constrained_mask({0: [0, 1, 2], 1: [0]}) for m in meshes]
    fm = meshes[-1]
    mu, lam, _ = lsmf.inclusion_material(fm.cell_centres(), dim, 1.0, 0.3, 4, 0.2, 10.0, 3)
    u, _ = lsmf.smooth_displacement(fm.coords, 0.002, 1)
    ref = GMG(meshes, cons, mu, lam, coarse_solver="direct")
    ref.setup(u)
    r = np.random.default_rng(1).random(fm.n_dofs)
    r[cons[-1] == 1] = 0
    z_ref = ref.precondition(r)
    _, its_ref, _ = ref.pcg(ref.levels[-1].op, r, 1e-8 * np.linalg.norm(r))
    rd = torch.tensor(r, device=device)
    out = {}
    for use_graph in (0, 1):
        o = lsmf.solver_options(dim=dim, degree=2, n_q_1d=3, n_cells_coarse=nc, n_levels=L, caching="tensor4",
                                mg_caching="tensor4", mg_precision="fp64", coarse_solver="direct",
                                use_graph=use_graph, dirichlet={0: ([0, 1, 2], [0.0] * 3), 1: ([0], [0.0])})
        S = lsmf.Solver(o, mu, lam)
        S.set_state(u)
        zs = []
        for rep in range(4):
            z = torch.zeros_like(rd)
            S.precondition(z, rd)
            zs.append(z.cpu().numpy())
            junk = torch.rand(1000, device=device).sum()  # eager work between replays
            torch.cuda.synchronize()
            del junk
        for z in zs:
            assert np.array_equal(z, zs[0])
        err = float(np.abs(zs[0] - z_ref).max() / np.abs(z_ref).max())
        log(f"  graph={use_graph}: V-cycle rel diff vs oracle {err:.2e}")
        assert err < 1e-8
        x = torch.zeros_like(rd)
        its, _ = S.cg(x, rd, 1e-8 * float(np.linalg.norm(r)))
        log(f"  graph={use_graph}: PCG its {its} oracle {its_ref}")
        assert abs(its - its_ref) <= 1
        out[use_graph] = zs[0]
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cg_coarse_solver_matches_oracle(prec, device):
    """LSMF_COARSE_CG (the reference's MGCoarseGridIterativeSolver: CG
    preconditioned by the coarse Chebyshev smoother to a 1e-3 reduction):
    the V-cycle equals the oracle's, and PCG needs the same iterations"""
    import torch
    from oracle.gmg import GMG
    dim, nc, L = 3, [2, 2, 2], 3
    log(f"[coarse cg] {prec} levels")
    meshes = [lsmf.Mesh(dim, 2, [k << l for k in nc]) for l in range(L)]
    cons = [m.constrained_mask({0: [0, 1, 2], 1: [0]}) for m in meshes]
    fm = meshes[-1]
    mu, lam, _ = lsmf.inclusion_material(fm.cell_centres(), dim, 1.0, 0.3, 4, 0.2, 10.0, 3)
    u, _ = lsmf.smooth_displacement(fm.coords, 0.002, 1)
    ref = GMG(meshes, cons, mu, lam, coarse_solver="cg")
    ref.setup(u)
    r = np.random.default_rng(1).random(fm.n_dofs)
    r[cons[-1] == 1] = 0
    z_ref = ref.precondition(r)
    _, its_ref, _ = ref.pcg(ref.levels[-1].op, r, 1e-8 * np.linalg.norm(r))
    o = lsmf.solver_options(dim=dim, degree=2, n_q_1d=3, n_cells_coarse=nc, n_levels=L, caching="tensor4",
                            mg_caching="tensor4", mg_precision=prec, coarse_solver="cg",
                            dirichlet={0: ([0, 1, 2], [0.0] * 3), 1: ([0], [0.0])})
    S = lsmf.Solver(o, mu, lam)
    S.set_state(u)
    rd = torch.tensor(r, device=device)
    z = torch.zeros_like(rd)
    S.precondition(z, rd)
    err = float(np.abs(z.cpu().numpy() - z_ref).max() / np.abs(z_ref).max())
    log(f"  V-cycle rel diff vs oracle {err:.2e}")
    assert err < (1e-8 if prec == "fp64" else 1e-4)
    x = torch.zeros_like(rd)
    its, _ = S.cg(x, rd, 1e-8 * float(np.linalg.norm(r)))
    log(f"  PCG its {its} oracle {its_ref}")
    assert abs(its - its_ref) <= 1


@pytest.mark.parametrize("dim,p,nc", [(3, 2, [2, 2, 2]), (3, 4, [2, 2, 2]), (2, 2, [3, 3]), (2, 1, [5, 4])])
def test_direct_coarse_inverse(dim, p, nc, device):
    """one-level hierarchy: the preconditioner is the exact coarse solve, so
    it equals numpy's inverse of the probed, symmetrised matrix applied to r
    (2,187 DoFs for 2^3 Q4 cells: 68 full blocks of 32 pivots and a ragged
    one)"""
    import torch
    from oracle.gmg import GMG
    log(f"[coarse inverse] dim={dim} p={p} cells={nc}")
    mesh = lsmf.Mesh(dim, p, nc)
    cons = mesh.constrained_mask({0: list(range(dim)), 1: [0]})
    mu, lam, _ = lsmf.inclusion_material(mesh.cell_centres(), dim, 1.0, 0.3, 2, 0.3, 10.0, 5)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.01, 2)
    ref = GMG([mesh], [cons], mu, lam, coarse_solver="direct")
    ref.setup(u)
    r = np.random.default_rng(7).standard_normal(mesh.n_dofs)
    r[cons == 1] = 0
    z_ref = ref.precondition(r)
    o = lsmf.solver_options(dim=dim, degree=p, n_q_1d=p + 1, n_cells_coarse=nc, n_levels=1, caching="tensor4",
                            mg_caching="tensor4", mg_precision="fp64", coarse_solver="direct",
                            dirichlet={0: (list(range(dim)), [0.0] * dim), 1: ([0], [0.0])})
    S = lsmf.Solver(o, mu, lam)
    S.set_state(u)
    z = torch.zeros(mesh.n_dofs, dtype=torch.float64, device=device)
    S.precondition(z, torch.tensor(r, device=device))
    err = float(np.abs(z.cpu().numpy() - z_ref).max() / np.abs(z_ref).max())
    log(f"  rel diff vs numpy inverse {err:.2e}")
    assert err < 1e-10


def cook_gmg_options(prec, caching):
    prm = GOLD["cook_mf_gmg_form_1"]["prm"]
    mu = float(prm["Shear modulus"])
    nu = float(prm["Poisson's ratio"])
    o = lsmf.solver_options(dim=2, degree=1, n_q_1d=2, n_cells_coarse=[2, 2],
                            n_levels=int(prm["Global refinement"]) + 1, geometry="cook", lo=[0, 0], hi=[48, 44],
                            scale=float(prm["Grid scale"]), caching=caching, mg_caching=caching, mg_precision=prec,
                            tol_lin=float(prm["Residual"]), tol_f=float(prm["Tolerance force"]),
                            tol_u=float(prm["Tolerance displacement"]),
                            max_newton=int(prm["Max iterations Newton-Raphson"]), dt=float(prm["Time step size"]),
                            t_end=float(prm["End time"]), dirichlet={0: ([0, 1], [0.0, 0.0])},
                            neumann={1: [0.0, 1.0 / 16e-6]})
    return o, mu, 2 * mu * nu / (1 - 2 * nu)


@pytest.mark.parametrize("prec,caching", [("fp32", "scalar"), ("fp64", "tensor4"), ("fp32", "tensor2"),
                                         ("fp32", "tensor4_ns"), ("fp32", "scalar_referential")])
def test_cook_gmg_newton_matches_blessed_output(prec, caching, device):
    log(f"[cook gmg newton] {prec} {caching}")
    out = GOLD["cook_mf_gmg_form_1"]["output"]
    o, mu, lam = cook_gmg_options(prec, caching)
    n_cells = 32 * 32
    S = lsmf.Solver(o, np.full(n_cells, mu), np.full(n_cells, lam))
    assert S.n_dofs() == out["n_dofs"]
    stats = S.run()
    u = S.solution()
    mesh = lsmf.Mesh(2, 1, [32, 32], lo=[0, 0], hi=[48, 44], geometry="cook", scale=1e-3)
    corner = int(np.argmin(np.sum((mesh.coords - np.array(out["steps"][0]["point"])) ** 2, axis=1)))
    assert len(stats) == len(out["steps"])
    assert stats[-1]["newton_iterations"] == out["steps"][-1]["newton"]
    for k, st in enumerate(out["steps"]):
        assert stats[k]["newton_iterations"] == st["newton"], k
    got = [float(f"{v:.6g}") for v in u[2 * corner:2 * corner + 2]]
    assert got == out["steps"][-1]["displacement"], (got, out["steps"][-1]["displacement"])
    log(f"  CG iterations per step: {[s['linear_iterations'] for s in stats]}")


def test_cook_gmg_every_step_displacement(device):
    """displacement history: rerun step by step with t_end = k dt"""
    out = GOLD["cook_mf_gmg_form_1"]["output"]
    mesh = lsmf.Mesh(2, 1, [32, 32], lo=[0, 0], hi=[48, 44], geometry="cook", scale=1e-3)
    corner = int(np.argmin(np.sum((mesh.coords - np.array(out["steps"][0]["point"])) ** 2, axis=1)))
    o, mu, lam = cook_gmg_options("fp32", "tensor4")
    for k in (1, 4, 7):
        o.t_end = 0.1 * k + 1e-9
        S = lsmf.Solver(o, np.full(1024, mu), np.full(1024, lam))
        stats = S.run()
        u = S.solution()
        assert len(stats) == k
        got = [float(f"{v:.6g}") for v in u[2 * corner:2 * corner + 2]]
        assert got == out["steps"][k - 1]["displacement"], (k, got)


def test_config0_newton_matches_oracle(device):
    """2D unit square 32^2 Q2, 8 inclusions, 5 levels, Chebyshev degree 4,
    x=0 clamped, u_x = 0.01 per step at x=1 over 5 steps"""
    log("[config0] GPU Newton vs oracle")
    cfg = lsmf.CONFIGS[0]
    mesh, mu, lam = lsmf.build_config_mesh(0)
    o = lsmf.solver_options(dim=2, degree=2, n_q_1d=3, n_cells_coarse=[2, 2], n_levels=cfg["mg_levels"],
                            caching="tensor4", mg_caching="tensor4", mg_precision="fp32",
                            smoother_degree=cfg["cheb_degree"], dt=0.2, t_end=1.0, tol_lin=1e-6,
                            dirichlet={0: ([0, 1], [0.0, 0.0]), 1: ([0], [cfg["increment"]])})
    S = lsmf.Solver(o, mu, lam)
    stats = S.run()
    u = S.solution()
    assert len(stats) == cfg["load_steps"]
    cons = mesh.constrained_mask({0: [0, 1], 1: [0]})
    inc = np.zeros(mesh.n_dofs)
    inc[mesh.boundary_nodes(1) * 2] = cfg["increment"]
    ref = oracle.OracleOperator(2, 2, mesh.cell_nodes, mesh.coords, cons, caching="tensor4")
    ref.set_material(mu, lam)
    its, _, u_ref = ref.newton(0.2, 1.0, dirichlet_inc=inc, tol_lin=1e-6)
    log(f"  GPU Newton {[s['newton_iterations'] for s in stats]} oracle {list(its)}")
    assert all(abs(a["newton_iterations"] - b) <= 1 for a, b in zip(stats, its))
    assert np.abs(u - u_ref).max() < 1e-7 * np.abs(u_ref).max()
    assert abs(u[mesh.boundary_nodes(1)[0] * 2] - 0.05) < 1e-12


def test_config2_problem_fp32_levels_against_fp64_oracle(device):
    """the 128^3 configuration's problem (nu = 0.45, 200 inclusions of
    radius 0.03 and 10x stiffness, tensor4, x=0 clamped, 3 load steps of
    u_x = 0.03, CG to 1e-8) on 8^3 cells: the FP32-V-cycle Newton against
    the all-FP64 oracle's Newton (Jacobi CG): Newton counts within 1 and the
    displacement to 1e-7 of its maximum"""
    log("[config2 8^3] FP32-GMG Newton vs FP64 oracle")
    cfg = lsmf.CONFIGS[2]
    mesh, mu, lam = lsmf.build_config_mesh(2, n=8)
    o = lsmf.solver_options(dim=3, degree=2, n_q_1d=3, n_cells_coarse=[2, 2, 2], n_levels=3, caching="tensor4",
                            mg_caching="tensor4", mg_precision="fp32", smoother_degree=cfg["cheb_degree"],
                            tol_lin=1e-8, tol_f=1e-8, dt=1.0 / 3, t_end=1.0 + 1e-9, coarse_solver="direct",
                            dirichlet={0: ([0, 1, 2], [0.0] * 3), 1: ([0], [cfg["increment"]])})
    S = lsmf.Solver(o, mu, lam)
    stats = S.run()
    u = S.solution()
    assert len(stats) == 3
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    inc = np.zeros(mesh.n_dofs)
    inc[mesh.boundary_nodes(1) * 3] = cfg["increment"]
    ref = oracle.OracleOperator(3, 2, mesh.cell_nodes, mesh.coords, cons, caching="tensor4")
    ref.set_material(mu, lam)
    its, _, u_ref = ref.newton(1.0 / 3, 1.0 + 1e-9, dirichlet_inc=inc, tol_f=1e-8, tol_lin=1e-8)
    err = float(np.abs(u - u_ref).max() / np.abs(u_ref).max())
    log(f"  GPU Newton {[s['newton_iterations'] for s in stats]} oracle {list(its)}, |u - u_oracle| / |u| = {err:.2e}")
    assert len(its) == 3
    assert all(abs(a["newton_iterations"] - b) <= 1 for a, b in zip(stats, its))
    assert err < 1e-7


def test_config1_problem_against_fp64_oracle(device):
    """the 32^3 configuration's problem (nu = 0.3, 20 inclusions of radius
    0.08 and 10x stiffness, tensor4, x=0 clamped, 5 load steps of u_x =
    0.01, GMG-CG to 1e-8, FP64 levels) on 8^3 cells against the oracle's
    Newton (Jacobi CG): Newton counts within 1 and the displacement to 1e-7
    of its maximum"""
    log("[config1 8^3] FP64-GMG Newton vs FP64 oracle")
    cfg = lsmf.CONFIGS[1]
    mesh, mu, lam = lsmf.build_config_mesh(1, n=8)
    o = lsmf.solver_options(dim=3, degree=2, n_q_1d=3, n_cells_coarse=[2, 2, 2], n_levels=3, caching="tensor4",
                            mg_caching="tensor4", mg_precision="fp64", tol_lin=1e-8, tol_f=1e-8, dt=1.0 / 5,
                            t_end=1.0 + 1e-9, coarse_solver="direct",
                            dirichlet={0: ([0, 1, 2], [0.0] * 3), 1: ([0], [cfg["increment"]])})
    S = lsmf.Solver(o, mu, lam)
    stats = S.run()
    u = S.solution()
    assert len(stats) == 5
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    inc = np.zeros(mesh.n_dofs)
    inc[mesh.boundary_nodes(1) * 3] = cfg["increment"]
    ref = oracle.OracleOperator(3, 2, mesh.cell_nodes, mesh.coords, cons, caching="tensor4")
    ref.set_material(mu, lam)
    its, _, u_ref = ref.newton(1.0 / 5, 1.0 + 1e-9, dirichlet_inc=inc, tol_f=1e-8, tol_lin=1e-8)
    err = float(np.abs(u - u_ref).max() / np.abs(u_ref).max())
    log(f"  GPU Newton {[s['newton_iterations'] for s in stats]} oracle {list(its)}, |u - u_oracle| / |u| = {err:.2e}")
    assert len(its) == 5
    assert all(abs(a["newton_iterations"] - b) <= 1 for a, b in zip(stats, its))
    assert err < 1e-7


def test_config4_problem_fp32_levels_against_fp64_oracle(device):
    """the 64^3 Q4 configuration's problem (nu = 0.45, log-normal mu/lambda
    with sigma = 0.3, tensor2, x=0 clamped, 3 load steps of u_x = 0.03, CG
    to 1e-8) on 4^3 Q4 cells with two levels (coarse 2^3 Q4, exact coarse
    solve): the FP32-V-cycle Newton against the all-FP64 oracle's Newton
    (Jacobi CG): Newton counts within 1 and the displacement to 1e-7 of its
    maximum"""
    log("[config4 4^3] FP32-GMG Newton vs FP64 oracle")
    cfg = lsmf.CONFIGS[4]
    mesh, mu, lam = lsmf.build_config_mesh(4, n=4)
    o = lsmf.solver_options(dim=3, degree=4, n_q_1d=5, n_cells_coarse=[2, 2, 2], n_levels=2, caching="tensor2",
                            mg_caching="tensor2", mg_precision="fp32", tol_lin=1e-8, tol_f=1e-8, dt=1.0 / 3,
                            t_end=1.0 + 1e-9, coarse_solver="direct",
                            dirichlet={0: ([0, 1, 2], [0.0] * 3), 1: ([0], [cfg["increment"]])})
    S = lsmf.Solver(o, mu, lam)
    stats = S.run()
    u = S.solution()
    assert len(stats) == 3
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    inc = np.zeros(mesh.n_dofs)
    inc[mesh.boundary_nodes(1) * 3] = cfg["increment"]
    ref = oracle.OracleOperator(3, 4, mesh.cell_nodes, mesh.coords, cons, caching="tensor2")
    ref.set_material(mu, lam)
    its, _, u_ref = ref.newton(1.0 / 3, 1.0 + 1e-9, dirichlet_inc=inc, tol_f=1e-8, tol_lin=1e-8)
    err = float(np.abs(u - u_ref).max() / np.abs(u_ref).max())
    log(f"  GPU Newton {[s['newton_iterations'] for s in stats]} CG {[s['linear_iterations'] for s in stats]} "
        f"oracle {list(its)}, |u - u_oracle| / |u| = {err:.2e}")
    assert len(its) == 3
    assert all(abs(a["newton_iterations"] - b) <= 1 for a, b in zip(stats, its))
    assert err < 1e-7


def test_cook_jacobi_newton_matches_blessed_output_per_step(device):
    """tests/cook_mf_form_1.prm: 32x32 Q1 Cook membrane, MF_CG with
    PreconditionJacobi(0.65); driven one load step at a time through
    lsmf_solver_step, every step's Newton count and corner displacement
    against the blessed output"""
    log("[cook jacobi] step by step")
    gold = GOLD["cook_mf_form_1"]
    prm, out = gold["prm"], gold["output"]
    mu = float(prm["Shear modulus"])
    nu = float(prm["Poisson's ratio"])
    lam = 2 * mu * nu / (1 - 2 * nu)
    o = lsmf.solver_options(dim=2, degree=1, n_q_1d=2, n_cells_coarse=[32, 32], n_levels=1, geometry="cook",
                            lo=[0, 0], hi=[48, 44], scale=float(prm["Grid scale"]), caching="tensor4",
                            preconditioner="jacobi", tol_lin=float(prm["Residual"]),
                            max_lin_iterations=out["n_dofs"], tol_f=float(prm["Tolerance force"]),
                            tol_u=float(prm["Tolerance displacement"]),
                            max_newton=int(prm["Max iterations Newton-Raphson"]), dt=float(prm["Time step size"]),
                            t_end=float(prm["End time"]), dirichlet={0: ([0, 1], [0.0, 0.0])},
                            neumann={1: [0.0, 1.0 / 16e-6]})
    assert abs(o.preconditioner_relaxation - 0.65) < 1e-15
    S = lsmf.Solver(o, np.full(1024, mu), np.full(1024, lam))
    assert S.n_dofs() == out["n_dofs"]
    mesh = lsmf.Mesh(2, 1, [32, 32], lo=[0, 0], hi=[48, 44], geometry="cook", scale=1e-3)
    corner = int(np.argmin(np.sum((mesh.coords - np.array(out["steps"][0]["point"])) ** 2, axis=1)))
    k = 0
    while True:
        r = S.step()
        if r is None:
            break
        t, st = r
        ref = out["steps"][k]
        assert abs(t - ref["time"]) < 1e-12
        assert st["newton_iterations"] == ref["newton"], k
        u = S.solution()
        got = [float(f"{v:.6g}") for v in u[2 * corner:2 * corner + 2]]
        assert got == ref["displacement"], (k, got, ref["displacement"])
        k += 1
    assert k == len(out["steps"])
    assert S.step() is None


@pytest.mark.parametrize("pre", ["none", "jacobi"])
def test_unpreconditioned_and_jacobi_newton_agree_with_gmg(pre, device):
    """same 3D Q2 problem solved with GMG, Jacobi and no preconditioner:
    identical Newton counts, solutions within the linear tolerance"""
    nc, L = [2, 2, 2], 2
    mesh = lsmf.Mesh(3, 2, [k << (L - 1) for k in nc])
    mu, lam = lsmf.lognormal_material(mesh.n_cells, 1.0, 0.3, 0.5, 5)

    def solve(p, levels):
        o = lsmf.solver_options(dim=3, degree=2, n_q_1d=3, n_cells_coarse=[k << (L - levels) for k in nc],
                                n_levels=levels, caching="tensor4", mg_precision="fp64", preconditioner=p,
                                tol_lin=1e-10, max_lin_iterations=5000, dt=0.5, t_end=1.0,
                                dirichlet={0: ([0, 1, 2], [0.0] * 3), 1: ([0], [0.02])})
        S = lsmf.Solver(o, mu, lam)
        return S.run(), S.solution()

    st_g, u_g = solve("gmg", L)
    st_p, u_p = solve(pre, 1)
    log(f"  {pre}: Newton {[s['newton_iterations'] for s in st_p]} CG {[s['linear_iterations'] for s in st_p]};"
        f" GMG CG {[s['linear_iterations'] for s in st_g]}")
    assert [s["newton_iterations"] for s in st_p] == [s["newton_iterations"] for s in st_g]
    assert np.abs(u_p - u_g).max() < 1e-8 * np.abs(u_g).max()


@pytest.mark.parametrize("name,pre,prec,caching", [("cook_mf", "jacobi", "fp64", "scalar"),
                                                   ("cook_mf_gmg", "gmg", "fp32", "scalar"),
                                                   ("cook_mf_gmg", "gmg", "fp32", "tensor4"),
                                                   ("cook_3d_quad_mf", "jacobi", "fp64", "tensor4")])
def test_cook_formulation0_newton_matches_blessed_output(name, pre, prec, caching, device):
    """formulation 0 through the solver API: every step's Newton count and
    point displacement against the reference's blessed output"""
    log(f"[cook formulation 0] {name} {pre} {prec} {caching}")
    gold = GOLD[name]
    prm, out = gold["prm"], gold["output"]
    dim = int(prm.get("Dimension", "2"))
    p = int(prm["Polynomial degree"])
    E, R = int(prm["Elements per edge"]), int(prm.get("Global refinement", "0"))
    mu = float(prm["Shear modulus"])
    nu = float(prm["Poisson's ratio"])
    lam = 2 * mu * nu / (1 - 2 * nu)
    gmg = pre == "gmg"
    nc = [E, E] + ([1] if dim == 3 else [])
    fine = [k << R for k in nc[:2]] + nc[2:]
    lo, hi = ([0, 0], [48, 44]) if dim == 2 else ([0, 0, -0.5], [48, 44, 0.5])
    o = lsmf.solver_options(dim=dim, degree=p, n_q_1d=p + 1, n_cells_coarse=nc if gmg else fine,
                            n_levels=R + 1 if gmg else 1, geometry="cook", lo=lo, hi=hi,
                            scale=float(prm["Grid scale"]), caching=caching, mg_caching=caching, mg_precision=prec,
                            preconditioner=pre, formulation=0, tol_lin=float(prm["Residual"]),
                            max_lin_iterations=out["n_dofs"] * int(prm.get("Max iteration multiplier", "1")),
                            tol_f=float(prm["Tolerance force"]), tol_u=float(prm["Tolerance displacement"]),
                            max_newton=int(prm["Max iterations Newton-Raphson"]), dt=float(prm["Time step size"]),
                            t_end=float(prm["End time"]), dirichlet={0: (list(range(dim)), [0.0] * dim)},
                            neumann={1: [0.0, 1.0 / 16e-6] + [0.0] * (dim - 2)})
    mesh = lsmf.Mesh(dim, p, fine, lo=lo, hi=hi, geometry="cook", scale=1e-3)
    S = lsmf.Solver(o, np.full(mesh.n_cells, mu), np.full(mesh.n_cells, lam))
    assert S.n_dofs() == out["n_dofs"]
    node = int(np.argmin(np.sum((mesh.coords - np.array(out["steps"][0]["point"])) ** 2, axis=1)))
    for k, ref in enumerate(out["steps"]):
        t, st = S.step()
        assert st["newton_iterations"] == ref["newton"], k
        u = S.solution()
        got = [float(f"{v:.6g}") for v in u[dim * node:dim * node + dim]]
        assert got == ref["displacement"], (k, got, ref["displacement"])
    assert S.step() is None


def test_condition_number_estimate(device):
    """the CG condition-number estimate run before each linear solve
    (include/mf_elasticity.h:3029-3046): with enough iterations it equals
    lambda_max / lambda_min of the tangent's free block (eigvalsh of the
    oracle's tangent at the converged state); with 3 iterations the Ritz
    values lie inside the spectrum, so the estimate is in [1, cond]; 0
    iterations report 0"""
    log("[cond] 6x6 Q1 unit square, x=0 clamped, u_x = 0.02 at x=1")
    mesh = lsmf.Mesh(2, 1, [6, 6])
    mu, lam = lsmf.lognormal_material(mesh.n_cells, 1.0, 0.3, 0.5, 7)
    cons = mesh.constrained_mask({0: [0, 1], 1: [0]})

    def solve(its):
        o = lsmf.solver_options(dim=2, degree=1, n_q_1d=2, n_cells_coarse=[6, 6], n_levels=1, caching="tensor4",
                                preconditioner="jacobi", tol_lin=1e-12, max_lin_iterations=1000, dt=1.0, t_end=1.0,
                                cond_number_cg_iterations=its, dirichlet={0: ([0, 1], [0.0, 0.0]), 1: ([0], [0.02])})
        S = lsmf.Solver(o, mu, lam)
        stats = S.run()
        assert len(stats) == 1
        return stats[0]["cond_number"], S.solution()

    est, u = solve(400)
    ref = oracle.OracleOperator(2, 1, mesh.cell_nodes, mesh.coords, cons, caching="tensor4")
    ref.set_material(mu, lam)
    ref.set_linearization(u)
    ref.cache()
    K = np.stack([ref.vmult(e) for e in np.eye(mesh.n_dofs)], axis=1)
    free = np.where(cons == 0)[0]
    ev = np.linalg.eigvalsh(0.5 * (K + K.T)[np.ix_(free, free)])
    exact = ev[-1] / ev[0]
    log(f"  estimate {est:.10e} eigvalsh {exact:.10e}")
    assert abs(est - exact) < 1e-6 * exact
    est3, _ = solve(3)
    log(f"  3 iterations: {est3:.6e}")
    assert 1.0 <= est3 <= exact * (1 + 1e-10)
    est0, _ = solve(0)
    assert est0 == 0.0
