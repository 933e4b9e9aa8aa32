"""Hand-written residual adjoint (VJP w.r.t. u and per-cell mu, lambda).

  * (dR/du)^T w equals -A w of the tensor4 tangent (unconstrained), i.e. the
    VJP matches the reference's tangent vmult; JVP = VJP by symmetry
  * central finite differences (h = 1e-6) on 64 seeded directions in
    (u, mu, lambda) on the 48^3 Q2 configuration with log-normal per-cell
    mu, lambda (sigma = 0.5): relative error <= 1e-6; the same on 16
    directions for the Q4 element of the 64^3 Q4 configuration
  * the torch.autograd.Function gives the same gradient as differences of
    a scalar loss
  * gradient of a face-misfit objective through a Newton solve (one adjoint
    GMG-CG solve + one VJP) against differences of two full solves, also
    with Q4 tensor2 and FP32 multigrid levels, and on the 48^3 Q2
    configuration at its full size
"""
import numpy as np
import pytest

import lsmf
from lsmf.autograd import misfit_gradient, residual

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def make_op(mesh, mu, lam, u, cons=None, caching="tensor4"):
    import torch
    op = lsmf.NeoHookOperator(mesh, precision="fp64", caching=caching, constrained=cons)
    op.set_material(mu, lam)
    op.set_linearization(torch.tensor(u, device="cuda"))
    return op


@pytest.mark.parametrize("dim,p,n", [(3, 2, 4), (2, 2, 6), (3, 4, 2), (2, 1, 8)])
def test_vjp_u_is_minus_tangent(dim, p, n, device):
    import torch
    log(f"[vjp=tangent] dim={dim} p={p}")
    mesh = lsmf.Mesh(dim, p, [n] * dim)
    mu, lam = lsmf.lognormal_material(mesh.n_cells, 1.0, 0.3, 0.5, 7)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.01, 2)
    op = make_op(mesh, mu, lam, u)
    op.cache()
    g = torch.Generator(device=device)
    g.manual_seed(3)
    w = torch.rand(mesh.n_dofs, dtype=torch.float64, device=device, generator=g)
    v = torch.rand(mesh.n_dofs, dtype=torch.float64, device=device, generator=g)
    du = torch.empty_like(w)
    dmu = torch.empty(mesh.n_cells, dtype=torch.float64, device=device)
    dla = torch.empty_like(dmu)
    op.residual_vjp(w, du, dmu, dla)
    Aw = torch.empty_like(w)
    op.vmult(Aw, w, mask="plain")
    assert float((du + Aw).abs().max()) < 1e-12 * float(Aw.abs().max())
    dv = torch.empty_like(v)
    op.residual_vjp(v, dv, dmu, dla)
    assert abs(float(v @ du) - float(w @ dv)) < 1e-12 * abs(float(v @ du))


def test_vjp_matches_finite_differences_config3(device):
    import torch
    log("[vjp FD] 48^3 Q2, log-normal mu/lambda, 64 directions")
    mesh, mu, lam = lsmf.build_config_mesh(3)
    assert mesh.n_dofs == 2738019
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.01, 11)
    op = make_op(mesh, mu, lam, u)
    rng = np.random.default_rng(64)
    w = torch.tensor(rng.standard_normal(mesh.n_dofs), device=device)
    du = torch.empty_like(w)
    dmu = torch.empty(mesh.n_cells, dtype=torch.float64, device=device)
    dla = torch.empty_like(dmu)
    op.residual_vjp(w, du, dmu, dla)
    du, dmu, dla = du.cpu().numpy(), dmu.cpu().numpy(), dla.cpu().numpy()
    ud = torch.tensor(u, device=device)
    r = torch.empty_like(ud)
    h = 1e-6
    worst = 0.0
    for k in range(64):
        a = rng.standard_normal(mesh.n_dofs) * 0.01
        b = rng.standard_normal(mesh.n_cells) * np.asarray(mu)
        c = rng.standard_normal(mesh.n_cells) * np.asarray(lam)
        vals = []
        for sgn in (1.0, -1.0):
            op.set_material(mu + sgn * h * b, lam + sgn * h * c)
            op.set_linearization(ud + sgn * h * torch.tensor(a, device=device))
            op.residual(r)
            vals.append(float(w @ r))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = float(du @ a + dmu @ b + dla @ c)
        err = abs(fd - an) / abs(an)
        worst = max(worst, err)
        assert err < 1e-6, (k, fd, an)
    log(f"  worst relative FD error {worst:.2e}")


def test_vjp_matches_finite_differences_q4(device):
    """config 4's element (Q4, 5^3 Gauss points, log-normal sigma = 0.3) on
    a 4^3 mesh: central differences on 16 seeded directions"""
    import torch
    log("[vjp FD] 4^3 Q4, log-normal mu/lambda, 16 directions")
    mesh = lsmf.Mesh(3, 4, [4, 4, 4])
    mu, lam = lsmf.lognormal_material(mesh.n_cells, 1.0, 0.45, 0.3, 41)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, 43)
    op = make_op(mesh, mu, lam, u, caching="tensor2")
    rng = np.random.default_rng(44)
    w = torch.tensor(rng.standard_normal(mesh.n_dofs), device=device)
    du = torch.empty_like(w)
    dmu = torch.empty(mesh.n_cells, dtype=torch.float64, device=device)
    dla = torch.empty_like(dmu)
    op.residual_vjp(w, du, dmu, dla)
    du, dmu, dla = du.cpu().numpy(), dmu.cpu().numpy(), dla.cpu().numpy()
    ud = torch.tensor(u, device=device)
    r = torch.empty_like(ud)
    h = 1e-6
    worst = 0.0
    for k in range(16):
        a = rng.standard_normal(mesh.n_dofs) * 0.01
        b = rng.standard_normal(mesh.n_cells) * np.asarray(mu)
        c = rng.standard_normal(mesh.n_cells) * np.asarray(lam)
        vals = []
        for sgn in (1.0, -1.0):
            op.set_material(mu + sgn * h * b, lam + sgn * h * c)
            op.set_linearization(ud + sgn * h * torch.tensor(a, device=device))
            op.residual(r)
            vals.append(float(w @ r))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = float(du @ a + dmu @ b + dla @ c)
        err = abs(fd - an) / abs(an)
        worst = max(worst, err)
        assert err < 1e-6, (k, fd, an)
    log(f"  worst relative FD error {worst:.2e}")


def test_autograd_function(device):
    import torch
    mesh = lsmf.Mesh(3, 2, [4, 4, 4])
    mu0, lam0 = lsmf.lognormal_material(mesh.n_cells, 1.0, 0.3, 0.5, 9)
    u0, _ = lsmf.smooth_displacement(mesh.coords, 0.01, 4)
    op = lsmf.NeoHookOperator(mesh, precision="fp64", caching="tensor4")
    u = torch.tensor(u0, device=device, requires_grad=True)
    mu = torch.tensor(mu0, requires_grad=True)
    lam = torch.tensor(lam0, requires_grad=True)

    def loss(u_, mu_, lam_):
        R = residual(u_, mu_, lam_, op)
        return 0.5 * (R * R).sum()

    L = loss(u, mu, lam)
    L.backward()
    rng = np.random.default_rng(5)
    a = torch.tensor(rng.standard_normal(mesh.n_dofs) * 0.01, device=device)
    b = torch.tensor(rng.standard_normal(mesh.n_cells) * mu0)
    c = torch.tensor(rng.standard_normal(mesh.n_cells) * lam0)
    h = 1e-6
    with torch.no_grad():
        Lp = float(loss(u + h * a, mu + h * b, lam + h * c))
        Lm = float(loss(u - h * a, mu - h * b, lam - h * c))
    fd = (Lp - Lm) / (2 * h)
    an = float((u.grad * a).sum()) + float((mu.grad * b).sum()) + float((lam.grad * c).sum())
    assert abs(fd - an) < 1e-6 * abs(an), (fd, an)


@pytest.mark.parametrize("degree,caching,mg_precision,sigma", [(2, "tensor4", "fp64", 0.5),
                                                                (4, "tensor2", "fp32", 0.3)])
def test_misfit_gradient_through_newton_solve(degree, caching, mg_precision, sigma, device):
    """x=0 clamped, u_x = 0.01 at x=1 (one load step); J = misfit of the x=1
    face displacement (free components) against the solution for a second
    seeded log-normal field; dJ/dmu, dJ/dlambda vs differences of solves.
    Q2 with FP64 levels on 8^3 cells; config 4's combination (Q4, tensor2,
    FP32 V-cycle inside the FP64 CG) on 4^3 cells."""
    nc = [2, 2, 2]
    L = 3 if degree == 2 else 2
    log(f"[adjoint solve] {nc[0] << (L - 1)}^3 Q{degree} {caching}, {mg_precision} levels")
    mesh = lsmf.Mesh(3, degree, [k << (L - 1) for k in nc])
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    o = lsmf.solver_options(dim=3, degree=degree, n_q_1d=degree + 1, n_cells_coarse=nc, n_levels=L,
                            caching=caching, mg_caching=caching, mg_precision=mg_precision,
                            coarse_solver="direct", tol_lin=1e-12, tol_f=1e-13, tol_u=1e-12, dt=1.0, t_end=1.0,
                            dirichlet={0: ([0, 1, 2], [0.0] * 3), 1: ([0], [0.01])})

    def solve(mu, lam):
        S = lsmf.Solver(o, mu, lam)
        S.run()
        return S, S.solution()

    mu_t, lam_t = lsmf.lognormal_material(mesh.n_cells, 1.0, 0.3, sigma, 31)
    _, target = solve(mu_t, lam_t)
    mu, lam = lsmf.lognormal_material(mesh.n_cells, 1.0, 0.3, sigma, 32)
    S, u_star = solve(mu, lam)
    face = mesh.boundary_nodes(1)
    J, gmu, gla = misfit_gradient(S, mesh, u_star, mu, lam, target, face, cons)
    assert J > 0
    rng = np.random.default_rng(3)
    for k in range(3):
        b = rng.standard_normal(mesh.n_cells) * mu
        c = rng.standard_normal(mesh.n_cells) * lam
        h = 1e-5
        Js = []
        for sgn in (1.0, -1.0):
            _, us = solve(mu + sgn * h * b, lam + sgn * h * c)
            dofs = (face[:, None] * 3 + np.arange(3)).reshape(-1)
            dofs = dofs[cons[dofs] == 0]
            d = us[dofs] - target[dofs]
            Js.append(0.5 * float(d @ d))
        fd = (Js[0] - Js[1]) / (2 * h)
        an = float(gmu @ b + gla @ c)
        log(f"  direction {k}: FD {fd:.8e} adjoint {an:.8e}")
        assert abs(fd - an) < 1e-5 * abs(an)


@pytest.mark.parametrize("cfg_id", [3, 4])
def test_misfit_gradient_at_full_size(cfg_id, device):
    """the differentiation workload at full size -- 48^3 Q2 tensor4
    (2,738,019 DoF, sigma = 0.5) and the 64^3 Q4 tensor2 configuration
    (50,923,779 DoF, sigma = 0.3) -- with log-normal mu, lambda, x=0
    clamped, u_x = 0.01 at x=1 in one load step, FP64 levels; J = misfit of
    the x=1 face displacement against the solution for a second seeded
    field.  dJ/dmu, dJ/dlambda from one adjoint solve + one VJP against
    central differences of two full Newton solves (h = 1e-5) on 2 seeded
    directions, to 1e-5 relative"""
    cfg = lsmf.CONFIGS[cfg_id]
    n, p = cfg["n"], cfg["degree"]
    log(f"[adjoint solve] {n}^3 Q{p} (config {cfg_id} at full size)")
    mesh = lsmf.Mesh(3, p, [n] * 3)
    assert mesh.n_dofs == {3: 2738019, 4: 50923779}[cfg_id]
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    nc = 3 if n % 3 == 0 else 2
    o = lsmf.solver_options(dim=3, degree=p, n_q_1d=p + 1, n_cells_coarse=[nc] * 3,
                            n_levels=int(round(np.log2(n // nc))) + 1, caching=cfg["caching"],
                            mg_caching=cfg["caching"], mg_precision="fp64", coarse_solver="direct", tol_lin=1e-12,
                            tol_f=1e-13, tol_u=1e-12, dt=1.0, t_end=1.0,
                            dirichlet={0: ([0, 1, 2], [0.0] * 3), 1: ([0], [0.01])})

    def solve(mu, lam, what):
        S = lsmf.Solver(o, mu, lam)
        stats = S.run()
        log(f"  solve {what}: Newton {[s['newton_iterations'] for s in stats]}")
        return S, S.solution()

    seed = lsmf.workload.SEED
    mu_t, lam_t = lsmf.lognormal_material(mesh.n_cells, 1.0, cfg["nu"], cfg["sigma"], seed + 31)
    S, target = solve(mu_t, lam_t, "target")
    del S
    mu, lam = lsmf.lognormal_material(mesh.n_cells, 1.0, cfg["nu"], cfg["sigma"], seed + cfg_id)
    S, u_star = solve(mu, lam, "forward")
    face = mesh.boundary_nodes(1)
    J, gmu, gla = misfit_gradient(S, mesh, u_star, mu, lam, target, face, cons)
    del S
    assert J > 0
    dofs = (face[:, None] * 3 + np.arange(3)).reshape(-1)
    dofs = dofs[cons[dofs] == 0]
    rng = np.random.default_rng(48)
    for k in range(2):
        b = rng.standard_normal(mesh.n_cells) * mu
        c = rng.standard_normal(mesh.n_cells) * lam
        h = 1e-5
        Js = []
        for sgn in (1.0, -1.0):
            S2, us = solve(mu + sgn * h * b, lam + sgn * h * c, f"direction {k} {'+' if sgn > 0 else '-'}")
            del S2
            d = us[dofs] - target[dofs]
            Js.append(0.5 * float(d @ d))
        fd = (Js[0] - Js[1]) / (2 * h)
        an = float(gmu @ b + gla @ c)
        log(f"  direction {k}: FD {fd:.8e} adjoint {an:.8e}")
        assert abs(fd - an) < 1e-5 * abs(an)
