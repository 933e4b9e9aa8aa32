"""GPU operator on curved meshes (quarter annulus, tests/curved_mesh.py)
against the oracle, through the C ABI: the spatial variants (scalar,
tensor2, tensor4) use the Eulerian Q_p metric of all support points, the
referential ones (tensor4_ns, scalar_referential) and the residual the Q1
vertex map, as the reference's MatrixFree data do
(include/mf_elasticity.h:1821-1831).

  * vmult (set, add), FP64 to 1e-12 and FP32 to 2e-5, every caching,
    tiled and cells layouts, 2D/3D Q2/Q4
  * diagonal and residual
  * the GPU operator itself passes the 2D patch test of the oracle tests
"""
import numpy as np
import pytest

import curved_mesh
import lsmf
import oracle

pytestmark = pytest.mark.gpu
ALL = ["scalar", "tensor2", "tensor4", "tensor4_ns", "scalar_referential"]


def log(msg):
    print(msg, flush=True)


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("dim,p,n", [(2, 2, [3, 4]), (2, 4, [2, 3]), (3, 2, [3, 4, 2]), (3, 4, [2, 2, 2])])
@pytest.mark.parametrize("caching", ALL)
def test_curved_operator_matches_oracle(dim, p, n, caching, device):
    import torch
    log(f"[curved] dim={dim} p={p} n={n} caching={caching}")
    mesh = curved_mesh.annulus(dim, p, n)
    rng = np.random.default_rng(4)
    mu = 1.0 + rng.random(mesh.n_cells)
    lam = 1.5 + rng.random(mesh.n_cells)
    x = mesh.coords
    u = 0.02 * np.sin(2.0 * x @ np.arange(1, dim + 1))[:, None] * np.linspace(1.0, 0.5, dim)
    u = u.reshape(-1)
    cons = mesh.constrained_mask({2: list(range(dim))})  # theta = 0 face clamped
    ref = oracle.OracleOperator(dim, p, mesh.cell_nodes, mesh.coords, cons, caching=caching)
    ref.set_material(mu, lam)
    ref.set_linearization(u)
    ref.cache()
    v = rng.standard_normal(mesh.n_dofs)
    w = rng.standard_normal(mesh.n_dofs)
    ud = torch.tensor(u, device=device)
    for layout in ("auto", "cells"):
        for prec, tol in [("fp64", 1e-12), ("fp32", 2e-5)]:
            op = lsmf.NeoHookOperator(mesh, precision=prec, caching=caching, constrained=cons, layout=layout)
            op.set_material(mu, lam)
            op.set_linearization(ud)
            op.cache()
            dt = torch.float64 if prec == "fp64" else torch.float32
            vv = v if prec == "fp64" else v.astype(np.float32).astype(np.float64)
            ww = w if prec == "fp64" else w.astype(np.float32).astype(np.float64)
            src = torch.tensor(vv, dtype=dt, device=device)
            dst = torch.full_like(src, np.nan)
            op.vmult(dst, src)
            assert rel(dst.cpu().double().numpy(), ref.vmult(vv)) < tol, (layout, prec)
            acc = torch.tensor(ww, dtype=dt, device=device)
            op.vmult_add(acc, src)
            assert rel(acc.cpu().double().numpy(), ww + ref.vmult(vv)) < tol, (layout, prec)
            op.compute_diagonal()
            assert rel(op.diagonal().cpu().double().numpy(), ref.diagonal()) < tol, (layout, prec)
    r = torch.full((mesh.n_dofs,), np.nan, dtype=torch.float64, device=device)
    op = lsmf.NeoHookOperator(mesh, precision="fp64", caching=caching, constrained=cons)
    op.set_material(mu, lam)
    op.set_linearization(ud)
    op.residual(r)
    assert rel(r.cpu().numpy(), ref.residual()) < 1e-12


@pytest.mark.parametrize("p", [2, 4])
def test_curved_patch_test_2d(p, device):
    """u = 0, v = B x (isoparametric linear field): A v vanishes at interior
    nodes (exact quadrature of the discrete divergence theorem in 2D)"""
    import torch
    mesh = curved_mesh.annulus(2, p, [3, 4])
    B = np.array([[0.3, -0.2], [0.5, 0.1]])
    v = (mesh.coords @ B.T).reshape(-1)
    for caching in ("scalar", "tensor2", "tensor4"):
        op = lsmf.NeoHookOperator(mesh, precision="fp64", caching=caching,
                                  constrained=np.zeros(mesh.n_dofs, np.uint8))
        op.set_material(np.full(mesh.n_cells, 0.7), np.full(mesh.n_cells, 1.3))
        op.set_linearization(torch.zeros(mesh.n_dofs, dtype=torch.float64, device=device))
        op.cache()
        y = torch.empty(mesh.n_dofs, dtype=torch.float64, device=device)
        op.vmult(y, torch.tensor(v, device=device))
        f = y.cpu().numpy().reshape(-1, 2)
        boundary = set()
        for s in range(4):
            boundary.update(mesh.boundary_nodes(s).tolist())
        interior = np.array([i for i in range(mesh.n_nodes) if i not in boundary])
        log(f"  p={p} {caching}: interior max {np.abs(f[interior]).max():.2e} of {np.abs(f).max():.2e}")
        assert np.abs(f[interior]).max() < 1e-12 * np.abs(f).max()
