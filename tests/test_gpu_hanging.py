"""GPU operator with hanging-node (general affine) constraints against the
oracle, through the C ABI: a coarse cell next to a once-refined region
(tests/hanging_mesh.py), 2D Q1/Q2/Q4, 3D Q1/Q2, x = 0 clamped.

  * vmult (set, add, plain), FP64 to 1e-12 and FP32 to 2e-5, for the
    tensor4, tensor2 and scalar caches
  * the probing diagonal and the residual (distributed to the masters)
  * the residual VJP's displacement part equals minus the plain operator
  * the cells layout is used (the mesh is not structured); the masters
    receive the constrained nodes' results after the colours
"""
import numpy as np
import pytest

import hanging_mesh
import lsmf
import oracle

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (2, 4), (3, 1), (3, 2)])
@pytest.mark.parametrize("caching", ["tensor4", "tensor2", "scalar"])
def test_hanging_operator_matches_oracle(dim, p, caching, device):
    import torch
    log(f"[hanging] dim={dim} p={p} caching={caching}")
    mesh, nc = hanging_mesh.build(dim, p)
    rng = np.random.default_rng(3)
    mu = 1.0 + rng.random(mesh.n_cells)
    lam = 1.5 + rng.random(mesh.n_cells)
    C = hanging_mesh.constraint_matrix(mesh, nc)
    x = mesh.coords
    u = C @ (0.02 * np.sin(np.pi * (x @ np.arange(1, dim + 1)))[:, None] * np.ones(dim)).reshape(-1)
    cons = np.zeros(mesh.n_dofs, np.uint8)
    left = np.where(np.abs(x[:, 0]) < 1e-12)[0]
    for c in range(dim):
        cons[left * dim + c] = 1
    ref = oracle.OracleOperator(dim, p, mesh.cell_nodes, mesh.coords, cons, caching=caching)
    ref.set_node_constraints(*nc)
    ref.set_material(mu, lam)
    ref.set_linearization(u)
    ref.cache()
    v = rng.standard_normal(mesh.n_dofs)
    w = rng.standard_normal(mesh.n_dofs)
    ud = torch.tensor(u, device=device)
    for prec, tol in [("fp64", 1e-12), ("fp32", 2e-5)]:
        op = lsmf.NeoHookOperator(mesh, precision=prec, caching=caching, constrained=cons,
                                  use_mesh_colouring=False, node_constraints=nc)
        op.set_material(mu, lam)
        op.set_linearization(ud)
        op.cache()
        dt = torch.float64 if prec == "fp64" else torch.float32
        vv = v if prec == "fp64" else v.astype(np.float32).astype(np.float64)
        ww = w if prec == "fp64" else w.astype(np.float32).astype(np.float64)
        src = torch.tensor(vv, dtype=dt, device=device)
        dst = torch.full_like(src, np.nan)
        op.vmult(dst, src)
        assert rel(dst.cpu().double().numpy(), ref.vmult(vv)) < tol, prec
        acc = torch.tensor(ww, dtype=dt, device=device)
        op.vmult_add(acc, src)
        assert rel(acc.cpu().double().numpy(), ww + ref.vmult(vv)) < tol, prec
        pl = torch.full_like(src, np.nan)
        op.vmult(pl, src, mask="plain")
        assert rel(pl.cpu().double().numpy(), ref.vmult(vv, plain=True)) < tol, prec
        op.compute_diagonal()
        assert rel(op.diagonal().cpu().double().numpy(), ref.diagonal()) < tol, prec
        # batched products (zeroed first, masters added) equal the single ones
        X = torch.stack([src, acc, pl])
        Y = torch.full_like(X, np.nan)
        op.vmult_batched(Y, X)
        for b in range(3):
            y = torch.full_like(src, np.nan)
            op.vmult(y, X[b].contiguous())
            assert torch.equal(Y[b], y), (prec, b)
    r = torch.full((mesh.n_dofs,), np.nan, dtype=torch.float64, device=device)
    op = lsmf.NeoHookOperator(mesh, precision="fp64", caching=caching, constrained=cons, use_mesh_colouring=False,
                              node_constraints=nc)
    op.set_material(mu, lam)
    op.set_linearization(ud)
    op.residual(r)
    assert rel(r.cpu().numpy(), ref.residual()) < 1e-12
    op.cache()
    wt = torch.tensor(w, device=device)
    du = torch.empty_like(wt)
    dmu = torch.empty(mesh.n_cells, dtype=torch.float64, device=device)
    dla = torch.empty_like(dmu)
    op.residual_vjp(wt, du, dmu, dla)
    pw = torch.empty_like(wt)
    op.vmult(pw, wt, mask="plain")
    assert rel(du.cpu().numpy(), -pw.cpu().numpy()) < 1e-12



@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("caching", ["tensor4", "tensor2"])
def test_half_refined_box_matches_oracle_and_is_deterministic(p, caching, device):
    """the hanging-node throughput mesh (lsmf.half_refined_box) at 2 coarse
    cells per side against the oracle, FP64 to 1e-12; the masters' sums are
    read-modify-writes in a fixed order (no atomics), so two products are
    bit-identical"""
    import torch
    log(f"[hanging] half_refined_box p={p} caching={caching}")
    mesh, nc = lsmf.half_refined_box(p, 2)
    rng = np.random.default_rng(11)
    mu, lam = 1.0 + rng.random(mesh.n_cells), 1.5 + rng.random(mesh.n_cells)
    C = hanging_mesh.constraint_matrix(mesh, nc)
    x = mesh.coords
    u = C @ (0.02 * np.sin(np.pi * (x @ np.arange(1, 4)))[:, None] * np.ones(3)).reshape(-1)
    cons = np.zeros(mesh.n_dofs, np.uint8)
    left = np.where(np.abs(x[:, 0]) < 1e-12)[0]
    for c in range(3):
        cons[left * 3 + c] = 1
    ref = oracle.OracleOperator(3, p, mesh.cell_nodes, mesh.coords, cons, caching=caching)
    ref.set_node_constraints(*nc)
    ref.set_material(mu, lam)
    ref.set_linearization(u)
    ref.cache()
    op = lsmf.NeoHookOperator(mesh, precision="fp64", caching=caching, constrained=cons, use_mesh_colouring=False,
                              node_constraints=nc)
    op.set_material(mu, lam)
    op.set_linearization(torch.tensor(u, device=device))
    op.cache()
    v = rng.standard_normal(mesh.n_dofs)
    src = torch.tensor(v, device=device)
    y1, y2 = torch.full_like(src, np.nan), torch.full_like(src, np.nan)
    op.vmult(y1, src)
    op.vmult(y2, src)
    assert torch.equal(y1, y2)
    assert rel(y1.cpu().numpy(), ref.vmult(v)) < 1e-12
    pl = torch.full_like(src, np.nan)
    op.vmult(pl, src, mask="plain")
    assert rel(pl.cpu().numpy(), ref.vmult(v, plain=True)) < 1e-12


def test_half_refined_box_batched_beyond_one_result_buffer(device):
    """batched products with constraints run in groups of the 8 vectors the
    constrained-node result buffer holds: 11 vectors give the 11 single
    products bit for bit"""
    import torch
    log("[hanging] half_refined_box batched, 11 vectors")
    mesh, nc = lsmf.half_refined_box(2, 2)
    rng = np.random.default_rng(13)
    mu, lam = 1.0 + rng.random(mesh.n_cells), 1.5 + rng.random(mesh.n_cells)
    C = hanging_mesh.constraint_matrix(mesh, nc)
    x = mesh.coords
    u = C @ (0.02 * np.sin(np.pi * (x @ np.arange(1, 4)))[:, None] * np.ones(3)).reshape(-1)
    cons = np.zeros(mesh.n_dofs, np.uint8)
    left = np.where(np.abs(x[:, 0]) < 1e-12)[0]
    for c in range(3):
        cons[left * 3 + c] = 1
    op = lsmf.NeoHookOperator(mesh, precision="fp64", caching="tensor4", constrained=cons, use_mesh_colouring=False,
                              node_constraints=nc)
    op.set_material(mu, lam)
    op.set_linearization(torch.tensor(u, device=device))
    op.cache()
    X = torch.tensor(rng.standard_normal((11, mesh.n_dofs)), device=device)
    Y = torch.full_like(X, np.nan)
    op.vmult_batched(Y, X)
    for b in range(11):
        y = torch.full_like(X[b], np.nan)
        op.vmult(y, X[b].contiguous())
        assert torch.equal(Y[b], y), b
