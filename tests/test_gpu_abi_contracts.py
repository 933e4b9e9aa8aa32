"""Error and edge-case contracts of the C ABI on the GPU.

  * the MFMask RW bit gates the cell loop and the identity on constrained
    rows, as in the reference (include/mf_nh_operator.h:703-705, :731-780):
    a mask without RW does no cell work, CellLoop equals Default
  * DoFs that no cell references (a slab operator on global vectors) have
    diagonal 0 and inverse diagonal 0, so the Jacobi preconditioner and a
    Chebyshev smoother stay finite there; every other entry is positive
  * a tangent with a non-positive diagonal entry is reported by
    compute_diagonal (the reference's Assert(diag > 0), :1543-1547)
  * the solver's CG reports non-convergence (also with max_it = 0) instead
    of returning an unconverged iterate as converged
"""
import numpy as np
import pytest
from types import SimpleNamespace

import lsmf

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def small_op(device, caching="tensor4", precision="fp64", mesh=None, mu=None, lam=None, cons=None):
    import torch
    mesh = mesh or lsmf.Mesh(3, 2, [3, 3, 3])
    n = mesh.n_cells
    mu = np.full(n, 1.0) if mu is None else mu
    lam = np.full(n, 1.5) if lam is None else lam
    if cons is None:
        cons = np.zeros(mesh.n_dofs, np.uint8)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, 3)
    op = lsmf.NeoHookOperator(mesh, precision=precision, caching=caching, constrained=cons)
    op.set_material(mu, lam)
    op.set_linearization(torch.tensor(u, device=device))
    op.cache()
    return op, mesh, cons


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mask_rw_gates_the_cell_loop(precision, device):
    import torch
    log(f"[mask RW] 3^3 Q2 tensor4 {precision}")
    mesh = lsmf.Mesh(3, 2, [3, 3, 3])
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    op, mesh, cons = small_op(device, precision=precision, mesh=mesh, cons=cons)
    dt = torch.float64 if precision == "fp64" else torch.float32
    x = torch.rand(mesh.n_dofs, dtype=dt, device=device)
    full = torch.full_like(x, float("nan"))
    op.vmult(full, x)
    cell = torch.full_like(x, float("nan"))
    op.vmult(cell, x, mask="cellloop")
    assert torch.equal(full, cell)
    # no RW: no cell work; set mode gives zero, add mode leaves dst alone
    for m in (lsmf.MASK["sf"] | lsmf.MASK["qd"], lsmf.MASK["zero"] | lsmf.MASK["mpi"], 0):
        y = torch.full_like(x, float("nan"))
        op.vmult(y, x, mask=m)
        assert torch.count_nonzero(y).item() == 0, m
        w = torch.rand_like(x)
        z = w.clone()
        op.vmult(z, x, add=True, mask=m)
        assert torch.equal(z, w), m
    # RW without SF (the reference's read/write timing loop): the cell loop
    # runs its gather and scatter only, and writes every DoF
    rw = torch.full_like(x, float("nan"))
    op.vmult(rw, x, mask="rw")
    assert torch.isfinite(rw).all()


def test_unreferenced_dofs_have_zero_inverse_diagonal(device):
    import torch
    log("[orphans] slab of a 4^3 Q2 mesh on global vectors")
    whole = lsmf.Mesh(3, 2, [4, 4, 4])
    cx = whole.cell_centres()[:, 0]
    sel = np.where(cx < 0.5)[0]
    slab = SimpleNamespace(dim=3, degree=2, coords=whole.coords, cell_nodes=np.ascontiguousarray(whole.cell_nodes[sel]),
                           n_cells=len(sel), n_nodes=whole.n_nodes, n_dofs=whole.n_dofs)
    u, _ = lsmf.smooth_displacement(whole.coords, 0.02, 3)
    op = lsmf.NeoHookOperator(slab, precision="fp64", caching="tensor4",
                              constrained=np.zeros(whole.n_dofs, np.uint8), layout="cells", use_mesh_colouring=False)
    op.set_material(np.ones(len(sel)), np.full(len(sel), 1.5))
    op.set_linearization(torch.tensor(u, device=device))
    op.cache()
    op.compute_diagonal()
    d = op.diagonal().cpu().numpy()
    inv = op.inverse_diagonal().cpu().numpy()
    touched = np.zeros(whole.n_nodes, bool)
    touched[slab.cell_nodes.reshape(-1)] = True
    orphan = ~np.repeat(touched, 3)
    assert orphan.sum() > 0
    assert np.all(d[orphan] == 0) and np.all(inv[orphan] == 0)
    assert np.all(d[~orphan] > 0) and np.allclose(inv[~orphan], 1.0 / d[~orphan], rtol=1e-15)
    r = torch.rand(whole.n_dofs, dtype=torch.float64, device=device)
    z = torch.full_like(r, float("nan"))
    op.precondition_Jacobi(z, r, 0.8)
    zn = z.cpu().numpy()
    assert np.isfinite(zn).all() and np.all(zn[orphan] == 0)


def test_nonpositive_diagonal_is_reported(device):
    """a negative shear modulus makes the tangent's diagonal negative"""
    log("[diag <= 0] 3^3 Q2 tensor4 with mu < 0")
    mesh = lsmf.Mesh(3, 2, [3, 3, 3])
    op, _, _ = small_op(device, mesh=mesh, mu=np.full(mesh.n_cells, -1.0), lam=np.zeros(mesh.n_cells))
    with pytest.raises(lsmf.LsmfError) as e:
        op.compute_diagonal()
    assert e.value.code == -9
    with pytest.raises(lsmf.LsmfError):
        op.diagonal()  # not available after the failed call


@pytest.mark.parametrize("max_it", [0, 1])
def test_cg_reports_non_convergence(max_it, device):
    import torch
    log(f"[cg max_it={max_it}] 2D Q1 GMG")
    o = lsmf.solver_options(dim=2, degree=1, n_q_1d=2, n_cells_coarse=[2, 2], n_levels=3, caching="tensor4",
                            mg_caching="tensor4", mg_precision="fp64",
                            dirichlet={0: ([0, 1], [0.0, 0.0])})
    n_cells = 8 * 8
    S = lsmf.Solver(o, np.ones(n_cells), np.full(n_cells, 1.5))
    S.set_state(np.zeros(S.n_dofs()))
    b = torch.rand(S.n_dofs(), dtype=torch.float64, device=device)
    x = torch.zeros_like(b)
    with pytest.raises(lsmf.LsmfError) as e:
        S.cg(x, b, 1e-12, max_it=max_it)
    assert e.value.code == -7
    its, res = S.cg(x, b, 1e-8 * float(b.norm()), max_it=500)
    assert 0 < its < 500 and res <= 1e-8 * float(b.norm())
