"""Oracle with hanging-node (general affine) constraints, against numpy.

Mesh: a coarse cell next to a once-refined region (tests/hanging_mesh.py),
2D Q1/Q2/Q4 and 3D Q1/Q2.  With C the constraint matrix (hanging rows hold
the coarse face's Lagrange weights) and D the projection that zeroes
Dirichlet DoFs, the operator is D C^T K C D on the free DoFs and the
identity on Dirichlet and hanging rows (K: numpy assembly of independent
element matrices); the plain product is C^T K C; the diagonal is the
operator's diagonal; the residual is distributed to the masters (C^T R).  Also: the constrained operator
reproduces a field that is continuous across the interface (a Q_p field of
the coarse cell evaluated on the children) -- a check of the weights.
"""
import numpy as np
import pytest

import hanging_mesh
import lsmf
import oracle
from test_oracle_math import assemble_numpy


def setup(dim, p, seed=3):
    mesh, nc = hanging_mesh.build(dim, p)
    rng = np.random.default_rng(seed)
    mu = 1.0 + rng.random(mesh.n_cells)
    lam = 1.5 + rng.random(mesh.n_cells)
    x = mesh.coords
    u = 0.02 * np.sin(np.pi * (x @ np.arange(1, dim + 1)))[:, None] * np.ones(dim)
    u = (u * (1 + 0.3 * np.arange(dim))).reshape(-1)
    C = hanging_mesh.constraint_matrix(mesh, nc)
    u = C @ u  # the linearisation point satisfies the constraints
    cons = np.zeros(mesh.n_dofs, np.uint8)
    left = np.where(np.abs(x[:, 0]) < 1e-12)[0]
    for c in range(dim):
        cons[left * dim + c] = 1
    op = oracle.OracleOperator(dim, p, mesh.cell_nodes, mesh.coords, cons, caching="tensor4")
    op.set_node_constraints(*nc)
    op.set_material(mu, lam)
    op.set_linearization(u)
    op.cache()
    return mesh, nc, C, cons, mu, lam, u, op


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (2, 4), (3, 1), (3, 2)])
def test_constrained_operator_equals_CtKC(dim, p):
    mesh, nc, C, cons, mu, lam, u, op = setup(dim, p)
    assert len(nc[0]) > 0
    K, R = assemble_numpy(mesh, mu, lam, u)
    D = np.diag(1.0 - cons)
    hang = np.zeros(mesh.n_dofs, bool)
    for h in nc[0]:
        hang[h * dim:(h + 1) * dim] = True
    A = D @ C.T @ K @ C @ D
    ident = (cons == 1) | hang
    A[ident, :] = 0.0
    A[np.ix_(ident, ident)] = np.diag(np.ones(ident.sum()))
    v = np.random.default_rng(1).standard_normal(mesh.n_dofs)
    got = op.vmult(v)
    assert np.allclose(got, A @ v, rtol=1e-12, atol=1e-12 * np.abs(A @ v).max())
    # the plain product (Dirichlet lifting) keeps the node constraints
    Ap = C.T @ K @ C
    assert np.allclose(op.vmult(v, plain=True), Ap @ v, rtol=1e-12, atol=1e-12 * np.abs(Ap @ v).max())
    d = op.diagonal()
    assert np.allclose(d, np.diag(A), rtol=1e-12, atol=1e-12 * np.abs(np.diag(A)).max())
    # residual: internal forces distributed to the masters, hanging rows zero
    r = op.residual()
    expect = C.T @ R
    assert np.allclose(r, expect, rtol=1e-11, atol=1e-12 * np.abs(expect).max())
    assert np.all(r[hang] == 0.0)


@pytest.mark.parametrize("dim,p", [(2, 2), (3, 2)])
def test_weights_reproduce_continuous_fields(dim, p):
    """u = C u holds for a field that is one polynomial of degree p in each
    direction on the whole mesh (continuous across the interface)"""
    mesh, nc = hanging_mesh.build(dim, p)
    C = hanging_mesh.constraint_matrix(mesh, nc)
    x = mesh.coords
    zf = (x[:, 2] ** p + 2) if dim == 3 else 1.0
    f = np.stack([(1 + c + x[:, 0] ** p) * (0.5 - x[:, 1] ** p) * zf for c in range(dim)], 1).reshape(-1)
    assert np.allclose(C @ f, f, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("p", [1, 2])
def test_half_refined_box_constraints(p):
    """lsmf.half_refined_box (the hanging-node throughput mesh): every node is
    used, the hanging nodes are the fine interface nodes that are not coarse
    nodes ((2np+1)^2 - (np+1)^2 of them), their weights reproduce fields of
    degree p on the coarse faces, and the constrained operator on it equals
    D C^T K C D on the free DoFs (x = 0 clamped)"""
    n = 2
    mesh, nc = lsmf.half_refined_box(p, n)
    assert mesh.n_cells == (n // 2) * n * n + n * (2 * n) ** 2
    used = np.zeros(mesh.n_nodes, bool)
    used[mesh.cell_nodes.ravel()] = True
    assert used.all()
    assert len(nc[0]) == (2 * n * p + 1) ** 2 - (n * p + 1) ** 2
    assert np.allclose(mesh.coords[nc[0], 0], 0.5)
    C = hanging_mesh.constraint_matrix(mesh, nc)
    x = mesh.coords
    f = np.stack([(1 + c + x[:, 0] ** p) * (0.5 - x[:, 1] ** p) * (x[:, 2] ** p + 2) for c in range(3)], 1)
    assert np.allclose(C @ f.reshape(-1), f.reshape(-1), rtol=1e-13, atol=1e-13)
    cons = np.zeros(mesh.n_dofs, np.uint8)
    left = np.where(np.abs(x[:, 0]) < 1e-12)[0]
    for c in range(3):
        cons[left * 3 + c] = 1
    rng = np.random.default_rng(5)
    mu, lam = 1.0 + rng.random(mesh.n_cells), 1.5 + rng.random(mesh.n_cells)
    u = C @ (0.02 * np.sin(np.pi * (x @ np.arange(1, 4)))[:, None] * np.ones(3)).reshape(-1)
    op = oracle.OracleOperator(3, p, mesh.cell_nodes, mesh.coords, cons, caching="tensor4")
    op.set_node_constraints(*nc)
    op.set_material(mu, lam)
    op.set_linearization(u)
    op.cache()
    K, _ = assemble_numpy(mesh, mu, lam, u)
    D = np.diag(1.0 - cons)
    hang = np.zeros(mesh.n_dofs, bool)
    for h in nc[0]:
        hang[h * 3:(h + 1) * 3] = True
    A = D @ C.T @ K @ C @ D
    ident = (cons == 1) | hang
    A[ident, :] = 0.0
    A[np.ix_(ident, ident)] = np.diag(np.ones(ident.sum()))
    v = rng.standard_normal(mesh.n_dofs)
    assert np.allclose(op.vmult(v), A @ v, rtol=1e-12, atol=1e-12 * np.abs(A @ v).max())
