"""Oracle checks fixed by the mathematics (multi-cell meshes):

  * assembled operator with Dirichlet rows = numpy assembly of independent
    element matrices (constrained rows identity, include/mf_nh_operator.h:703-705)
  * u = 0: linear elasticity with Lame pair (mu, 2 lambda) -- Voigt B^T C B
    element matrices, a different formula from the tangent path
  * symmetry  w.A v = v.A w
  * rigid translations are in the null space of the unconstrained operator
  * patch test: homogeneous deformation + homogeneous material, A (linear
    field) vanishes at interior DoFs
  * central finite differences of the residual equal -A v
  * diagonal = e_i.A e_i; all caching variants agree on a heterogeneous mesh
  * formulation 0 (volumetric/isochoric split): at u = 0 linear elasticity
    with (mu, kappa - 2 mu / d), kappa = lambda + 2 mu / 3; patch test; the
    residual's finite differences equal -A v
"""
import numpy as np
import pytest

import fe_numpy as fe
import lsmf
import oracle

VARIANTS = ["scalar", "tensor2", "tensor4", "tensor4_ns", "scalar_referential"]


def hetero_problem(dim=3, p=2, n=3, seed=5):
    mesh = lsmf.Mesh(dim, p, [n] * dim)
    mu, lam, _ = lsmf.inclusion_material(mesh.cell_centres(), dim, 1.0, 0.3, 3, 0.3, 10.0, seed)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, seed)
    cons = mesh.constrained_mask({0: list(range(dim)), 1: [0]})
    return mesh, mu, lam, u, cons


def make_op(mesh, mu, lam, u, cons, caching, formulation=1):
    op = oracle.OracleOperator(mesh.dim, mesh.degree, mesh.cell_nodes, mesh.coords, cons, caching=caching,
                               formulation=formulation)
    op.set_material(mu, lam)
    op.set_linearization(u)
    op.cache()
    return op


def assemble_numpy(mesh, mu, lam, u):
    dim, p = mesh.dim, mesh.degree
    n = mesh.n_dofs
    K = np.zeros((n, n))
    R = np.zeros(n)
    for c in range(mesh.n_cells):
        nodes = mesh.cell_nodes[c]
        corners = [nodes[sum(((v >> e) & 1) * p * (p + 1) ** e for e in range(dim))] for v in range(2 ** dim)]
        cfe = fe.CellFE(dim, p, p + 1, mesh.coords[corners])
        dofs = (dim * nodes[:, None] + np.arange(dim)[None, :]).reshape(-1)
        Ke, Re = cfe.element_matrix(u[dofs], mu[c], lam[c])
        K[np.ix_(dofs, dofs)] += Ke
        R[dofs] += Re
    return K, R


@pytest.mark.parametrize("dim,p,n", [(2, 2, 3), (3, 1, 2), (3, 2, 2)])
def test_assembled_operator_vs_numpy(dim, p, n):
    mesh, mu, lam, u, cons = hetero_problem(dim, p, n)
    K, R = assemble_numpy(mesh, mu, lam, u)
    free = cons == 0
    v = np.random.default_rng(3).random(mesh.n_dofs)
    expect = np.where(free, K[:, free] @ v[free], v)
    for caching in VARIANTS:
        op = make_op(mesh, mu, lam, u, cons, caching)
        got = op.vmult(v)
        assert np.allclose(got, expect, rtol=1e-11, atol=1e-11 * np.abs(expect).max()), caching
        assert np.allclose(op.vmult(v, plain=True), K @ v, rtol=1e-11, atol=1e-11 * np.abs(K @ v).max())
    assert np.allclose(op.residual(), R, rtol=1e-11, atol=1e-12 * np.abs(R).max())


def voigt_linear_elasticity(cfe, mu, lam_eff):
    dim = cfe.dim
    nv = 3 if dim == 2 else 6
    pairs = [(0, 0), (1, 1), (0, 1)] if dim == 2 else [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    C = np.zeros((nv, nv))
    for a in range(dim):
        for b in range(dim):
            C[a, b] = lam_eff
        C[a, a] += 2 * mu
    for k in range(dim, nv):
        C[k, k] = mu
    n = cfe.nloc * dim
    K = np.zeros((n, n))
    for GradN, jxw in cfe.qp_data():
        B = np.zeros((nv, n))
        for a in range(cfe.nloc):
            for k, (i, j) in enumerate(pairs):
                if i == j:
                    B[k, dim * a + i] = GradN[a, i]
                else:
                    B[k, dim * a + i] = GradN[a, j]
                    B[k, dim * a + j] = GradN[a, i]
        K += B.T @ C @ B * jxw
    return K


@pytest.mark.parametrize("dim,p", [(2, 2), (3, 2), (3, 4)])
def test_zero_displacement_is_linear_elasticity_with_2lambda(dim, p):
    cfe = fe.CellFE(dim, p, p + 1, fe.unit_cell_vertices(dim) * 0.7)
    X = cfe.support_points()
    mu, lam = 1.3, 2.1
    K = voigt_linear_elasticity(cfe, mu, 2 * lam)
    v = np.random.default_rng(7).random(len(X) * dim)
    for caching in VARIANTS:
        op = oracle.OracleOperator(dim, p, np.arange(len(X), dtype=np.int32)[None], X, None, caching=caching)
        op.set_material([mu], [lam])
        op.set_linearization(np.zeros(len(X) * dim))
        op.cache()
        assert np.allclose(op.vmult(v), K @ v, rtol=1e-12, atol=1e-12 * np.abs(K @ v).max()), caching


def test_symmetry_translation_diagonal_and_variants_agree():
    mesh, mu, lam, u, cons = hetero_problem(3, 2, 3)
    rng = np.random.default_rng(11)
    v, w = rng.random(mesh.n_dofs), rng.random(mesh.n_dofs)
    ref = None
    for caching in VARIANTS:
        op = make_op(mesh, mu, lam, u, cons, caching)
        Av, Aw = op.vmult(v), op.vmult(w)
        assert abs(w @ Av - v @ Aw) < 1e-13 * abs(w @ Av)
        t = np.tile([0.3, -0.2, 0.5], mesh.n_nodes)
        At = op.vmult(t, plain=True)
        assert np.abs(At).max() < 1e-11 * np.abs(op.vmult(v, plain=True)).max()
        if ref is None:
            ref = Av
        else:
            assert np.allclose(Av, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), caching
    op = make_op(mesh, mu, lam, u, cons, "tensor4")
    d = op.diagonal()
    for i in rng.choice(mesh.n_dofs, 12, replace=False):
        e = np.zeros(mesh.n_dofs)
        e[i] = 1.0
        expect = 1.0 if cons[i] else op.vmult(e)[i]
        assert abs(d[i] - expect) < 1e-12 * abs(expect)
    assert np.all(d > 0)


def test_patch_test_homogeneous_deformation():
    mesh = lsmf.Mesh(3, 2, [3, 3, 3])
    F0 = np.array([[1.05, 0.02, 0.0], [0.01, 0.97, 0.03], [0.0, -0.02, 1.01]])
    u = (mesh.coords @ (F0 - np.eye(3)).T).reshape(-1)
    mu = np.full(mesh.n_cells, 0.7)
    lam = np.full(mesh.n_cells, 1.1)
    B = np.array([[0.1, 0.2, -0.1], [0.0, 0.3, 0.05], [0.2, -0.1, 0.1]])
    v = (mesh.coords @ B.T).reshape(-1)
    interior = np.all((mesh.coords > 1e-12) & (mesh.coords < 1 - 1e-12), axis=1)
    idof = (3 * np.where(interior)[0][:, None] + np.arange(3)).reshape(-1)
    for caching in VARIANTS:
        op = make_op(mesh, mu, lam, u, None, caching)
        Av = op.vmult(v, plain=True)
        assert np.abs(Av[idof]).max() < 1e-12 * np.abs(Av).max(), caching
        # the residual of a homogeneous stress state also vanishes inside
        r = op.residual()
        assert np.abs(r[idof]).max() < 1e-12 * np.abs(r).max()
    assert abs(op.min_det_F() - np.linalg.det(F0)) < 1e-13


@pytest.mark.parametrize("formulation", [1, 0])
@pytest.mark.parametrize("dim,p", [(2, 2), (3, 2)])
def test_residual_finite_differences_equal_tangent(dim, p, formulation):
    mesh, mu, lam, u, cons = hetero_problem(dim, p, 2)
    v = np.random.default_rng(4).standard_normal(mesh.n_dofs)
    h = 1e-6
    op = make_op(mesh, mu, lam, u, cons, "tensor4", formulation)
    Av = op.vmult(v, plain=True)
    op.set_linearization(u + h * v)
    rp = op.residual()
    op.set_linearization(u - h * v)
    rm = op.residual()
    fd = (rp - rm) / (2 * h)
    assert np.allclose(fd, -Av, rtol=1e-6, atol=1e-7 * np.abs(Av).max())


@pytest.mark.parametrize("dim,p", [(2, 2), (3, 2), (3, 1)])
def test_formulation0_at_zero_displacement_is_linear_elasticity(dim, p):
    """tau_bar = mu I, tau = 0 and J c = 2 mu dev + kappa tr I at F = I"""
    cfe = fe.CellFE(dim, p, p + 1, fe.unit_cell_vertices(dim) * 0.7)
    X = cfe.support_points()
    mu, lam = 1.3, 2.1
    kappa = lam + 2 * mu / 3
    K = voigt_linear_elasticity(cfe, mu, kappa - 2 * mu / dim)
    v = np.random.default_rng(7).random(len(X) * dim)
    op = oracle.OracleOperator(dim, p, np.arange(len(X), dtype=np.int32)[None], X, None, caching="scalar",
                               formulation=0)
    op.set_material([mu], [lam])
    op.set_linearization(np.zeros(len(X) * dim))
    op.cache()
    assert np.allclose(op.vmult(v), K @ v, rtol=1e-12, atol=1e-12 * np.abs(K @ v).max())
    assert np.abs(op.residual()).max() < 1e-13


def test_formulation0_patch_test_and_symmetry():
    mesh = lsmf.Mesh(3, 2, [3, 3, 3])
    F0 = np.array([[1.05, 0.02, 0.0], [0.01, 0.97, 0.03], [0.0, -0.02, 1.01]])
    u = (mesh.coords @ (F0 - np.eye(3)).T).reshape(-1)
    mu, lam = np.full(mesh.n_cells, 0.7), np.full(mesh.n_cells, 1.1)
    B = np.array([[0.1, 0.2, -0.1], [0.0, 0.3, 0.05], [0.2, -0.1, 0.1]])
    v = (mesh.coords @ B.T).reshape(-1)
    interior = np.all((mesh.coords > 1e-12) & (mesh.coords < 1 - 1e-12), axis=1)
    idof = (3 * np.where(interior)[0][:, None] + np.arange(3)).reshape(-1)
    op = make_op(mesh, mu, lam, u, None, "scalar", formulation=0)
    Av = op.vmult(v, plain=True)
    assert np.abs(Av[idof]).max() < 1e-12 * np.abs(Av).max()
    r = op.residual()
    assert np.abs(r[idof]).max() < 1e-12 * np.abs(r).max()
    w = np.random.default_rng(2).random(mesh.n_dofs)
    assert abs(w @ Av - v @ op.vmult(w, plain=True)) < 1e-12 * abs(w @ Av)
