"""Oracle on curved meshes (quarter annulus, tests/curved_mesh.py): the
current configuration is the Eulerian Q_p map of all support points
(MappingQEulerian of degree p, include/mf_elasticity.h:1821-1831), the
reference quantities the Q1 vertex map (:1828-1829).  Pinned by closed
forms that fail with a Q1 current metric:

  * at u = 0 the tangent is linear elasticity with (mu, 2 lambda); for the
    isoparametric linear field v = B x (nodal values B X_a) the strain is
    exactly sym(B), so v^T A v = (sym B : C : sym B) |Omega_h| with
    |Omega_h| the volume of the Q_p geometry, which converges to the exact
    annulus volume at order 2p (a Q1 current metric misses it by ~1e-2)
  * 2D: A v vanishes at interior nodes to round-off (the discrete
    divergence theorem: cof(J)^T grad_xi N has degree 2p - 1 per direction,
    integrated exactly by p + 1 Gauss points)
  * straight-sided meshes (support points on the Q1 maps) give the same
    operator as before to round-off, and the residual (Q1 reference map in
    the reference's FEValues assembly) ignores the support-point positions
"""
import numpy as np
import pytest

import curved_mesh
import oracle


def elasticity_energy(B, mu, lam2):
    e = 0.5 * (B + B.T)
    return 2 * mu * np.sum(e * e) + lam2 * np.trace(e) ** 2


def volume_error(dim, p, n, straight):
    rng = np.random.default_rng(1)
    B = 0.1 * rng.standard_normal((dim, dim))
    mu, lam = 0.7, 1.3
    mesh = curved_mesh.annulus(dim, p, n, straight=straight)
    op = oracle.OracleOperator(dim, p, mesh.cell_nodes, mesh.coords, np.zeros(mesh.n_dofs, np.uint8),
                               caching="tensor4")
    op.set_material(np.full(mesh.n_cells, mu), np.full(mesh.n_cells, lam))
    op.set_linearization(np.zeros(mesh.n_dofs))
    op.cache()
    v = (mesh.coords @ B.T).reshape(-1)
    vol = float(v @ op.vmult(v)) / elasticity_energy(B, mu, 2 * lam)
    return abs(vol - curved_mesh.exact_volume(dim)) / curved_mesh.exact_volume(dim)


@pytest.mark.parametrize("dim,p,n", [(2, 2, [2, 4]), (2, 4, [2, 2]), (3, 2, [2, 4, 1])])
def test_constant_strain_energy_is_qp_volume(dim, p, n):
    """the energy of a constant strain measures the Q_p geometry's volume:
    its error against the exact annulus falls at order >= 2p when the angular
    cells are halved; the chords' polygon (straight-sided) misses by ~1e-2"""
    e1 = volume_error(dim, p, n, False)
    n2 = list(n)
    n2[1] *= 2
    e2 = volume_error(dim, p, n2, False)
    assert e2 < e1 / 2 ** (2 * p) * 1.5, (e1, e2)
    assert e2 < (1e-5 if p == 2 else 1e-9), e2
    assert volume_error(dim, p, n, True) > 1e-3


@pytest.mark.parametrize("p,n", [(2, [4, 3]), (4, [2, 3])])
def test_patch_2d_interior_rows_vanish(p, n):
    mesh = curved_mesh.annulus(2, p, n)
    op = oracle.OracleOperator(2, p, mesh.cell_nodes, mesh.coords, np.zeros(mesh.n_dofs, np.uint8),
                               caching="tensor2")
    op.set_material(np.full(mesh.n_cells, 0.7), np.full(mesh.n_cells, 1.3))
    op.set_linearization(np.zeros(mesh.n_dofs))
    op.cache()
    B = np.array([[0.3, -0.2], [0.5, 0.1]])
    f = op.vmult((mesh.coords @ B.T).reshape(-1)).reshape(-1, 2)
    boundary = set()
    for s in range(4):
        boundary.update(mesh.boundary_nodes(s).tolist())
    interior = np.array([i for i in range(mesh.n_nodes) if i not in boundary])
    assert len(interior) > 0
    assert np.abs(f[interior]).max() < 1e-12 * np.abs(f).max()


def test_straight_mesh_and_residual_independent_of_midpoints():
    """straight=True support points reproduce the Q1 metric operator (all
    caching variants); the residual only sees the vertices"""
    dim, p = 3, 2
    curved = curved_mesh.annulus(dim, p, [2, 3, 2])
    straight = curved_mesh.annulus(dim, p, [2, 3, 2], straight=True)
    rng = np.random.default_rng(2)
    u = 0.01 * np.sin(3 * straight.coords).reshape(-1)
    v = rng.standard_normal(straight.n_dofs)
    mu = 1 + rng.random(straight.n_cells)
    lam = 1 + rng.random(straight.n_cells)
    res = []
    for mesh in (curved, straight):
        op = oracle.OracleOperator(dim, p, mesh.cell_nodes, mesh.coords, np.zeros(mesh.n_dofs, np.uint8),
                                   caching="tensor4")
        op.set_material(mu, lam)
        op.set_linearization(u)
        res.append(op.residual())
    assert np.abs(res[0] - res[1]).max() < 1e-13 * np.abs(res[1]).max()
    # every variant agrees with the referential tensor4_ns form on the straight mesh
    out = {}
    for caching in oracle.CACHING:
        op = oracle.OracleOperator(dim, p, straight.cell_nodes, straight.coords,
                                   np.zeros(straight.n_dofs, np.uint8), caching=caching)
        op.set_material(mu, lam)
        op.set_linearization(u)
        op.cache()
        out[caching] = op.vmult(v)
    for k, y in out.items():
        assert np.abs(y - out["tensor4_ns"]).max() < 1e-11 * np.abs(y).max(), k
    # on the curved mesh the spatial variants differ from the referential ones
    op = oracle.OracleOperator(dim, p, curved.cell_nodes, curved.coords, np.zeros(curved.n_dofs, np.uint8),
                               caching="tensor4")
    op.set_material(mu, lam)
    op.set_linearization(u)
    op.cache()
    opr = oracle.OracleOperator(dim, p, curved.cell_nodes, curved.coords, np.zeros(curved.n_dofs, np.uint8),
                                caching="tensor4_ns")
    opr.set_material(mu, lam)
    opr.set_linearization(u)
    opr.cache()
    assert np.abs(op.vmult(v) - opr.vmult(v)).max() > 1e-4 * np.abs(opr.vmult(v)).max()
