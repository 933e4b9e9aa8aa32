"""Known-answer tests of the oracle, restating the reference's operator unit
test tests/mf_neo_hook_operator_form1.cc (formulation 1):

  * one hyper_cube cell, Q1 and Q2, 2D and 3D (:185-190, :596-621),
  * displacement fields Shear, Shear2, UniaxialTension, UniaxialTension2 and
    Rotation<3> interpolated at the support points (:37-176, :211-215),
  * mu = 0.4225e6, nu = 0.3 (:268-274),
  * per quadrature point: Grad u and grad v agree to 1e-12 (absolute), the Jc
    and geometric actions to 1e-12 relative (:536-541),
  * per cell: matrix-free vmult against the element matrix,
    |diff| < 1e-10 |value| entry by entry (:552-572).

The element matrix comes from tests/fe_numpy.py (independent numpy path).
"""
import numpy as np
import pytest

import fe_numpy as fe
import oracle

MU = 0.4225e6
NU = 0.3
LAM = 2 * MU * NU / (1 - 2 * NU)

CASES = [
    (2, 1, fe.field_shear), (2, 2, fe.field_shear2), (2, 1, fe.field_uniaxial), (2, 2, fe.field_uniaxial2),
    (3, 1, fe.field_shear), (3, 2, fe.field_shear2), (3, 1, fe.field_uniaxial), (3, 2, fe.field_uniaxial2),
    (3, 1, fe.field_rotation3d),
]


def single_cell(dim, p):
    cfe = fe.CellFE(dim, p, p + 1, fe.unit_cell_vertices(dim))
    X = cfe.support_points()
    cell_nodes = np.arange(len(X), dtype=np.int32)[None, :]
    return cfe, X, cell_nodes


@pytest.mark.parametrize("dim,p,field", CASES, ids=[f"{d}d-Q{p}-{f.__name__}" for d, p, f in CASES])
@pytest.mark.parametrize("caching", ["scalar", "tensor2", "tensor4", "tensor4_ns", "scalar_referential"])
def test_single_cell_vmult_vs_element_matrix(dim, p, field, caching):
    cfe, X, cn = single_cell(dim, p)
    u = field(X).reshape(-1)
    rng = np.random.default_rng(1234)
    src = rng.random(len(u))
    op = oracle.OracleOperator(dim, p, cn, X, None, caching=caching)
    op.set_material([MU], [LAM])
    op.set_linearization(u)
    op.cache()
    mf = op.vmult(src)
    K, _ = cfe.element_matrix(u, MU, LAM)
    mb = K @ src
    assert np.all(np.abs(mf - mb) < 1e-10 * np.abs(mb)), np.max(np.abs(mf - mb) / np.abs(mb))


@pytest.mark.parametrize("dim,p,field", CASES, ids=[f"{d}d-Q{p}-{f.__name__}" for d, p, f in CASES])
def test_single_cell_qp_quantities(dim, p, field):
    cfe, X, cn = single_cell(dim, p)
    u = field(X).reshape(-1)
    src = np.random.default_rng(99).random(len(u))
    op = oracle.OracleOperator(dim, p, cn, X, None, caching="tensor4")
    op.set_material([MU], [LAM])
    op.set_linearization(u)
    op.cache()
    st = op.qp_state(0, src)
    ref = cfe.qp_quantities(u, src, MU, LAM)
    dd = dim * dim
    for q, (Gu, gv, jc, geo, jxw, J) in enumerate(ref):
        row = st[q]
        assert np.linalg.norm(row[:dd].reshape(dim, dim) - Gu) < 1e-12
        assert np.linalg.norm(row[dd:2 * dd].reshape(dim, dim) - gv) < 1e-12
        assert np.linalg.norm(row[2 * dd:3 * dd].reshape(dim, dim) - jc) < 1e-12 * np.linalg.norm(jc)
        assert np.linalg.norm(row[3 * dd:4 * dd].reshape(dim, dim) - geo) < 1e-12 * np.linalg.norm(geo)
        assert abs(row[4 * dd] - jxw) < 1e-14
        # JxW_cur = J JxW_ref (the reference's Debug check, include/mf_nh_operator.h:1131-1156)
        assert abs(row[4 * dd + 1] / row[4 * dd] / J - 1.0) < 1e-12


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_single_cell_residual_vs_element_vector(dim, p):
    cfe, X, cn = single_cell(dim, p)
    u = fe.field_shear2(X).reshape(-1) + 0.05 * fe.field_uniaxial(X).reshape(-1)
    op = oracle.OracleOperator(dim, p, cn, X, None, caching="tensor4")
    op.set_material([MU], [LAM])
    op.set_linearization(u)
    r = op.residual()
    _, R = cfe.element_matrix(u, MU, LAM)
    assert np.allclose(r, R, rtol=1e-12, atol=1e-12 * np.abs(R).max())


@pytest.mark.parametrize("dim,p", [(2, 2), (3, 2)])
def test_single_cell_diagonal_vs_element_matrix(dim, p):
    cfe, X, cn = single_cell(dim, p)
    u = fe.field_shear2(X).reshape(-1)
    for caching in ["scalar", "tensor2", "tensor4", "tensor4_ns", "scalar_referential"]:
        op = oracle.OracleOperator(dim, p, cn, X, None, caching=caching)
        op.set_material([MU], [LAM])
        op.set_linearization(u)
        op.cache()
        K, _ = cfe.element_matrix(u, MU, LAM)
        assert np.allclose(op.diagonal(), np.diag(K), rtol=1e-12)
