"""Oracle pinned against the reference's blessed regression outputs.

The fixture tests/golden/reference_outputs.json holds the reference's own
expected outputs (tests/cook_mf_form_1.output, cook_mat_form_1.output,
cook_mf_gmg_form_1.output, cook_mf_form_1_tensor4_ns.output): the Cook
membrane (2D, Q1, 32x32 cells, formulation 1, Neumann shear load t/16e-6 on
the +x face, clamped -x face), 10 load steps, Newton iteration counts and the
displacement of the corner (48e-3, 60e-3) to 6 significant digits.  All MF
caching variants must reproduce them (the reference's outputs are identical
across variants, including the matrix-based solver).
"""
import json
import os

import numpy as np
import pytest

import lsmf
import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_outputs.json")))


def cook_problem(n=32, degree=1):
    mesh = lsmf.Mesh(2, degree, [n, n], lo=[0, 0], hi=[48, 44], geometry="cook", scale=1e-3)
    cons = mesh.constrained_mask({0: [0, 1]})
    return mesh, cons


def round6(x):
    return float(f"{x:.6g}")


@pytest.mark.parametrize("caching", ["scalar", "tensor2", "tensor4", "tensor4_ns", "scalar_referential"])
def test_cook_formulation1_matches_blessed_output(caching):
    ref = GOLD["cook_mf_form_1"]
    prm, out = ref["prm"], ref["output"]
    assert prm["Formulation"] == "1" and prm["Polynomial degree"] == "1"
    mesh, cons = cook_problem(int(prm["Elements per edge"]), int(prm["Polynomial degree"]))
    assert mesh.n_cells == out["n_cells"] and mesh.n_dofs == out["n_dofs"]
    op = oracle.OracleOperator(2, 1, mesh.cell_nodes, mesh.coords, cons, caching=caching)
    assert abs(op.volume() - out["reference_volume"]) < 1e-12
    mu = float(prm["Shear modulus"])
    nu = float(prm["Poisson's ratio"])
    lam = 2 * mu * nu / (1 - 2 * nu)
    op.set_material(np.full(mesh.n_cells, mu), np.full(mesh.n_cells, lam))
    pt = np.array(out["steps"][0]["point"])
    corner = int(np.argmin(np.sum((mesh.coords - pt) ** 2, axis=1)))
    assert np.allclose(mesh.coords[corner], pt)
    its, hist, _ = op.newton(float(prm["Time step size"]), float(prm["End time"]), faces=mesh.boundary_faces(1),
                             traction_per_t=[0.0, 1.0 / 16e-6], tol_f=float(prm["Tolerance force"]),
                             tol_u=float(prm["Tolerance displacement"]), tol_lin=float(prm["Residual"]),
                             max_newton=int(prm["Max iterations Newton-Raphson"]),
                             out_dofs=[2 * corner, 2 * corner + 1])
    assert len(its) == len(out["steps"])
    for k, st in enumerate(out["steps"]):
        assert its[k] == st["newton"], f"step {k + 1}: Newton iterations"
        got = [round6(v) for v in hist[k]]
        assert got == st["displacement"], f"step {k + 1}: {got} vs {st['displacement']}"


def test_blessed_outputs_identical_across_solvers():
    """the fixture itself: MF (scalar), matrix-based, GMG and tensor4_ns runs agree"""
    base = GOLD["cook_mf_form_1"]["output"]
    for name in ["cook_mat_form_1", "cook_mf_gmg_form_1", "cook_mf_form_1_tensor4_ns"]:
        assert GOLD[name]["output"] == base


def cook_case(name):
    """mesh, constrained mask, traction and output DoFs of a Cook fixture:
    2D, or 3D with one element through the thickness z in [-0.5, 0.5] (before
    the 1e-3 scaling); -x face clamped, shear load t/16e-6 on the +x face"""
    ref = GOLD[name]
    prm, out = ref["prm"], ref["output"]
    dim = int(prm.get("Dimension", "2"))
    p = int(prm["Polynomial degree"])
    E = int(prm["Elements per edge"]) << int(prm.get("Global refinement", "0"))
    if dim == 2:
        mesh = lsmf.Mesh(2, p, [E, E], lo=[0, 0], hi=[48, 44], geometry="cook", scale=1e-3)
    else:
        mesh = lsmf.Mesh(3, p, [E, E, 1], lo=[0, 0, -0.5], hi=[48, 44, 0.5], geometry="cook", scale=1e-3)
    cons = mesh.constrained_mask({0: list(range(dim))})
    traction = [0.0, 1.0 / 16e-6] + [0.0] * (dim - 2)
    pt = np.array(out["steps"][0]["point"])
    node = int(np.argmin(np.sum((mesh.coords - pt) ** 2, axis=1)))
    assert np.allclose(mesh.coords[node], pt)
    return mesh, cons, traction, [dim * node + c for c in range(dim)]


@pytest.mark.parametrize("name", ["cook_mf", "cook_3d_lin_mf", "cook_3d_quad_mf"])
def test_cook_formulation0_matches_blessed_output(name):
    """formulation 0 (volumetric/isochoric split, include/material.h:116-141,
    171-226): tests/cook_mf.output (2D Q1; cook.output, the matrix-based
    run, and cook_mf_gmg.output are the same file), cook_3d_lin_mf.output
    (3D Q1, 30x30x1) and cook_3d_quad_mf.output (3D Q2, 12x12x1)"""
    ref = GOLD[name]
    prm, out = ref["prm"], ref["output"]
    assert prm.get("Formulation", "0") == "0"
    mesh, cons, traction, dofs = cook_case(name)
    assert mesh.n_dofs == out["n_dofs"] and mesh.n_cells == out["n_cells"]
    op = oracle.OracleOperator(mesh.dim, mesh.degree, mesh.cell_nodes, mesh.coords, cons, caching="scalar",
                               formulation=0)
    assert abs(op.volume() - out["reference_volume"]) < 1e-12 * max(1.0, out["reference_volume"]) + 1e-15
    mu = float(prm["Shear modulus"])
    nu = float(prm["Poisson's ratio"])
    op.set_material(np.full(mesh.n_cells, mu), np.full(mesh.n_cells, 2 * mu * nu / (1 - 2 * nu)))
    its, hist, _ = op.newton(float(prm["Time step size"]), float(prm["End time"]), faces=mesh.boundary_faces(1),
                             traction_per_t=traction, tol_f=float(prm["Tolerance force"]),
                             tol_u=float(prm["Tolerance displacement"]), tol_lin=float(prm["Residual"]),
                             max_newton=int(prm["Max iterations Newton-Raphson"]), out_dofs=dofs)
    assert len(its) == len(out["steps"])
    for k, st in enumerate(out["steps"]):
        assert its[k] == st["newton"], f"step {k + 1}: Newton iterations"
        got = [round6(v) for v in hist[k]]
        assert got == st["displacement"], f"step {k + 1}: {got} vs {st['displacement']}"


def test_formulation0_fixtures_agree_across_solvers():
    assert GOLD["cook_mf"]["output"] == GOLD["cook_mf_gmg"]["output"] == GOLD["cook"]["output"]
