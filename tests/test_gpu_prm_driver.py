"""lsmf_run on the GPU: the reference's own parameter files reproduce their
blessed output files byte for byte.

  * tests/cook_mf_form_1.prm (32x32 Q1 Cook membrane, MF_CG + Jacobi 0.65)
  * tests/cook_mf_gmg_form_1.prm (2x2 coarse mesh refined 4 times, GMG)
  * tests/cook_mat_form_1.prm (assembled CG in the reference; the same
    operator here, and the reference's blessed output is the same file)
  * tests/cook_mf_form_1_tensor4_ns.prm (referential dP/dF cache)
  * formulation 0: tests/cook_mf.prm, cook_mf_gmg.prm, cook.prm (2D Q1),
    cook_3d_lin_mf.prm (3D Q1, 30x30x1), cook_3d_quad_mf.prm (3D Q2)
and a unit-square run whose point values agree with the Python solver API
on the same problem.
"""
import os
import subprocess

import numpy as np
import pytest

import lsmf

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PRM = os.path.join(HERE, "golden", "prm")
EXE = os.path.join(os.path.dirname(HERE), "large-strain-matrix-free_amd", "lsmf", "lsmf_run")


def run_driver(prm, cwd):
    print(f"[lsmf_run] {os.path.basename(prm)}", flush=True)
    p = subprocess.run([EXE, prm], capture_output=True, text=True, cwd=cwd, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    return open(os.path.join(cwd, "output")).read(), open(os.path.join(cwd, "timings.txt")).read()


@pytest.mark.parametrize("name", ["cook_mf_form_1", "cook_mf_gmg_form_1", "cook_mat_form_1",
                                  "cook_mf_form_1_tensor4_ns", "cook_mf", "cook_mf_gmg", "cook", "cook_3d_lin_mf",
                                  "cook_3d_quad_mf"])
def test_blessed_output_reproduced(name, tmp_path, device):
    out, timings = run_driver(os.path.join(PRM, name + ".prm"), str(tmp_path))
    blessed = open(os.path.join(PRM, name + ".output")).read()
    assert out == blessed
    n_steps = blessed.count("Converged in")
    assert f"Total CG solve = {4 * n_steps}" in timings  # 4 Newton iterations per step
    assert "| Linear solver" in timings


def test_cg_coarse_solver_from_the_prm_file(tmp_path, device):
    """cook_mf_gmg_form_1 with "MF Chebyshev coarse = false" (the
    reference's CG coarse solver preconditioned by the coarse smoother):
    only the preconditioner changes, so the blessed output is reproduced"""
    src = open(os.path.join(PRM, "cook_mf_gmg_form_1.prm")).read()
    assert "MF Chebyshev coarse" not in src
    prm = src.replace("  set Solver type               = MF_CG\n",
                      "  set Solver type               = MF_CG\n  set MF Chebyshev coarse = false\n")
    assert prm != src
    path = tmp_path / "cook_cg_coarse.prm"
    path.write_text(prm)
    out, timings = run_driver(str(path), str(tmp_path))
    assert out == open(os.path.join(PRM, "cook_mf_gmg_form_1.output")).read()
    assert "| Setup MF: GMG setup" in timings


def test_ad_operator_and_global_coarsening_names(tmp_path, device):
    """cook_mf_gmg_form_1 with "Solver type = MF_AD_CG" (the reference's
    automatic-differentiation operator: the same tangent) and
    "Preconditioner type = gmg-gc" (global coarsening: on uniformly refined
    meshes the same levels) reproduces the blessed output"""
    src = open(os.path.join(PRM, "cook_mf_gmg_form_1.prm")).read()
    prm = src.replace("set Solver type               = MF_CG", "set Solver type               = MF_AD_CG")
    prm = prm.replace("set Preconditioner type        = gmg", "set Preconditioner type        = gmg-gc")
    assert prm.count("MF_AD_CG") == 1 and prm.count("gmg-gc") == 1
    path = tmp_path / "cook_ad_gc.prm"
    path.write_text(prm)
    out, _ = run_driver(str(path), str(tmp_path))
    assert out == open(os.path.join(PRM, "cook_mf_gmg_form_1.output")).read()


def test_unit_square_driver_matches_solver_api(tmp_path, device):
    """2D unit square, 2x2 coarse mesh refined 3 times, Q2, GMG, log-normal
    materials; x = 0 clamped, u_x = 0.01 per step at x = 1, two steps;
    displacements at three points against the Python API's solution"""
    pts = [(1.0, 0.5), (0.5, 0.5), (0.25, 0.875)]
    prm = f"""
subsection Geometry
  set Type = Box
  set Elements per edge = 2
  set Global refinement = 3
  set Grid scale = 1
end
subsection Finite element system
  set Polynomial degree = 2
  set Quadrature order = 3
end
subsection Material properties
  set Formulation = 1
  set Young's modulus = 1
  set Poisson's ratio = 0.3
  set Material field = lognormal
  set Log-normal sigma = 0.5
  set Random seed = 5
end
subsection Linear solver
  set Solver type = MF_CG
  set Preconditioner type = gmg
  set MF caching = tensor4
  set Residual = 1e-10
end
subsection Boundary conditions
  set Dirichlet IDs and expressions = 0:0,0;1:0.01,0
  set Dirichlet IDs and component mask = 0:true,true;1:true,false
end
subsection Time
  set Time step size = 0.5
  set End time = 1
end
subsection Misc
  set Output points = {'; '.join(f'{x}, {y}' for x, y in pts)}
end
"""
    path = tmp_path / "box.prm"
    path.write_text(prm)
    out, _ = run_driver(str(path), str(tmp_path))
    assert out.count("Converged in") == 2
    last = out.split("Timestep 2")[1]
    got = [[float(v) for v in l.split(":")[1].split()] for l in last.splitlines() if "displacement:" in l]

    mesh = lsmf.Mesh(2, 2, [16, 16])
    mu, lam = lsmf.lognormal_material(mesh.n_cells, 1.0, 0.3, 0.5, 5)
    o = lsmf.solver_options(dim=2, degree=2, n_q_1d=3, n_cells_coarse=[2, 2], n_levels=4, caching="tensor4",
                            mg_caching="tensor4", tol_lin=1e-10, dt=0.5, t_end=1.0,
                            dirichlet={0: ([0, 1], [0.0, 0.0]), 1: ([0], [0.01])})
    S = lsmf.Solver(o, mu, lam)
    S.run()
    u = S.solution().reshape(-1, 2)
    for (x, y), g in zip(pts, got):
        node = int(np.argmin(np.sum((mesh.coords - np.array([x, y])) ** 2, axis=1)))
        assert np.allclose(mesh.coords[node], [x, y])
        ref = [float(f"{v:.6g}") for v in u[node]]
        assert np.allclose(g, ref, rtol=2e-5, atol=1e-12), (g, ref)


def test_benchmark_vmult_timings(tmp_path, device):
    """Misc/Benchmark vmult: the reference's per-step vmult timing loop adds
    the vmult (MF) rows of its timings table"""
    src = open(os.path.join(PRM, "cook_mf_gmg_form_1.prm")).read()
    src = src.replace("subsection Time", "subsection Misc\n  set Benchmark vmult = 3\nend\n\nsubsection Time")
    src = src.replace("set End time       = 1", "set End time       = 0.2")
    path = tmp_path / "bench.prm"
    path.write_text(src)
    out, timings = run_driver(str(path), str(tmp_path))
    assert out.count("Converged in 4 Newton iterations") == 2
    for row in ("| vmult (MF) ", "| vmult (MF) Cell loop", "| vmult (MF) RW ", "| vmult (MF) RWSF"):
        line = [l for l in timings.splitlines() if l.startswith(row)]
        assert line and line[0].split("|")[2].strip() == "6", (row, line)
