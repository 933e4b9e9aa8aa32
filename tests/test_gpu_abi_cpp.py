"""The C ABI driven from C++ with no Python or torch in the process
(lsmf/lsmf_abi_check, built by the package Makefile): initialize ->
set_material -> set_linearization -> cache -> compute_diagonal -> vmult /
vmult_add -> get_diagonal (get_matrix_diagonal_inverse, el) ->
precondition_Jacobi -> m / n / memory_consumption -> destroy, the sequence
the reference's NeoHookOperator users run (include/mf_elasticity.h:1833-1836,
2032-2033, 2109-2110, 3146-3180).  Inputs come from a file; the outputs are
compared with the oracle: vmult and vmult_add within 1e-12 (FP64) / 2e-5
(FP32), the diagonal within 1e-12 / 2e-5, the inverse diagonal and
omega D^-1 src against those diagonals.
"""
import os
import subprocess

import numpy as np
import pytest

import lsmf
import oracle

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "large-strain-matrix-free_amd",
                   "lsmf", "lsmf_abi_check")


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("dim,p,n,caching,precision", [(3, 2, 4, "tensor4", "fp64"), (3, 4, 3, "tensor2", "fp64"),
                                                       (2, 2, 6, "scalar", "fp64"), (3, 2, 3, "tensor4", "fp32")])
def test_cpp_caller_matches_oracle(dim, p, n, caching, precision, tmp_path, device):
    print(f"[abi c++] dim={dim} p={p} n={n} {caching} {precision}", flush=True)
    mesh = lsmf.Mesh(dim, p, [n] * dim)
    mu, lam, _ = lsmf.inclusion_material(mesh.cell_centres(), dim, 1.0, 0.3, 3, 0.25, 10.0, 4)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, 6)
    cons = mesh.constrained_mask({0: list(range(dim)), 1: [0]})
    rng = np.random.default_rng(9)
    src, add0 = rng.random(mesh.n_dofs), rng.random(mesh.n_dofs)
    omega = 0.65
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([dim, p, mesh.n_cells, mesh.n_nodes, lsmf.CACHING[caching], lsmf.PRECISION[precision]],
                 np.int32).tofile(f)
        np.ascontiguousarray(mesh.cell_nodes, np.int32).tofile(f)
        np.ascontiguousarray(mesh.coords, np.float64).tofile(f)
        np.ascontiguousarray(cons, np.uint8).tofile(f)
        for a in (mu, lam, u, src, add0, np.array([omega])):
            np.ascontiguousarray(a, np.float64).tofile(f)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    nd = mesh.n_dofs
    vals = np.fromfile(out, np.float64, count=5 * nd + 1)
    mn = np.fromfile(out, np.int64, offset=(5 * nd + 1) * 8)
    y, z, d, inv, jac = (vals[k * nd:(k + 1) * nd] for k in range(5))
    ref = oracle.OracleOperator(dim, p, mesh.cell_nodes, mesh.coords, cons, caching=caching)
    ref.set_material(mu, lam)
    ref.set_linearization(u)
    ref.cache()
    tol = 1e-12 if precision == "fp64" else 2e-5
    xs = src if precision == "fp64" else src.astype(np.float32).astype(np.float64)
    expect = ref.vmult(xs)
    assert rel(y, expect) < tol
    assert rel(z, add0 + expect) < tol
    dref = ref.diagonal()
    assert rel(d, dref) < tol
    assert np.all(d[cons == 1] == 1.0)
    assert rel(inv, 1.0 / d) < (1e-15 if precision == "fp64" else 1e-6)
    assert rel(jac, omega * xs / d) < (1e-15 if precision == "fp64" else 1e-6)
    assert vals[-1] == d[0]
    assert mn[0] == nd and mn[1] == nd and mn[2] > 0
