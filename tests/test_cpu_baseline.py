"""The CPU baseline that bench.py times (oracle/cpu_baseline.cpp: the
reference's MatrixFree algorithm with cell batches as SIMD lanes) computes
the same operator as the oracle: vmult within 1e-12 relative for tensor2 and
tensor4 caching, Q2 and Q4, with Dirichlet rows and partly filled batches.
"""
import numpy as np
import pytest

import lsmf
import oracle
from oracle import cpu_baseline as cb


@pytest.mark.parametrize("p,n", [(2, 3), (2, 5), (4, 2), (1, 4)])
@pytest.mark.parametrize("caching", ["tensor2", "tensor4"])
def test_batch_vmult_matches_oracle(p, n, caching):
    mesh = lsmf.Mesh(3, p, [n, n + 1, n])
    mu, lam, _ = lsmf.inclusion_material(mesh.cell_centres(), 3, 1.0, 0.45, 3, 0.3, 10.0, 11)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.03, 5)
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    ref = oracle.OracleOperator(3, p, mesh.cell_nodes, mesh.coords, cons, caching=caching)
    ref.set_material(mu, lam)
    ref.set_linearization(u)
    ref.cache()
    v = np.random.default_rng(2).random(mesh.n_dofs)
    expect = ref.vmult(v)
    op = cb.BatchOperator(mesh, cons, caching=caching, threads=2)
    op.set_material(mu, lam)
    op.set_linearization(u)
    op.cache()
    got = op.vmult(np.full(mesh.n_dofs, np.nan), v)
    assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max()


def test_lanes_match_the_isa():
    assert cb.lib().cpb_lanes() in (4, 8)
