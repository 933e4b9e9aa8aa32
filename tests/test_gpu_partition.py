"""1-vs-K-partition identity of the GPU operator, emulated on one GPU: the
cells are split into K slabs (as the reference's MPI ranks own cell ranges,
include/mf_nh_operator.h:656-705), one operator per slab works on the global
vectors (cells layout: a slab is not the whole structured mesh; the nodes outside
it are referenced by no cell and get zero rows), and the sum
of the slabs' products, diagonals and residuals equals the whole mesh's
(tiled layout) to round-off -- the owner-sum that a multi-GPU vmult's
ghost exchange and reverse add would perform.  Unconstrained operators
(constrained rows would add the identity once per slab).
"""
import numpy as np
import pytest
from types import SimpleNamespace

import lsmf

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def slab_meshes(mesh, k):
    """k slabs of cells along x, each with the global node numbering"""
    cx = mesh.cell_centres()[:, 0]
    edges = np.linspace(cx.min() - 1e-9, cx.max() + 1e-9, k + 1)
    out = []
    for j in range(k):
        sel = np.where((cx >= edges[j]) & (cx < edges[j + 1]))[0]
        out.append((sel, SimpleNamespace(dim=mesh.dim, degree=mesh.degree, coords=mesh.coords,
                                         cell_nodes=np.ascontiguousarray(mesh.cell_nodes[sel]), n_cells=len(sel),
                                         n_nodes=mesh.n_nodes, n_dofs=mesh.n_dofs)))
    return out


@pytest.mark.parametrize("dim,p,n,k", [(3, 2, 6, 3), (3, 4, 3, 2), (2, 2, 8, 4)])
@pytest.mark.parametrize("caching", ["tensor4", "tensor2", "scalar"])
def test_partition_sum_equals_whole(dim, p, n, k, caching, device):
    import torch
    log(f"[partition] dim={dim} p={p} n={n} slabs={k} caching={caching}")
    mesh = lsmf.Mesh(dim, p, [n] * dim)
    rng = np.random.default_rng(8)
    mu = 1.0 + rng.random(mesh.n_cells)
    lam = 1.0 + rng.random(mesh.n_cells)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, 3)
    ud = torch.tensor(u, device=device)
    free = np.zeros(mesh.n_dofs, np.uint8)
    v = torch.tensor(rng.standard_normal(mesh.n_dofs), device=device)

    def run(m, mu_, lam_, layout):
        op = lsmf.NeoHookOperator(m, precision="fp64", caching=caching, constrained=free, layout=layout,
                                  use_mesh_colouring=False)
        op.set_material(mu_, lam_)
        op.set_linearization(ud)
        op.cache()
        y = torch.full_like(v, float("nan"))  # nodes outside the slab must come back zero
        op.vmult(y, v)
        op.compute_diagonal()
        r = torch.full_like(v, float("nan"))
        op.residual(r)
        return y.cpu().numpy(), op.diagonal().cpu().numpy(), r.cpu().numpy()

    whole = run(mesh, mu, lam, "auto")
    parts = [run(sm, mu[sel], lam[sel], "cells") for sel, sm in slab_meshes(mesh, k)]
    for i, name in enumerate(("vmult", "diagonal", "residual")):
        s = sum(pp[i] for pp in parts)
        err = rel(s, whole[i])
        log(f"  {name}: rel diff {err:.2e}")
        assert err < 1e-13, name
