"""Workload helpers used by the benchmark and the GPU tests (CPU only):
the Cuthill-McKee numbering of a box mesh is a renumbering of the same mesh
and narrows the node-index spread of the cells against a random order."""
import numpy as np

import lsmf


def test_cuthill_mckee_numbering_is_a_renumbering():
    mesh = lsmf.Mesh(3, 2, (6, 5, 4))
    cm, perm, order = lsmf.cuthill_mckee_numbering(mesh)
    assert sorted(perm.tolist()) == list(range(mesh.n_nodes))
    assert sorted(order.tolist()) == list(range(mesh.n_cells))
    # the same cells with the same geometry, only renamed
    assert np.array_equal(cm.coords[cm.cell_nodes], mesh.coords[mesh.cell_nodes[order]])
    # Morton order starts at the origin cell and visits its 2x2x2 block first
    assert order[0] == 0 and sorted(order[:8].tolist()) == [0, 1, 6, 7, 30, 31, 36, 37]
    # bandwidth: a cell's nodes span far fewer positions than under a random order
    rng = np.random.default_rng(0)
    rnd = lsmf.renumber_mesh(mesh, rng.permutation(mesh.n_nodes).astype(np.int32), np.arange(mesh.n_cells))
    spread = lambda m: (m.cell_nodes.max(1) - m.cell_nodes.min(1)).mean()
    assert spread(cm) < 0.5 * spread(rnd)
    # Cuthill-McKee starts at a corner of the grid
    lo, hi = mesh.coords.min(0), mesh.coords.max(0)
    assert all(np.isclose(x, a) or np.isclose(x, b) for x, a, b in zip(cm.coords[0], lo, hi))
