"""Meshes with hanging nodes for the node-constraint tests.

A unit coarse cell next to a unit region refined once (2^dim children): the
children's nodes on the interface that are not nodes of the coarse cell
hang on the coarse face, with the coarse face's Lagrange weights (support
points at the Gauss-Lobatto points, as FE_Q).
"""
import itertools
from types import SimpleNamespace

import numpy as np


def _gll(p):
    if p == 1:
        return np.array([0.0, 1.0])
    # Gauss-Lobatto points on [0, 1]: ends and the roots of P'_p
    from numpy.polynomial import legendre as L
    c = np.zeros(p + 1)
    c[-1] = 1
    r = np.sort(L.legroots(L.legder(c)))
    return np.concatenate([[0.0], 0.5 * (r + 1), [1.0]])


def lagrange(nodes, k, x):
    v = 1.0
    for j, xj in enumerate(nodes):
        if j != k:
            v *= (x - xj) / (nodes[k] - xj)
    return v


def build(dim, p):
    """cells: one coarse cell [0,1]^dim, children of [1,2]x[0,1]^(dim-1)"""
    g = _gll(p)
    cells = [(np.zeros(dim), 1.0)]
    for off in itertools.product([0, 1], repeat=dim):
        o = np.array(off, dtype=float) * 0.5
        o[0] += 1.0
        cells.append((o, 0.5))
    key = {}
    coords = []
    cell_nodes = []
    for origin, h in cells:
        ids = []
        for a in itertools.product(range(p + 1), repeat=dim):
            a = a[::-1]  # lexicographic, x fastest
            x = origin + h * np.array([g[a[d]] for d in range(dim)])
            k = tuple(np.round(x, 12))
            if k not in key:
                key[k] = len(coords)
                coords.append(x)
            ids.append(key[k])
        cell_nodes.append(ids)
    coords = np.array(coords)
    cell_nodes = np.array(cell_nodes, dtype=np.int32)
    coarse = set(cell_nodes[0].tolist())
    # hanging: interface nodes (x = 1) of the children that the coarse cell lacks
    face = [n for n in range(len(coords)) if abs(coords[n, 0] - 1.0) < 1e-12 and n in coarse]
    nodes, ptr, masters, weights = [], [0], [], []
    for n in range(len(coords)):
        if abs(coords[n, 0] - 1.0) > 1e-12 or n in coarse:
            continue
        nodes.append(n)
        for m in face:
            w = 1.0
            for d in range(1, dim):
                k = int(np.argmin(np.abs(g - coords[m, d])))
                w *= lagrange(g, k, coords[n, d])
            if abs(w) > 1e-14:
                masters.append(m)
                weights.append(w)
        ptr.append(len(masters))
    mesh = SimpleNamespace(dim=dim, degree=p, coords=coords, cell_nodes=cell_nodes, n_cells=len(cell_nodes),
                           n_nodes=len(coords), n_dofs=len(coords) * dim)
    return mesh, (np.array(nodes, np.int32), np.array(ptr, np.int32), np.array(masters, np.int32),
                  np.array(weights))


def constraint_matrix(mesh, nc):
    """C: u = C u (hanging rows are the weights of their masters, zero diagonal)"""
    nodes, ptr, masters, weights = nc
    n = mesh.n_dofs
    C = np.eye(n)
    for k, h in enumerate(nodes):
        for c in range(mesh.dim):
            C[h * mesh.dim + c, :] = 0.0
            for j in range(ptr[k], ptr[k + 1]):
                C[h * mesh.dim + c, masters[j] * mesh.dim + c] = weights[j]
    return C
