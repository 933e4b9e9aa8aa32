"""Curved meshes for the Eulerian Q_p metric tests.

A quarter annulus r in [1, 2], theta in [0, pi/2] (2D), extruded over
z in [0, 1] (3D), meshed as a structured grid in (r, theta, z) with the
library's lexicographic numbering.  Every support point sits at the exact
polar position of its Gauss-Lobatto coordinate, so the cells are curved
(their edge midpoints are off the chords); with straight=True the support
points are instead placed on the cells' Q1 maps of their vertices (a
straight-sided mesh of the same vertices).
"""
import numpy as np

import lsmf

R0, R1, ANGLE = 1.0, 2.0, 0.5 * np.pi


def polar(xi, dim):
    r = R0 + (R1 - R0) * xi[:, 0]
    th = ANGLE * xi[:, 1]
    X = np.zeros_like(xi)
    X[:, 0] = r * np.cos(th)
    X[:, 1] = r * np.sin(th)
    if dim == 3:
        X[:, 2] = xi[:, 2]
    return X


def annulus(dim, p, n, straight=False):
    mesh = lsmf.Mesh(dim, p, n)  # unit box: support points at the GLL coordinates
    X = polar(mesh.coords, dim)
    if straight:
        nv = 1 << dim
        corner = [sum(((v >> d) & 1) * p * (p + 1) ** d for d in range(dim)) for v in range(nv)]
        for c in range(mesh.n_cells):
            nodes = mesh.cell_nodes[c]
            Xv = X[nodes[corner]]
            xi_box = mesh.coords[nodes]
            lo = xi_box[corner[0]]
            h = xi_box[corner[-1]] - lo
            t = (xi_box - lo) / h  # local coordinates in [0,1]^dim
            w = np.ones((len(nodes), nv))
            for v in range(nv):
                for d in range(dim):
                    w[:, v] *= t[:, d] if (v >> d) & 1 else 1.0 - t[:, d]
            X[nodes] = w @ Xv
    mesh.coords = np.ascontiguousarray(X)
    return mesh


def exact_volume(dim):
    return 0.5 * ANGLE * (R1 ** 2 - R0 ** 2)  # times a unit height in 3D
