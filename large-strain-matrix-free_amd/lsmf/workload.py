"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

Materials follow the reference's Material_Compressible_Neo_Hook_One_Field
(include/material.h:82-94): lambda = 2 mu nu / (1 - 2 nu).  The reference
selects between a matrix and an inclusion material by material_id
(include/mf_nh_operator.h:930-932); here every cell carries its own (mu,
lambda), with seeded ball inclusions or log-normal fields.

The random stream is SplitMix64 (Steele, Lea, Flood 2014) so the same seed
gives the same workload in Python and C++.
"""
from __future__ import annotations

import math
import numpy as np

__all__ = ["SplitMix64", "lame_from_E_nu", "inclusion_material", "lognormal_material",
           "smooth_displacement", "CONFIGS", "build_config_mesh", "half_refined_box", "renumber_mesh",
           "cuthill_mckee_numbering"]

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, n: int) -> np.ndarray:
        return np.array([(self.next_u64() >> 11) * (1.0 / (1 << 53)) for _ in range(n)])

    def uniform_fast(self, n: int) -> np.ndarray:
        """vectorised SplitMix64 stream, identical values to uniform()"""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = (np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15))
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & _M64
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller on pairs of uniforms"""
        m = (n + 1) // 2
        u = self.uniform_fast(2 * m)
        u1, u2 = u[0::2], u[1::2]
        r = np.sqrt(-2.0 * np.log(1.0 - u1))
        z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
        return z[:n]


def lame_from_E_nu(E, nu):
    mu = E / (2.0 * (1.0 + nu))
    lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
    return mu, lam


def inclusion_material(centres_cells, dim, E, nu, n_inclusions, radius, factor, seed):
    """matrix (E, nu) with seeded ball inclusions of stiffness factor*(mu, lambda).

    A cell belongs to an inclusion when its centre lies inside a ball.
    Returns (mu[n_cells], lambda[n_cells], is_inclusion[n_cells]).
    """
    rng = SplitMix64(seed)
    centres = rng.uniform_fast(n_inclusions * dim).reshape(n_inclusions, dim)
    mu0, lam0 = lame_from_E_nu(E, nu)
    x = np.asarray(centres_cells)
    inside = np.zeros(len(x), dtype=bool)
    chunk = 1 << 16
    r2 = radius * radius
    for s in range(0, len(x), chunk):
        d = x[s:s + chunk, None, :] - centres[None, :, :]
        inside[s:s + chunk] = (np.einsum("cid,cid->ci", d, d) <= r2).any(axis=1)
    mu = np.where(inside, factor * mu0, mu0)
    lam = np.where(inside, factor * lam0, lam0)
    return mu, lam, inside


def lognormal_material(n_cells, E, nu, sigma, seed):
    """log-normal per-cell mu and lambda with medians from (E, nu)"""
    rng = SplitMix64(seed)
    mu0, lam0 = lame_from_E_nu(E, nu)
    z = rng.normal(2 * n_cells)
    return mu0 * np.exp(sigma * z[:n_cells]), lam0 * np.exp(sigma * z[n_cells:])


def smooth_displacement(coords, amplitude, seed, kmax=3):
    """u_i = amplitude * sin(pi k_i . x), integer k_i in [1, kmax]^dim, seeded"""
    dim = coords.shape[1]
    rng = SplitMix64(seed)
    k = 1 + np.floor(rng.uniform_fast(dim * dim) * kmax).astype(np.int64).reshape(dim, dim)
    u = np.empty((coords.shape[0], dim))
    for i in range(dim):
        u[:, i] = amplitude * np.sin(math.pi * coords @ k[i].astype(np.float64))
    return u.reshape(-1), k


SEED = 20261202

# Benchmark configurations: mesh, element and material of each.
CONFIGS = {
    0: dict(dim=2, n=32, degree=2, nu=0.3, n_incl=8, radius=0.05, factor=10.0, caching="tensor4",
            load_steps=5, increment=0.01, mg_levels=5, cheb_degree=4),
    1: dict(dim=3, n=32, degree=2, nu=0.3, n_incl=20, radius=0.08, factor=10.0, caching="tensor4",
            load_steps=5, increment=0.01, amplitude=0.02),
    2: dict(dim=3, n=128, degree=2, nu=0.45, n_incl=200, radius=0.03, factor=10.0, caching="tensor4",
            load_steps=3, increment=0.03, amplitude=0.02, cheb_degree=4),
    3: dict(dim=3, n=48, degree=2, nu=0.3, sigma=0.5, caching="tensor4"),
    4: dict(dim=3, n=64, degree=4, nu=0.45, sigma=0.3, caching="tensor2", load_steps=3, increment=0.03),
}


def build_config_mesh(cfg_id, n=None):
    """mesh + materials of a configuration (n overrides the cells per edge)"""
    from . import Mesh
    cfg = CONFIGS[cfg_id]
    n = cfg["n"] if n is None else n
    dim = cfg["dim"]
    mesh = Mesh(dim, cfg["degree"], [n] * dim)
    if "n_incl" in cfg:
        mu, lam, _ = inclusion_material(mesh.cell_centres(), dim, 1.0, cfg["nu"], cfg["n_incl"], cfg["radius"],
                                        cfg["factor"], SEED + cfg_id)
    else:
        mu, lam = lognormal_material(mesh.n_cells, 1.0, cfg["nu"], cfg["sigma"], SEED + cfg_id)
    return mesh, mu, lam


def half_refined_box(p, n):
    """unit cube whose half x > 1/2 is refined once: n/2 x n x n cells of
    size 1/n and n x 2n x 2n of size 1/(2n), Q_p with Gauss-Lobatto support
    points.  The fine cells' nodes on x = 1/2 that are not nodes of the
    coarse cells hang on the coarse faces: each carries the coarse face's
    Lagrange weights (the hanging-node lines of deal.II's
    make_hanging_node_constraints).  Returns (mesh, (nodes, ptr, masters,
    weights)) in the format of NeoHookOperator's node_constraints."""
    from types import SimpleNamespace
    assert n % 2 == 0
    g = _gll01(p)
    ncx, nf = n // 2, 2 * n
    Ncx, Nc = ncx * p + 1, n * p + 1
    Nfx, Nf = n * p + 1, nf * p + 1
    n_coarse = Ncx * Nc * Nc

    def pos(i, cells, h):
        c = np.minimum(i // p, cells - 1)
        return (c + g[i - c * p]) * h

    # coarse nodes
    ix, iy, iz = np.meshgrid(np.arange(Ncx), np.arange(Nc), np.arange(Nc), indexing="ij")
    cxyz = np.stack([pos(ix, ncx, 1.0 / n), pos(iy, n, 1.0 / n), pos(iz, n, 1.0 / n)], axis=-1)
    cxyz = cxyz.transpose(2, 1, 0, 3).reshape(-1, 3)  # x fastest
    # fine nodes: local id jx + Nfx (jy + Nf jz); x = 1/2 + ...
    jx, jy, jz = np.meshgrid(np.arange(Nfx), np.arange(Nf), np.arange(Nf), indexing="ij")
    fxyz = np.stack([0.5 + pos(jx, n, 0.5 / n), pos(jy, nf, 0.5 / n), pos(jz, nf, 0.5 / n)], axis=-1)
    fxyz = fxyz.transpose(2, 1, 0, 3).reshape(-1, 3)
    gid = np.full(Nfx * Nf * Nf, -1, np.int64)
    interior = np.ones(Nfx * Nf * Nf, bool)
    # interface nodes (jx = 0): coincide with a coarse node, or hang
    hang_local, hang_masters, hang_weights = [], [], []

    def coarse_coord(j):
        """fine interface index -> (coarse cell, local coordinate in [0, 1])"""
        fcell = min(j // p, nf - 1)
        xi = ((fcell % 2) + g[j - fcell * p]) / 2.0
        return fcell // 2, xi

    def lag(k, x):
        v = 1.0
        for m in range(p + 1):
            if m != k:
                v *= (x - g[m]) / (g[k] - g[m])
        return v

    for kz in range(Nf):
        cz, xz = coarse_coord(kz)
        for ky in range(Nf):
            cy, xy = coarse_coord(ky)
            loc = Nfx * (ky + Nf * kz)
            interior[loc] = False
            wy = [lag(b, xy) for b in range(p + 1)]
            wz = [lag(c, xz) for c in range(p + 1)]
            ms, ws = [], []
            for c in range(p + 1):
                for b in range(p + 1):
                    w = wy[b] * wz[c]
                    if abs(w) > 1e-14:
                        ms.append((Ncx - 1) + Ncx * ((cy * p + b) + Nc * (cz * p + c)))
                        ws.append(w)
            if len(ms) == 1 and abs(ws[0] - 1.0) < 1e-12:
                gid[loc] = ms[0]  # a coarse node
            else:
                hang_local.append(loc)
                hang_masters.append(ms)
                hang_weights.append(ws)
    # number the remaining fine nodes after the coarse ones
    rest = np.where(gid < 0)[0]
    gid[rest] = n_coarse + np.arange(len(rest))
    coords = np.empty((n_coarse + len(rest), 3))
    coords[:n_coarse] = cxyz
    coords[n_coarse:] = fxyz[rest]
    N = p + 1
    a = np.arange(N)
    loc3 = (a[None, None, :] + Ncx * (a[None, :, None] + Nc * a[:, None, None])).reshape(-1)
    cc = np.stack(np.meshgrid(np.arange(ncx), np.arange(n), np.arange(n), indexing="ij"), -1).transpose(2, 1, 0, 3)
    cc = cc.reshape(-1, 3)
    cn_c = (cc[:, 0] * p + Ncx * (cc[:, 1] * p + Nc * cc[:, 2] * p))[:, None] + loc3[None, :]
    loc3f = (a[None, None, :] + Nfx * (a[None, :, None] + Nf * a[:, None, None])).reshape(-1)
    fc = np.stack(np.meshgrid(np.arange(n), np.arange(nf), np.arange(nf), indexing="ij"), -1).transpose(2, 1, 0, 3)
    fc = fc.reshape(-1, 3)
    cn_f = gid[(fc[:, 0] * p + Nfx * (fc[:, 1] * p + Nf * fc[:, 2] * p))[:, None] + loc3f[None, :]]
    cell_nodes = np.ascontiguousarray(np.concatenate([cn_c, cn_f]).astype(np.int32))
    nodes = gid[np.array(hang_local, np.int64)].astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum([len(m) for m in hang_masters])]).astype(np.int32)
    masters = np.concatenate([np.array(m, np.int32) for m in hang_masters]).astype(np.int32)
    weights = np.concatenate([np.array(w) for w in hang_weights])
    mesh = SimpleNamespace(dim=3, degree=p, coords=coords, cell_nodes=cell_nodes, n_cells=len(cell_nodes),
                           n_nodes=len(coords), n_dofs=3 * len(coords), cell_colour=None)
    return mesh, (nodes, ptr, masters, weights)


def _gll01(p):
    """Gauss-Lobatto points on [0, 1] (FE_Q's support points)"""
    if p == 1:
        return np.array([0.0, 1.0])
    from numpy.polynomial import legendre as L
    c = np.zeros(p + 1)
    c[-1] = 1
    r = np.sort(L.legroots(L.legder(c)).real)
    return np.concatenate([[0.0], 0.5 * (r + 1), [1.0]])


def renumber_mesh(mesh, perm, order):
    """the same mesh with node k renamed perm[k] and its cells in the order
    `order` (cell i of the result is cell order[i])"""
    from types import SimpleNamespace
    coords = np.empty_like(mesh.coords)
    coords[perm] = mesh.coords
    return SimpleNamespace(dim=mesh.dim, degree=mesh.degree, n_cells=mesh.n_cells, n_nodes=mesh.n_nodes,
                           n_dofs=mesh.n_dofs, cell_nodes=np.ascontiguousarray(perm[mesh.cell_nodes[order]]),
                           coords=coords, cell_colour=None)


def cuthill_mckee_numbering(mesh):
    """a lexicographic 3D box mesh (lsmf.Mesh) numbered as the reference
    numbers its unknowns: nodes in Cuthill-McKee order
    (DoFRenumbering::Cuthill_McKee, include/mf_elasticity.h:1620) -- here
    scipy's reverse Cuthill-McKee of the graph joining nodes adjacent along
    the grid lines, read backwards; the reference orders the full DoF
    coupling graph, whose wavefronts have another shape -- and cells in
    Morton order, the order in which a refined hypercube's active cells are
    visited.  Returns (mesh, perm, order) as renumber_mesh takes them."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    assert mesh.dim == 3
    nn = [c * mesh.degree + 1 for c in mesh.n_cells_dir]
    idx = np.arange(mesh.n_nodes, dtype=np.int64).reshape(nn[2], nn[1], nn[0])
    a, b = [], []
    for ax in range(3):
        v = np.moveaxis(idx, ax, 0)
        a.append(v[:-1].ravel())
        b.append(v[1:].ravel())
    r, c = np.concatenate(a + b), np.concatenate(b + a)
    g = sp.csr_matrix((np.ones(len(r), np.int8), (r.astype(np.int32), c.astype(np.int32))),
                      shape=(mesh.n_nodes, mesh.n_nodes))
    cm = reverse_cuthill_mckee(g, symmetric_mode=True)[::-1]  # position -> old node
    perm = np.empty(mesh.n_nodes, np.int32)
    perm[cm] = np.arange(mesh.n_nodes, dtype=np.int32)
    del g, r, c, a, b, idx
    nc = mesh.n_cells_dir
    k = np.arange(mesh.n_cells, dtype=np.int64)
    xyz = (k % nc[0], (k // nc[0]) % nc[1], k // (nc[0] * nc[1]))
    code = np.zeros(mesh.n_cells, np.int64)
    for bit in range(21):
        for d in range(3):
            code |= ((xyz[d] >> bit) & 1) << (3 * bit + d)
    order = np.argsort(code, kind="stable")
    return renumber_mesh(mesh, perm, order), perm, order
