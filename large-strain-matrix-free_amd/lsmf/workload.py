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
           "smooth_displacement", "CONFIGS", "build_config_mesh"]

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
