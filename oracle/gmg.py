"""FP64 CPU restatement of the geometric multigrid preconditioned CG solve.

TEST INFRASTRUCTURE ONLY (checker of the GPU solver; never the product).

Restates, on structured box hierarchies, what the reference wires in
include/mf_elasticity.h:2078-2238 and solves in :3012-3214 with deal.II's
Multigrid / PreconditionChebyshev / MGTransferMatrixFree / SolverCG:
  * level operators linearised at the fine displacement interpolated to each
    level (interpolate_to_mg, :1926-1942), homogeneous Dirichlet rows,
  * Chebyshev smoother with Jacobi preconditioner on
    [1.2 lambda_max / range, 1.2 lambda_max]; lambda_max from a CG/Lanczos
    estimate (eig_cg_n_iterations); vmult from zero uses degree-1 operator
    applications, step uses degree,
  * coarse level: Chebyshev as a solver with the degree for a given
    reduction, eigenvalue bounds from a CG run over all coarse DoFs
    ("MF Chebyshev coarse = true"), or deal.II's MGCoarseGridIterativeSolver
    ("= false", :2115-2127): SolverCG from zero preconditioned by the coarse
    level's Chebyshev smoother, ReductionControl(max(n, 100), 1e-10, 1e-3),
  * V-cycle: pre-smooth, residual, restrict (P^T), coarse correction,
    prolongate-add, post-smooth,
  * FP64 PCG stopping at |r| <= tol.
Transfer matrices are built here from numpy Lagrange polynomials,
independently of the product.
"""
from __future__ import annotations

import math

import numpy as np

from . import OracleOperator


def _gll(p):
    from numpy.polynomial import legendre as L
    if p == 1:
        return np.array([0.0, 1.0])
    c = np.zeros(p + 1)
    c[-1] = 1
    r = np.sort(L.legroots(L.legder(c)).real)
    return 0.5 * (np.concatenate([[-1.0], r, [1.0]]) + 1.0)


def _lagrange(nodes, i, x):
    v = 1.0
    for j, xj in enumerate(nodes):
        if j != i:
            v *= (x - xj) / (nodes[i] - xj)
    return v


def _positions(p, n_cells):
    gl = _gll(p)
    out = []
    for i in range(n_cells * p + 1):
        k = min(i // p, n_cells - 1)
        out.append(k + gl[i - k * p])
    return np.array(out)


def prolongation_1d(p, n_c):
    """fine nodal values of a coarse Q_p function: (2 n_c p + 1) x (n_c p + 1)"""
    gl = _gll(p)
    xf = _positions(p, 2 * n_c) / 2.0
    P = np.zeros((len(xf), n_c * p + 1))
    for i, x in enumerate(xf):
        k = min(int(math.floor(x)), n_c - 1)
        for a in range(p + 1):
            P[i, k * p + a] = _lagrange(gl, a, x - k)
    return P


def interpolation_1d(p, n_c):
    """coarse nodal values of a fine Q_p function"""
    gl = _gll(p)
    xc = _positions(p, n_c) * 2.0
    n_f = 2 * n_c
    R = np.zeros((len(xc), n_f * p + 1))
    for I, x in enumerate(xc):
        k = min(int(math.floor(x + 1e-14)), n_f - 1)
        for a in range(p + 1):
            R[I, k * p + a] = _lagrange(gl, a, x - k)
    return R


def apply_tensor(mats, v, nin, dim):
    """apply 1D matrices per direction to a node-major vector (x fastest, dim comps)"""
    shape = tuple(reversed(nin)) + (dim,)
    a = v.reshape(shape)
    if dim == 2:
        a = np.einsum("Yy,Xx,yxc->YXc", mats[1], mats[0], a)
    else:
        a = np.einsum("Zz,Yy,Xx,zyxc->ZYXc", mats[2], mats[1], mats[0], a)
    return a.reshape(-1)


def pseudo_random(n, seed=20261202):
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def coarsen(dim, n_fine, f):
    a = np.asarray(f).reshape(tuple(reversed(n_fine)))
    if dim == 2:
        return a.reshape(n_fine[1] // 2, 2, n_fine[0] // 2, 2).mean(axis=(1, 3)).reshape(-1)
    return a.reshape(n_fine[2] // 2, 2, n_fine[1] // 2, 2, n_fine[0] // 2, 2).mean(axis=(1, 3, 5)).reshape(-1)


class Level:
    pass


class GMG:
    """levels: list of (mesh, constrained mask) coarsest .. finest"""

    def __init__(self, meshes, cons, mu, lam, caching="tensor4", smoother_degree=5, smoothing_range=20.0,
                 eig_its=30, coarse_reduction=1e-3, coarse_eig_its=0, coarse_solver="chebyshev"):
        self.dim = meshes[-1].dim
        self.p = meshes[-1].degree
        self.opts = dict(smoother_degree=smoother_degree, smoothing_range=smoothing_range, eig_its=eig_its,
                         coarse_reduction=coarse_reduction, coarse_eig_its=coarse_eig_its)
        self.coarse_solver = coarse_solver
        self.levels = []
        mus, lams = [np.asarray(mu, float)], [np.asarray(lam, float)]
        for l in range(len(meshes) - 1, 0, -1):
            mus.insert(0, coarsen(self.dim, meshes[l].n_cells_dir, mus[0]))
            lams.insert(0, coarsen(self.dim, meshes[l].n_cells_dir, lams[0]))
        for l, (m, c) in enumerate(zip(meshes, cons)):
            lv = Level()
            lv.mesh, lv.cons = m, np.asarray(c, np.uint8)
            lv.nn = [k * self.p + 1 for k in m.n_cells_dir]
            lv.op = OracleOperator(self.dim, self.p, m.cell_nodes, m.coords, lv.cons, caching=caching)
            lv.op.set_material(mus[l], lams[l])
            if l > 0:
                nc = meshes[l - 1].n_cells_dir
                lv.P = [prolongation_1d(self.p, nc[d]) for d in range(self.dim)]
                lv.R = [interpolation_1d(self.p, nc[d]) for d in range(self.dim)]
            self.levels.append(lv)

    def setup(self, u):
        L = len(self.levels) - 1
        us = [None] * (L + 1)
        us[L] = np.asarray(u, float)
        for l in range(L, 0, -1):
            us[l - 1] = apply_tensor(self.levels[l].R, us[l], self.levels[l].nn, self.dim)
        for l, lv in enumerate(self.levels):
            lv.op.set_linearization(us[l])
            lv.op.cache()
            lv.invd = 1.0 / lv.op.diagonal()
            if l == 0 and self.coarse_solver == "direct":
                n0 = lv.mesh.n_dofs
                A0 = np.column_stack([lv.op.vmult(np.eye(n0)[:, i]) for i in range(n0)])
                lv.Ainv = np.linalg.inv(0.5 * (A0 + A0.T))
                lv.degree, lv.lmax_est, lv.lmin_est = 0, 0.0, 0.0
                continue
            if l == 0 and self.coarse_solver == "chebyshev":
                its = self.opts["coarse_eig_its"] or lv.mesh.n_dofs
                lv.lmax_est, lv.lmin_est = self.eigen(lv, its)
                lmax = 1.2 * lv.lmax_est
                alpha = min(0.9 * lmax, lv.lmin_est)
                rng = lmax / alpha
                sigma = (1 - math.sqrt(1 / rng)) / (1 + math.sqrt(1 / rng))
                eps = self.opts["coarse_reduction"]
                lv.degree = min(max(1 + int(math.log(1 / eps + math.sqrt(1 / eps / eps - 1)) / math.log(1 / sigma)),
                                    1), 2000)
            else:
                lv.lmax_est, lv.lmin_est = self.eigen(lv, self.opts["eig_its"])
                lmax = 1.2 * lv.lmax_est
                alpha = lmax / self.opts["smoothing_range"]
                lv.degree = self.opts["smoother_degree"]
            lv.theta, lv.delta = 0.5 * (lmax + alpha), 0.5 * (lmax - alpha)

    def eigen(self, lv, its):
        n = lv.mesh.n_dofs
        r = pseudo_random(n)
        r[lv.cons == 1] = 0
        x = np.zeros(n)
        z = lv.invd * r
        p = z.copy()
        rz = r @ z
        r0 = math.sqrt(r @ r)
        alphas, betas = [], []
        for _ in range(its):
            Ap = lv.op.vmult(p)
            pAp = p @ Ap
            if not pAp > 0:
                break
            a = rz / pAp
            alphas.append(a)
            x += a * p
            r -= a * Ap
            if math.sqrt(r @ r) <= 1e-10 * r0:
                break
            z = lv.invd * r
            rzn = r @ z
            b = rzn / rz
            betas.append(b)
            rz = rzn
            p = z + b * p
        m = len(alphas)
        T = np.zeros((m, m))
        T[0, 0] = 1 / alphas[0]
        for k in range(1, m):
            T[k, k] = 1 / alphas[k] + betas[k - 1] / alphas[k - 1]
            T[k, k - 1] = T[k - 1, k] = math.sqrt(betas[k - 1]) / alphas[k - 1]
        ev = np.linalg.eigvalsh(T)
        return ev[-1], ev[0]

    def chebyshev(self, lv, x, b, x_zero):
        if x_zero:
            d = lv.invd * b / lv.theta
            x = d.copy()
        else:
            d = lv.invd * (b - lv.op.vmult(x)) / lv.theta
            x = x + d
        if lv.degree < 2:
            return x
        rhok = lv.delta / lv.theta
        sigma = lv.theta / lv.delta
        for _ in range(lv.degree - 1):
            t = lv.op.vmult(x)
            rhokp = 1 / (2 * sigma - rhok)
            f1, f2 = rhokp * rhok, 2 * rhokp / lv.delta
            rhok = rhokp
            d = f1 * d + f2 * lv.invd * (b - t)
            x = x + d
        return x

    def coarse_cg(self, lv, b):
        """SolverCG with ReductionControl(max(n, 100), 1e-10, 1e-3) from x = 0,
        preconditioned by the level's Chebyshev smoother (vmult from zero)"""
        n = lv.mesh.n_dofs
        x = np.zeros(n)
        r = b.copy()
        r0 = math.sqrt(r @ r)
        tol = max(1e-10, 1e-3 * r0)
        if r0 <= tol:
            return x
        z = self.chebyshev(lv, None, r, True)
        p = z.copy()
        rz = r @ z
        for _ in range(max(n, 100)):
            q = lv.op.vmult(p)
            a = rz / (p @ q)
            x += a * p
            r -= a * q
            if math.sqrt(r @ r) <= tol:
                return x
            z = self.chebyshev(lv, None, r, True)
            rzn = r @ z
            p = z + (rzn / rz) * p
            rz = rzn
        raise RuntimeError("coarse CG: no convergence")

    def vcycle(self, l, b):
        lv = self.levels[l]
        if l == 0:
            if self.coarse_solver == "direct":
                return lv.Ainv @ b
            if self.coarse_solver == "cg":
                return self.coarse_cg(lv, b)
            return self.chebyshev(lv, None, b, True)
        x = self.chebyshev(lv, None, b, True)
        t = b - lv.op.vmult(x)
        c = self.levels[l - 1]
        rc = apply_tensor([P.T for P in lv.P], t, lv.nn, self.dim)
        rc[c.cons == 1] = 0
        xc = self.vcycle(l - 1, rc)
        x = x + apply_tensor(lv.P, xc, c.nn, self.dim)
        x[lv.cons == 1] = 0
        return self.chebyshev(lv, x, b, False)

    def precondition(self, r):
        return self.vcycle(len(self.levels) - 1, r)

    def pcg(self, fine_op, b, tol, max_it=1000):
        x = np.zeros_like(b)
        r = b.copy()
        rn = math.sqrt(r @ r)
        if rn <= tol:
            return x, 0, rn
        z = self.precondition(r)
        p = z.copy()
        rz = r @ z
        it = 0
        while it < max_it:
            it += 1
            Ap = fine_op.vmult(p)
            a = rz / (p @ Ap)
            x += a * p
            r -= a * Ap
            rn = math.sqrt(r @ r)
            if rn <= tol:
                break
            z = self.precondition(r)
            rzn = r @ z
            p = z + (rzn / rz) * p
            rz = rzn
        return x, it, rn
