"""Independent numpy finite-element path used to pin the oracle.

Restates the matrix-based side of the reference's known-answer test
tests/mf_neo_hook_operator_form1.cc:301-458 (FEValues on the reference
configuration, grad_Nx = Grad N F^-1, cell_matrix(i,j) += sym(grad_Nx_i) :
act_Jc(sym(grad_Nx_j)) JxW + grad_Nx_i : (grad_Nx_j tau) JxW) and the
residual of include/mf_elasticity.h:2769-2804.  Shape functions, quadrature
and Q1 geometry are computed here from numpy.polynomial, independently of
both the oracle and the product.
"""
from __future__ import annotations

import itertools
import numpy as np
from numpy.polynomial import legendre as L


def gauss01(n):
    x, w = L.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def gll01(p):
    if p == 1:
        return np.array([0.0, 1.0])
    c = np.zeros(p + 1)
    c[-1] = 1.0
    inner = np.sort(L.legroots(L.legder(c)).real)
    return 0.5 * (np.concatenate([[-1.0], inner, [1.0]]) + 1.0)


def lagrange_1d(nodes, x):
    """values V[i] and derivatives Dv[i] of the Lagrange basis at x"""
    n = len(nodes)
    V = np.ones(n)
    Dv = np.zeros(n)
    for i in range(n):
        for j in range(n):
            if j != i:
                V[i] *= (x - nodes[j]) / (nodes[i] - nodes[j])
        s = 0.0
        for k in range(n):
            if k == i:
                continue
            t = 1.0 / (nodes[i] - nodes[k])
            for j in range(n):
                if j != i and j != k:
                    t *= (x - nodes[j]) / (nodes[i] - nodes[j])
            s += t
        Dv[i] = s
    return V, Dv


class CellFE:
    """Q_p vector element on one cell with Q1 (vertex) geometry."""

    def __init__(self, dim, p, nq, vertices):
        self.dim, self.p, self.nq = dim, p, nq
        self.nodes1 = gll01(p)
        self.qp1, self.qw1 = gauss01(nq)
        self.vertices = np.asarray(vertices, float)  # [2^dim][dim], lexicographic
        self.nloc = (p + 1) ** dim
        self.qpts = [np.array(q[::-1]) for q in itertools.product(range(nq), repeat=dim)]  # x fastest
        self.lex = [np.array(a[::-1]) for a in itertools.product(range(p + 1), repeat=dim)]

    def shape_grads_unit(self, xi):
        """[nloc][dim] gradients of scalar shape functions w.r.t. unit coords"""
        vd = [lagrange_1d(self.nodes1, xi[d]) for d in range(self.dim)]
        G = np.zeros((self.nloc, self.dim))
        Vv = np.ones(self.nloc)
        for a, idx in enumerate(self.lex):
            for d in range(self.dim):
                g = 1.0
                for e in range(self.dim):
                    g *= vd[e][1][idx[e]] if e == d else vd[e][0][idx[e]]
                G[a, d] = g
            Vv[a] = np.prod([vd[e][0][idx[e]] for e in range(self.dim)])
        return G, Vv

    def q1_jacobian(self, xi):
        J = np.zeros((self.dim, self.dim))
        for v in range(2 ** self.dim):
            bits = [(v >> e) & 1 for e in range(self.dim)]
            for d in range(self.dim):
                g = 1.0
                for e in range(self.dim):
                    if e == d:
                        g *= 1.0 if bits[e] else -1.0
                    else:
                        g *= xi[e] if bits[e] else 1.0 - xi[e]
                J[:, d] += self.vertices[v] * g
        return J

    def support_points(self):
        """reference coordinates of the support points (Q1 map of GLL points)"""
        pts = []
        for idx in self.lex:
            xi = np.array([self.nodes1[i] for i in idx])
            X = np.zeros(self.dim)
            for v in range(2 ** self.dim):
                bits = [(v >> e) & 1 for e in range(self.dim)]
                w = np.prod([xi[e] if bits[e] else 1 - xi[e] for e in range(self.dim)])
                X += w * self.vertices[v]
            pts.append(X)
        return np.array(pts)

    def qp_data(self):
        """per qp: (Grad N [nloc][dim] on the reference config, JxW_ref)"""
        out = []
        for q in self.qpts:
            xi = np.array([self.qp1[i] for i in q])
            w = np.prod([self.qw1[i] for i in q])
            Gu, _ = self.shape_grads_unit(xi)
            J = self.q1_jacobian(xi)
            out.append((Gu @ np.linalg.inv(J), np.linalg.det(J) * w))
        return out

    # local DoF (a, c) -> index dim*a + c, matching global dof = dim*node + c
    def element_matrix(self, u_loc, mu, lam):
        """tangent of formulation 1 (include/material.h:171-250, act_Jc) and
        the residual -sum tau : sym(grad_Nx) JxW_ref"""
        dim = self.dim
        n = self.nloc * dim
        K = np.zeros((n, n))
        R = np.zeros(n)
        U = u_loc.reshape(self.nloc, dim)
        for GradN, jxw in self.qp_data():
            F = np.eye(dim) + U.T @ GradN
            J = np.linalg.det(F)
            Fi = np.linalg.inv(F)
            b = F @ F.T
            c = mu - 2.0 * lam * np.log(J)
            tau = mu * b - c * np.eye(dim)
            grads = np.zeros((n, dim, dim))
            for a in range(self.nloc):
                gx = GradN[a] @ Fi
                for comp in range(dim):
                    grads[dim * a + comp, comp, :] = gx
            sym = 0.5 * (grads + grads.transpose(0, 2, 1))
            tr = np.trace(sym, axis1=1, axis2=2)
            jc = 2.0 * c * sym + 2.0 * lam * tr[:, None, None] * np.eye(dim)[None]
            geo = np.einsum("jab,bc->jac", grads, tau)
            K += (np.einsum("iab,jab->ij", sym, jc) + np.einsum("iab,jab->ij", grads, geo)) * jxw
            R -= np.einsum("iab,ab->i", sym, tau) * jxw
        return K, R

    def qp_quantities(self, u_loc, v_loc, mu, lam):
        """per qp: Grad u, grad v = Grad v F^-1, Jc action, geo action (the
        quantities compared at 1e-12 in the reference's unit test)"""
        dim = self.dim
        U = u_loc.reshape(self.nloc, dim)
        V = v_loc.reshape(self.nloc, dim)
        res = []
        for GradN, jxw in self.qp_data():
            Gu = U.T @ GradN
            F = np.eye(dim) + Gu
            J = np.linalg.det(F)
            Fi = np.linalg.inv(F)
            c = mu - 2.0 * lam * np.log(J)
            tau = mu * F @ F.T - c * np.eye(dim)
            gv = V.T @ GradN @ Fi
            s = 0.5 * (gv + gv.T)
            jc = 2.0 * c * s + 2.0 * lam * np.trace(s) * np.eye(dim)
            res.append((Gu, gv, jc, gv @ tau, jxw, J))
        return res


def unit_cell_vertices(dim):
    return np.array([[(v >> e) & 1 for e in range(dim)] for v in range(2 ** dim)], float)


# analytic displacement fields of tests/mf_neo_hook_operator_form1.cc:37-176
def field_shear(X):
    u = np.zeros_like(X)
    u[:, 0] = 0.1 * X[:, 1]
    return u


def field_shear2(X):
    u = np.zeros_like(X)
    u[:, 0] = 0.1 * X[:, 1] ** 2
    return u


def field_uniaxial(X):
    u = np.zeros_like(X)
    u[:, 0] = 0.1 * X[:, 0]
    return u


def field_uniaxial2(X):
    u = np.zeros_like(X)
    u[:, 0] = 0.1 * X[:, 0] ** 2
    return u


def field_rotation3d(X):
    """Rotation<3>: rotation about z by angle 0.6 z; the reference's function
    returns R(0.6 z) p (a position), interpolated as the displacement."""
    th = 0.6 * X[:, 2]
    c, s = np.cos(th), np.sin(th)
    out = np.empty_like(X)
    out[:, 0] = c * X[:, 0] - s * X[:, 1]
    out[:, 1] = s * X[:, 0] + c * X[:, 1]
    out[:, 2] = X[:, 2]
    return out
