"""ctypes wrapper of the FP64 CPU oracle (lsmf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker (and the timed CPU
baseline), never the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
CACHING = {"scalar": 0, "tensor2": 1, "tensor4": 2, "tensor4_ns": 3, "scalar_referential": 4}


def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            txt = f.read()
        return all(flag in txt for flag in ("avx512f", "avx512bw", "avx512vl", "avx512dq"))
    except OSError:
        return False


def library_path():
    # LSMF_ORACLE_LIB selects another build of the same source (the
    # AddressSanitizer / UBSan build that the sanitizer test loads)
    override = os.environ.get("LSMF_ORACLE_LIB")
    if override:
        return override
    v4 = os.path.join(_HERE, "lsmf_oracle_v4.so")
    v3 = os.path.join(_HERE, "lsmf_oracle_v3.so")
    if _cpu_has_avx512() and os.path.exists(v4):
        return v4
    return v3


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        sig = {
            "orc_fe_tables": (i, [i, i, vp, vp, vp, vp, vp]),
            "orc_op_create": (vp, [i, i, i, i, i, vp, vp, vp, i]),
            "orc_op_destroy": (None, [vp]),
            "orc_op_set_material": (i, [vp, vp, vp]),
            "orc_op_set_linearization": (i, [vp, vp]),
            "orc_op_cache": (i, [vp]),
            "orc_op_set_formulation": (i, [vp, i]),
            "orc_op_set_node_constraints": (i, [vp, i, vp, vp, vp, vp]),
            "orc_op_vmult": (i, [vp, vp, vp, i, i]),
            "orc_op_diagonal": (i, [vp, vp]),
            "orc_op_residual": (i, [vp, vp]),
            "orc_op_neumann": (i, [vp, vp, vp, i, vp, vp]),
            "orc_op_qp_state": (i, [vp, i, vp, vp]),
            "orc_op_volume": (d, [vp]),
            "orc_op_n_colours": (i, [vp]),
            "orc_op_min_detF": (i, [vp, vp]),
            "orc_num_threads": (i, []),
            "orc_newton": (i, [vp, d, d, vp, i, vp, vp, vp, i, d, d, d, i, vp, vp, vp, vp, i, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def fe_tables(p, nq):
    nodes = np.zeros(p + 1)
    qp = np.zeros(nq)
    qw = np.zeros(nq)
    S = np.zeros((nq, p + 1))
    D = np.zeros((nq, p + 1))
    lib().orc_fe_tables(p, nq, _p(nodes), _p(qp), _p(qw), _p(S), _p(D))
    return nodes, qp, qw, S, D


def num_threads():
    return lib().orc_num_threads()


class OracleOperator:
    """FP64 CPU restatement of NeoHookOperator (formulation 1 or 0)."""

    def __init__(self, dim, degree, cell_nodes, coords, constrained=None, caching="tensor4", n_q_1d=None,
                 formulation=1):
        self.dim, self.degree = dim, degree
        self.nq = degree + 1 if n_q_1d is None else n_q_1d
        self.cell_nodes = np.ascontiguousarray(cell_nodes, dtype=np.int32)
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.n_cells = self.cell_nodes.shape[0]
        self.n_nodes = self.coords.shape[0]
        self.n_dofs = self.n_nodes * dim
        self.constrained = (np.zeros(self.n_dofs, np.uint8) if constrained is None
                            else np.ascontiguousarray(constrained, dtype=np.uint8))
        self._h = lib().orc_op_create(dim, degree, self.nq, self.n_cells, self.n_nodes, _p(self.cell_nodes),
                                      _p(self.coords), _p(self.constrained), CACHING[caching])
        if not self._h:
            raise ValueError(f"oracle: unsupported (dim, degree, nq) = ({dim}, {degree}, {self.nq})")
        self._destroy = lib().orc_op_destroy
        if lib().orc_op_set_formulation(self._h, int(formulation)) != 0:
            raise ValueError(f"oracle: unknown formulation {formulation}")

    def __del__(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def set_node_constraints(self, nodes, ptr, masters, weights):
        """node nodes[k] = sum_j weights[j] * masters[j], j in [ptr[k], ptr[k+1])"""
        self._nc = [np.ascontiguousarray(nodes, np.int32), np.ascontiguousarray(ptr, np.int32),
                    np.ascontiguousarray(masters, np.int32), np.ascontiguousarray(weights, np.float64)]
        rc = lib().orc_op_set_node_constraints(self._h, len(self._nc[0]), *[_p(a) for a in self._nc])
        if rc != 0:
            raise ValueError("oracle: masters must be unconstrained nodes")

    def set_material(self, mu, lam):
        self._mu = np.ascontiguousarray(mu, dtype=np.float64)
        self._lam = np.ascontiguousarray(lam, dtype=np.float64)
        lib().orc_op_set_material(self._h, _p(self._mu), _p(self._lam))

    def set_linearization(self, u):
        self._u = np.ascontiguousarray(u, dtype=np.float64)
        lib().orc_op_set_linearization(self._h, _p(self._u))

    def cache(self):
        rc = lib().orc_op_cache(self._h)
        if rc != 0:
            raise RuntimeError(f"oracle cache failed ({rc}): det F <= 0")

    def vmult(self, src, plain=False):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros(self.n_dofs)
        rc = lib().orc_op_vmult(self._h, _p(dst), _p(src), 0, 1 if plain else 0)
        if rc != 0:
            raise RuntimeError(f"oracle vmult failed ({rc})")
        return dst

    def diagonal(self):
        d = np.zeros(self.n_dofs)
        lib().orc_op_diagonal(self._h, _p(d))
        return d

    def residual(self):
        r = np.zeros(self.n_dofs)
        lib().orc_op_residual(self._h, _p(r))
        return r

    def neumann(self, face_cells, face_ids, traction, rhs=None):
        rhs = np.zeros(self.n_dofs) if rhs is None else rhs
        fc = np.ascontiguousarray(face_cells, dtype=np.int32)
        fi = np.ascontiguousarray(face_ids, dtype=np.int32)
        t = np.ascontiguousarray(list(traction) + [0.0] * (3 - len(traction)), dtype=np.float64)
        lib().orc_op_neumann(self._h, _p(fc), _p(fi), len(fc), _p(t), _p(rhs))
        return rhs

    def qp_state(self, cell, src):
        dd = self.dim * self.dim
        nqp = self.nq ** self.dim
        out = np.zeros((nqp, 4 * dd + 3))
        src = np.ascontiguousarray(src, dtype=np.float64)
        lib().orc_op_qp_state(self._h, cell, _p(src), _p(out))
        return out

    def volume(self):
        return lib().orc_op_volume(self._h)

    def n_colours(self):
        return lib().orc_op_n_colours(self._h)

    def min_det_F(self):
        v = np.zeros(1)
        lib().orc_op_min_detF(self._h, _p(v))
        return float(v[0])

    def newton(self, dt, t_end, dirichlet_inc=None, faces=None, traction_per_t=None, max_newton=10,
               tol_f=1e-9, tol_u=1e-6, tol_lin=1e-6, out_dofs=(), max_steps=1000, verbose=False):
        """Load-stepping Newton with MF_CG + Jacobi (restates Solid::run)."""
        g = np.zeros(self.n_dofs) if dirichlet_inc is None else np.ascontiguousarray(dirichlet_inc, np.float64)
        if faces is None:
            fc = np.zeros(0, np.int32)
            fi = np.zeros(0, np.int32)
        else:
            fc = np.ascontiguousarray(faces[0], np.int32)
            fi = np.ascontiguousarray(faces[1], np.int32)
        tr = np.zeros(3)
        if traction_per_t is not None:
            tr[:len(traction_per_t)] = traction_per_t
        od = np.ascontiguousarray(out_dofs, np.int32)
        n_steps_max = int(round(t_end / dt)) + 2
        its = np.zeros(n_steps_max, np.int32)
        hist = np.zeros((n_steps_max, max(len(od), 1)))
        u = np.zeros(self.n_dofs)
        rc = lib().orc_newton(self._h, dt, t_end, _p(g), len(fc), _p(fc), _p(fi), _p(tr), max_newton, tol_f,
                              tol_u, tol_lin, len(od), _p(od), _p(its), _p(hist), _p(u), min(max_steps, n_steps_max),
                              1 if verbose else 0)
        if rc < 0:
            raise RuntimeError(f"oracle newton failed ({rc})")
        return its[:rc].copy(), hist[:rc, :len(od)].copy(), u
