"""ctypes wrapper of the CPU baseline (cpu_baseline.cpp): the reference's
MatrixFree algorithm with cell batches as SIMD lanes, timed by bench.py.

TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it and
tests/test_cpu_baseline.py checks it against the oracle; never the product.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _cpu_has_avx512, fe_tables

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
CACHING = {"tensor2": 1, "tensor4": 2}


def library_path():
    v4 = os.path.join(_HERE, "cpu_baseline_v4.so")
    v3 = os.path.join(_HERE, "cpu_baseline_v3.so")
    return v4 if _cpu_has_avx512() and os.path.exists(v4) else v3


def lib():
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(f"CPU baseline library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        vp, i = ctypes.c_void_p, ctypes.c_int
        L.cpb_lanes.restype = i
        L.cpb_lanes.argtypes = []
        L.cpb_create.restype = vp
        L.cpb_create.argtypes = [i, i, i, vp, vp, vp, vp, i, i, vp, vp, vp]
        for name, args in (("cpb_destroy", [vp]), ("cpb_set_material", [vp, vp, vp]),
                           ("cpb_set_linearization", [vp, vp]), ("cpb_cache", [vp]), ("cpb_vmult", [vp, vp, vp])):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = args
        _lib = L
    return _lib


class BatchOperator:
    """3D tangent operator (tensor2 or tensor4 caching, formulation 1) on
    cell batches of `lanes` cells; vectors are host numpy arrays"""

    def __init__(self, mesh, cons, caching="tensor4", threads=1):
        if mesh.dim != 3:
            raise ValueError("the CPU baseline covers 3D meshes")
        L = lib()
        p = mesh.degree
        _, qp, qw, S, D = fe_tables(p, p + 1)
        self._tabs = [np.ascontiguousarray(S), np.ascontiguousarray(D), np.concatenate([qw, qp])]
        self._cn = np.ascontiguousarray(mesh.cell_nodes, np.int32)
        self._xyz = np.ascontiguousarray(mesh.coords, np.float64)
        self._cons = np.ascontiguousarray(cons, np.uint8)
        self._col = np.ascontiguousarray(mesh.cell_colour, np.int32)
        self.n_dofs = mesh.n_dofs
        self.lanes = L.cpb_lanes()
        self._h = L.cpb_create(p, mesh.n_cells, mesh.n_nodes, self._cn.ctypes.data, self._xyz.ctypes.data,
                               self._cons.ctypes.data, self._col.ctypes.data, CACHING[caching], int(threads),
                               *[t.ctypes.data for t in self._tabs])
        if not self._h:
            raise ValueError("cpb_create rejected the mesh or caching")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.cpb_destroy(h)
            self._h = None

    def set_material(self, mu, lam):
        self._mu = np.ascontiguousarray(mu, np.float64)
        self._lam = np.ascontiguousarray(lam, np.float64)
        lib().cpb_set_material(self._h, self._mu.ctypes.data, self._lam.ctypes.data)

    def set_linearization(self, u):
        u = np.ascontiguousarray(u, np.float64)
        lib().cpb_set_linearization(self._h, u.ctypes.data)

    def cache(self):
        lib().cpb_cache(self._h)

    def vmult(self, dst, src):
        src = np.ascontiguousarray(src, np.float64)
        assert dst.dtype == np.float64 and dst.flags.c_contiguous and dst.shape == (self.n_dofs,)
        lib().cpb_vmult(self._h, dst.ctypes.data, src.ctypes.data)
        return dst
