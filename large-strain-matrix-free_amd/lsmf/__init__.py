"""Python binding of liblsmf.so, the MI355X matrix-free Neo-Hookean tangent.

This is a thin ctypes layer over the C ABI declared in ``include/lsmf.h``; the
HIP kernels and the host logic are in the shared library.  It mirrors the
reference's ``NeoHookOperator`` member set (reference:
include/mf_nh_operator.h:147-296):

    op = NeoHookOperator(mesh, n_q_1d=3, precision="fp64", caching="tensor4")
    op.set_material(mu, lam)        # set_material (per-cell arrays)
    op.set_linearization(u)         # the `displacement` of initialize()
    op.cache()                      # cache()
    op.compute_diagonal()           # compute_diagonal()
    op.vmult(dst, src)              # vmult / vmult_add / Tvmult / Tvmult_add

Device vectors are torch tensors on the HIP device (torch is only used for
device memory and streams).  Importing this module never falls back to a CPU
path: if liblsmf.so is missing the import raises.
"""
from __future__ import annotations

import ctypes
import os
import numpy as np

# PyTorch-ROCm ships its own libamdhip64.so.7 / libhsa-runtime64.so.1 (same
# sonames as /opt/rocm).  Loading torch first makes liblsmf.so bind to those
# already-loaded objects, so the process holds a single HIP runtime; loading
# liblsmf.so first would make torch load a second one, and the two do not
# share devices.
try:
    import torch  # noqa: F401
except ImportError:  # CPU-only host logic still works without torch
    pass

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liblsmf.so")

if not os.path.exists(_LIBPATH):
    raise ImportError(
        f"liblsmf.so not found at {_LIBPATH}; build it with `make -C large-strain-matrix-free_amd` "
        "or __graft_entry__.build()")

_lib = ctypes.CDLL(_LIBPATH)

LSMF_OK = 0
ERRORS = {
    -1: "invalid argument", -2: "unknown caching", -3: "out of memory",
    -4: "det F <= 0", -5: "not cached", -6: "HIP error", -7: "not converged",
    -8: "unsupported", -9: "non-positive diagonal",
}
CACHING = {"scalar": 0, "tensor2": 1, "tensor4": 2, "tensor4_ns": 3, "scalar_referential": 4}
PRECISION = {"fp64": 0, "fp32": 1}
MASK = {"zero": 1, "mpi": 2, "rw": 4, "sf": 8, "qd": 16, "default": 0x1F,
        "cellloop": 4 | 8 | 16, "rwsf": 4 | 8, "plain": 32}
GEOMETRY = {"box": 0, "cook": 1}


class LsmfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _MeshDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("degree", ctypes.c_int), ("n_cells", ctypes.c_int),
                ("n_nodes", ctypes.c_int), ("cell_nodes", ctypes.c_void_p), ("coords", ctypes.c_void_p),
                ("constrained", ctypes.c_void_p), ("cell_colour", ctypes.c_void_p),
                ("n_node_constraints", ctypes.c_int), ("nc_nodes", ctypes.c_void_p), ("nc_ptr", ctypes.c_void_p),
                ("nc_masters", ctypes.c_void_p), ("nc_weights", ctypes.c_void_p)]


class _OpOptions(ctypes.Structure):
    _fields_ = [("n_q_1d", ctypes.c_int), ("precision", ctypes.c_int), ("caching", ctypes.c_int),
                ("layout", ctypes.c_int)]


class _Mesh(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("degree", ctypes.c_int), ("n_cells_dir", ctypes.c_int * 3),
                ("n_cells", ctypes.c_int), ("n_nodes", ctypes.c_int),
                ("cell_nodes", ctypes.POINTER(ctypes.c_int32)), ("coords", ctypes.POINTER(ctypes.c_double)),
                ("cell_colour", ctypes.POINTER(ctypes.c_int32))]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_sig = {
    "lsmf_op_create": (_i, [ctypes.POINTER(_MeshDesc), ctypes.POINTER(_OpOptions), ctypes.POINTER(_vp)]),
    "lsmf_op_destroy": (_i, [_vp]),
    "lsmf_set_material_per_cell": (_i, [_vp, _vp, _vp]),
    "lsmf_set_material_two": (_i, [_vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "lsmf_set_linearization": (_i, [_vp, _vp, _vp]),
    "lsmf_set_formulation": (_i, [_vp, _i]),
    "lsmf_cache": (_i, [_vp, _vp]),
    "lsmf_cache_noalloc": (_i, [_vp, _vp]),
    "lsmf_compute_diagonal": (_i, [_vp, _vp]),
    "lsmf_vmult": (_i, [_vp, _vp, _vp, _i, ctypes.c_uint, _vp]),
    "lsmf_vmult_batched": (_i, [_vp, _vp, _vp, _i, _vp]),
    "lsmf_get_diagonal": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "lsmf_precondition_jacobi": (_i, [_vp, _vp, _vp, ctypes.c_double, _vp]),
    "lsmf_op_size": (ctypes.c_size_t, [_vp]),
    "lsmf_memory_consumption": (ctypes.c_size_t, [_vp]),
    "lsmf_residual": (_i, [_vp, _vp, _vp]),
    "lsmf_residual_vjp": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "lsmf_last_error": (ctypes.c_char_p, []),
    "lsmf_profile_enable": (_i, [_vp, _i]),
    "lsmf_profile_read": (_i, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i)]),
    "lsmf_mesh_create": (_i, [_i, _i, _vp, _vp, _vp, _i, ctypes.c_double, ctypes.POINTER(ctypes.POINTER(_Mesh))]),
    "lsmf_mesh_destroy": (_i, [ctypes.POINTER(_Mesh)]),
    "lsmf_mesh_boundary_faces": (_i, [ctypes.POINTER(_Mesh), _i, _vp, _vp]),
    "lsmf_mesh_boundary_nodes": (_i, [ctypes.POINTER(_Mesh), _i, _vp]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = sorted(_sig)


def _check(code):
    if code < 0:
        raise LsmfError(code, _lib.lsmf_last_error().decode())
    return code


def _ptr(a):
    """device or host address of a torch tensor / numpy array"""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def _stream(stream=None):
    if stream is not None:
        return stream
    import torch
    return torch.cuda.current_stream().cuda_stream


class Mesh:
    """Structured mesh built by the library (unit square/cube, Cook membrane).

    Numbering: node(ix,iy,iz) = ix + Nx*(iy + Ny*iz), dof = dim*node + c.
    """

    def __init__(self, dim, degree, n_cells_dir, lo=None, hi=None, geometry="box", scale=1.0):
        n = (ctypes.c_int * 3)(*(list(n_cells_dir) + [1] * (3 - len(n_cells_dir))))
        lo = lo if lo is not None else [0.0] * dim
        hi = hi if hi is not None else [1.0] * dim
        lo_a = (ctypes.c_double * 3)(*(list(lo) + [0.0] * (3 - len(lo))))
        hi_a = (ctypes.c_double * 3)(*(list(hi) + [0.0] * (3 - len(hi))))
        out = ctypes.POINTER(_Mesh)()
        _check(_lib.lsmf_mesh_create(dim, degree, ctypes.addressof(n), ctypes.addressof(lo_a),
                                     ctypes.addressof(hi_a), GEOMETRY[geometry], scale, ctypes.byref(out)))
        m = out.contents
        self.dim, self.degree = m.dim, m.degree
        self.n_cells_dir = tuple(m.n_cells_dir[:dim])
        self.n_cells, self.n_nodes = m.n_cells, m.n_nodes
        npc = (degree + 1) ** dim
        self.cell_nodes = np.ctypeslib.as_array(m.cell_nodes, shape=(self.n_cells, npc)).copy()
        self.coords = np.ctypeslib.as_array(m.coords, shape=(self.n_nodes, dim)).copy()
        self.cell_colour = np.ctypeslib.as_array(m.cell_colour, shape=(self.n_cells,)).copy()
        self._sides_f, self._sides_n = {}, {}
        for s in range(2 * dim):
            nf = _check(_lib.lsmf_mesh_boundary_faces(out, s, None, None))
            cells = np.zeros(nf, np.int32)
            faces = np.zeros(nf, np.int32)
            _lib.lsmf_mesh_boundary_faces(out, s, cells.ctypes.data, faces.ctypes.data)
            self._sides_f[s] = (cells, faces)
            nn = _check(_lib.lsmf_mesh_boundary_nodes(out, s, None))
            nodes = np.zeros(nn, np.int32)
            _lib.lsmf_mesh_boundary_nodes(out, s, nodes.ctypes.data)
            self._sides_n[s] = nodes
        _lib.lsmf_mesh_destroy(out)

    @property
    def n_dofs(self):
        return self.n_nodes * self.dim

    def boundary_faces(self, side):
        return self._sides_f[side]

    def boundary_nodes(self, side):
        return self._sides_n[side]

    def cell_centres(self):
        return self.coords[self.cell_nodes].mean(axis=1)

    def constrained_mask(self, sides_components):
        """sides_components: {side: [components]} -> uint8 mask per DoF"""
        m = np.zeros(self.n_dofs, np.uint8)
        for side, comps in sides_components.items():
            nodes = self.boundary_nodes(side)
            for c in comps:
                m[nodes * self.dim + c] = 1
        return m


class NeoHookOperator:
    """Opaque lsmf_op handle with the reference's member functions."""

    def __init__(self, mesh, n_q_1d=None, precision="fp64", caching="tensor4", constrained=None,
                 use_mesh_colouring=True, formulation=1, layout="auto", node_constraints=None):
        self.dim, self.degree = mesh.dim, mesh.degree
        self.n_q_1d = n_q_1d if n_q_1d is not None else mesh.degree + 1
        self.precision = precision
        self.caching = caching
        self.n_cells, self.n_nodes = mesh.n_cells, mesh.n_nodes
        self._cn = np.ascontiguousarray(mesh.cell_nodes, dtype=np.int32)
        self._xyz = np.ascontiguousarray(mesh.coords, dtype=np.float64)
        self._cons = None if constrained is None else np.ascontiguousarray(constrained, dtype=np.uint8)
        self._col = np.ascontiguousarray(mesh.cell_colour, dtype=np.int32) if use_mesh_colouring else None
        desc = _MeshDesc(self.dim, self.degree, self.n_cells, self.n_nodes, self._cn.ctypes.data,
                         self._xyz.ctypes.data, None if self._cons is None else self._cons.ctypes.data,
                         None if self._col is None else self._col.ctypes.data)
        if node_constraints is not None:
            # (nodes, ptr, masters, weights): hanging-node constraints
            nodes, ptr, masters, weights = node_constraints
            self._nc = [np.ascontiguousarray(nodes, np.int32), np.ascontiguousarray(ptr, np.int32),
                        np.ascontiguousarray(masters, np.int32), np.ascontiguousarray(weights, np.float64)]
            desc.n_node_constraints = len(self._nc[0])
            desc.nc_nodes, desc.nc_ptr, desc.nc_masters, desc.nc_weights = [a.ctypes.data for a in self._nc]
        if caching not in CACHING:
            raise LsmfError(-2, "Unknown caching")
        opt = _OpOptions(self.n_q_1d, PRECISION[precision], CACHING[caching], {"auto": 0, "cells": 1}[layout])
        h = _vp()
        _check(_lib.lsmf_op_create(ctypes.byref(desc), ctypes.byref(opt), ctypes.byref(h)))
        self._h = h
        self.formulation = formulation
        if formulation != 1:
            _check(_lib.lsmf_set_formulation(h, formulation))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lsmf_op_destroy(h)
            self._h = None

    @property
    def dtype(self):
        import torch
        return torch.float64 if self.precision == "fp64" else torch.float32

    def m(self):
        return int(_lib.lsmf_op_size(self._h))

    n = m

    def memory_consumption(self):
        return int(_lib.lsmf_memory_consumption(self._h))

    def set_material(self, mu, lam):
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        assert mu.shape == (self.n_cells,) and lam.shape == (self.n_cells,)
        _check(_lib.lsmf_set_material_per_cell(self._h, mu.ctypes.data, lam.ctypes.data))

    def set_material_two(self, material_id, matrix, inclusion):
        """the reference's set_material(matrix, inclusion): (mu, lambda) of
        the inclusion where material_id != 0, of the matrix elsewhere"""
        mid = np.ascontiguousarray(material_id, dtype=np.uint8)
        assert mid.shape == (self.n_cells,)
        _check(_lib.lsmf_set_material_two(self._h, mid.ctypes.data, float(matrix[0]), float(matrix[1]),
                                          float(inclusion[0]), float(inclusion[1])))

    def set_linearization(self, u, stream=None):
        import torch
        assert u.dtype == torch.float64 and u.is_cuda and u.numel() == self.m()
        _check(_lib.lsmf_set_linearization(self._h, _ptr(u.contiguous()), _stream(stream)))

    def cache(self, stream=None):
        _check(_lib.lsmf_cache(self._h, _stream(stream)))

    def compute_diagonal(self, stream=None):
        _check(_lib.lsmf_compute_diagonal(self._h, _stream(stream)))

    def vmult(self, dst, src, add=False, mask="default", stream=None):
        m = MASK[mask] if isinstance(mask, str) else int(mask)
        assert dst.dtype == self.dtype and src.dtype == self.dtype
        assert dst.numel() == self.m() and src.numel() == self.m()
        _check(_lib.lsmf_vmult(self._h, _ptr(dst), _ptr(src), 1 if add else 0, m, _stream(stream)))

    def vmult_add(self, dst, src, stream=None):
        self.vmult(dst, src, add=True, stream=stream)

    def vmult_batched(self, dst, src, stream=None):
        """dst[b] = A src[b] for every row b of two contiguous (k, m()) tensors,
        in one launch per colour"""
        assert dst.dtype == self.dtype and src.dtype == self.dtype
        assert dst.dim() == 2 and dst.shape == src.shape and dst.shape[1] == self.m()
        assert dst.is_contiguous() and src.is_contiguous()
        _check(_lib.lsmf_vmult_batched(self._h, _ptr(dst), _ptr(src), int(dst.shape[0]), _stream(stream)))

    Tvmult = vmult

    def Tvmult_add(self, dst, src, stream=None):
        self.vmult(dst, src, add=True, stream=stream)

    def _diag_ptrs(self):
        d, inv = _vp(), _vp()
        _check(_lib.lsmf_get_diagonal(self._h, ctypes.byref(d), ctypes.byref(inv)))
        return d.value, inv.value

    def _wrap(self, ptr):
        import torch
        n = self.m()
        es = 8 if self.precision == "fp64" else 4
        out = torch.empty(n, dtype=self.dtype, device="cuda")
        torch.cuda.current_stream().synchronize()
        rc = _hiprt().hipMemcpy(out.data_ptr(), ptr, n * es, 3)  # device to device
        if rc != 0:
            raise LsmfError(-6, f"hipMemcpy failed ({rc})")
        return out

    def diagonal(self):
        """copy of the diagonal (constrained entries 1)"""
        return self._wrap(self._diag_ptrs()[0])

    def inverse_diagonal(self):
        return self._wrap(self._diag_ptrs()[1])

    def el(self, row, col):
        assert row == col, "only diagonal entries are available"
        return float(self.diagonal()[row])

    def precondition_Jacobi(self, dst, src, omega, stream=None):
        _check(_lib.lsmf_precondition_jacobi(self._h, _ptr(dst), _ptr(src), float(omega), _stream(stream)))

    def profile_enable(self, enable=True):
        _check(_lib.lsmf_profile_enable(self._h, 1 if enable else 0))

    def profile_read(self):
        """(summed kernel milliseconds, number of cell-kernel launches) since enable"""
        ms = ctypes.c_double()
        n = ctypes.c_int()
        _check(_lib.lsmf_profile_read(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def residual(self, rhs, stream=None):
        import torch
        assert rhs.dtype == torch.float64 and rhs.numel() == self.m()
        _check(_lib.lsmf_residual(self._h, _ptr(rhs), _stream(stream)))

    def residual_vjp(self, w, du, dmu, dlam, stream=None):
        """du = (dR/du)^T w, dmu = (dR/dmu)^T w, dlam = (dR/dlambda)^T w (device FP64)"""
        import torch
        for t, n in ((w, self.m()), (du, self.m()), (dmu, self.n_cells), (dlam, self.n_cells)):
            assert t.dtype == torch.float64 and t.is_cuda and t.numel() == n and t.is_contiguous()
        _check(_lib.lsmf_residual_vjp(self._h, _ptr(w), _ptr(du), _ptr(dmu), _ptr(dlam), _stream(stream)))


_HIPRT = None


def _hiprt():
    global _HIPRT
    if _HIPRT is None:
        _HIPRT = ctypes.CDLL("libamdhip64.so")
        _HIPRT.hipMemcpy.restype = ctypes.c_int
        _HIPRT.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _HIPRT


from .workload import *  # noqa: E402,F401,F403


# ---------------------------------------------------------------------------
# Newton / GMG-CG solver (lsmf_solver_* in include/lsmf.h)
# ---------------------------------------------------------------------------
class SolverOptions(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("degree", ctypes.c_int), ("n_q_1d", ctypes.c_int),
                ("n_cells_coarse", ctypes.c_int * 3), ("n_levels", ctypes.c_int), ("geometry", ctypes.c_int),
                ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3), ("scale", ctypes.c_double),
                ("caching", ctypes.c_int), ("mg_precision", ctypes.c_int), ("mg_caching", ctypes.c_int),
                ("smoother_degree", ctypes.c_int), ("smoothing_range", ctypes.c_double),
                ("eig_cg_iterations", ctypes.c_int), ("coarse_reduction", ctypes.c_double),
                ("coarse_eig_iterations", ctypes.c_int), ("tol_lin", ctypes.c_double),
                ("max_lin_iterations", ctypes.c_int), ("max_newton", ctypes.c_int), ("tol_f", ctypes.c_double),
                ("tol_u", ctypes.c_double), ("dt", ctypes.c_double), ("t_end", ctypes.c_double),
                ("dirichlet_components", ctypes.c_uint * 6), ("dirichlet_increment", (ctypes.c_double * 3) * 6),
                ("neumann_sides", ctypes.c_uint), ("neumann_traction", (ctypes.c_double * 3) * 6),
                ("verbose", ctypes.c_int), ("coarse_solver", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("preconditioner", ctypes.c_int), ("preconditioner_relaxation", ctypes.c_double),
                ("tol_f_abs", ctypes.c_double), ("tol_u_abs", ctypes.c_double), ("formulation", ctypes.c_int),
                ("cond_number_cg_iterations", ctypes.c_int)]


COARSE = {"chebyshev": 0, "direct": 1, "cg": 2}
PRECONDITIONER = {"gmg": 0, "jacobi": 1, "none": 2}


class StepStats(ctypes.Structure):
    _fields_ = [("newton_iterations", ctypes.c_int), ("linear_iterations", ctypes.c_int),
                ("max_linear_iterations", ctypes.c_int), ("residual_0", ctypes.c_double),
                ("residual_final", ctypes.c_double), ("time_total_s", ctypes.c_double),
                ("time_setup_s", ctypes.c_double), ("time_linear_s", ctypes.c_double),
                ("time_residual_s", ctypes.c_double), ("cond_number", ctypes.c_double),
                ("time_setup_cache_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_ssig = {
    "lsmf_solver_options_default": (_i, [ctypes.POINTER(SolverOptions)]),
    "lsmf_solver_create": (_i, [ctypes.POINTER(SolverOptions), _vp, _vp, ctypes.POINTER(_vp)]),
    "lsmf_solver_destroy": (_i, [_vp]),
    "lsmf_solver_run": (_i, [_vp, _i, ctypes.POINTER(StepStats), ctypes.POINTER(_i)]),
    "lsmf_solver_step": (_i, [_vp, ctypes.POINTER(StepStats), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i)]),
    "lsmf_solver_set_dirichlet_increment": (_i, [_vp, _vp]),
    "lsmf_solver_set_traction": (_i, [_vp, ctypes.POINTER((ctypes.c_double * 3) * 6)]),
    "lsmf_solver_get_solution": (_i, [_vp, _vp]),
    "lsmf_solver_n_dofs": (ctypes.c_size_t, [_vp, _i]),
    "lsmf_solver_set_state": (_i, [_vp, _vp]),
    "lsmf_solver_precondition": (_i, [_vp, _vp, _vp]),
    "lsmf_solver_cg": (_i, [_vp, _vp, _vp, ctypes.c_double, _i, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_double)]),
    "lsmf_solver_level_info": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(_i)]),
    "lsmf_solver_transfer": (_i, [_vp, _i, _i, _vp, _vp]),
    "lsmf_solver_fine_operator": (_vp, [_vp]),
}
for _name, (_res, _args) in _ssig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args
_sig.update(_ssig)
EXPORTED = sorted(_sig)


def solver_options(**kw):
    """lsmf_solver_options with the reference defaults, overridden by kw.
    dirichlet: {side: (components, increments)}, neumann: {side: traction}"""
    o = SolverOptions()
    _check(_lib.lsmf_solver_options_default(ctypes.byref(o)))
    dirichlet = kw.pop("dirichlet", {})
    neumann = kw.pop("neumann", {})
    for k, v in kw.items():
        if k in ("caching", "mg_caching") and isinstance(v, str):
            v = CACHING[v]
        if k == "mg_precision" and isinstance(v, str):
            v = PRECISION[v]
        if k == "geometry" and isinstance(v, str):
            v = GEOMETRY[v]
        if k == "coarse_solver" and isinstance(v, str):
            v = COARSE[v]
        if k == "preconditioner" and isinstance(v, str):
            v = PRECONDITIONER[v]
        if k in ("n_cells_coarse", "lo", "hi"):
            arr = getattr(o, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(o, k, v)
    for side, (comps, incs) in dirichlet.items():
        o.dirichlet_components[side] = sum(1 << c for c in comps)
        for c, val in zip(comps, incs):
            o.dirichlet_increment[side][c] = val
    for side, tr in neumann.items():
        o.neumann_sides |= 1 << side
        for c, val in enumerate(tr):
            o.neumann_traction[side][c] = val
    return o


class Solver:
    """Load-stepping Newton with FP64 CG preconditioned by a GMG V-cycle."""

    def __init__(self, options, mu, lam):
        self.options = options
        self._mu = np.ascontiguousarray(mu, dtype=np.float64)
        self._lam = np.ascontiguousarray(lam, dtype=np.float64)
        h = _vp()
        _check(_lib.lsmf_solver_create(ctypes.byref(options), self._mu.ctypes.data, self._lam.ctypes.data,
                                       ctypes.byref(h)))
        self._h = h
        self.n_levels = options.n_levels

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lsmf_solver_destroy(h)
            self._h = None

    def n_dofs(self, level=-1):
        return int(_lib.lsmf_solver_n_dofs(self._h, level))

    def run(self, max_steps=1000):
        stats = (StepStats * max_steps)()
        n = ctypes.c_int()
        _check(_lib.lsmf_solver_run(self._h, max_steps, stats, ctypes.byref(n)))
        return [stats[k].as_dict() for k in range(n.value)]

    def step(self):
        """one load step; None once the next time exceeds t_end, else
        (t, stats)"""
        st = StepStats()
        t = ctypes.c_double()
        done = ctypes.c_int()
        _check(_lib.lsmf_solver_step(self._h, ctypes.byref(st), ctypes.byref(t), ctypes.byref(done)))
        return None if done.value else (t.value, st.as_dict())

    def set_dirichlet_increment(self, inc):
        """per-DoF increments (host, all finest DoFs) for the next steps"""
        inc = np.ascontiguousarray(inc, dtype=np.float64)
        assert inc.shape == (self.n_dofs(),)
        _check(_lib.lsmf_solver_set_dirichlet_increment(self._h, inc.ctypes.data))

    def set_traction(self, traction):
        """{side: vector} surface tractions applied as given (not scaled by t)"""
        tr = ((ctypes.c_double * 3) * 6)()
        for side, v in traction.items():
            for c, x in enumerate(v):
                tr[side][c] = x
        _check(_lib.lsmf_solver_set_traction(self._h, ctypes.byref(tr)))

    def solution(self):
        u = np.zeros(self.n_dofs())
        _check(_lib.lsmf_solver_get_solution(self._h, u.ctypes.data))
        return u

    def set_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        _check(_lib.lsmf_solver_set_state(self._h, u.ctypes.data))

    def precondition(self, z, r):
        _check(_lib.lsmf_solver_precondition(self._h, _ptr(z), _ptr(r)))

    def cg(self, x, b, tol, max_it=1000):
        its = ctypes.c_int()
        res = ctypes.c_double()
        _check(_lib.lsmf_solver_cg(self._h, _ptr(x), _ptr(b), tol, max_it, ctypes.byref(its), ctypes.byref(res)))
        return its.value, res.value

    def level_info(self, level):
        a, b, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(_lib.lsmf_solver_level_info(self._h, level, ctypes.byref(a), ctypes.byref(b), ctypes.byref(d)))
        return a.value, b.value, d.value

    def transfer(self, level, mode, dst, src):
        _check(_lib.lsmf_solver_transfer(self._h, level, mode, _ptr(dst), _ptr(src)))
