"""CPU checks of the C ABI library and of the host logic (no GPU calls).

  * liblsmf.so loads and exports every function include/lsmf.h declares
  * structured meshes: counts, numbering, coordinates, colouring
  * argument errors are reported before any device work (status codes +
    lsmf_last_error, the counterpart of the reference's AssertThrow paths)
"""
import ctypes
import os
import re

import numpy as np
import pytest

import lsmf
import oracle
import fe_numpy as fe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "lsmf.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lsmf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    decl = declared_functions()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(lsmf._lib, name), name
    assert set(decl) == set(lsmf.EXPORTED), set(decl) ^ set(lsmf.EXPORTED)


@pytest.mark.parametrize("dim,p,n", [(2, 1, [3, 4]), (2, 2, [4, 3]), (3, 2, [2, 3, 4]), (3, 4, [2, 2, 2])])
def test_box_mesh_numbering_and_colouring(dim, p, n):
    m = lsmf.Mesh(dim, p, n)
    nn = [k * p + 1 for k in n]
    assert m.n_nodes == int(np.prod(nn)) and m.n_cells == int(np.prod(n))
    gl = fe.gll01(p)
    # node coordinates: lexicographic, Gauss-Lobatto points inside each cell
    idx = np.unravel_index(np.arange(m.n_nodes), nn[::-1])[::-1]
    for d in range(dim):
        i = idx[d]
        k = np.minimum(i // p, n[d] - 1)
        assert np.allclose(m.coords[:, d], (k + gl[i - k * p]) / n[d])
    # colouring: no node shared inside a colour, 2^dim colours
    assert set(m.cell_colour) == set(range(2 ** dim))
    for col in range(2 ** dim):
        nodes = m.cell_nodes[m.cell_colour == col].reshape(-1)
        assert len(np.unique(nodes)) == len(nodes)
    # every node belongs to some cell; cell volumes sum to 1
    assert len(np.unique(m.cell_nodes)) == m.n_nodes
    op = oracle.OracleOperator(dim, p, m.cell_nodes, m.coords)
    assert abs(op.volume() - 1.0) < 1e-13
    # boundary sets
    for s in range(2 * dim):
        nodes = m.boundary_nodes(s)
        assert np.allclose(m.coords[nodes, s // 2], s % 2)
        cells, faces = m.boundary_faces(s)
        assert len(cells) == m.n_cells // n[s // 2] and np.all(faces == s)


def test_cook_mesh_matches_reference_geometry():
    m = lsmf.Mesh(2, 1, [32, 32], lo=[0, 0], hi=[48, 44], geometry="cook", scale=1e-3)
    assert m.n_cells == 1024 and m.n_dofs == 2178
    corners = [(0, 0), (48e-3, 44e-3), (48e-3, 60e-3), (0, 44e-3)]
    for c in corners:
        assert np.min(np.sum((m.coords - np.array(c)) ** 2, axis=1)) < 1e-24


def test_fe_tables_match_numpy():
    for p, nq in [(1, 2), (2, 3), (3, 4), (4, 5)]:
        nodes, qp, qw, S, D = oracle.fe_tables(p, nq)
        x, w = fe.gauss01(nq)
        assert np.allclose(qp, x, atol=1e-15) and np.allclose(qw, w, atol=1e-15)
        assert np.allclose(nodes, fe.gll01(p), atol=1e-15)
        for q in range(nq):
            V, Dv = fe.lagrange_1d(fe.gll01(p), x[q])
            assert np.allclose(S[q], V, atol=1e-14) and np.allclose(D[q], Dv, atol=1e-13)


def test_argument_errors_before_device_work():
    m = lsmf.Mesh(3, 2, [2, 2, 2])
    with pytest.raises(lsmf.LsmfError) as e:
        lsmf.NeoHookOperator(m, caching="nonsense")
    assert e.value.code == -2
    desc = lsmf._MeshDesc(3, 2, m.n_cells, m.n_nodes, m.cell_nodes.ctypes.data, m.coords.ctypes.data, None, None)
    h = ctypes.c_void_p()
    bad_caching = lsmf._OpOptions(3, 0, 17)
    assert lsmf._lib.lsmf_op_create(ctypes.byref(desc), ctypes.byref(bad_caching), ctypes.byref(h)) == -2
    assert b"Unknown caching" in lsmf._lib.lsmf_last_error()
    unsupported = lsmf._OpOptions(4, 0, 2)  # n_q_1d != degree + 1
    assert lsmf._lib.lsmf_op_create(ctypes.byref(desc), ctypes.byref(unsupported), ctypes.byref(h)) == -8
    cn = m.cell_nodes.copy()
    cn[0, 0] = m.n_nodes + 5
    desc_bad = lsmf._MeshDesc(3, 2, m.n_cells, m.n_nodes, cn.ctypes.data, m.coords.ctypes.data, None, None)
    ok = lsmf._OpOptions(3, 0, 2)
    assert lsmf._lib.lsmf_op_create(ctypes.byref(desc_bad), ctypes.byref(ok), ctypes.byref(h)) == -1
    col = m.cell_colour.copy()
    col[1] = col[0]  # neighbours in one colour share nodes
    desc_col = lsmf._MeshDesc(3, 2, m.n_cells, m.n_nodes, m.cell_nodes.ctypes.data, m.coords.ctypes.data, None,
                              col.ctypes.data)
    assert lsmf._lib.lsmf_op_create(ctypes.byref(desc_col), ctypes.byref(ok), ctypes.byref(h)) == -1
    assert b"share a node" in lsmf._lib.lsmf_last_error()
    assert lsmf._lib.lsmf_vmult(None, None, None, 0, 31, None) == -1
    assert lsmf._lib.lsmf_set_material_two(None, None, 1.0, 1.0, 2.0, 2.0) == -1
    assert lsmf._lib.lsmf_vmult_batched(None, None, None, 1, None) == -1


def test_workload_generators_are_seeded():
    r1 = lsmf.SplitMix64(20261202)
    r2 = lsmf.SplitMix64(20261202)
    assert np.array_equal(r1.uniform(50), r2.uniform_fast(50))
    m = lsmf.Mesh(3, 2, [8, 8, 8])
    mu, lam, inc = lsmf.inclusion_material(m.cell_centres(), 3, 1.0, 0.3, 20, 0.08, 10.0, 7)
    mu0, lam0 = lsmf.lame_from_E_nu(1.0, 0.3)
    assert abs(mu0 - 0.3846153846) < 1e-9 and abs(lam0 - 0.5769230769) < 1e-9
    assert np.allclose(mu[~inc], mu0) and np.allclose(mu[inc], 10 * mu0) and inc.any()
    z = lsmf.SplitMix64(3).normal(200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_node_constraint_errors_before_device_work():
    import hanging_mesh
    mesh, (nodes, ptr, masters, weights) = hanging_mesh.build(2, 2)
    bad = masters.copy()
    bad[0] = nodes[1]  # a master that is constrained itself
    with pytest.raises(lsmf.LsmfError) as e:
        lsmf.NeoHookOperator(mesh, use_mesh_colouring=False, node_constraints=(nodes, ptr, bad, weights))
    assert e.value.code == -1 and b"masters must be unconstrained" in lsmf._lib.lsmf_last_error()
    dup = nodes.copy()
    dup[1] = dup[0]
    with pytest.raises(lsmf.LsmfError):
        lsmf.NeoHookOperator(mesh, use_mesh_colouring=False, node_constraints=(dup, ptr, masters, weights))
    cons = np.zeros(mesh.n_dofs, np.uint8)
    cons[nodes[0] * 2] = 1
    with pytest.raises(lsmf.LsmfError):
        lsmf.NeoHookOperator(mesh, use_mesh_colouring=False, constrained=cons,
                             node_constraints=(nodes, ptr, masters, weights))
    assert b"both constrained and Dirichlet" in lsmf._lib.lsmf_last_error()
