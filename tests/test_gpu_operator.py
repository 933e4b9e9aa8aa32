"""GPU parity of the HIP operator against the FP64 oracle, through the C ABI.

Tolerances: FP64 vmult / diagonal / residual within 1e-12 relative (max-norm,
the correctness target for vmult); FP32 level operators within
2e-5 relative (the cache is rounded to FP32 after FP64 kinematics).
Full-size configurations are checked through size-independent properties
(symmetry, null space, agreement between caching variants, add = 2x set).
"""
import numpy as np
import pytest

import lsmf
import oracle

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def problem(dim, p, n, seed=5, amplitude=0.02, n_incl=4):
    mesh = lsmf.Mesh(dim, p, [n] * dim)
    mu, lam, _ = lsmf.inclusion_material(mesh.cell_centres(), dim, 1.0, 0.3, n_incl, 0.25, 10.0, seed)
    u, _ = lsmf.smooth_displacement(mesh.coords, amplitude, seed)
    cons = mesh.constrained_mask({0: list(range(dim)), 1: [0]})
    return mesh, mu, lam, u, cons


def gpu_op(mesh, mu, lam, u, cons, caching, precision="fp64"):
    import torch
    op = lsmf.NeoHookOperator(mesh, precision=precision, caching=caching, constrained=cons)
    op.set_material(mu, lam)
    op.set_linearization(torch.tensor(u, dtype=torch.float64, device="cuda"))
    op.cache()
    return op


def orc_op(mesh, mu, lam, u, cons, caching="tensor4"):
    op = oracle.OracleOperator(mesh.dim, mesh.degree, mesh.cell_nodes, mesh.coords, cons, caching=caching)
    op.set_material(mu, lam)
    op.set_linearization(u)
    op.cache()
    return op


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


CASES = [(3, 2, 4), (3, 2, 5), (2, 1, 6), (2, 2, 5), (3, 1, 4), (3, 4, 2), (2, 4, 3)]


ALL_CACHING = ["scalar", "tensor2", "tensor4", "tensor4_ns", "scalar_referential"]


@pytest.mark.parametrize("dim,p,n", CASES, ids=[f"{d}d-Q{p}-n{n}" for d, p, n in CASES])
@pytest.mark.parametrize("caching", ALL_CACHING)
def test_vmult_fp64_matches_oracle(dim, p, n, caching, device):
    import torch
    log(f"[vmult fp64] dim={dim} p={p} n={n} caching={caching}")
    mesh, mu, lam, u, cons = problem(dim, p, n)
    ref = orc_op(mesh, mu, lam, u, cons, caching)
    op = gpu_op(mesh, mu, lam, u, cons, caching)
    v = np.random.default_rng(1).random(mesh.n_dofs)
    src = torch.tensor(v, device=device)
    dst = torch.full_like(src, np.nan)
    op.vmult(dst, src)
    expect = ref.vmult(v)
    assert rel(dst.cpu().numpy(), expect) < 1e-12
    # vmult_add: dst += A src
    w = np.random.default_rng(2).random(mesh.n_dofs)
    dst2 = torch.tensor(w, device=device)
    op.vmult_add(dst2, src)
    assert rel(dst2.cpu().numpy(), w + expect) < 1e-12
    # plain (no constraints), used for the Dirichlet lifting
    dst3 = torch.full_like(src, np.nan)
    op.vmult(dst3, src, mask="plain")
    assert rel(dst3.cpu().numpy(), ref.vmult(v, plain=True)) < 1e-12


@pytest.mark.parametrize("dim,p,n", [(3, 2, 4), (2, 2, 5), (3, 4, 2)])
@pytest.mark.parametrize("caching", ALL_CACHING)
def test_vmult_fp32_level_operator(dim, p, n, caching, device):
    import torch
    log(f"[vmult fp32] dim={dim} p={p} n={n} caching={caching}")
    mesh, mu, lam, u, cons = problem(dim, p, n)
    ref = orc_op(mesh, mu, lam, u, cons, caching)
    op = gpu_op(mesh, mu, lam, u, cons, caching, precision="fp32")
    v = np.random.default_rng(1).random(mesh.n_dofs)
    dst = torch.zeros(mesh.n_dofs, dtype=torch.float32, device=device)
    op.vmult(dst, torch.tensor(v, dtype=torch.float32, device=device))
    assert rel(dst.cpu().double().numpy(), ref.vmult(v.astype(np.float32).astype(np.float64))) < 2e-5


@pytest.mark.parametrize("dim,p,n", [(3, 2, 3), (2, 2, 4), (3, 1, 3), (3, 4, 2)])
@pytest.mark.parametrize("caching", ALL_CACHING)
def test_diagonal_matches_oracle(dim, p, n, caching, device):
    log(f"[diagonal] dim={dim} p={p} n={n} caching={caching}")
    mesh, mu, lam, u, cons = problem(dim, p, n)
    ref = orc_op(mesh, mu, lam, u, cons, caching).diagonal()
    for prec, tol in [("fp64", 1e-12), ("fp32", 2e-5)]:
        op = gpu_op(mesh, mu, lam, u, cons, caching, precision=prec)
        op.compute_diagonal()
        d = op.diagonal().cpu().double().numpy()
        assert rel(d, ref) < tol, prec
        inv = op.inverse_diagonal().cpu().double().numpy()
        assert rel(inv, 1.0 / ref) < tol, prec
        assert np.all(d[cons == 1] == 1.0)


@pytest.mark.parametrize("dim,p,n", [(3, 2, 4), (2, 1, 6), (2, 2, 4), (3, 4, 2)])
def test_residual_matches_oracle(dim, p, n, device):
    import torch
    log(f"[residual] dim={dim} p={p} n={n}")
    mesh, mu, lam, u, cons = problem(dim, p, n, amplitude=0.05)
    ref = orc_op(mesh, mu, lam, u, cons).residual()
    op = gpu_op(mesh, mu, lam, u, cons, "tensor4")
    r = torch.full((mesh.n_dofs,), np.nan, dtype=torch.float64, device=device)
    op.residual(r)
    assert rel(r.cpu().numpy(), ref) < 1e-12


def test_cook_membrane_operator(device):
    """the reference's Cook geometry (non-affine Q1 cells), Q1 and Q2"""
    import torch
    for p in (1, 2):
        log(f"[cook] p={p}")
        mesh = lsmf.Mesh(2, p, [16, 16], lo=[0, 0], hi=[48, 44], geometry="cook", scale=1e-3)
        cons = mesh.constrained_mask({0: [0, 1]})
        mu = np.full(mesh.n_cells, 0.4225e6)
        lam = 2 * mu * 0.3 / 0.4
        u, _ = lsmf.smooth_displacement(mesh.coords * 20, 0.001, 3)
        for caching in ("scalar", "tensor2", "tensor4"):
            ref = orc_op(mesh, mu, lam, u, cons, caching)
            op = gpu_op(mesh, mu, lam, u, cons, caching)
            v = np.random.default_rng(5).random(mesh.n_dofs)
            dst = torch.zeros(mesh.n_dofs, dtype=torch.float64, device=device)
            op.vmult(dst, torch.tensor(v, device=device))
            assert rel(dst.cpu().numpy(), ref.vmult(v)) < 1e-12, caching


def test_config1_full_size_tensor4(device):
    """32^3 Q2 (823,875 DoF), 20 inclusions, smooth linearisation state"""
    import torch
    log("[config1] 32^3 Q2 tensor4 fp64 vs oracle")
    mesh, mu, lam = lsmf.build_config_mesh(1)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, lsmf.workload.SEED + 1)
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    assert mesh.n_dofs == 823875
    ref = orc_op(mesh, mu, lam, u, cons, "tensor4")
    v = np.random.default_rng(8).random(mesh.n_dofs)
    v[cons == 1] = 0.0
    expect = ref.vmult(v)
    del ref
    for caching in ("tensor4", "tensor2", "scalar"):
        op = gpu_op(mesh, mu, lam, u, cons, caching)
        dst = torch.zeros(mesh.n_dofs, dtype=torch.float64, device=device)
        op.vmult(dst, torch.tensor(v, device=device))
        assert rel(dst.cpu().numpy(), expect) < 1e-12, caching
        del op


def test_nonpositive_det_F_is_reported(device):
    import torch
    mesh, mu, lam, u, cons = problem(3, 2, 2)
    op = lsmf.NeoHookOperator(mesh, caching="tensor4", constrained=cons)
    op.set_material(mu, lam)
    op.set_linearization(torch.tensor(-2.0 * mesh.coords.reshape(-1), device=device))
    with pytest.raises(lsmf.LsmfError) as e:
        op.cache()
    assert e.value.code == -4
    with pytest.raises(lsmf.LsmfError) as e:
        op.vmult(torch.zeros(mesh.n_dofs, dtype=torch.float64, device=device),
                 torch.zeros(mesh.n_dofs, dtype=torch.float64, device=device))
    assert e.value.code == -5


@pytest.mark.parametrize("cfg_id", [2, 4], ids=["config2-128^3-Q2", "config4-64^3-Q4"])
def test_full_size_properties(cfg_id, device):
    """the two 50,923,779-DoF configurations (128^3 Q2 and 64^3 Q4) at full
    size: symmetry, translations in the null space of the unconstrained
    operator, tensor2 = tensor4 = scalar, add = set + set, and the FP32
    operator within 2e-5 of the FP64 one, for the configuration's caching"""
    import torch
    cfg = lsmf.CONFIGS[cfg_id]
    log(f"[config{cfg_id}] {cfg['n']}^3 Q{cfg['degree']} properties")
    mesh, mu, lam = lsmf.build_config_mesh(cfg_id)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, lsmf.workload.SEED + cfg_id)
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    assert mesh.n_dofs == 50923779
    ud = torch.tensor(u, device=device)
    del u
    g = torch.Generator(device=device)
    g.manual_seed(4)
    v = torch.rand(mesh.n_dofs, dtype=torch.float64, device=device, generator=g)
    w = torch.rand(mesh.n_dofs, dtype=torch.float64, device=device, generator=g)
    base_caching = cfg["caching"]
    res = {}
    for caching in [base_caching] + [c for c in ("tensor4", "tensor2", "scalar") if c != base_caching]:
        log(f"[config{cfg_id}] {caching}")
        op = lsmf.NeoHookOperator(mesh, caching=caching, constrained=cons)
        op.set_material(mu, lam)
        op.set_linearization(ud)
        op.cache()
        Av = torch.empty_like(v)
        op.vmult(Av, v)
        if caching == base_caching:
            Aw = torch.empty_like(v)
            op.vmult(Aw, w)
            s1, s2 = float(w @ Av), float(v @ Aw)
            assert abs(s1 - s2) < 1e-12 * abs(s1)
            t = torch.tensor([0.3, -0.2, 0.5], dtype=torch.float64, device=device).repeat(mesh.n_nodes)
            At = torch.empty_like(v)
            op.vmult(At, t, mask="plain")
            Avp = torch.empty_like(v)
            op.vmult(Avp, v, mask="plain")
            assert float(At.abs().max()) < 1e-11 * float(Avp.abs().max())
            acc = Av.clone()
            op.vmult_add(acc, v)
            assert float((acc - 2 * Av).abs().max()) <= 1e-14 * float(Av.abs().max())
        res[caching] = Av
        del op
    base = res[base_caching]
    for k in res:
        assert float((res[k] - base).abs().max()) < 1e-12 * float(base.abs().max()), k
    log(f"[config{cfg_id}] {base_caching} fp32")
    op = lsmf.NeoHookOperator(mesh, precision="fp32", caching=base_caching, constrained=cons)
    op.set_material(mu, lam)
    op.set_linearization(ud)
    op.cache()
    v32 = v.float()
    Av32 = torch.empty_like(v32)
    op.vmult(Av32, v32)
    del op
    ref = torch.empty_like(v)
    op = lsmf.NeoHookOperator(mesh, caching=base_caching, constrained=cons)
    op.set_material(mu, lam)
    op.set_linearization(ud)
    op.cache()
    op.vmult(ref, v32.double())
    assert float((Av32.double() - ref).abs().max()) < 2e-5 * float(ref.abs().max())


@pytest.mark.parametrize("caching", ["tensor2", "tensor4", "scalar"])
def test_config4_problem_small_mesh_matches_oracle(caching, device):
    """the 64^3 Q4 configuration's inputs (nu = 0.45, log-normal mu/lambda
    with sigma = 0.3, smooth linearisation state) on 4^3 Q4 cells: FP64
    vmult, diagonal and residual against the oracle to 1e-12, FP32 to 2e-5"""
    import torch
    log(f"[config4 4^3] {caching}")
    mesh, mu, lam = lsmf.build_config_mesh(4, n=4)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, lsmf.workload.SEED + 4)
    cons = mesh.constrained_mask({0: [0, 1, 2], 1: [0]})
    ref = orc_op(mesh, mu, lam, u, cons, caching)
    v = np.random.default_rng(3).random(mesh.n_dofs)
    v[cons == 1] = 0.0
    expect = ref.vmult(v)
    for prec, tol in [("fp64", 1e-12), ("fp32", 2e-5)]:
        op = gpu_op(mesh, mu, lam, u, cons, caching, precision=prec)
        dt = torch.float64 if prec == "fp64" else torch.float32
        dst = torch.full((mesh.n_dofs,), np.nan, dtype=dt, device=device)
        op.vmult(dst, torch.tensor(v, dtype=dt, device=device))
        vv = v if prec == "fp64" else v.astype(np.float32).astype(np.float64)
        assert rel(dst.cpu().double().numpy(), expect if prec == "fp64" else ref.vmult(vv)) < tol, prec
        op.compute_diagonal()
        assert rel(op.diagonal().cpu().double().numpy(), ref.diagonal()) < tol, prec
    r = torch.full((mesh.n_dofs,), np.nan, dtype=torch.float64, device=device)
    gpu_op(mesh, mu, lam, u, cons, caching).residual(r)
    assert rel(r.cpu().numpy(), ref.residual()) < 1e-12


F0_CASES = [(3, 2, 4), (2, 1, 6), (2, 2, 5), (3, 1, 4), (3, 4, 2)]


@pytest.mark.parametrize("dim,p,n", F0_CASES, ids=[f"{d}d-Q{p}-n{n}" for d, p, n in F0_CASES])
@pytest.mark.parametrize("caching", ["scalar", "tensor4"])
def test_formulation0_matches_oracle(dim, p, n, caching, device):
    """formulation 0 (include/mf_nh_operator.h:467-487, 1001-1181): vmult,
    diagonal and residual in FP64 to 1e-12, FP32 level operator to 2e-5"""
    import torch
    log(f"[formulation 0] dim={dim} p={p} n={n} caching={caching}")
    mesh, mu, lam, u, cons = problem(dim, p, n, amplitude=0.03)
    ref = oracle.OracleOperator(mesh.dim, mesh.degree, mesh.cell_nodes, mesh.coords, cons, caching="scalar",
                                formulation=0)
    ref.set_material(mu, lam)
    ref.set_linearization(u)
    ref.cache()
    v = np.random.default_rng(1).random(mesh.n_dofs)
    expect = ref.vmult(v)
    for prec, tol in [("fp64", 1e-12), ("fp32", 2e-5)]:
        op = lsmf.NeoHookOperator(mesh, precision=prec, caching=caching, constrained=cons, formulation=0)
        op.set_material(mu, lam)
        op.set_linearization(torch.tensor(u, device=device))
        op.cache()
        dt = torch.float64 if prec == "fp64" else torch.float32
        dst = torch.full((mesh.n_dofs,), np.nan, dtype=dt, device=device)
        op.vmult(dst, torch.tensor(v, dtype=dt, device=device))
        vv = v if prec == "fp64" else v.astype(np.float32).astype(np.float64)
        assert rel(dst.cpu().double().numpy(), expect if prec == "fp64" else ref.vmult(vv)) < tol, prec
        op.compute_diagonal()
        assert rel(op.diagonal().cpu().double().numpy(), ref.diagonal()) < tol, prec
    r = torch.full((mesh.n_dofs,), np.nan, dtype=torch.float64, device=device)
    op64 = lsmf.NeoHookOperator(mesh, precision="fp64", caching=caching, constrained=cons, formulation=0)
    op64.set_material(mu, lam)
    op64.set_linearization(torch.tensor(u, device=device))
    op64.residual(r)
    assert rel(r.cpu().numpy(), ref.residual()) < 1e-12


def test_formulation0_rejected_where_unsupported(device):
    mesh = lsmf.Mesh(2, 1, [2, 2])
    for caching in ("tensor2", "tensor4_ns", "scalar_referential"):
        with pytest.raises(lsmf.LsmfError):
            lsmf.NeoHookOperator(mesh, caching=caching, formulation=0)


LAYOUT_CASES = [(3, 2, 4), (3, 2, 5), (3, 2, 7), (3, 2, 11), (2, 1, 9), (2, 2, 9), (3, 1, 5), (3, 4, 3), (3, 4, 5),
                (3, 4, 6), (2, 4, 7)]


@pytest.mark.parametrize("dim,p,n", LAYOUT_CASES, ids=[f"{d}d-Q{p}-n{n}" for d, p, n in LAYOUT_CASES])
def test_tiled_layout_equals_cell_layout(dim, p, n, device):
    """tiles (shared nodes summed in LDS, ragged tiles at the mesh end)
    against one-cell-per-lane-group colouring: vmult (set, add, plain),
    diagonal, residual and its VJP agree to round-off"""
    import torch
    log(f"[layout] dim={dim} p={p} n={n}")
    mesh, mu, lam, u, cons = problem(dim, p, n)
    ud = torch.tensor(u, device=device)
    ops = {}
    for layout in ("auto", "cells"):
        op = lsmf.NeoHookOperator(mesh, precision="fp64", caching="tensor4", constrained=cons, layout=layout)
        op.set_material(mu, lam)
        op.set_linearization(ud)
        op.cache()
        ops[layout] = op
    v = torch.tensor(np.random.default_rng(3).random(mesh.n_dofs), device=device)
    w = torch.tensor(np.random.default_rng(4).random(mesh.n_dofs), device=device)
    out = {}
    for layout, op in ops.items():
        a = torch.full_like(v, np.nan)
        op.vmult(a, v)
        b = w.clone()
        op.vmult_add(b, v)
        c = torch.full_like(v, np.nan)
        op.vmult(c, v, mask="plain")
        op.compute_diagonal()
        d = op.diagonal().clone()
        r = torch.full_like(v, np.nan)
        op.residual(r)
        du = torch.empty_like(v)
        dmu = torch.empty(mesh.n_cells, dtype=torch.float64, device=device)
        dla = torch.empty_like(dmu)
        op.residual_vjp(w, du, dmu, dla)
        out[layout] = [x.cpu().numpy() for x in (a, b, c, d, r, du, dmu, dla)]
    for x, y in zip(out["auto"], out["cells"]):
        assert rel(x, y) < 1e-13
    # every variant runs tiled against the oracle elsewhere; FP32 too
    op32 = {}
    for layout in ("auto", "cells"):
        op = lsmf.NeoHookOperator(mesh, precision="fp32", caching="scalar", constrained=cons, layout=layout)
        op.set_material(mu, lam)
        op.set_linearization(ud)
        op.cache()
        y = torch.empty(mesh.n_dofs, dtype=torch.float32, device=device)
        op.vmult(y, v.float())
        op32[layout] = y.cpu().double().numpy()
    assert rel(op32["auto"], op32["cells"]) < 1e-5


def test_two_material_setter_equals_per_cell_arrays(device):
    """lsmf_set_material_two (the reference's set_material(matrix,
    inclusion), material_id != 0 -> inclusion) gives the same operator as
    the equivalent per-cell arrays, bit for bit"""
    import torch
    mesh, _, _, u, cons = problem(3, 2, 4)
    mid = (mesh.cell_centres()[:, 0] > 0.5).astype(np.uint8)
    m0, l0, m1, l1 = 0.4225, 0.5769, 42.25, 57.69
    ud = torch.tensor(u, device=device)
    v = torch.tensor(np.random.default_rng(3).standard_normal(mesh.n_dofs), device=device)
    out = []
    for two in (True, False):
        op = lsmf.NeoHookOperator(mesh, caching="tensor4", constrained=cons)
        if two:
            op.set_material_two(mid, (m0, l0), (m1, l1))
        else:
            op.set_material(np.where(mid, m1, m0), np.where(mid, l1, l0))
        op.set_linearization(ud)
        op.cache()
        y = torch.empty_like(v)
        op.vmult(y, v)
        out.append(y.cpu().numpy())
    assert np.array_equal(out[0], out[1])


BATCH_CASES = [(3, 2, 3, "auto", "tensor4", "fp64"), (3, 2, 3, "cells", "tensor4", "fp64"),
               (3, 4, 2, "auto", "tensor2", "fp64"), (2, 2, 5, "auto", "scalar", "fp32"),
               (3, 1, 4, "auto", "tensor4_ns", "fp32"), (2, 4, 3, "cells", "scalar_referential", "fp64")]


@pytest.mark.parametrize("dim,p,n,layout,caching,precision", BATCH_CASES,
                         ids=["-".join(map(str, c)) for c in BATCH_CASES])
def test_batched_vmult_equals_single_products(dim, p, n, layout, caching, precision, device):
    """lsmf_vmult_batched: every row of the batch bit-identical to
    lsmf_vmult of that row (constrained rows as identity); the unit vectors
    of the batch probe the columns of the operator"""
    import torch
    log(f"[batched] dim={dim} p={p} n={n} {layout} {caching} {precision}")
    mesh, mu, lam, u, cons = problem(dim, p, n)
    op = lsmf.NeoHookOperator(mesh, precision=precision, caching=caching, constrained=cons, layout=layout)
    op.set_material(mu, lam)
    op.set_linearization(torch.tensor(u, device=device))
    op.cache()
    dt = torch.float64 if precision == "fp64" else torch.float32
    k = 37
    X = torch.tensor(np.random.default_rng(8).standard_normal((k, mesh.n_dofs)), dtype=dt, device=device)
    X[5].zero_()
    X[6].zero_()
    X[6, 3 * dim + 1] = 1.0  # a unit vector
    Y = torch.full_like(X, np.nan)
    op.vmult_batched(Y, X)
    for b in range(k):
        y = torch.full((mesh.n_dofs,), np.nan, dtype=dt, device=device)
        op.vmult(y, X[b].contiguous())
        assert torch.equal(Y[b], y), b
    assert float(Y[5].abs().max()) == 0.0
    if layout == "auto" and caching == "tensor4":
        # argument checks: no vectors, overlapping source and destination
        import ctypes
        assert lsmf._lib.lsmf_vmult_batched(op._h, ctypes.c_void_p(Y.data_ptr()), ctypes.c_void_p(X.data_ptr()),
                                            0, None) == -1
        assert lsmf._lib.lsmf_vmult_batched(op._h, ctypes.c_void_p(X[1].data_ptr()),
                                            ctypes.c_void_p(X.data_ptr()), 2, None) == -1
        assert b"overlap" in lsmf._lib.lsmf_last_error()


def renumbered(mesh, seed, rotate_cells=()):
    """the same mesh with its nodes and cells in a seeded random order (the
    reference's DoF renumbering); cells in `rotate_cells` list their nodes
    turned by 180 degrees about z (local (a0, a1) -> (p - a0, p - a1)), a
    valid cell that is not in standard orientation"""
    from types import SimpleNamespace
    rng = np.random.default_rng(seed)
    perm = rng.permutation(mesh.n_nodes).astype(np.int32)
    order = rng.permutation(mesh.n_cells)
    coords = np.empty_like(mesh.coords)
    coords[perm] = mesh.coords
    cn = perm[mesh.cell_nodes[order]]
    N = mesh.degree + 1
    loc = np.arange(N ** mesh.dim)
    a0, a1, rest = loc % N, (loc // N) % N, loc // (N * N)
    turned = (N - 1 - a0) + N * (N - 1 - a1) + N * N * rest
    for c in rotate_cells:
        cn[c] = cn[c][turned]
    return SimpleNamespace(dim=mesh.dim, degree=mesh.degree, n_cells=mesh.n_cells, n_nodes=mesh.n_nodes,
                           n_dofs=mesh.n_dofs, cell_nodes=np.ascontiguousarray(cn), coords=coords,
                           cell_colour=None), perm, order


@pytest.mark.parametrize("dim,p,n,rotate", [(3, 2, 5, False), (3, 4, 3, False), (2, 2, 7, False), (3, 1, 4, False),
                                             (3, 2, 4, True), (2, 4, 3, True)])
@pytest.mark.parametrize("caching", ["tensor4", "tensor2", "scalar"])
def test_renumbered_mesh_gives_the_same_operator(dim, p, n, rotate, caching, device):
    """a structured mesh with its nodes and cells numbered at random is
    still recognised as logically structured (tiles from the cells' face
    neighbours) and gives the lexicographic mesh's operator; with one cell
    not in standard orientation it is served by the cells layout: vmult,
    diagonal and residual agree with the lexicographic mesh to 1e-13"""
    import torch
    log(f"[renumbered] dim={dim} p={p} n={n} rotate={rotate} {caching}")
    mesh, mu, lam, u, cons = problem(dim, p, n)
    rm, perm, order = renumbered(mesh, 11, rotate_cells=(3,) if rotate else ())
    pu = np.empty_like(u)
    pu.reshape(-1, dim)[perm] = u.reshape(-1, dim)
    pcons = np.empty_like(cons)
    pcons.reshape(-1, dim)[perm] = cons.reshape(-1, dim)
    v = np.random.default_rng(3).random(mesh.n_dofs)
    pv = np.empty_like(v)
    pv.reshape(-1, dim)[perm] = v.reshape(-1, dim)
    out = []
    for m, mu_, lam_, uu, cc, vv in ((mesh, mu, lam, u, cons, v), (rm, mu[order], lam[order], pu, pcons, pv)):
        op = lsmf.NeoHookOperator(m, precision="fp64", caching=caching, constrained=cc,
                                  use_mesh_colouring=m is mesh)
        op.set_material(mu_, lam_)
        op.set_linearization(torch.tensor(uu, device=device))
        op.cache()
        y = torch.full((m.n_dofs,), float("nan"), dtype=torch.float64, device=device)
        op.vmult(y, torch.tensor(vv, device=device))
        op.compute_diagonal()
        r = torch.full_like(y, float("nan"))
        op.residual(r)
        out.append([y.cpu().numpy(), op.diagonal().cpu().numpy(), r.cpu().numpy()])
    for k, name in enumerate(("vmult", "diagonal", "residual")):
        back = out[1][k].reshape(-1, dim)[perm].reshape(-1)  # renumbered -> lexicographic
        assert rel(back, out[0][k]) < 1e-13, name


@pytest.mark.parametrize("p,n", [(2, 6), (4, 3)])
@pytest.mark.parametrize("caching", ["tensor4", "tensor2"])
def test_cuthill_mckee_numbering_gives_the_same_operator(p, n, caching, device):
    """a box mesh numbered like the reference's DoFs (Cuthill-McKee nodes,
    Morton-ordered cells, lsmf.cuthill_mckee_numbering) runs on tiles
    recovered from the face topology -- the same tile tables as the
    lexicographic mesh, so the same memory -- and gives the lexicographic
    mesh's vmult, diagonal and residual to 1e-13"""
    import torch
    log(f"[cuthill-mckee] p={p} n={n} {caching}")
    mesh, mu, lam, u, cons = problem(3, p, n)
    cm, perm, order = lsmf.cuthill_mckee_numbering(mesh)
    pu = np.empty_like(u)
    pu.reshape(-1, 3)[perm] = u.reshape(-1, 3)
    pcons = np.empty_like(cons)
    pcons.reshape(-1, 3)[perm] = cons.reshape(-1, 3)
    v = np.random.default_rng(5).random(mesh.n_dofs)
    pv = np.empty_like(v)
    pv.reshape(-1, 3)[perm] = v.reshape(-1, 3)
    out, mem = [], []
    for m, mu_, lam_, uu, cc, vv in ((mesh, mu, lam, u, cons, v), (cm, mu[order], lam[order], pu, pcons, pv)):
        op = lsmf.NeoHookOperator(m, precision="fp64", caching=caching, constrained=cc, use_mesh_colouring=m is mesh)
        op.set_material(mu_, lam_)
        op.set_linearization(torch.tensor(uu, device=device))
        op.cache()
        y = torch.full((m.n_dofs,), float("nan"), dtype=torch.float64, device=device)
        op.vmult(y, torch.tensor(vv, device=device))
        op.compute_diagonal()
        r = torch.full_like(y, float("nan"))
        op.residual(r)
        out.append([y.cpu().numpy(), op.diagonal().cpu().numpy(), r.cpu().numpy()])
        mem.append(op.memory_consumption())
    assert mem[0] == mem[1]
    for k, name in enumerate(("vmult", "diagonal", "residual")):
        back = out[1][k].reshape(-1, 3)[perm].reshape(-1)  # renumbered -> lexicographic
        assert rel(back, out[0][k]) < 1e-13, name
