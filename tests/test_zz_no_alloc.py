"""lsmf_cache_noalloc, the cache() variant that allocates nothing (its det
F <= 0 counter lives in a file-scope device array of the library), against
lsmf_cache through the C ABI: two calls in a row, the second on a smaller
operator, give caches whose products are bit-identical to lsmf_cache's; a
call with det F <= 0 reports it, and the next call starts from a cleared
counter.
"""
import ctypes

import numpy as np
import pytest

import lsmf

pytestmark = pytest.mark.gpu


def make_op(mesh, caching, precision, u, device):
    import torch
    cons = mesh.constrained_mask({0: list(range(mesh.dim)), 1: [0]})
    mu, lam, _ = lsmf.inclusion_material(mesh.cell_centres(), mesh.dim, 1.0, 0.3, 3, 0.25, 10.0, 2)
    op = lsmf.NeoHookOperator(mesh, precision=precision, caching=caching, constrained=cons)
    op.set_material(mu, lam)
    op.set_linearization(torch.tensor(u, dtype=torch.float64, device=device))
    return op


def noalloc(op):
    return lsmf._lib.lsmf_cache_noalloc(op._h, ctypes.c_void_p(lsmf._stream()))


def product(op, mesh, device):
    import torch
    dt = torch.float64 if op.precision == "fp64" else torch.float32
    x = torch.tensor(np.random.default_rng(3).random(mesh.n_dofs), dtype=dt, device=device)
    y = torch.full_like(x, float("nan"))
    op.vmult(y, x)
    return y.cpu().numpy()


@pytest.mark.parametrize("first,second", [((3, 2, 4, "tensor4", "fp64"), (3, 2, 2, "tensor2", "fp32")),
                                          ((3, 4, 2, "tensor2", "fp64"), (2, 1, 3, "scalar", "fp64"))])
def test_noalloc_cache_equals_cache(first, second, device):
    for dim, p, n, caching, prec in (first, second):
        print(f"[cache_noalloc] dim={dim} p={p} n={n} {caching} {prec}", flush=True)
        mesh = lsmf.Mesh(dim, p, [n] * dim)
        u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, 5)
        ref = make_op(mesh, caching, prec, u, device)
        ref.cache()
        op = make_op(mesh, caching, prec, u, device)
        assert noalloc(op) == lsmf.LSMF_OK, lsmf._lib.lsmf_last_error()
        assert np.array_equal(product(op, mesh, device), product(ref, mesh, device))


def test_noalloc_cache_reports_nonpositive_det_F_and_clears(device):
    print("[cache_noalloc] det F <= 0, then a good operator", flush=True)
    mesh = lsmf.Mesh(3, 2, [2, 2, 2])
    bad = make_op(mesh, "tensor4", "fp64", -2.0 * mesh.coords.reshape(-1), device)
    assert noalloc(bad) == -4
    assert b"det_F is not positive" in lsmf._lib.lsmf_last_error()
    with pytest.raises(lsmf.LsmfError):  # not cached: vmult refuses
        product(bad, mesh, device)
    u, _ = lsmf.smooth_displacement(mesh.coords, 0.02, 5)
    good = make_op(mesh, "tensor4", "fp64", u, device)
    assert noalloc(good) == lsmf.LSMF_OK
    ref = make_op(mesh, "tensor4", "fp64", u, device)
    ref.cache()
    assert np.array_equal(product(good, mesh, device), product(ref, mesh, device))
