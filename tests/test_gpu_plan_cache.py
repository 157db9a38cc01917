"""The C ABI as an FFI handler would drive it, with no PyTorch plan objects in between:
nufftx_plan_acquire once per (parameters, stream), then setpts + execute(_sign) + release per
call, on two streams; the cached plans are reused, a reused plan's setpts + execute of the same
sizes makes no device allocation, and the results equal the direct sum (FP64 oracle, 3x eps)
and each other.  Also: the public transforms are functions of their inputs only (coordinates
changed in place through .data between two calls give the new points' result)."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import nufft_ref as R

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(autouse=True)
def _empty_cache():
    from nufftax_amd import _lib

    yield
    torch.cuda.synchronize()
    _lib.lib().nufftx_plan_cache_clear()


def _rel(a, b):
    a = torch.as_tensor(a).to(torch.complex128).cpu()
    b = torch.as_tensor(b).to(torch.complex128).cpu()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def _stats(L):
    v = [ctypes.c_int64() for _ in range(4)]
    assert L.nufftx_plan_cache_stats(*(ctypes.byref(x) for x in v)) == 0
    return tuple(x.value for x in v)  # idle, held, made, reused


def _points(rng, M):
    x = torch.as_tensor(rng.uniform(-np.pi, np.pi, M), dtype=torch.float32, device=DEV)
    y = torch.as_tensor(rng.uniform(-np.pi, np.pi, M), dtype=torch.float32, device=DEV)
    c = torch.as_tensor(rng.standard_normal(M) + 1j * rng.standard_normal(M), dtype=torch.complex64, device=DEV)
    return x, y, c


def _handler_type1(L, stream, x, y, c, n1, n2, isign, eps, upsampfac):
    """What a type-1 FFI handler does per call: acquire, setpts, execute_sign, release."""
    nm = (ctypes.c_int64 * 3)(n1, n2, 1)
    f = torch.empty((n2, n1), dtype=torch.complex64, device=DEV)
    h = ctypes.c_void_p()
    assert L.nufftx_plan_acquire(1, 2, nm, isign, 1, eps, upsampfac, 0, 0, ctypes.c_void_p(stream.cuda_stream),
                                 ctypes.byref(h)) == 0, L.nufftx_last_error()
    try:
        assert L.nufftx_setpts(h, x.numel(), x.data_ptr(), y.data_ptr(), None, 0, None, None, None) == 0
        assert L.nufftx_execute_sign(h, c.data_ptr(), f.data_ptr(), isign) == 0, L.nufftx_last_error()
    finally:
        assert L.nufftx_plan_release(h) == 0, L.nufftx_last_error()
    return h.value, f


def _handler_type2(L, stream, x, y, f, isign, eps, upsampfac):
    """The transpose of type 1 (same points, same sign, upsampfac_T): a type-2 handler."""
    n2, n1 = f.shape
    nm = (ctypes.c_int64 * 3)(n1, n2, 1)
    c = torch.empty(x.numel(), dtype=torch.complex64, device=DEV)
    h = ctypes.c_void_p()
    assert L.nufftx_plan_acquire(2, 2, nm, isign, 1, eps, upsampfac, 0, 0, ctypes.c_void_p(stream.cuda_stream),
                                 ctypes.byref(h)) == 0, L.nufftx_last_error()
    try:
        assert L.nufftx_setpts(h, x.numel(), x.data_ptr(), y.data_ptr(), None, 0, None, None, None) == 0
        assert L.nufftx_execute_sign(h, c.data_ptr(), f.data_ptr(), isign) == 0, L.nufftx_last_error()
    finally:
        assert L.nufftx_plan_release(h) == 0
    return h.value, c


def test_plan_cache_two_streams_steady_state():
    from nufftax_amd import _lib

    L = _lib.lib()
    L.nufftx_plan_cache_clear()
    rng = np.random.default_rng(11)
    n1, n2, eps, M = 96, 80, 1e-6, 60_000
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    x, y, c = _points(rng, M)
    torch.cuda.synchronize()
    made0 = _stats(L)[2]
    ha, fa = _handler_type1(L, sa, x, y, c, n1, n2, 1, eps, 2.0)
    hb, fb = _handler_type1(L, sb, x, y, c, n1, n2, 1, eps, 2.0)
    torch.cuda.synchronize()
    assert ha != hb  # one plan per stream
    assert _stats(L)[2] - made0 == 2
    # steady state: the same sizes again on both streams reuse the plans and allocate nothing
    x2, y2, c2 = _points(rng, M)
    torch.cuda.synchronize()
    allocs = L.nufftx_device_allocations()
    reused0 = _stats(L)[3]
    for _ in range(3):
        ha2, fa2 = _handler_type1(L, sa, x2, y2, c2, n1, n2, 1, eps, 2.0)
        hb2, fb2 = _handler_type1(L, sb, x2, y2, c2, n1, n2, 1, eps, 2.0)
    torch.cuda.synchronize()
    assert (ha2, hb2) == (ha, hb)
    assert L.nufftx_device_allocations() == allocs
    assert _stats(L)[3] - reused0 == 6
    idle, held, _, _ = _stats(L)
    assert held == 0 and idle >= 2
    # results: both streams agree, and match the FP64 direct sum on a mode subset
    assert _rel(fa, fb) < 1e-6 and _rel(fa2, fb2) < 1e-6
    coords = [x2.double().cpu().numpy(), y2.double().cpu().numpy()]
    cc = c2.to(torch.complex128).cpu().numpy()
    i2, i1 = rng.integers(0, n2, 48), rng.integers(0, n1, 48)
    modes = np.stack([R.mode_grid(n1)[i1], R.mode_grid(n2)[i2]], axis=1)
    ref = R.direct_type1(coords, cc, (n1, n2), isign=1, mode_subset=modes)
    got = fa2.cpu().numpy()[i2, i1]
    assert R.rel_l2(got, ref) < 3 * eps
    # concurrent holders of one key get distinct plans
    nm = (ctypes.c_int64 * 3)(n1, n2, 1)
    h1, h2 = ctypes.c_void_p(), ctypes.c_void_p()
    st = ctypes.c_void_p(sa.cuda_stream)
    assert L.nufftx_plan_acquire(1, 2, nm, 1, 1, eps, 2.0, 0, 0, st, ctypes.byref(h1)) == 0
    assert L.nufftx_plan_acquire(1, 2, nm, 1, 1, eps, 2.0, 0, 0, st, ctypes.byref(h2)) == 0
    assert h1.value != h2.value
    assert L.nufftx_destroy(h1) != 0 and L.nufftx_set_stream(h1, None) != 0  # cached plans are released
    assert L.nufftx_plan_release(h1) == 0 and L.nufftx_plan_release(h2) == 0
    assert L.nufftx_plan_release(h2) != 0  # twice
    torch.cuda.synchronize()
    assert L.nufftx_plan_cache_clear() >= 3
    assert _stats(L)[:2] == (0, 0)


def test_plan_cache_transpose_handler_upsampfac_T():
    """<T1 c, f> = <c, T2 f> with T2 the transpose handler's plan: type 2, the same sign, the
    backward upsampfac (1.25 here, 2.0 forward); the type-2 plan of the forward sign is reused
    for the other sign."""
    from nufftax_amd import _lib

    L = _lib.lib()
    L.nufftx_plan_cache_clear()
    rng = np.random.default_rng(5)
    n1, n2, eps, M = 64, 48, 1e-6, 30_000
    s = torch.cuda.Stream()
    x, y, c = _points(rng, M)
    f = torch.as_tensor(rng.standard_normal((n2, n1)) + 1j * rng.standard_normal((n2, n1)),
                        dtype=torch.complex64, device=DEV)
    torch.cuda.synchronize()
    _, t1 = _handler_type1(L, s, x, y, c, n1, n2, 1, eps, 2.0)
    h_fwd, _ = _handler_type2(L, s, x, y, f, -1, eps, 1.25)   # a forward type-2 call, isign -1
    h_t, t2 = _handler_type2(L, s, x, y, f, 1, eps, 1.25)     # the transpose of type 1, isign +1
    torch.cuda.synchronize()
    assert h_t == h_fwd
    lhs = complex((t1.to(torch.complex128) * f.to(torch.complex128)).sum())
    rhs = complex((c.to(torch.complex128) * t2.to(torch.complex128)).sum())
    # each side carries its transform's relative error: bound by the Cauchy-Schwarz scale
    scale = float(torch.linalg.vector_norm(t1.to(torch.complex128)) * torch.linalg.vector_norm(f.to(torch.complex128)))
    assert abs(lhs - rhs) / scale < 10 * eps
    # and the transpose result itself against the direct sum, within twice the error of the
    # reference's algorithm (the FP64 oracle) at the same kernel: at upsampfac 1.25 the
    # reference keeps the eps kernel width, so both sit near 5e-5, not near eps
    coords = [x.double().cpu().numpy(), y.double().cpu().numpy()]
    fn = f.to(torch.complex128).cpu().numpy()
    sub = rng.integers(0, M, 64)
    exact = R.direct_type2(coords, fn, isign=1, point_subset=sub)
    oracle = R.nufft_type2([xc[sub] for xc in coords], fn, eps=eps, isign=1, upsampfac=1.25)
    err_ref = R.rel_l2(oracle, exact)
    assert R.rel_l2(t2.cpu().numpy()[sub], exact) <= 2 * err_ref + 1e-6


def test_public_transform_follows_coordinates_changed_in_place():
    """The public transforms keep no point state across calls: writing new coordinates into
    the same tensors (through .data) changes the result to the new points' transform."""
    import nufftax_amd as N

    rng = np.random.default_rng(3)
    M, n1, n2, eps = 20_000, 40, 36, 1e-6
    x, y, c = _points(rng, M)
    f_old = N.nufft2d1(x, y, c, (n1, n2), eps=eps)
    xn, yn, _ = _points(rng, M)
    x.data.copy_(xn)
    y.data.copy_(yn)
    f_new = N.nufft2d1(x, y, c, (n1, n2), eps=eps)
    c2 = N.nufft2d2(x, y, f_new, eps=eps)
    torch.cuda.synchronize()
    coords = [xn.double().cpu().numpy(), yn.double().cpu().numpy()]
    cc = c.to(torch.complex128).cpu().numpy()
    ref = R.direct_type1(coords, cc, (n1, n2), isign=1)
    assert R.rel_l2(f_new.cpu().numpy(), ref) < 3 * eps
    assert _rel(f_old, torch.as_tensor(ref)) > 0.1  # the old points' result is far from it
    sub = rng.integers(0, M, 64)
    ref2 = R.direct_type2(coords, f_new.to(torch.complex128).cpu().numpy(), isign=-1, point_subset=sub)
    assert R.rel_l2(c2.cpu().numpy()[sub], ref2) < 3 * eps
