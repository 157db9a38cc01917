"""CPU checks of the C-ABI library and the host-side logic (no GPU compute calls)."""

import ctypes
import math
import os
import re

import numpy as np
import pytest

from oracle import nufft_ref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nufftx.h")


def _declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nufftx_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from nufftax_amd import _lib

    handle = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared_symbols()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(handle, s)]
    assert not missing, missing
    # the Python binding covers the whole header, with explicit signatures
    assert set(declared) == set(_lib.SIGNATURES), set(declared) ^ set(_lib.SIGNATURES)
    _lib.lib()  # binds every signature


def test_version_and_errors_without_gpu():
    from nufftax_amd import _lib

    L = _lib.lib()
    assert L.nufftx_version() == 100
    h = ctypes.c_void_p()
    nm = (ctypes.c_int64 * 3)(16, 16, 1)
    # argument validation happens before any device call
    assert L.nufftx_makeplan(7, 2, nm, 1, 1, 1e-6, 2.0, 0, 0, None, ctypes.byref(h)) == 1
    assert "type" in _lib.last_error()
    assert L.nufftx_makeplan(1, 4, nm, 1, 1, 1e-6, 2.0, 0, 0, None, ctypes.byref(h)) == 1
    assert L.nufftx_makeplan(1, 2, nm, 1, 0, 1e-6, 2.0, 0, 0, None, ctypes.byref(h)) == 1
    assert L.nufftx_makeplan(1, 2, nm, 1, 1, -1.0, 2.0, 0, 0, None, ctypes.byref(h)) == 1
    assert L.nufftx_makeplan(1, 2, nm, 1, 1, 1e-6, 0.9, 0, 0, None, ctypes.byref(h)) == 1
    assert L.nufftx_makeplan(1, 2, nm, 1, 1, 1e-6, 2.0, 5, 0, None, ctypes.byref(h)) == 1
    assert L.nufftx_setpts(None, 10, None, None, None, 0, None, None, None) == 1
    assert L.nufftx_execute(None, None, None) == 1
    assert L.nufftx_destroy(None) == 0
    assert L.nufftx_grid_size(-1, 2.0, 8) < 0


def test_plan_cache_host_checks_without_gpu():
    """The plan cache's argument checks and bookkeeping run before any device call."""
    from nufftax_amd import _lib

    L = _lib.lib()
    h = ctypes.c_void_p()
    nm = (ctypes.c_int64 * 3)(16, 16, 1)
    assert L.nufftx_plan_acquire(7, 2, nm, 1, 1, 1e-6, 2.0, 0, 0, None, ctypes.byref(h)) == 1
    assert "type" in _lib.last_error() and not h.value
    assert L.nufftx_plan_acquire(1, 2, nm, 1, 1, 1e-6, 0.9, 0, 0, None, ctypes.byref(h)) == 1
    assert L.nufftx_plan_acquire(1, 2, None, 1, 1, 1e-6, 2.0, 0, 0, None, ctypes.byref(h)) == 1
    assert L.nufftx_plan_acquire(1, 2, nm, 1, 1, 1e-6, 2.0, 0, 0, None, None) == 1
    assert L.nufftx_plan_acquire_grid(2, nm, 7, 18.0, 0.5, 1, 0, None, ctypes.byref(h)) == 1
    assert "nspread" in _lib.last_error()
    assert L.nufftx_plan_acquire_grid(2, nm, 8, 18.0, 0.5, 0, 0, None, ctypes.byref(h)) == 1
    # release only takes plans that acquire handed out (pointers are compared, never read)
    assert L.nufftx_plan_release(None) == 1
    assert L.nufftx_plan_release(ctypes.c_void_p(0x1000)) == 1
    assert "not acquired" in _lib.last_error()
    idle, held, made, reused = (ctypes.c_int64() for _ in range(4))
    assert L.nufftx_plan_cache_stats(ctypes.byref(idle), ctypes.byref(held), ctypes.byref(made),
                                     ctypes.byref(reused)) == 0
    assert (idle.value, held.value, made.value, reused.value) == (0, 0, 0, 0)
    assert L.nufftx_plan_cache_clear() == 0
    assert L.nufftx_device_allocations() >= 0


@pytest.mark.parametrize("eps", [1e-1, 1e-3, 1e-4, 1e-6, 1e-9, 1e-12, 1e-15])
@pytest.mark.parametrize("upsampfac", [2.0, 1.25, 1.5])
def test_kernel_params_match_oracle(eps, upsampfac):
    import nufftax_amd as N

    kp = N.compute_kernel_params(eps, upsampfac)
    ref = R.compute_kernel_params(eps, upsampfac)
    assert kp.nspread == ref.nspread and kp.beta == pytest.approx(ref.beta, rel=1e-15) and kp.c == ref.c
    for n in (1, 7, 32, 100, 512, 1000, 1023):
        assert N.compute_grid_size(n, upsampfac, kp.nspread) == R.compute_grid_size(n, upsampfac, kp.nspread)


def test_type3_grid_helpers_match_oracle(rng):
    import nufftax_amd as N

    x = rng.uniform(-3, 1, 100)
    s = rng.uniform(-40, 10, 80)
    assert N.compute_type3_grid_size(x, s, eps=1e-9) == R.compute_type3_grid_size(x, s, eps=1e-9)
    assert N.compute_type3_grid_size(math.pi, 64.0, eps=1e-9) == 270
    assert N.compute_type3_grid_sizes_2d(math.pi, 1.0, 64.0, 5.0, eps=1e-6) == (
        R.compute_type3_grid_size(math.pi, 64.0, eps=1e-6), R.compute_type3_grid_size(1.0, 5.0, eps=1e-6))
    nf3 = N.compute_type3_grid_sizes_3d(1, 2, 3, 4, 5, 6, eps=1e-4)
    assert all(v % 2 == 0 for v in nf3)


def test_public_api_surface():
    import nufftax_amd as N

    for name in ["nufft1d1", "nufft1d2", "nufft1d3", "nufft2d1", "nufft2d2", "nufft2d3", "nufft3d1", "nufft3d2",
                 "nufft3d3", "compute_type3_grid_size", "compute_type3_grid_sizes_2d", "compute_type3_grid_sizes_3d",
                 "fftc", "ifftc", "to_complex", "from_complex"]:
        assert callable(getattr(N, name))
        assert callable(getattr(N, name + "_jvp", getattr(N, name)))


def test_centered_fft_utils_match_numpy(rng):
    import torch

    import nufftax_amd as N

    a = rng.standard_normal((6, 8)) + 1j * rng.standard_normal((6, 8))
    t = torch.from_numpy(a)
    ref = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(a)))
    np.testing.assert_allclose(N.fftc(t).numpy(), ref, rtol=1e-12)
    np.testing.assert_allclose(N.ifftc(N.fftc(t)).numpy(), a, rtol=1e-12, atol=1e-12)
    two = N.from_complex(t)
    assert two.shape == (6, 8, 2)
    np.testing.assert_allclose(N.to_complex(two).numpy(), a)


def test_no_cpu_fallback():
    """Without a GPU the transforms raise instead of computing on the CPU."""
    import torch

    import nufftax_amd as N

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(Exception):
        N.nufft1d1(np.zeros(4), np.ones(4, complex), 8)


def test_workload_generators_are_seeded():
    from nufftax_amd import workloads

    a = workloads.radial_2d(seed=3, spokes=8, samples=10, n_modes=(4, 4))
    b = workloads.radial_2d(seed=3, spokes=8, samples=10, n_modes=(4, 4))
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    x, y, c, f = a
    assert x.shape == (80,) and f.shape == (4, 4) and np.all(np.abs(x) <= np.pi) and np.all(np.abs(y) <= np.pi)


@pytest.mark.parametrize("nf,eps", [(64, 1e-6), (270, 1e-9), (1000, 1e-4), (2048, 1e-12), (7, 1e-2)])
def test_kernel_fourier_series_matches_oracle(nf, eps):
    from nufftax_amd import core

    kp = R.compute_kernel_params(eps)
    got = core.kernel_fourier_series(nf, kp.nspread, kp.beta, kp.c, dtype=torch_f64())
    ref = R.kernel_fourier_series(nf, kp.nspread, kp.beta, kp.c)
    np.testing.assert_allclose(got.numpy(), ref, rtol=1e-13, atol=1e-15)


def torch_f64():
    import torch

    return torch.float64


def test_core_kernel_helpers_match_oracle(rng):
    import torch

    from nufftax_amd import core

    kp = R.compute_kernel_params(1e-6)
    z = np.concatenate([rng.uniform(-5, 5, 200), [0.0, kp.nspread / 2, -kp.nspread / 2, 4.0000001]])
    for dt in (np.float32, np.float64):
        zz = z.astype(dt)
        phi = core.es_kernel(torch.from_numpy(zz), kp.beta, kp.c).numpy()
        np.testing.assert_allclose(phi, R.es_kernel(zz, kp.beta, kp.c), rtol=2e-6 if dt == np.float32 else 1e-14,
                                   atol=1e-30)
        p2, d2 = core.es_kernel_with_derivative(torch.from_numpy(zz), kp.beta, kp.c)
        rp, rd = R.es_kernel_with_derivative(zz, kp.beta, kp.c)
        tol = 2e-5 if dt == np.float32 else 1e-13
        np.testing.assert_allclose(p2.numpy(), rp, rtol=tol, atol=1e-30)
        np.testing.assert_allclose(d2.numpy(), rd, rtol=tol, atol=1e-25)
    x = rng.uniform(-10, 10, 1000)
    np.testing.assert_allclose(core.fold_rescale(torch.from_numpy(x), 512).numpy(), R.fold_rescale(x, 512),
                               rtol=1e-12, atol=1e-9)


def test_core_rejects_custom_kernels_and_needs_gpu():
    import torch

    from nufftax_amd import core

    kp = core.compute_kernel_params(1e-6)
    custom = core.Kernel(nspread=8, phi=lambda z: torch.exp(-z * z))
    with pytest.raises(NotImplementedError):
        core._es_params(custom)
    assert core._es_params(core.es_kernel_spec(kp)) == kp
    if not torch.cuda.is_available():
        with pytest.raises(Exception):
            core.spread_1d(np.zeros(4), np.ones(4, complex), 16, kp)


def test_serving_host_logic():
    import torch

    from nufftax_amd import serving as S

    b1 = S.request_bytes(1, 2, 10**5, (128, 128), 1e-6)
    b2 = S.request_bytes(1, 2, 10**6, (128, 128), 1e-6)
    b3 = S.request_bytes(1, 2, 10**6, (512, 512), 1e-6)
    assert 0 < b1 < b2 < b3
    # 1e6 points: x, y (4 B) + 4 B sort index + 8 B strengths; 512^2 modes and the 1024^2 fine grid
    assert b3 == 10**6 * 12 + (10**6 + 512 * 512) * 8 + 1024 * 1024 * 8
    recs = []
    for i in range(100):
        r = S.ServeRequest(1, (), None, (8,), 1e-6, 1)
        r.t_submit, r.t_done = float(i), float(i) + 0.001 * (i + 1)
        recs.append(r)
    st = S.latency_stats(recs)
    assert st["n"] == 100 and abs(st["p50_ms"] - 50.5) < 1 and st["p99_ms"] > 98
    if not torch.cuda.is_available():
        with pytest.raises(Exception):
            S.NufftServer()


def test_serving_group_formation():
    """The dispatcher's grouping on host-side requests: same points and signature form one
    unit, other small point sets of the same kind/dim/eps/sign join as further units (up to
    max_units), everything else stays pending in arrival order."""
    import torch

    from nufftax_amd import serving as S

    srv = object.__new__(S.NufftServer)
    srv.max_group, srv.max_units, srv.batch_max_points, srv.scan_depth, srv.batch_min_queue = 64, 3, 1000, 256, 4
    srv.budget = 1 << 30

    def req(kind, pts, nm=(16, 16), big=False):
        M = pts[0].shape[0]
        data = torch.zeros(M, dtype=torch.complex64) if kind == 1 else torch.zeros(nm, dtype=torch.complex64)
        r = S.ServeRequest(kind, pts, data, nm, 1e-6, 1 if kind == 1 else -1)
        r.nbytes = 100
        return r

    p1 = (torch.zeros(100), torch.zeros(100))
    p2 = (torch.zeros(200), torch.zeros(200))
    p3 = (torch.zeros(300), torch.zeros(300))
    p4 = (torch.zeros(400), torch.zeros(400))
    big = (torch.zeros(5000), torch.zeros(5000))
    a, b = req(1, p1), req(1, p1)              # one unit (shared points)
    t2 = req(2, p2)                            # other kind: pending
    c = req(1, p2, nm=(32, 32))                # second unit (other grid size)
    d = req(1, big)                            # too many points to share a launch: pending
    e = req(1, p3, nm=(8, 8))                  # third unit
    f = req(1, p4, nm=(4, 4))                  # over max_units: pending
    g = req(1, p4)                             # a second unit of a's shape: pending
    units, held = srv._take_group([a, t2, b, g, c, d, e, f])
    assert [[id(r) for r in u] for u in units] == [[id(a), id(b)], [id(c)], [id(e)]]
    assert [id(r) for r in held] == [id(t2), id(g), id(d), id(f)]
    # a large first request keeps its own points only
    units, held = srv._take_group([d, a, b])
    assert [[id(r) for r in u] for u in units] == [[id(d)]] and len(held) == 2
    # a short queue: only the same points group
    units, held = srv._take_group([a, c, b])
    assert [[id(r) for r in u] for u in units] == [[id(a), id(b)]] and [id(r) for r in held] == [id(c)]
    # the budget caps the group
    srv.budget = 250
    units, held = srv._take_group([a, b, c, e])
    assert sum(len(u) for u in units) == 2 and [id(r) for r in held] == [id(c), id(e)]


def test_serving_admission_skips_requests_that_do_not_fit():
    """Admission on host-side requests: a request that does not fit the budget left by the
    work in flight does not block later ones that fit (they are admitted ahead of it), at
    most max_skips times; then the oldest request waits for the budget first."""
    import threading

    import torch

    from nufftax_amd import serving as S

    srv = object.__new__(S.NufftServer)
    srv.max_group, srv.max_units, srv.batch_max_points, srv.scan_depth, srv.batch_min_queue = 64, 1, 1000, 16, 4
    srv.budget, srv.max_skips, srv.skipped_admissions, srv._in_use = 1000, 2, 0, 0
    srv._cv = threading.Condition()

    def req(nbytes, M):
        pts = (torch.zeros(M), torch.zeros(M))
        r = S.ServeRequest(1, pts, torch.zeros(M, dtype=torch.complex64), (16, 16), 1e-6, 1)
        r.nbytes = nbytes
        return r

    big, s1, s2, s3 = req(800, 10), req(100, 11), req(100, 12), req(100, 13)
    srv._in_use = 500  # work in flight: big (800) does not fit, the small ones do
    units, pending, nb = srv._next_group([big, s1, s2, s3])
    assert units == [[s1]] and nb == 100 and pending == [big, s2, s3] and srv._in_use == 600
    assert big.skipped == 1 and srv.skipped_admissions == 1
    units, pending, nb = srv._next_group(pending)
    assert units == [[s2]] and pending == [big, s3] and big.skipped == 2
    # big has been passed over max_skips times: it goes first, so admission waits for the
    # budget that completions release
    released = []

    def complete():
        with srv._cv:
            srv._in_use = 0
            released.append(True)
            srv._cv.notify_all()

    t = threading.Timer(0.2, complete)
    t.start()
    units, pending, nb = srv._next_group(pending)
    t.join()
    assert released and units == [[big]] and pending == [s3] and srv._in_use == 800


def test_autodiff_dispatch_predicate():
    """The public transforms skip their autograd Function only when nothing can differentiate
    through the call: no input requiring grad under grad mode, no forward-mode dual level, no
    torch.func transform (vmap, jvp, grad)."""
    import torch
    from torch.autograd import forward_ad as fwad

    from nufftax_amd.transforms import _autodiff_active

    x = torch.zeros(4)
    assert not _autodiff_active([x])
    w = torch.zeros(4, requires_grad=True)
    assert _autodiff_active([x, w])
    with torch.no_grad():
        assert not _autodiff_active([x, w])
    with torch.inference_mode():
        assert not _autodiff_active([x])
    with fwad.dual_level():
        assert _autodiff_active([x])
    assert not _autodiff_active([x])
    seen = []

    def probe(v):
        seen.append(_autodiff_active([v]))
        return v * 2

    torch.func.vmap(probe)(torch.zeros(3, 4))
    torch.func.jvp(probe, (x,), (torch.ones(4),))
    torch.func.grad(lambda v: probe(v).sum())(x)
    assert seen == [True, True, True]
