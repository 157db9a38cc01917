"""GPU gradients: finite differences, the reference's k-weighted JVP formulas, torch.func jvp/vmap.

Tolerances: FP64 transforms at eps 1e-12 with central differences h = 1e-6 ->
rtol 1e-6 (the reference's own bar, tests/test_adjoint.py:221-676).  The
ES-derivative coordinate gradient matches the reference's k-weighted formula
to 1e-8 relative at eps 1e-12.
"""

import numpy as np
import pytest
import torch

from oracle import nufft_ref as R

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(DEV)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _fd_real(loss, arr, idx, h=1e-6):
    a = arr.copy()
    a[idx] += h
    lp = loss(a)
    a[idx] -= 2 * h
    lm = loss(a)
    return (lp - lm) / (2 * h)


def _fd_complex(loss, arr, idx, h=1e-6):
    """dL/dRe + i dL/dIm (PyTorch's gradient convention for complex inputs)."""
    a = arr.astype(complex).copy()
    a[idx] += h
    lp = loss(a)
    a[idx] -= 2 * h
    lm = loss(a)
    gre = (lp - lm) / (2 * h)
    a[idx] += h + 1j * h
    lp = loss(a)
    a[idx] -= 2j * h
    lm = loss(a)
    return gre + 1j * (lp - lm) / (2 * h)


@pytest.mark.parametrize("isign", [1, -1])
def test_type1_2d_grads_fd(rng, isign):
    import nufftax_amd as N

    M, nm = 40, (12, 10)
    x, y = rng.uniform(-np.pi, np.pi, M), rng.uniform(-np.pi, np.pi, M)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    w = rng.standard_normal((nm[1], nm[0])) + 1j * rng.standard_normal((nm[1], nm[0]))

    def loss_np(xv, yv, cv):
        f = N.nufft2d1(_t(xv), _t(yv), _t(cv), nm, eps=1e-12, isign=isign)
        return float(torch.sum(torch.abs(f - _t(w)) ** 2))

    xt, yt, ct = (_t(a).requires_grad_(True) for a in (x, y, c))
    f = N.nufft2d1(xt, yt, ct, nm, eps=1e-12, isign=isign)
    torch.sum(torch.abs(f - _t(w)) ** 2).backward()
    for j in range(0, M, 7):
        assert _np(xt.grad)[j] == pytest.approx(_fd_real(lambda a: loss_np(a, y, c), x, j), rel=1e-6, abs=1e-6)
        assert _np(yt.grad)[j] == pytest.approx(_fd_real(lambda a: loss_np(x, a, c), y, j), rel=1e-6, abs=1e-6)
        assert _np(ct.grad)[j] == pytest.approx(_fd_complex(lambda a: loss_np(x, y, a), c, j), rel=1e-6, abs=1e-6)


def test_type2_3d_grads_fd(rng):
    import nufftax_amd as N

    M, nm = 25, (6, 8, 4)
    xs = [rng.uniform(-np.pi, np.pi, M) for _ in range(3)]
    f = rng.standard_normal((4, 8, 6)) + 1j * rng.standard_normal((4, 8, 6))
    w = rng.standard_normal(M) + 1j * rng.standard_normal(M)

    def loss_np(x0, x1, x2, fv):
        out = N.nufft3d2(_t(x0), _t(x1), _t(x2), _t(fv), eps=1e-12)
        return float(torch.sum(torch.abs(out - _t(w)) ** 2))

    ts = [_t(a).requires_grad_(True) for a in xs]
    ft = _t(f).requires_grad_(True)
    out = N.nufft3d2(*ts, ft, eps=1e-12)
    torch.sum(torch.abs(out - _t(w)) ** 2).backward()
    for ax in range(3):
        for j in (0, 11, 24):
            def l(a, ax=ax):
                args = list(xs)
                args[ax] = a
                return loss_np(*args, f)
            assert _np(ts[ax].grad)[j] == pytest.approx(_fd_real(l, xs[ax], j), rel=1e-6, abs=1e-6)
    for idx in [(0, 0, 0), (3, 7, 5), (2, 4, 1)]:
        assert _np(ft.grad)[idx] == pytest.approx(_fd_complex(lambda a: loss_np(*xs, a), f, idx), rel=1e-6, abs=1e-6)


@pytest.mark.parametrize("dim", [1, 2])
def test_coordinate_grad_matches_k_weighted_reference_formula(rng, dim):
    """dL/dx via ES-derivative weights == the reference's JVP/transpose route
    (primitives.py:170-194): Re(conj(c) * T2_{-isign}(i isign k1 * g)) ... evaluated with the oracle."""
    import nufftax_amd as N

    M = 300
    nm = (30,) if dim == 1 else (20, 14)
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    g = rng.standard_normal(tuple(reversed(nm))) + 1j * rng.standard_normal(tuple(reversed(nm)))
    isign = 1
    ts = [_t(a).requires_grad_(True) for a in coords]
    fn = N.nufft1d1 if dim == 1 else N.nufft2d1
    out = fn(*ts, _t(c), nm[0] if dim == 1 else nm, eps=1e-12, isign=isign)
    torch.sum(torch.real(torch.conj(_t(g)) * out)).backward()  # L = Re<g, f>: cotangent of f is g
    for ax in range(dim):
        k = np.arange(-(nm[ax] // 2), (nm[ax] + 1) // 2)
        shape = [1] * dim
        shape[dim - 1 - ax] = -1
        kg = g * k.reshape(shape)
        # dL/dx_j = Re(sum_k conj(g_k) i isign k c_j e^{i isign k x_j}) = Re(conj(i isign c_j) * T2_{-isign}(k g)_j)
        t2 = R.nufft_type2(coords, kg, eps=1e-12, isign=-isign)
        ref = np.real(np.conj(1j * isign * c) * t2)
        np.testing.assert_allclose(_np(ts[ax].grad), ref, rtol=1e-8, atol=1e-8 * np.abs(ref).max())


def test_type3_grads_fd(rng):
    import nufftax_amd as N

    M, Nt = 20, 15
    x = rng.uniform(-2, 2, M)
    s = rng.uniform(-6, 6, Nt)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    w = rng.standard_normal(Nt) + 1j * rng.standard_normal(Nt)
    nf = N.compute_type3_grid_size(x, s, eps=1e-12)

    def loss_np(xv, cv, sv):
        out = N.nufft1d3(_t(xv), _t(cv), _t(sv), nf, eps=1e-12)
        return float(torch.sum(torch.abs(out - _t(w)) ** 2))

    xt, ct, st = _t(x).requires_grad_(True), _t(c).requires_grad_(True), _t(s).requires_grad_(True)
    out = N.nufft1d3(xt, ct, st, nf, eps=1e-12)
    torch.sum(torch.abs(out - _t(w)) ** 2).backward()
    for j in (0, 7, 14):
        assert _np(xt.grad)[j] == pytest.approx(_fd_real(lambda a: loss_np(a, c, s), x, j), rel=1e-6, abs=1e-6)
        assert _np(st.grad)[j] == pytest.approx(_fd_real(lambda a: loss_np(x, c, a), s, j), rel=1e-6, abs=1e-6)
        assert _np(ct.grad)[j] == pytest.approx(_fd_complex(lambda a: loss_np(x, a, s), c, j), rel=1e-6, abs=1e-6)


def test_roundtrip_loss_grads_c64_vs_fp64(rng):
    """Differentiated 2D type 2 -> type 1 round trip (the benchmark's differentiated workload, small):
    complex64 gradients agree with the FP64 gradients to FP32 accuracy; FP64 agrees with FD."""
    import nufftax_amd as N

    M, nm = 2000, (16, 16)
    x, y = rng.uniform(-np.pi, np.pi, M), rng.uniform(-np.pi, np.pi, M)
    f = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    g = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))

    def grads(rd, cd, eps):
        xt, yt, ft = _t(x, rd).requires_grad_(True), _t(y, rd).requires_grad_(True), _t(f, cd).requires_grad_(True)
        out = N.nufft2d1(xt, yt, N.nufft2d2(xt, yt, ft, eps=eps), nm, eps=eps)
        loss = torch.sum(torch.abs(out - _t(g, cd)) ** 2)
        loss.backward()
        return _np(xt.grad).astype(np.float64), _np(yt.grad).astype(np.float64), _np(ft.grad).astype(np.complex128)

    gx64, gy64, gf64 = grads(torch.float64, torch.complex128, 1e-12)
    gx32, gy32, gf32 = grads(torch.float32, torch.complex64, 1e-6)
    assert R.rel_l2(gx32, gx64) < 1e-3 and R.rel_l2(gy32, gy64) < 1e-3 and R.rel_l2(gf32, gf64) < 1e-4

    def loss_np(xv):
        out = N.nufft2d1(_t(xv), _t(y), N.nufft2d2(_t(xv), _t(y), _t(f), eps=1e-12), nm, eps=1e-12)
        return float(torch.sum(torch.abs(out - _t(g)) ** 2))

    for j in (3, 999):
        assert gx64[j] == pytest.approx(_fd_real(loss_np, x, j, h=1e-5), rel=1e-5, abs=1e-5)


def test_forward_mode_jvp(rng):
    import nufftax_amd as N
    from torch.func import jvp

    M, nm = 50, (10, 12)
    x, y = _t(rng.uniform(-np.pi, np.pi, M)), _t(rng.uniform(-np.pi, np.pi, M))
    c = _t(rng.standard_normal(M) + 1j * rng.standard_normal(M))
    dx, dy = _t(rng.standard_normal(M)), _t(rng.standard_normal(M))
    dc = _t(rng.standard_normal(M) + 1j * rng.standard_normal(M))

    def f(a, b, cc):
        return N.nufft2d1(a, b, cc, nm, eps=1e-12)

    _, tan = jvp(f, (x, y, c), (dx, dy, dc))
    h = 1e-6
    fd = (f(x + h * dx, y + h * dy, c + h * dc) - f(x - h * dx, y - h * dy, c - h * dc)) / (2 * h)
    assert R.rel_l2(_np(tan), _np(fd)) < 1e-7

    f2 = rng.standard_normal((12, 10)) + 1j * rng.standard_normal((12, 10))
    df2 = rng.standard_normal((12, 10)) + 1j * rng.standard_normal((12, 10))

    def g(a, b, ff):
        return N.nufft2d2(a, b, ff, eps=1e-12)

    _, tan2 = jvp(g, (x, y, _t(f2)), (dx, dy, _t(df2)))
    fd2 = (g(x + h * dx, y + h * dy, _t(f2 + h * df2)) - g(x - h * dx, y - h * dy, _t(f2 - h * df2))) / (2 * h)
    assert R.rel_l2(_np(tan2), _np(fd2)) < 1e-7


def test_vmap_fast_path_and_fallback(rng):
    import nufftax_amd as N
    from torch.func import vmap

    M, nm = 300, (16, 12)
    x, y = _t(rng.uniform(-np.pi, np.pi, M)), _t(rng.uniform(-np.pi, np.pi, M))
    cs = _t(rng.standard_normal((4, M)) + 1j * rng.standard_normal((4, M)))
    out = vmap(lambda cc: N.nufft2d1(x, y, cc, nm))(cs)
    loop = torch.stack([N.nufft2d1(x, y, cs[i], nm) for i in range(4)])
    assert R.rel_l2(_np(out), _np(loop)) < 1e-12
    xb = _t(rng.uniform(-np.pi, np.pi, (3, M)))
    out2 = vmap(lambda xx, cc: N.nufft2d1(xx, y, cc, nm))(xb, cs[:3])
    loop2 = torch.stack([N.nufft2d1(xb[i], y, cs[i], nm) for i in range(3)])
    assert R.rel_l2(_np(out2), _np(loop2)) < 1e-12


@pytest.mark.parametrize("dim,nm", [(2, (40, 36)), (3, (16, 12, 10))])
@pytest.mark.parametrize("kind", [1, 2])
def test_c64_coordinate_grads_wide_reads_match(rng, dim, nm, kind, monkeypatch):
    """complex64 value and coordinate gradients through the gather with 16-byte tile reads
    (default) equal those of the 8-byte-read gather (NUFFTX_INTERP_WIDE=0) to 1e-5 relative:
    the two differ only in FP32 summation order."""
    import nufftax_amd as N

    M = 6000
    xs0 = [np.concatenate([rng.normal(0.5, 0.1, M - 1000), rng.uniform(-np.pi, np.pi, 1000)]) for _ in range(dim)]
    if kind == 1:
        d0 = rng.standard_normal(M) + 1j * rng.standard_normal(M)
        fn = [None, N.nufft2d1, N.nufft3d1][dim - 1]
    else:
        shp = tuple(reversed(nm))
        d0 = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        fn = [None, N.nufft2d2, N.nufft3d2][dim - 1]
    g0 = None

    def grads(wide):
        nonlocal g0
        monkeypatch.setenv("NUFFTX_INTERP_WIDE", wide)
        xs = [_t(x, torch.float32).requires_grad_(True) for x in xs0]
        d = _t(d0, torch.complex64).requires_grad_(True)
        out = fn(*xs, d, nm, eps=1e-6) if kind == 1 else fn(*xs, d, eps=1e-6)
        if g0 is None:
            g0 = torch.randn(out.shape, dtype=out.dtype, device=out.device)
        torch.sum((out * g0.conj()).real).backward()
        return [_np(t.grad) for t in xs + [d]]

    a = grads("1")
    b = grads("0")
    for u, v in zip(a, b):
        assert R.rel_l2(u, v) < 1e-5
