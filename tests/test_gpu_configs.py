"""Parity at the full sizes of the measured workloads, through size-independent checks.

At these sizes the oracle does not finish in seconds, so each case is checked
against the FP64 direct NUDFT on a seeded subset of modes / points (evaluated
on the GPU in FP64 by plain torch arithmetic, the mathematical definition of
the transform) and, where it applies, through adjointness.

Bars (relative l2 on the subset):
  * 1D type 1, M = 100,000, 1024 modes, complex64, eps 1e-6: <= 2x the error of
    the complex64 restatement of the reference algorithm (oracle, full input) + 1e-7;
  * 2D radial, M = 3,999,744, 512^2 modes, complex64, eps 1e-6: type 1 and type 2 each
    <= 2x the error of the complex64 restatement of the reference algorithm run on the
    full input (the reference-error proxy, same 64 modes / 64 points) + 1e-7, and < 1e-5;
    <T1 c, f> = <c, T2 f> to 1e-5 relative;
  * 2D round trip, M = 1e6, 256^2 modes: complex64 gradients within 1e-4 of the
    complex128 (eps 1e-12) gradients; complex128 coordinate gradients on 32 coordinates
    match central differences at h = 1e-3 to 1e-2 (the O(h^2) truncation) and their
    Richardson extrapolation with h / 2 to 2e-4;
  * 3D Gaussian clusters, M = 8e6, 128^3 modes, complex64: type 1 < 10 eps at eps 1e-4
    and 1e-6; type 2 on 64 seeded points <= 2x the complex64 proxy at those points + 1e-7;
  * 3D complex128, M = 5e5, 64^3 modes, eps 1e-12: type 1 and type 2 < 1e-10;
  * 3D type 3, M = N = 2e5, complex128, eps 1e-9: < 1e-8.
"""

import numpy as np
import pytest
import torch

from nufftax_amd import workloads
from oracle import nufft_ref as R

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _t(a, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(DEV)


def _direct_modes(coords, c, modes, isign, chunk=1 << 22):
    """sum_j c_j exp(i isign k . x_j) for each mode tuple k, FP64 on the GPU."""
    xs = [_t(x, torch.float64) for x in coords]
    cd = _t(c, torch.complex128)
    out = torch.zeros(len(modes), dtype=torch.complex128, device=DEV)
    for s in range(0, cd.shape[0], chunk):
        e = min(cd.shape[0], s + chunk)
        for i, k in enumerate(modes):
            ph = sum(float(kd) * x[s:e] for kd, x in zip(k, xs)) * isign
            out[i] += torch.sum(cd[s:e] * torch.polar(torch.ones_like(ph), ph))
    return out.cpu().numpy()


def _direct_points(coords, f, idx, isign):
    """sum_k f_k exp(i isign k . x_j) at the points idx, FP64 on the GPU (f is (n_d, .., n_1))."""
    dim = len(coords)
    fd = _t(f, torch.complex128)
    ks = [torch.arange(-(n // 2), (n + 1) // 2, device=DEV, dtype=torch.float64) for n in reversed(f.shape)]
    out = []
    for j in idx:
        ph = torch.zeros(f.shape, dtype=torch.float64, device=DEV)
        for d in range(dim):
            shape = [1] * dim
            shape[dim - 1 - d] = f.shape[dim - 1 - d]
            ph = ph + ks[d].reshape(shape) * float(coords[d][j])
        ph = ph * isign
        out.append(torch.sum(fd * torch.polar(torch.ones_like(ph), ph)))
    return torch.stack(out).cpu().numpy()


def _pick(out, modes, nm):
    """Entries of a CMCL-ordered (n_d, .., n_1) array at mode tuples (k1, .., kd)."""
    return np.array([out[tuple(k[d] + nm[d] // 2 for d in reversed(range(len(nm))))] for k in modes])


def _rand_modes(rng, nm, n=64):
    return [tuple(int(rng.integers(-(m // 2), (m + 1) // 2)) for m in nm) for _ in range(n)]


def test_1d_type1_100k():
    import nufftax_amd as N

    x, c = workloads.uniform_1d_type1(seed=0)
    x32, c64 = x.astype(np.float32), c.astype(np.complex64)
    out = N.nufft1d1(_t(x32), _t(c64), 1024, eps=1e-6).cpu().numpy()
    modes = [(k,) for k in range(-512, -448)]  # output indices 0..63
    truth = _direct_modes([x32.astype(np.float64)], c64.astype(np.complex128), modes, 1)
    err = R.rel_l2(out[:64], truth)
    proxy_out = R.nufft_type1([x32], c64, (1024,), eps=1e-6)
    proxy = R.rel_l2(proxy_out[:64], truth)
    assert err <= 2 * proxy + 1e-7, (err, proxy)


def test_2d_radial_4m_type1_type2_adjoint():
    import nufftax_amd as N

    x, y, c, f = workloads.radial_2d(seed=0)
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    c64, f64 = c.astype(np.complex64), f.astype(np.complex64)
    xd, yd, cd, fd = _t(x32), _t(y32), _t(c64), _t(f64)
    F = N.nufft2d1(xd, yd, cd, (512, 512), eps=1e-6, isign=1)
    C = N.nufft2d2(xd, yd, fd, eps=1e-6, isign=-1)
    rng = np.random.default_rng(1)
    modes = _rand_modes(rng, (512, 512))
    coords64 = [x32.astype(np.float64), y32.astype(np.float64)]
    truth1 = _direct_modes(coords64, c64.astype(np.complex128), modes, 1)
    err1 = R.rel_l2(_pick(F.cpu().numpy(), modes, (512, 512)), truth1)
    idx = rng.integers(0, x.size, 64)
    truth2 = _direct_points(coords64, f64.astype(np.complex128), idx, -1)
    err2 = R.rel_l2(C.cpu().numpy()[idx], truth2)
    # the reference-error proxy: the complex64 restatement on the full input
    proxy1 = R.rel_l2(_pick(R.nufft_type1([x32, y32], c64, (512, 512), eps=1e-6, isign=1), modes, (512, 512)),
                      truth1)
    proxy2 = R.rel_l2(R.nufft_type2([x32[idx], y32[idx]], f64, eps=1e-6, isign=-1), truth2)
    print(f"radial 4M: type 1 {err1:.2e} (proxy {proxy1:.2e}), type 2 {err2:.2e} (proxy {proxy2:.2e})")
    assert err1 <= 2 * proxy1 + 1e-7 and err1 < 1e-5, (err1, proxy1)
    assert err2 <= 2 * proxy2 + 1e-7 and err2 < 1e-5, (err2, proxy2)
    lhs = torch.vdot(F.reshape(-1).to(torch.complex128), fd.reshape(-1).to(torch.complex128))
    rhs = torch.vdot(cd.to(torch.complex128), C.to(torch.complex128))
    assert abs(complex(lhs - rhs)) <= 1e-5 * abs(complex(lhs))


def test_2d_roundtrip_1m_gradients():
    import nufftax_amd as N

    (x, y), _ = workloads.uniform_nd(seed=0, M=1_000_000, dim=2)
    rng = np.random.default_rng(2)
    f0 = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))
    g = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))

    def grads(cdt, rdt, eps):
        xs = _t(x, rdt).requires_grad_(True)
        ys = _t(y, rdt).requires_grad_(True)
        f = _t(f0, cdt).requires_grad_(True)
        out = N.nufft2d1(xs, ys, N.nufft2d2(xs, ys, f, eps=eps), (256, 256), eps=eps)
        loss = torch.sum(torch.abs(out - _t(g, cdt)) ** 2) / 1e12
        loss.backward()
        return [t.grad.detach().to(torch.complex128 if t.is_complex() else torch.float64) for t in (f, xs, ys)]

    g64 = grads(torch.complex128, torch.float64, 1e-12)
    g32 = grads(torch.complex64, torch.float32, 1e-6)
    for a, b in zip(g32, g64):
        assert R.rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-4

    def loss64(xv):
        with torch.no_grad():
            xs = _t(xv, torch.float64)
            ys = _t(y, torch.float64)
            out = N.nufft2d1(xs, ys, N.nufft2d2(xs, ys, _t(f0), eps=1e-12), (256, 256), eps=1e-12)
            return float(torch.sum(torch.abs(out - _t(g)) ** 2) / 1e12)

    gx = g64[1].cpu().numpy()[idx_fd := rng.choice(x.size, 32, replace=False)]

    def central(h):
        out = []
        for j in idx_fd:
            xp, xm = x.copy(), x.copy()
            xp[j] += h
            xm[j] -= h
            out.append((loss64(xp) - loss64(xm)) / (2 * h))
        return np.array(out)

    # central differences at the configured step h = 1e-3 carry a truncation error of about
    # (k h)^2 / 6 for modes up to |k| ~ 128 sqrt(2) (about 5e-3); Richardson extrapolation with
    # h / 2 cancels that term, leaving the O(h^4) term and the FP64 rounding of the loss
    d1, d2 = central(1e-3), central(5e-4)
    rich = (4 * d2 - d1) / 3
    print(f"FD rel l2: h=1e-3 {R.rel_l2(gx, d1):.2e}, Richardson {R.rel_l2(gx, rich):.2e}")
    assert R.rel_l2(gx, d1) < 1e-2, R.rel_l2(gx, d1)
    assert R.rel_l2(gx, rich) < 2e-4, R.rel_l2(gx, rich)


@pytest.mark.parametrize("eps", [1e-4, 1e-6])
def test_3d_gaussian_clusters_8m_c64(eps):
    import nufftax_amd as N

    coords, c = workloads.gaussian_mixture_3d(seed=0)
    c32 = [x.astype(np.float32) for x in coords]
    c64 = c.astype(np.complex64)
    out = N.nufft3d1(*[_t(x) for x in c32], _t(c64), (128, 128, 128), eps=eps).cpu().numpy()
    rng = np.random.default_rng(3)
    modes = _rand_modes(rng, (128, 128, 128))
    truth = _direct_modes([x.astype(np.float64) for x in c32], c64.astype(np.complex128), modes, 1)
    assert R.rel_l2(_pick(out, modes, (128, 128, 128)), truth) < 10 * eps


@pytest.mark.parametrize("eps", [1e-4, 1e-6])
def test_3d_gaussian_clusters_8m_c64_type2(eps):
    import nufftax_amd as N

    coords, _ = workloads.gaussian_mixture_3d(seed=0)
    c32 = [x.astype(np.float32) for x in coords]
    rng = np.random.default_rng(6)
    f = (rng.standard_normal((128, 128, 128)) + 1j * rng.standard_normal((128, 128, 128))).astype(np.complex64)
    out = N.nufft3d2(*[_t(x) for x in c32], _t(f), eps=eps).cpu().numpy()
    idx = rng.integers(0, c32[0].size, 64)
    truth = _direct_points([x.astype(np.float64) for x in c32], f.astype(np.complex128), idx, -1)
    err = R.rel_l2(out[idx], truth)
    # reference-error proxy at the same points: complex64 restatement (its interpolation is per point)
    proxy = R.rel_l2(R.nufft_type2([x[idx] for x in c32], f, eps=eps, isign=-1), truth)
    print(f"3D clusters 8M type 2 eps {eps}: {err:.2e} (proxy {proxy:.2e})")
    assert err <= 2 * proxy + 1e-7, (err, proxy)


def test_3d_c128_eps1e12_500k():
    import nufftax_amd as N

    coords, c = workloads.gaussian_mixture_3d(seed=1, M=500_000)
    out = N.nufft3d1(*[_t(x) for x in coords], _t(c), (64, 64, 64), eps=1e-12).cpu().numpy()
    rng = np.random.default_rng(4)
    modes = _rand_modes(rng, (64, 64, 64))
    truth = _direct_modes(coords, c, modes, 1)
    assert R.rel_l2(_pick(out, modes, (64, 64, 64)), truth) < 1e-10
    f = rng.standard_normal((64, 64, 64)) + 1j * rng.standard_normal((64, 64, 64))
    out2 = N.nufft3d2(*[_t(x) for x in coords], _t(f), eps=1e-12).cpu().numpy()
    idx = rng.integers(0, coords[0].size, 64)
    truth2 = _direct_points(coords, f, idx, -1)
    assert R.rel_l2(out2[idx], truth2) < 1e-10


def test_3d_type3_200k_c128():
    import nufftax_amd as N

    xs, ss, c = workloads.type3_3d(seed=0)
    nf = tuple(N.compute_type3_grid_size(xs[d], ss[d], eps=1e-9) for d in range(3))
    assert nf == (270, 270, 270)
    out = N.nufft3d3(*[_t(x) for x in xs], _t(c), *[_t(s) for s in ss], nf, eps=1e-9).cpu().numpy()
    rng = np.random.default_rng(5)
    idx = rng.integers(0, ss[0].size, 64)
    modes = [tuple(float(s[j]) for s in ss) for j in idx]
    truth = _direct_modes(xs, c, modes, 1)
    assert R.rel_l2(out[idx], truth) < 1e-8
