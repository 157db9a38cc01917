"""Pins the CPU oracle (oracle/nufft_ref.py) before anything is checked against it.

Each check is chosen so that a plausible slip in the restatement (a dropped
(-1)^k, a wrong isign mapping, swapped x/y axes, an off-by-one in i0, a wrong
derivative scale) fails: closed-form outputs from the reference's own edge-case
tests, the FP64 direct NUDFT, adjoint identities, brute-force loops in the
reference's spreading convention and central finite differences.
"""

import math

import numpy as np
import pytest
from scipy import integrate

from oracle import nufft_ref as R


# ---------------------------------------------------------------- parameters


@pytest.mark.parametrize(
    "eps,ns,beta",
    [(1e-4, 6, 13.8551), (1e-6, 8, 18.6389), (1e-9, 10, 23.3938), (1e-12, 14, 32.8668), (1e-3, 4, None)],
)
def test_kernel_params_table(eps, ns, beta):
    kp = R.compute_kernel_params(eps, 2.0)
    assert kp.nspread == ns
    assert kp.c == pytest.approx(4.0 / ns**2)
    if beta is not None:
        assert kp.beta == pytest.approx(beta, abs=1e-4)


def test_kernel_params_ranges():
    # tests/test_kernel.py:152-168: ns in [2, 16], even
    for e in [1e-1, 1e-2, 1e-5, 1e-7, 1e-14, 1e-20]:
        kp = R.compute_kernel_params(e)
        assert 2 <= kp.nspread <= 16 and kp.nspread % 2 == 0


def test_grid_sizes():
    assert [R.compute_grid_size(n, 2.0, 8) for n in (32, 64, 128, 256, 512, 1024)] == [64, 128, 256, 512, 1024, 2048]
    assert [R.next_smooth_int(n) for n in (1, 7, 11, 13, 97, 1021, 1025)] == [1, 8, 12, 15, 100, 1024, 1080]
    assert R.compute_grid_size(3, 2.0, 8) == 16  # 2*ns floor
    assert R.compute_grid_size(100, 1.25, 8) == 125
    # config-4 type 3: sources U[-pi,pi)^3, targets U[-64,64]^3, eps 1e-9 -> 270 per axis
    assert R.compute_type3_grid_size(math.pi, 64.0, eps=1e-9) == 270
    assert R.compute_grid_size(270, 2.0, 10) == 540
    for n in (3, 17, 101, 271):
        v = R.next_smooth_even(n)
        assert v % 2 == 0 and v >= n and R.next_smooth_int(v) == v


# ---------------------------------------------------------------- kernel


def test_es_kernel_properties():
    kp = R.compute_kernel_params(1e-6)
    z = np.linspace(-4, 4, 401)
    phi = R.es_kernel(z, kp.beta, kp.c)
    assert R.es_kernel(np.array([0.0]), kp.beta, kp.c)[0] == 1.0
    np.testing.assert_allclose(phi, phi[::-1], rtol=1e-13, atol=0)
    assert np.all(np.diff(phi[200:]) <= 0)
    assert np.all(R.es_kernel(np.array([4.01, -5.0, 10.0]), kp.beta, kp.c) == 0)
    assert R.es_kernel(np.array([4.0]), kp.beta, kp.c)[0] == pytest.approx(math.exp(-kp.beta))


def test_es_kernel_derivative_fd():
    kp = R.compute_kernel_params(1e-6)
    z = np.linspace(-3.9, 3.9, 57)
    _, d = R.es_kernel_with_derivative(z, kp.beta, kp.c)
    h = 1e-6
    fd = (R.es_kernel(z + h, kp.beta, kp.c) - R.es_kernel(z - h, kp.beta, kp.c)) / (2 * h)
    np.testing.assert_allclose(d, fd, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("eps", [1e-4, 1e-6, 1e-12])
def test_kernel_fourier_series_vs_adaptive_quadrature(eps):
    kp = R.compute_kernel_params(eps)
    nf = 64
    ph = R.kernel_fourier_series(nf, kp.nspread, kp.beta, kp.c)
    assert ph.shape == (nf // 2 + 1,)
    j2 = kp.nspread / 2
    for k in (0, 1, 7, 16, 31):
        exact, _ = integrate.quad(
            lambda z: R.es_kernel(np.array([z]), kp.beta, kp.c)[0] * math.cos(2 * math.pi * k * z / nf),
            -j2, j2, limit=200, epsabs=1e-14,
        )
        # the trapezoid rule is the reference's choice; its error is far below eps at used modes
        assert abs(ph[k] - exact) <= max(1e-2 * eps, 1e-12) * abs(ph[0])
    assert np.all(ph > 0) and np.all(np.diff(ph) < 0)
    np.testing.assert_allclose(
        R.kernel_ft_at_point(2 * np.pi * np.arange(nf // 2 + 1) / nf, kp.nspread, kp.beta, kp.c), ph, rtol=1e-13
    )


# ---------------------------------------------------------------- spread/interp


def _brute_spread(coords, c, nfs, kp):
    """Pure-Python loops in the reference's convention (tests/test_custom_kernel.py:64-104)."""
    d = len(nfs)
    fw = np.zeros(tuple(reversed(nfs)), dtype=complex)
    ns = kp.nspread
    xs = [R.fold_rescale(np.asarray(coords[a], dtype=np.float64), nfs[a]) for a in range(d)]
    i0 = [np.ceil(xs[a] - ns / 2).astype(int) for a in range(d)]
    for j in range(len(c)):
        for k in np.ndindex(*([ns] * d)):
            w = 1.0
            idx = []
            for a in range(d):
                z = (i0[a][j] + k[a]) - xs[a][j]
                w *= math.exp(kp.beta * (math.sqrt(max(1 - kp.c * z * z, 0.0)) - 1))
                idx.append((i0[a][j] + k[a]) % nfs[a])
            fw[tuple(reversed(idx))] += c[j] * w
    return fw


@pytest.mark.parametrize("nfs", [(20,), (16, 12), (12, 10, 16)])
def test_spread_matches_brute_force(rng, nfs):
    kp = R.compute_kernel_params(1e-4)
    d = len(nfs)
    M = 13
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(d)]
    coords[0][0] = -np.pi  # boundary point
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    np.testing.assert_allclose(R.spread(coords, c, nfs, kp), _brute_spread(coords, c, nfs, kp), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("nfs", [(32,), (24, 20), (12, 16, 10)])
def test_spread_interp_adjoint(rng, nfs):
    kp = R.compute_kernel_params(1e-6)
    d = len(nfs)
    M = 40
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(d)]
    c = rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))
    fw = rng.standard_normal((2,) + tuple(reversed(nfs))) + 1j * rng.standard_normal((2,) + tuple(reversed(nfs)))
    lhs = np.vdot(R.spread(coords, c, nfs, kp), fw)
    rhs = np.vdot(c, R.interp(coords, fw, kp))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_spread_coordinate_gradient_fd(rng):
    # real strengths and a real grid cotangent (the reference's own FD setting,
    # tests/test_custom_kernel.py:194-218)
    kp = R.compute_kernel_params(1e-6)
    nfs = (16, 12)
    M = 6
    coords = [rng.uniform(-2.0, 2.0, M) for _ in range(2)]
    c = rng.standard_normal(M).astype(complex)
    g = rng.standard_normal((12, 16))

    def loss(cc):
        return float(np.sum(g * np.real(R.spread(cc, c, nfs, kp))))

    gx, gy = R.spread_grad_coords(coords, c, g.astype(complex), kp)
    h = 1e-6
    for ax, gr in enumerate((gx, gy)):
        for j in range(M):
            cp = [a.copy() for a in coords]
            cm = [a.copy() for a in coords]
            cp[ax][j] += h
            cm[ax][j] -= h
            fd = (loss(cp) - loss(cm)) / (2 * h)
            assert gr[j] == pytest.approx(fd, rel=1e-5, abs=1e-7)


def test_interp_coordinate_gradient_fd(rng):
    kp = R.compute_kernel_params(1e-6)
    nfs = (10, 14, 12)
    M = 4
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(3)]
    fw = rng.standard_normal((12, 14, 10)) + 1j * rng.standard_normal((12, 14, 10))
    g = rng.standard_normal(M) + 1j * rng.standard_normal(M)

    def loss(cc):
        return float(np.real(np.vdot(g, R.interp(cc, fw, kp))))  # sum Re(conj(g) c)

    grads = R.interp_grad_coords(coords, fw, g, kp)
    h = 1e-6
    for ax in range(3):
        for j in range(M):
            cp = [a.copy() for a in coords]
            cm = [a.copy() for a in coords]
            cp[ax][j] += h
            cm[ax][j] -= h
            assert grads[ax][j] == pytest.approx((loss(cp) - loss(cm)) / (2 * h), rel=1e-5, abs=1e-7)


# ---------------------------------------------------------------- deconvolution


def test_deconv_pad_shuffle_consistency(rng):
    kp = R.compute_kernel_params(1e-6)
    nm = (7, 10)  # odd and even, non-square
    nfs = [R.compute_grid_size(n, 2.0, kp.nspread) for n in nm]
    ph = [R.kernel_fourier_series(nf, kp.nspread, kp.beta, kp.c) for nf in nfs]
    f = rng.standard_normal((3, nm[1], nm[0])) + 1j * rng.standard_normal((3, nm[1], nm[0]))
    padded = R.deconvolve_pad(f, ph, nfs)
    back = R.deconvolve_shuffle(padded, ph, nm)
    fac = np.outer(R.deconv_factors(ph[1], nm[1]), R.deconv_factors(ph[0], nm[0]))
    np.testing.assert_allclose(back, f * fac**2, rtol=1e-13)
    # everything outside the kept corners is zero
    assert np.count_nonzero(padded[0]) == nm[0] * nm[1]
    # CMCL: mode 0 sits at fine-grid index 0, mode -1 at nf-1
    assert padded[0, 0, 0] == pytest.approx(f[0, nm[1] // 2, nm[0] // 2] * fac[nm[1] // 2, nm[0] // 2])


# ---------------------------------------------------------------- transforms


def test_type1_single_point_closed_form():
    # tests/test_edge_cases.py:17-28
    f = R.nufft_type1([np.array([0.5])], np.array([1.0 + 1j]), 16, eps=1e-6)
    k = np.arange(-8, 8)
    np.testing.assert_allclose(f, (1 + 1j) * np.exp(1j * k * 0.5), rtol=1e-5)


def test_type1_points_at_0_and_pi():
    # tests/test_edge_cases.py:30-70
    k = np.arange(-16, 16)
    f0 = R.nufft_type1([np.zeros(10)], np.ones(10, complex), 32)
    np.testing.assert_allclose(f0, np.full(32, 10.0), rtol=1e-5)
    fpi = R.nufft_type1([np.full(10, np.pi)], np.ones(10, complex), 32)
    np.testing.assert_allclose(fpi, 10 * np.exp(1j * k * np.pi), rtol=1e-5, atol=1e-5)
    f2 = R.nufft_type1([np.array([0.0, np.pi])], np.ones(2, complex), 8)
    np.testing.assert_allclose(f2.real, 1 + np.cos(np.arange(-4, 4) * np.pi), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(f2.imag, 0, atol=1e-6)


def test_type1_uniform_grid_delta():
    # tests/test_edge_cases.py:72-87
    x = np.linspace(-np.pi, np.pi, 32, endpoint=False)
    f = R.nufft_type1([x], np.ones(32, complex), 32)
    assert abs(f[16]) == pytest.approx(32, rel=1e-5)
    assert np.max(np.abs(np.delete(f, 16))) < 1e-4


def test_type1_2d_3d_single_point_shapes():
    # tests/test_edge_cases.py:332-346, :362-380: output (n2, n1) and (n3, n2, n1)
    f = R.nufft_type1([np.array([0.3]), np.array([-1.1])], np.array([2.0 + 0j]), (6, 4))
    assert f.shape == (4, 6)
    k1, k2 = np.arange(-3, 3), np.arange(-2, 2)
    np.testing.assert_allclose(f, 2 * np.exp(1j * (k2[:, None] * -1.1 + k1[None, :] * 0.3)), rtol=1e-5)
    f3 = R.nufft_type1([np.array([0.3]), np.array([-1.1]), np.array([2.0])], np.array([1.0 + 0j]), (4, 6, 8))
    assert f3.shape == (8, 6, 4)
    k3, k2b, k1b = np.arange(-4, 4), np.arange(-3, 3), np.arange(-2, 2)
    exp3 = np.exp(1j * (k3[:, None, None] * 2.0 + k2b[None, :, None] * -1.1 + k1b[None, None, :] * 0.3))
    np.testing.assert_allclose(f3, exp3, rtol=1e-5)


def test_wrap_invariance(rng):
    # tests/test_edge_cases.py:105-120
    x = rng.uniform(np.pi, 3 * np.pi, 50)
    c = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    np.testing.assert_allclose(R.nufft_type1([x], c, 32), R.nufft_type1([R.wrap_pi(x)], c, 32), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("d,nm", [(1, (33,)), (2, (12, 9)), (3, (8, 6, 10))])
@pytest.mark.parametrize("isign", [1, -1])
def test_type1_type2_vs_direct(rng, d, nm, isign):
    M = 60
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(d)]
    c = rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))
    f1 = R.nufft_type1(coords, c, nm, eps=1e-9, isign=isign)
    assert R.rel_l2(f1, R.direct_type1(coords, c, nm, isign=isign)) < 1e-8
    fm = rng.standard_normal((2,) + tuple(reversed(nm))) + 1j * rng.standard_normal((2,) + tuple(reversed(nm)))
    c2 = R.nufft_type2(coords, fm, eps=1e-9, isign=isign)
    assert R.rel_l2(c2, R.direct_type2(coords, fm, isign=isign)) < 1e-8


def test_complex64_proxy_accuracy(rng):
    # tests/test_dtypes.py:66-98: complex64 rel err < 1e-4, complex128 < 1e-5 at eps 1e-6
    M, N = 50, 32
    x = rng.uniform(-np.pi, np.pi, M)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    exp = R.direct_type1([x], c, N)
    f32 = R.nufft_type1([x.astype(np.float32)], c.astype(np.complex64), N)
    assert f32.dtype == np.complex64 and R.rel_l2(f32, exp) < 1e-4
    assert R.rel_l2(R.nufft_type1([x], c, N), exp) < 1e-5


@pytest.mark.parametrize("d", [1, 2, 3])
def test_type1_type2_adjoint(rng, d):
    # tests/test_nufft2.py:307-421: <T1 c, f> = <c, T2 f> with isign +1 vs -1
    nm = [(16,), (10, 8), (6, 8, 4)][d - 1]
    M = 30
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(d)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    f = rng.standard_normal(tuple(reversed(nm))) + 1j * rng.standard_normal(tuple(reversed(nm)))
    lhs = np.vdot(f, R.nufft_type1(coords, c, nm, eps=1e-9, isign=1))
    rhs = np.vdot(R.nufft_type2(coords, f, eps=1e-9, isign=-1), c)
    assert abs(lhs - rhs) < 1e-6 * abs(lhs)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_type3_vs_direct(rng, d):
    M, N = 40, 35
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(d)]
    targets = [rng.uniform(-12, 12, N) for _ in range(d)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    nfs = tuple(R.compute_type3_grid_size(coords[a], targets[a], eps=1e-9) for a in range(d))
    out = R.nufft_type3(coords, c, targets, nfs, eps=1e-9, isign=1)
    # tests/test_nufft3.py:105-292: rel < 1e-4 (1e-3 in 3D) -- eps 1e-9 is far tighter
    assert R.rel_l2(out, R.direct_type3(coords, c, targets, isign=1)) < 1e-6


def test_type3_swap_adjoint(rng):
    # primitives.py:376-380: T3 is transposed by swapping sources and targets
    M, N = 25, 30
    x = rng.uniform(-2, 2, M)
    s = rng.uniform(-5, 5, N)
    nf = R.compute_type3_grid_size(x, s, eps=1e-9)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    g = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    lhs = np.sum(g * R.nufft_type3([x], c, [s], nf, eps=1e-9))
    rhs = np.sum(c * R.nufft_type3([s], g, [x], nf, eps=1e-9))
    assert abs(lhs - rhs) < 1e-6 * abs(lhs)


# ---------------------------------------------------------------- bin sort


def test_bin_keys_and_sort(rng):
    nfs, bins = (64, 48), (16, 16)
    x = rng.uniform(-np.pi, np.pi, 500)
    y = rng.uniform(-np.pi, np.pi, 500)
    x[:3] = [-np.pi, np.pi - 1e-16, 0.0]
    keys = R.bin_keys([x, y], nfs, bins)
    xs = R.fold_rescale(x, 64)
    ys = R.fold_rescale(y, 48)
    np.testing.assert_array_equal(keys, (xs // 16).astype(int) + 4 * (ys // 16).astype(int))
    perm, offs = R.bin_sort(keys, 4 * 3)
    assert offs[-1] == 500 and np.all(np.diff(keys[perm]) >= 0)
    for b in range(12):
        seg = perm[offs[b]:offs[b + 1]]
        assert np.all(keys[seg] == b) and np.all(np.diff(seg) > 0)  # stable

