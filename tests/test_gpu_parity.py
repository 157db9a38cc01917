"""GPU parity of the HIP path (through the C ABI) against the oracle.

Tolerances:
  * bin-sort: bit-exact (perm == argsort(keys, stable), bin_start == cumsum(bincount));
  * spread/interp vs the FP64 oracle: rel l2 < 2e-6 in FP32, < 1e-12 in FP64;
  * transforms: relative l2 error against the FP64 direct NUDFT no worse than
    2x the error of the complex64 restatement of the reference algorithm (the
    reference-error proxy) plus 1e-7, and < 1e-10 * 10 in FP64 at eps 1e-12.
"""

import math

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
    import nufftax_amd  # noqa: F401


# ------------------------------------------------------------------ bin sort


@pytest.mark.parametrize("dim,n_modes,M", [(1, (1000,), 50000), (2, (100, 70), 30000), (2, (512, 512), 200000),
                                           (2, (1024, 1024), 300000), (3, (40, 32, 24), 40000),
                                           (3, (64, 64, 64), 200000), (3, (128, 128, 128), 300000)])
@pytest.mark.parametrize("dist", ["uniform", "clustered"])
def test_bin_sort_bit_exact(rng, dim, n_modes, M, dist):
    from nufftax_amd._plan import F32, transform_plan

    if dist == "uniform":
        coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    else:  # a dense spot plus a radial line: many equal keys in a row
        coords = [np.concatenate([rng.normal(0.0, 0.01, M // 2), np.linspace(-np.pi, np.pi, M - M // 2)])
                  for _ in range(dim)]
    coords[0][:5] = [-np.pi, np.pi, 0.0, -1e-300, 3 * np.pi]  # edges and wrap
    xs = [_t(c.astype(np.float32)) for c in coords]
    plan = transform_plan(1, dim, n_modes, 1, 1e-6, 1, 2.0, F32, torch.device(DEV))
    plan.setpts(xs)
    perm, bst, nbins, bs, nbd = plan.sort_state()
    torch.cuda.synchronize()
    nfs = [plan.grid_size(d) for d in range(dim)]
    keys = R.bin_keys([c.astype(np.float32).astype(np.float64) for c in coords], nfs, bs[:dim])
    assert int(np.prod(nbd[:dim])) == nbins
    ref_perm, ref_off = R.bin_sort(keys, nbins)
    np.testing.assert_array_equal(_np(perm).astype(np.int64), ref_perm)
    np.testing.assert_array_equal(_np(bst).astype(np.int64), ref_off)


def _check_sort(coords, n_modes, kind=1):
    """perm and bin offsets of a plan's sort equal the oracle's stable sort of the same keys."""
    from nufftax_amd._plan import F32, transform_plan

    dim = len(coords)
    xs = [_t(c.astype(np.float32)) for c in coords]
    plan = transform_plan(kind, dim, n_modes, 1, 1e-6, 1 if kind == 1 else -1, 2.0, F32, torch.device(DEV))
    plan.setpts(xs)
    perm, bst, nbins, bs, nbd = plan.sort_state()
    torch.cuda.synchronize()
    nfs = [plan.grid_size(d) for d in range(dim)]
    keys = R.bin_keys([c.astype(np.float32).astype(np.float64) for c in coords], nfs, bs[:dim])
    ref_perm, ref_off = R.bin_sort(keys, nbins)
    np.testing.assert_array_equal(_np(perm).astype(np.int64), ref_perm)
    np.testing.assert_array_equal(_np(bst).astype(np.int64), ref_off)


@pytest.mark.parametrize("M", [1, 63, 64, 65, 8191, 8193, 16384 + 5])
@pytest.mark.parametrize("kind", [1, 2])
def test_bin_sort_ragged_sizes(rng, M, kind):
    """sizes around the 64-key ranking step and the 8192-point chunk; type 2 also writes records."""
    _check_sort([rng.uniform(-np.pi, np.pi, M) for _ in range(2)], (64, 48), kind)


def test_bin_sort_one_bin_many_points(rng):
    """every point in one bin (the one-comparison ranking path), plus a second bin at the end."""
    M = 100_000
    x = np.full(M, 0.25)
    y = np.full(M, -1.0)
    x[-3:] = 2.5
    _check_sort([x, y], (128, 128), kind=2)


def test_bin_sort_clusters_8m_full_size():
    """the 3D benchmark's point set (8M clustered points, 256^3 grid, 16384 bins: the 16-bit
    wave-counter scatter), type-1 plan."""
    from nufftax_amd import workloads

    coords, _ = workloads.gaussian_mixture_3d(seed=0)
    _check_sort(coords, (128, 128, 128), kind=1)


def test_bin_sort_radial_4m_full_size():
    """the benchmark's point set (3,999,744 radial samples, 1024^2 grid), type-2 plan (records)."""
    from nufftax_amd import workloads

    x, y, _, _ = workloads.radial_2d(seed=0)
    _check_sort([x, y], (512, 512), kind=2)


# ------------------------------------------------------------------ spread / interp


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dim,nfs", [(1, (300,)), (2, (96, 64)), (3, (40, 36, 32))])
def test_spread_interp_vs_oracle(rng, prec, dim, nfs):
    from nufftax_amd._plan import F32, F64, grid_plan

    rdt = np.float32 if prec == "f32" else np.float64
    cdt = np.complex64 if prec == "f32" else np.complex128
    kp = R.compute_kernel_params(1e-6 if prec == "f32" else 1e-12)
    M = 5000
    coords = [rng.uniform(-np.pi, np.pi, M).astype(rdt) for _ in range(dim)]
    c = (rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))).astype(cdt)
    plan = grid_plan(dim, nfs, kp.nspread, kp.beta, kp.c, 2, F32 if prec == "f32" else F64, torch.device(DEV))
    xs = [_t(x) for x in coords]
    plan.setpts(xs)
    fw = torch.empty((2,) + tuple(reversed(nfs)), dtype=torch.complex64 if prec == "f32" else torch.complex128,
                     device=DEV)
    plan.spread(_t(c), fw)
    ref = R.spread([x.astype(np.float64) for x in coords], c.astype(np.complex128), nfs, kp)
    tol = 2e-6 if prec == "f32" else 1e-12
    assert R.rel_l2(_np(fw), ref) < tol
    g = (rng.standard_normal((2,) + tuple(reversed(nfs))) + 1j * rng.standard_normal((2,) + tuple(reversed(nfs))))
    out = torch.empty((2, M), dtype=fw.dtype, device=DEV)
    plan.interp(_t(g.astype(cdt)), out)
    refi = R.interp([x.astype(np.float64) for x in coords], g.astype(cdt).astype(np.complex128), kp)
    assert R.rel_l2(_np(out), refi) < tol


def test_spread_global_path_matches_tiled(rng, monkeypatch):
    """A tile larger than the LDS budget takes the global-memory path: same result."""
    from nufftax_amd._plan import F64, grid_plan

    monkeypatch.setenv("NUFFTX_LDS_TILE_MAX", "65536")  # ns = 16 in 3D FP64: 128 KB tiles
    kp = R.compute_kernel_params(1e-14)
    nfs = (40, 36, 34)
    M = 3000
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(3)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    plan = grid_plan(3, nfs, kp.nspread, kp.beta, kp.c, 1, F64, torch.device(DEV))
    plan.setpts([_t(x) for x in coords])
    fw = torch.empty(tuple(reversed(nfs)), dtype=torch.complex128, device=DEV)
    plan.spread(_t(c), fw)
    ref = R.spread(coords, c, nfs, kp)
    assert R.rel_l2(_np(fw), ref) < 1e-12


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("eps", [1e-3, 1e-6, 1e-10])
@pytest.mark.parametrize("dim", [2, 3])
def test_global_path_wave_vs_thread(rng, dim, prec, eps, monkeypatch):
    """2D and 3D spreading and interpolation on the global-memory path (tiles above
    NUFFTX_LDS_TILE_MAX): one wave per point (default) and one thread per point
    (NUFFTX_GLOBAL_WAVE=0), each against the FP64 oracle, and against each other; two
    transforms, clustered and spread-out points, grids with wrapped supports."""
    from nufftax_amd._plan import F32, F64, grid_plan

    if prec == "f32" and eps < 1e-7:
        pytest.skip("FP32 plans take eps >= 1e-7")
    monkeypatch.setenv("NUFFTX_LDS_TILE_MAX", "4096")
    kp = R.compute_kernel_params(eps)
    nfs = (36, 30, 26)[:dim] if dim == 3 else (120, 90)
    M = 4000
    coords = [np.concatenate([rng.normal(3.0, 0.1, M - 1500), rng.uniform(-np.pi, np.pi, 1500)]) for _ in range(dim)]
    c = rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))
    g = rng.standard_normal((2,) + tuple(reversed(nfs))) + 1j * rng.standard_normal((2,) + tuple(reversed(nfs)))
    cdt, rdt, pr = (np.complex64, np.float32, F32) if prec == "f32" else (np.complex128, np.float64, F64)
    tol = 2e-6 if prec == "f32" else 1e-12
    ref_s = R.spread(coords, c, nfs, kp)
    ref_i = R.interp(coords, g, kp)
    got = {}
    for wave in ("1", "0"):
        monkeypatch.setenv("NUFFTX_GLOBAL_WAVE", wave)
        plan = grid_plan(dim, nfs, kp.nspread, kp.beta, kp.c, 2, pr, torch.device(DEV))
        plan.setpts([_t(x.astype(rdt)) for x in coords])
        fw = torch.full((2,) + tuple(reversed(nfs)), float("nan"), dtype=_t(c.astype(cdt)).dtype, device=DEV)
        plan.spread(_t(c.astype(cdt)), fw)
        out = torch.empty((2, M), dtype=fw.dtype, device=DEV)
        plan.interp(_t(g.astype(cdt)), out)
        got[wave] = (_np(fw), _np(out))
        assert R.rel_l2(got[wave][0], ref_s) < tol
        assert R.rel_l2(got[wave][1], ref_i) < tol
    assert R.rel_l2(got["1"][0], got["0"][0]) < tol
    assert R.rel_l2(got["1"][1], got["0"][1]) < tol


# ------------------------------------------------------------------ transforms vs direct NUDFT


def _proxy_err_type1(coords, c, nm, isign, eps):
    """Error of the complex64 restatement of the reference algorithm vs the FP64 NUDFT."""
    f32 = R.nufft_type1([x.astype(np.float32) for x in coords], c.astype(np.complex64), nm, eps=eps, isign=isign)
    return R.rel_l2(f32, R.direct_type1(coords, c, nm, isign=isign))


@pytest.mark.parametrize("dim,nm", [(1, (64,)), (1, (33,)), (2, (32, 24)), (2, (17, 20)), (3, (12, 10, 16))])
@pytest.mark.parametrize("isign", [1, -1])
def test_type1_vs_direct_c64(rng, dim, nm, isign):
    import nufftax_amd as N

    M = 2000
    coords = [rng.uniform(-np.pi, np.pi, M).astype(np.float32).astype(np.float64) for _ in range(dim)]
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64).astype(np.complex128)
    fn = [N.nufft1d1, N.nufft2d1, N.nufft3d1][dim - 1]
    args = [_t(x, torch.float32) for x in coords] + [_t(c, torch.complex64)]
    out = fn(*args, nm if dim > 1 else nm[0], eps=1e-6, isign=isign)
    assert out.dtype == torch.complex64 and tuple(out.shape) == tuple(reversed(nm))
    truth = R.direct_type1(coords, c, nm, isign=isign)
    err = R.rel_l2(_np(out), truth)
    proxy = _proxy_err_type1(coords, c, nm, isign, 1e-6)
    assert err <= 2 * proxy + 1e-7, (err, proxy)
    assert err < 1e-5


@pytest.mark.parametrize("dim,nm", [(2, (48, 40)), (3, (20, 16, 12))])
@pytest.mark.parametrize("eps", [1e-1, 1e-3, 1e-4])
@pytest.mark.parametrize("dist", ["uniform", "clustered"])
def test_type1_packed_tiles_c64(rng, dim, nm, eps, dist, monkeypatch):
    """complex64 type 1 at eps >= 1e-4 spreads into packed tiles (real and imaginary 32-bit
    fixed point in one 64-bit word, a unit <= 2^-19 max|c| for 1024-point subproblems).
    Bars: error vs the FP64 direct sum <= 2x the oracle's complex64 error + 2e-5, and packed
    vs 64-bit-per-part tiles (NUFFTX_NO_PACKED) within 3e-5 relative l2."""
    import nufftax_amd as N

    M = 20000
    if dist == "uniform":
        coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    else:
        coords = [np.concatenate([rng.normal(0.3, 0.02, M - 500), rng.uniform(-np.pi, np.pi, 500)]) for _ in range(dim)]
    coords = [x.astype(np.float32).astype(np.float64) for x in coords]
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64).astype(np.complex128)
    fn = [N.nufft1d1, N.nufft2d1, N.nufft3d1][dim - 1]
    args = [_t(x, torch.float32) for x in coords] + [_t(c, torch.complex64)]
    packed = _np(fn(*args, nm, eps=eps))
    monkeypatch.setenv("NUFFTX_NO_PACKED", "1")
    full = _np(fn(*args, nm, eps=eps))
    monkeypatch.delenv("NUFFTX_NO_PACKED")
    assert R.rel_l2(packed, full) < 3e-5
    truth = R.direct_type1(coords, c, nm, isign=1)
    proxy = _proxy_err_type1(coords, c, nm, 1, eps)
    err = R.rel_l2(packed, truth)
    assert err <= 2 * proxy + 2e-5, (err, proxy)


@pytest.mark.parametrize("dim,nm", [(1, (40,)), (2, (24, 18)), (3, (10, 12, 8))])
@pytest.mark.parametrize("isign", [1, -1])
def test_type2_vs_direct_c64(rng, dim, nm, isign):
    import nufftax_amd as N

    M = 1500
    coords = [rng.uniform(-np.pi, np.pi, M).astype(np.float32).astype(np.float64) for _ in range(dim)]
    f = (rng.standard_normal(tuple(reversed(nm))) + 1j * rng.standard_normal(tuple(reversed(nm))))
    f = f.astype(np.complex64).astype(np.complex128)
    fn = [N.nufft1d2, N.nufft2d2, N.nufft3d2][dim - 1]
    out = fn(*[_t(x, torch.float32) for x in coords], _t(f, torch.complex64), eps=1e-6, isign=isign)
    assert out.dtype == torch.complex64 and tuple(out.shape) == (M,)
    truth = R.direct_type2(coords, f, isign=isign)
    err = R.rel_l2(_np(out), truth)
    proxy = R.rel_l2(R.nufft_type2([x.astype(np.float32) for x in coords], f.astype(np.complex64), eps=1e-6,
                                   isign=isign), truth)
    assert err <= 2 * proxy + 1e-7, (err, proxy)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_type12_c128_high_precision(rng, dim):
    import nufftax_amd as N

    nm = [(50,), (20, 16), (10, 8, 12)][dim - 1]
    M = 800
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    f1 = [N.nufft1d1, N.nufft2d1, N.nufft3d1][dim - 1]
    out = f1(*[_t(x) for x in coords], _t(c), nm if dim > 1 else nm[0], eps=1e-12)
    assert out.dtype == torch.complex128
    assert R.rel_l2(_np(out), R.direct_type1(coords, c, nm)) < 1e-10
    f = rng.standard_normal(tuple(reversed(nm))) + 1j * rng.standard_normal(tuple(reversed(nm)))
    f2 = [N.nufft1d2, N.nufft2d2, N.nufft3d2][dim - 1]
    out2 = f2(*[_t(x) for x in coords], _t(f), eps=1e-12)
    assert R.rel_l2(_np(out2), R.direct_type2(coords, f)) < 1e-10


def test_complex64_small_eps_runs_in_fp64(rng):
    import nufftax_amd as N

    M, n = 500, 40
    x = rng.uniform(-np.pi, np.pi, M).astype(np.float32).astype(np.float64)
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64).astype(np.complex128)
    out = N.nufft1d1(_t(x, torch.float32), _t(c, torch.complex64), n, eps=1e-9)
    assert out.dtype == torch.complex64
    # FP64 pipeline, rounded to complex64 at the end
    assert R.rel_l2(_np(out), R.direct_type1([x], c, (n,))) < 3e-7


def test_matches_oracle_restatement_fp64(rng):
    """Same algorithm in FP64: GPU == oracle to rounding (pins deconvolution/FFT mapping)."""
    import nufftax_amd as N

    M = 3000
    x, y = (rng.uniform(-4, 4, M) for _ in range(2))
    c = rng.standard_normal((3, M)) + 1j * rng.standard_normal((3, M))
    out = N.nufft2d1(_t(x), _t(y), _t(c), (30, 22), eps=1e-6, isign=-1, upsampfac=1.25)
    ref = R.nufft_type1([x, y], c, (30, 22), eps=1e-6, isign=-1, upsampfac=1.25)
    assert out.shape == (3, 22, 30)
    assert R.rel_l2(_np(out), ref) < 1e-12
    f = rng.standard_normal((2, 22, 30)) + 1j * rng.standard_normal((2, 22, 30))
    out2 = N.nufft2d2(_t(x), _t(y), _t(f), eps=1e-6, isign=1, upsampfac=1.25)
    ref2 = R.nufft_type2([x, y], f, eps=1e-6, isign=1, upsampfac=1.25)
    assert R.rel_l2(_np(out2), ref2) < 1e-12


# ------------------------------------------------------------------ closed forms and edge cases


def test_closed_forms():
    import nufftax_amd as N

    k = np.arange(-8, 8)
    out = N.nufft1d1(_t(np.array([0.5])), _t(np.array([1.0 + 1j])), 16, eps=1e-6)
    np.testing.assert_allclose(_np(out), (1 + 1j) * np.exp(1j * k * 0.5), rtol=1e-5)
    out = N.nufft1d1(_t(np.full(10, np.pi, np.float32)), _t(np.ones(10, np.complex64)), 32, eps=1e-6)
    np.testing.assert_allclose(_np(out), 10 * np.exp(1j * np.arange(-16, 16) * np.pi), rtol=1e-5, atol=1e-5)
    x = np.linspace(-np.pi, np.pi, 32, endpoint=False)
    out = _np(N.nufft1d1(_t(x), _t(np.ones(32, complex)), 32, eps=1e-6))
    assert abs(out[16]) == pytest.approx(32, rel=1e-5) and np.max(np.abs(np.delete(out, 16))) < 1e-4
    out2 = N.nufft2d1(_t(np.array([0.3])), _t(np.array([-1.1])), _t(np.array([2.0 + 0j])), (6, 4))
    k1, k2 = np.arange(-3, 3), np.arange(-2, 2)
    np.testing.assert_allclose(_np(out2), 2 * np.exp(1j * (k2[:, None] * -1.1 + k1[None, :] * 0.3)), rtol=1e-5)


def test_wrap_invariance_and_empty(rng):
    import nufftax_amd as N

    x = rng.uniform(np.pi, 3 * np.pi, 200)
    c = rng.standard_normal(200) + 1j * rng.standard_normal(200)
    a = N.nufft1d1(_t(x), _t(c), 32)
    b = N.nufft1d1(_t(R.wrap_pi(x)), _t(c), 32)
    np.testing.assert_allclose(_np(a), _np(b), rtol=1e-9, atol=1e-9)
    z = N.nufft2d1(_t(np.zeros(0)), _t(np.zeros(0)), _t(np.zeros(0, complex)), (8, 6))
    assert z.shape == (6, 8) and float(z.abs().max()) == 0.0
    e = N.nufft2d2(_t(np.zeros(0)), _t(np.zeros(0)), _t(np.ones((6, 8), complex)))
    assert e.shape == (0,)


@pytest.mark.parametrize("dim,nm", [(1, (64,)), (2, (64, 64)), (3, (32, 32, 32))])
def test_special_strengths_and_coordinates(rng, dim, nm):
    """tests/test_edge_cases.py:123-197, :386-425 through each dimension's default complex64
    path (2D, 3D: the matrix-core window spreads, whose fixed-point scale follows the largest
    strength): all-zero strengths give exact zeros; one nonzero strength equals the one-point
    transform (rtol 1e-5); strengths and modes scaled by 1e-15 and 1e10 (the reference's) and by
    1e-30 and 1e30 give finite results equal to the scaled ones (relative l2 < 1e-6); points at
    +-(pi - 1e-6) match the FP64 direct sum."""
    import nufftax_amd as N

    fn1 = {1: N.nufft1d1, 2: N.nufft2d1, 3: N.nufft3d1}[dim]
    fn2 = {1: N.nufft1d2, 2: N.nufft2d2, 3: N.nufft3d2}[dim]
    n_modes = nm[0] if dim == 1 else nm
    M = 100
    x = [rng.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(dim)]
    xs = [_t(v) for v in x]
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    shape = tuple(reversed(nm))
    f = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)

    zero = fn1(*xs, _t(np.zeros(M, np.complex64)), n_modes, eps=1e-6)
    assert float(zero.abs().max()) == 0.0
    one = np.zeros(M, np.complex64)
    one[0] = 1 + 2j
    a = _np(fn1(*xs, _t(one), n_modes, eps=1e-6))
    b = _np(fn1(*[v[:1] for v in xs], _t(one[:1]), n_modes, eps=1e-6))
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-5 * np.abs(b).max())

    base1 = _np(fn1(*xs, _t(c), n_modes, eps=1e-6))
    base2 = _np(fn2(*xs, _t(f), eps=1e-6))
    for sc in (1e-15, 1e10, 1e-30, 1e30):
        o1 = _np(fn1(*xs, _t((c * sc).astype(np.complex64)), n_modes, eps=1e-6)).astype(np.complex128)
        o2 = _np(fn2(*xs, _t((f * sc).astype(np.complex64)), eps=1e-6)).astype(np.complex128)
        assert np.all(np.isfinite(o1)) and np.all(np.isfinite(o2)), sc
        assert R.rel_l2(o1 / sc, base1) < 1e-6, sc
        assert R.rel_l2(o2 / sc, base2) < 1e-6, sc

    edge = np.concatenate([np.full(M // 2, np.pi - 1e-6), np.full(M - M // 2, -np.pi + 1e-6)]).astype(np.float32)
    xe = [edge] + [rng.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(dim - 1)]
    o = _np(fn1(*[_t(v) for v in xe], _t(c), n_modes, eps=1e-6))
    assert np.all(np.isfinite(o))
    ref = R.direct_type1([v.astype(np.float64) for v in xe], c.astype(np.complex128), nm, isign=1)
    assert R.rel_l2(o, ref) < 1e-5


@pytest.mark.parametrize("dim,nm", [(1, (64,)), (2, (64, 64)), (3, (32, 32, 32))])
@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_non_finite_inputs(rng, dim, nm, bad):
    """A non-finite strength poisons a type-1 result as in the reference (spread.py:466-489
    scatters it into the grid, the FFT carries it into every mode): every mode is NaN.  A
    non-finite coordinate is handled without a fault and leaves the other points' type-2
    values equal to the all-finite result; the value of that point itself, and its type-1
    contribution, are not pinned (the reference's ES weights come out 0 for it,
    kernel.py:101-111; here it depends on the path)."""
    import nufftax_amd as N

    fn1 = {1: N.nufft1d1, 2: N.nufft2d1, 3: N.nufft3d1}[dim]
    fn2 = {1: N.nufft1d2, 2: N.nufft2d2, 3: N.nufft3d2}[dim]
    n_modes = nm[0] if dim == 1 else nm
    M = 2000
    x = [rng.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(dim)]
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    shape = tuple(reversed(nm))
    f = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    cb = c.copy()
    cb[11] = complex(bad, 0.0)
    assert np.all(np.isnan(_np(fn1(*[_t(v) for v in x], _t(cb), n_modes, eps=1e-6))))

    xb = [v.copy() for v in x]
    xb[dim - 1][7] = bad
    o1 = fn1(*[_t(v) for v in xb], _t(c), n_modes, eps=1e-6)
    assert o1.shape == shape
    good = _np(fn2(*[_t(v) for v in x], _t(f), eps=1e-6))
    got = _np(fn2(*[_t(v) for v in xb], _t(f), eps=1e-6))
    keep = np.arange(M) != 7
    assert np.all(np.isfinite(got[keep]))
    assert R.rel_l2(got[keep], good[keep]) < 1e-6


def test_3d_single_point_closed_form():
    """tests/test_edge_cases.py:362-380 in complex64: f[k3, k2, k1] = c exp(i (k1 x + k2 y + k3 z))."""
    import nufftax_amd as N

    v = lambda a: _t(np.array([a], np.float32))  # noqa: E731
    out = _np(N.nufft3d1(v(0.5), v(0.3), v(0.1), _t(np.array([1 + 1j], np.complex64)), (4, 4, 4), eps=1e-6))
    k = np.arange(-2, 2)
    expected = (1 + 1j) * np.exp(1j * (k[None, None, :] * 0.5 + k[None, :, None] * 0.3 + k[:, None, None] * 0.1))
    assert out.shape == (4, 4, 4)
    np.testing.assert_allclose(out, expected, rtol=1e-5, atol=1e-5)


def test_batched_n_trans_matches_loop(rng):
    import nufftax_amd as N

    M = 4000
    x, y = (_t(rng.uniform(-np.pi, np.pi, M).astype(np.float32)) for _ in range(2))
    c = _t((rng.standard_normal((2, 3, M)) + 1j * rng.standard_normal((2, 3, M))).astype(np.complex64))
    out = N.nufft2d1(x, y, c, (40, 30))
    assert out.shape == (2, 3, 30, 40)
    for i in range(2):
        for j in range(3):
            single = N.nufft2d1(x, y, c[i, j], (40, 30))
            assert R.rel_l2(_np(out[i, j]), _np(single)) < 1e-6


@pytest.mark.parametrize("dim,nm", [(1, (77,)), (2, (40, 36)), (3, (16, 12, 10))])
def test_plain_transpose_identity_upsampfac_pair(rng, dim, nm):
    """The reference's transpose rule: the linear transpose of type 1 (isign s, upsampfac
    sigma) is type 2 with the same isign and the backward upsampfac (primitives.py:191-194,
    235-241; tests/test_linear_solve.py:23-46), so sum_k f_k T1(c)_k = sum_j c_j T2(f)_j
    (bilinear, no conjugation) with T1 at upsampfac 2.0 and T2 at 1.25, the type-2 plan
    made for isign -1 and run through nufftx_execute_sign with +1."""
    from nufftax_amd._plan import F32, transform_plan

    M = 20000
    xs = [_t(rng.uniform(-np.pi, np.pi, M), torch.float32) for _ in range(dim)]
    c = _t(rng.standard_normal(M) + 1j * rng.standard_normal(M), torch.complex64)
    f = _t(rng.standard_normal(tuple(reversed(nm))) + 1j * rng.standard_normal(tuple(reversed(nm))),
           torch.complex64)
    p1 = transform_plan(1, dim, nm, 1, 1e-6, 1, 2.0, F32, DEV)
    p2 = transform_plan(2, dim, nm, 1, 1e-6, -1, 1.25, F32, DEV)
    assert p2.grid_size(0) != p1.grid_size(0)  # the two plans really use different grids
    F = torch.empty(tuple(reversed(nm)), dtype=torch.complex64, device=DEV)
    C = torch.empty(M, dtype=torch.complex64, device=DEV)
    with p1.lock:
        p1.setpts(xs)
        p1.execute(c, F, 1)
    with p2.lock:
        p2.setpts(xs)
        p2.execute(C, f, 1)
    lhs = torch.sum(f.to(torch.complex128) * F.to(torch.complex128))
    rhs = torch.sum(c.to(torch.complex128) * C.to(torch.complex128))
    scale = float(torch.linalg.vector_norm(f.to(torch.complex128)) * torch.linalg.vector_norm(F.to(torch.complex128)))
    assert abs(complex(lhs - rhs)) <= 1e-5 * scale, (complex(lhs), complex(rhs))
    # and each side equals the exact bilinear form (FP64 direct sums of both transforms)
    xn = [_np(x).astype(np.float64) for x in xs]
    T1 = R.direct_type1(xn, _np(c).astype(np.complex128), nm, isign=1)
    exact = np.sum(_np(f).astype(np.complex128) * T1)
    assert abs(complex(lhs) - exact) <= 1e-5 * scale


def test_type1_type2_adjoint(rng):
    import nufftax_amd as N

    M = 3000
    x, y, z = (rng.uniform(-np.pi, np.pi, M) for _ in range(3))
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    f = rng.standard_normal((6, 8, 10)) + 1j * rng.standard_normal((6, 8, 10))
    lhs = np.vdot(f, _np(N.nufft3d1(_t(x), _t(y), _t(z), _t(c), (10, 8, 6), eps=1e-12, isign=1)))
    rhs = np.vdot(_np(N.nufft3d2(_t(x), _t(y), _t(z), _t(f), eps=1e-12, isign=-1)), c)
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


# ------------------------------------------------------------------ type 3


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_type3_vs_direct(rng, dim, prec):
    import nufftax_amd as N

    M, Nt = 2000, 1500
    xs = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    ss = [rng.uniform(-20, 20, Nt) for _ in range(dim)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    eps = 1e-6 if prec == "c64" else 1e-9
    rd, cd = (torch.float32, torch.complex64) if prec == "c64" else (torch.float64, torch.complex128)
    if prec == "c64":
        xs = [x.astype(np.float32).astype(np.float64) for x in xs]
        ss = [s.astype(np.float32).astype(np.float64) for s in ss]
        c = c.astype(np.complex64).astype(np.complex128)
    if dim == 1:
        nm = N.compute_type3_grid_size(xs[0], ss[0], eps=eps)
        out = N.nufft1d3(_t(xs[0], rd), _t(c, cd), _t(ss[0], rd), nm, eps=eps)
    elif dim == 2:
        ext = [float((a.max() - a.min()) / 2) for a in xs + ss]
        nm = N.compute_type3_grid_sizes_2d(ext[0], ext[1], ext[2], ext[3], eps=eps)
        out = N.nufft2d3(_t(xs[0], rd), _t(xs[1], rd), _t(c, cd), _t(ss[0], rd), _t(ss[1], rd), nm, eps=eps)
    else:
        ext = [float((a.max() - a.min()) / 2) for a in xs + ss]
        nm = N.compute_type3_grid_sizes_3d(*ext, eps=eps)
        out = N.nufft3d3(*[_t(x, rd) for x in xs], _t(c, cd), *[_t(s, rd) for s in ss], nm, eps=eps)
    truth = R.direct_type3(xs, c, ss)
    err = R.rel_l2(_np(out), truth)
    if prec == "c64":
        nfs = nm if isinstance(nm, tuple) else (nm,)
        proxy = R.rel_l2(R.nufft_type3([x.astype(np.float32) for x in xs], c.astype(np.complex64),
                                       [s.astype(np.float32) for s in ss], nfs, eps=eps), truth)
        assert err <= 2 * proxy + 1e-6, (err, proxy)
    else:
        assert err < 1e-7


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_type3_small_grids_vs_direct(rng, dim, prec):
    """Grid sizes below 2 * nspread, as compute_type3_grid_size returns them for a small
    space-bandwidth product (x in [-pi, pi], s in [-1, 1]: nf = 12 at eps 1e-6, 10 at 1e-4);
    the reference accepts any n_modes (nufft3.py:266)."""
    import nufftax_amd as N

    M, Nt = 3000, 700
    for eps in (1e-4, 1e-6):
        xs = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
        ss = [rng.uniform(-1, 1, Nt) for _ in range(dim)]
        c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
        rd, cd = (torch.float32, torch.complex64) if prec == "c64" else (torch.float64, torch.complex128)
        if prec == "c64":
            xs = [x.astype(np.float32).astype(np.float64) for x in xs]
            ss = [s.astype(np.float32).astype(np.float64) for s in ss]
            c = c.astype(np.complex64).astype(np.complex128)
        nfs = tuple(N.compute_type3_grid_size(xs[d], ss[d], eps=eps) for d in range(dim))
        ns = N.compute_kernel_params(eps).nspread
        assert all(n < 2 * ns for n in nfs), nfs
        if dim == 1:
            out = N.nufft1d3(_t(xs[0], rd), _t(c, cd), _t(ss[0], rd), nfs[0], eps=eps)
        else:
            out = N.nufft2d3(_t(xs[0], rd), _t(xs[1], rd), _t(c, cd), _t(ss[0], rd), _t(ss[1], rd), nfs, eps=eps)
        truth = R.direct_type3(xs, c, ss)
        err = R.rel_l2(_np(out), truth)
        proxy = R.rel_l2(R.nufft_type3(xs, c, ss, nfs, eps=eps), truth)  # FP64 restatement, same grid
        assert err <= 2 * proxy + 10 * eps, (eps, nfs, err, proxy)


@pytest.mark.parametrize("eps", [1e-1, 1e-3, 1e-4, 1e-6])
@pytest.mark.parametrize("dist", ["uniform", "clustered"])
@pytest.mark.parametrize("mfma", ["1", None, "exp", "f32win", "bank"])
def test_spread2d_f32_every_nspread(rng, eps, dist, mfma, monkeypatch):
    """2D FP32 spreading vs the FP64 oracle, every even ns: the default (fp16-split window kernel
    for ns <= 8, 4096-point subproblems, polynomial kernel weights at ns = 8), the same with the
    exp form of the weights (NUFFTX_ES_POLY=0), the FP32 window kernel (NUFFTX_WIN_F32 with
    1024-point subproblems), the bank-spread kernel (NUFFTX_SPREAD_WIN=0) and the records-based
    matrix-core path (NUFFTX_MFMA)."""
    from nufftax_amd._plan import F32, grid_plan

    if mfma == "exp":
        monkeypatch.setenv("NUFFTX_ES_POLY", "0")
    elif mfma == "f32win":
        monkeypatch.setenv("NUFFTX_WIN_F32", "1")
        monkeypatch.setenv("NUFFTX_MAXSUB", "1024")
    elif mfma == "bank":
        monkeypatch.setenv("NUFFTX_SPREAD_WIN", "0")
    elif mfma:
        monkeypatch.setenv("NUFFTX_MFMA", mfma)
    kp = R.compute_kernel_params(eps)
    nfs = (4 * (16 - kp.nspread) * 3 + 5, 4 * (16 - kp.nspread) * 2 + 3)  # several bins, ragged edge
    M = 20000
    if dist == "uniform":
        coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(2)]
    else:
        coords = [np.concatenate([rng.normal(0.3, 0.02, M - 500), rng.uniform(-np.pi, np.pi, 500)]) for _ in range(2)]
    coords = [x.astype(np.float32) for x in coords]
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    plan = grid_plan(2, nfs, kp.nspread, kp.beta, kp.c, 1, F32, torch.device(DEV))
    plan.setpts([_t(x) for x in coords])
    fw = torch.empty((nfs[1], nfs[0]), dtype=torch.complex64, device=DEV)
    plan.spread(_t(c), fw)
    ref = R.spread([x.astype(np.float64) for x in coords], c.astype(np.complex128), nfs, kp)
    assert R.rel_l2(_np(fw), ref) < 2e-6


@pytest.mark.parametrize("eps", [1e-1, 1e-3, 1e-4, 1e-6])
@pytest.mark.parametrize("dist", ["uniform", "clustered"])
@pytest.mark.parametrize("mfma", ["1", None, "exp", "win"])
def test_interp2d_f32_every_nspread(rng, eps, dist, mfma, monkeypatch):
    """2D FP32 interpolation (LDS-tiled, polynomial kernel weights at ns = 8; the same with the
    exp form, NUFFTX_ES_POLY=0; the records-based matrix-core kernel, NUFFTX_MFMA; the fp16-split
    window interpolation, NUFFTX_INTERP_WIN, for ns <= 8) vs the FP64 oracle, every even ns,
    batched."""
    from nufftax_amd._plan import F32, grid_plan

    if mfma == "exp":
        monkeypatch.setenv("NUFFTX_ES_POLY", "0")
    elif mfma == "win":
        monkeypatch.setenv("NUFFTX_INTERP_WIN", "1")
    elif mfma:
        monkeypatch.setenv("NUFFTX_MFMA", mfma)
    kp = R.compute_kernel_params(eps)
    nfs = (4 * (16 - kp.nspread) * 3 + 5, 4 * (16 - kp.nspread) * 2 + 3)
    M = 20000
    if dist == "uniform":
        coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(2)]
    else:
        coords = [np.concatenate([rng.normal(0.3, 0.02, M - 500), rng.uniform(-np.pi, np.pi, 500)]) for _ in range(2)]
    coords = [x.astype(np.float32) for x in coords]
    fw = (rng.standard_normal((3, nfs[1], nfs[0])) + 1j * rng.standard_normal((3, nfs[1], nfs[0]))).astype(np.complex64)
    plan = grid_plan(2, nfs, kp.nspread, kp.beta, kp.c, 3, F32, torch.device(DEV))
    plan.setpts([_t(x) for x in coords])
    c = torch.empty((3, M), dtype=torch.complex64, device=DEV)
    plan.interp(_t(fw), c)
    ref = R.interp([x.astype(np.float64) for x in coords], fw.astype(np.complex128), kp)
    assert R.rel_l2(_np(c), ref) < 2e-6


@pytest.mark.parametrize("eps", [1e-1, 1e-3, 1e-4, 1e-6])
@pytest.mark.parametrize("dist", ["uniform", "clustered"])
@pytest.mark.parametrize("path", ["window", "packed", "bank"])
def test_spread3d_f32_every_nspread(rng, eps, dist, path, monkeypatch):
    """3D FP32 spreading vs the FP64 oracle, two transforms, rel l2 < 2e-6: the default paths
    (the matrix-core window kernel with exact tiles at ns = 8, eps 1e-6: bins of 16 x 8 x 8
    cells, 8 x 8 x 4 sub-bins, ragged last bins; the bank-spread kernel below ns = 8), the
    window kernel with packed tiles (NUFFTX_WIN3_PACKED=1) and the bank-spread kernel at every
    ns (NUFFTX_SPREAD_WIN=0)."""
    from nufftax_amd._plan import F32, grid_plan

    if path == "bank":
        monkeypatch.setenv("NUFFTX_SPREAD_WIN", "0")
    if path == "packed":
        if eps != 1e-6:
            pytest.skip("the window kernel takes ns = 8 only")
        monkeypatch.setenv("NUFFTX_WIN3_PACKED", "1")
    kp = R.compute_kernel_params(eps)
    nfs = (44, 40, 28)
    M = 20000
    if dist == "uniform":
        coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(3)]
    else:
        coords = [np.concatenate([rng.normal(-0.5, 0.08, M - 3000), rng.uniform(-np.pi, np.pi, 3000)]) for _ in range(3)]
    coords = [x.astype(np.float32) for x in coords]
    c = (rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))).astype(np.complex64)
    plan = grid_plan(3, nfs, kp.nspread, kp.beta, kp.c, 2, F32, torch.device(DEV))
    plan.setpts([_t(x) for x in coords])
    fw = torch.full((2,) + tuple(reversed(nfs)), float("nan"), dtype=torch.complex64, device=DEV)
    plan.spread(_t(c), fw)
    ref = R.spread([x.astype(np.float64) for x in coords], c.astype(np.complex128), nfs, kp)
    assert R.rel_l2(_np(fw), ref) < 2e-6


@pytest.mark.parametrize("eps", [1e-2, 1e-4, 1e-6])
@pytest.mark.parametrize("wide", ["1", "0"])
def test_interp3d_f32_wide_reads(rng, eps, wide, monkeypatch):
    """3D FP32 interpolation with 16-byte tile reads (default) and 8-byte reads
    (NUFFTX_INTERP_WIDE=0) vs the FP64 oracle, rel l2 < 2e-6, two transforms."""
    from nufftax_amd._plan import F32, grid_plan

    monkeypatch.setenv("NUFFTX_INTERP_WIDE", wide)
    kp = R.compute_kernel_params(eps)
    nfs = (41, 36, 34)
    M = 20000
    coords = [np.concatenate([rng.normal(0.3, 0.05, M - 3000), rng.uniform(-np.pi, np.pi, 3000)]).astype(np.float32)
              for _ in range(3)]
    shape = (2,) + tuple(reversed(nfs))
    fw = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    plan = grid_plan(3, nfs, kp.nspread, kp.beta, kp.c, 2, F32, torch.device(DEV))
    plan.setpts([_t(x) for x in coords])
    c = torch.empty((2, M), dtype=torch.complex64, device=DEV)
    plan.interp(_t(fw), c)
    ref = R.interp([x.astype(np.float64) for x in coords], fw.astype(np.complex128), kp)
    assert R.rel_l2(_np(c), ref) < 2e-6


@pytest.mark.parametrize("upsampfac", [2.0, 1.25])
@pytest.mark.parametrize("dim", [2, 3])
def test_es_poly_weights_vs_exp(rng, upsampfac, dim, monkeypatch):
    """The polynomial ES kernel weights (ns = 8: degree-9 even/odd Horner forms fitted on the
    host) against the exp form (NUFFTX_ES_POLY=0) and the FP64 oracle, at both upsampling
    factors: spreading and interpolation within 2e-6 of the oracle, and the two forms within
    1e-6 of each other."""
    from nufftax_amd._plan import F32, grid_plan

    kp = R.compute_kernel_params(1e-6, upsampfac)
    assert kp.nspread == 8
    nfs = (100, 84) if dim == 2 else (44, 40, 28)
    M = 20000
    coords = [np.concatenate([rng.normal(0.4, 0.1, M - 4000), rng.uniform(-np.pi, np.pi, 4000)]).astype(np.float32)
              for _ in range(dim)]
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    shape = tuple(reversed(nfs))
    fw = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ref_s = R.spread([x.astype(np.float64) for x in coords], c.astype(np.complex128), nfs, kp)
    ref_i = R.interp([x.astype(np.float64) for x in coords], fw.astype(np.complex128), kp)
    got = {}
    for form in ("poly", "exp"):
        monkeypatch.setenv("NUFFTX_ES_POLY", "1" if form == "poly" else "0")
        plan = grid_plan(dim, nfs, kp.nspread, kp.beta, kp.c, 1, F32, torch.device(DEV))
        plan.setpts([_t(x) for x in coords])
        g = torch.full(shape, float("nan"), dtype=torch.complex64, device=DEV)
        plan.spread(_t(c), g)
        v = torch.empty(M, dtype=torch.complex64, device=DEV)
        plan.interp(_t(fw), v)
        got[form] = (_np(g), _np(v))
        assert R.rel_l2(got[form][0], ref_s) < 2e-6
        assert R.rel_l2(got[form][1], ref_i) < 2e-6
    assert R.rel_l2(got["poly"][0], got["exp"][0]) < 1e-6
    assert R.rel_l2(got["poly"][1], got["exp"][1]) < 1e-6


@pytest.mark.parametrize("nfs", [(77, 64), (45, 27)])
@pytest.mark.parametrize("eps", [1e-2, 1e-4, 1e-6])
@pytest.mark.parametrize("wide", ["1", "0"])
def test_interp2d_f32_wide_reads(rng, nfs, eps, wide, monkeypatch):
    """2D FP32 interpolation from point records with 16-byte tile reads (default; tile rows of an
    even cell count, odd and even support starts) and with 8-byte reads (NUFFTX_INTERP_WIDE=0),
    and a grid whose tile rows have an odd cell count at ns = 4 (8-byte path), vs the FP64
    oracle, < 2e-6."""
    from nufftax_amd._plan import F32, grid_plan

    monkeypatch.setenv("NUFFTX_INTERP_WIDE", wide)
    kp = R.compute_kernel_params(eps)
    M = 20000
    coords = [np.concatenate([rng.normal(-0.4, 0.05, M - 4000), rng.uniform(-np.pi, np.pi, 4000)]).astype(np.float32)
              for _ in range(2)]
    fw = (rng.standard_normal((2, nfs[1], nfs[0])) + 1j * rng.standard_normal((2, nfs[1], nfs[0]))).astype(np.complex64)
    plan = grid_plan(2, nfs, kp.nspread, kp.beta, kp.c, 2, F32, torch.device(DEV))
    plan.setpts([_t(x) for x in coords])
    c = torch.empty((2, M), dtype=torch.complex64, device=DEV)
    plan.interp(_t(fw), c)
    ref = R.interp([x.astype(np.float64) for x in coords], fw.astype(np.complex128), kp)
    assert R.rel_l2(_np(c), ref) < 2e-6


# ------------------------------------------------------------------ fused FFT + deconvolution


def _fft_stage_refs(p1, p2, c, f, nm, isign, sigma, modeord):
    """FP64 references for the FFT + deconvolution stage alone: type 1 from the FP32 spread
    grid of plan p1 itself (FP64 FFT and deconvolution of it), type 2 from the FP64 padded and
    transformed modes, interpolated by plan p2.  The deconvolution factors at upsampfac 1.25
    span ~30x per axis (2.7x at upsampfac 2), so FP32 FFT rounding, which is uniform in
    absolute size, shows up amplified by up to ~30^d at the corners of the mode band."""
    dim = len(nm)
    nf = [p1.grid_size(d) for d in range(dim)]
    kp = R.compute_kernel_params(1e-6, sigma)
    ph = [R.kernel_fourier_series(n, kp.nspread, kp.beta, kp.c) for n in nf]
    g = torch.empty((c.shape[0],) + tuple(reversed(nf)), dtype=torch.complex64, device=DEV)
    p1.spread(c, g)
    torch.cuda.synchronize()
    ref1 = R.deconvolve_shuffle(R.fft_isign(_np(g).astype(np.complex128), dim, isign), ph, nm, modeord)
    fw = R.fft_isign(R.deconvolve_pad(_np(f).astype(np.complex128), ph, nf, modeord), dim, -isign)
    o = torch.empty((c.shape[0], c.shape[1]), dtype=torch.complex64, device=DEV)
    p2.interp(_t(np.ascontiguousarray(fw.astype(np.complex64))), o)
    torch.cuda.synchronize()
    return ref1, _np(o)


def _check_fused_vs_rocfft(fused, rocfft, refs, sigma):
    """upsampfac 2: fused and rocFFT results within 2e-6 of each other.  Otherwise (rounding
    amplified by the deconvolution, see _fft_stage_refs): the fused error against the FP64
    stage reference at most 1.5x rocFFT's + 1e-7, and below 1e-4."""
    for a, b, ref in zip(fused, rocfft, refs):
        if sigma == 2.0:
            assert R.rel_l2(a, b) < 2e-6
        else:
            ef, er = R.rel_l2(a, ref), R.rel_l2(b, ref)
            assert ef <= 1.5 * er + 1e-7 and ef < 1e-4, (ef, er)


@pytest.mark.parametrize(
    "nm,sigma", [((32,), 2.0), ((63,), 2.0), ((1024,), 2.0), ((2047,), 2.0), ((51,), 1.25), ((409,), 1.25)]
)
@pytest.mark.parametrize("isign", [1, -1])
@pytest.mark.parametrize("modeord", [0, 1])
def test_fused_fft1_matches_rocfft(rng, monkeypatch, nm, sigma, isign, modeord):
    """1D FP32 plans on power-of-two grids (64 .. 4096 cells here) run FFT + deconvolution as one
    fused Stockham pass per transform (default); the same plans built with NUFFTX_ROCFFT take
    rocFFT + the separate deconvolution kernels.  Type 1 and type 2, two transforms per call,
    odd and even mode counts, upsampfac 2 and 1.25; bars as in the 2D and 3D tests."""
    from nufftax_amd._plan import F32, Plan

    M = 20000
    xs = [_t(rng.uniform(-np.pi, np.pi, M).astype(np.float32))]
    c = _t((rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))).astype(np.complex64))
    f = _t((rng.standard_normal((2, nm[0])) + 1j * rng.standard_normal((2, nm[0]))).astype(np.complex64))

    def run(rocfft):
        if rocfft:
            monkeypatch.setenv("NUFFTX_ROCFFT", "1")
        else:
            monkeypatch.delenv("NUFFTX_ROCFFT", raising=False)
        p1 = Plan.transform(1, 1, nm, 2, 1e-6, isign, sigma, F32, modeord, torch.device(DEV))
        p2 = Plan.transform(2, 1, nm, 2, 1e-6, -isign, sigma, F32, modeord, torch.device(DEV))
        assert p1.grid_size(0) == 2 ** int(np.log2(p1.grid_size(0)))
        p1.setpts(xs)
        p2.setpts(xs)
        o1 = torch.empty((2, nm[0]), dtype=torch.complex64, device=DEV)
        o2 = torch.empty((2, M), dtype=torch.complex64, device=DEV)
        p1.execute(c, o1)
        p2.execute(o2, f)
        torch.cuda.synchronize()
        refs = None if sigma == 2.0 else _fft_stage_refs(p1, p2, c, f, nm, isign, sigma, modeord)
        return (_np(o1), _np(o2)), refs

    fused, refs = run(False)
    rocfft, _ = run(True)
    _check_fused_vs_rocfft(fused, rocfft, refs or (None, None), sigma)


@pytest.mark.parametrize(
    "nm,sigma",
    [((32, 32, 64), 2.0), ((31, 63, 64), 2.0), ((128, 32, 256), 2.0), ((127, 255, 32), 2.0),
     ((51, 102, 204), 1.25)],
)
@pytest.mark.parametrize("isign", [1, -1])
@pytest.mark.parametrize("modeord", [0, 1])
def test_fused_fft3_matches_rocfft(rng, monkeypatch, nm, sigma, isign, modeord):
    """3D FP32 plans on power-of-two grids run FFT + deconvolution as three fused Stockham
    passes (default: along x, y, z, each on the kept modes of the axes done before); the same
    plans built with NUFFTX_ROCFFT take rocFFT + the separate deconvolution kernels.  Type 1 and
    type 2, two transforms per call, odd and even mode counts, axes of 64 .. 512 cells (each
    thread layout of the passes: 8, 16 and 64 threads per sequence), upsampfac 2 and 1.25 (kept
    modes over more than half of each axis); relative l2 < 2e-6."""
    from nufftax_amd._plan import F32, Plan

    M = 20000
    xs = [_t(rng.uniform(-np.pi, np.pi, M).astype(np.float32)) for _ in range(3)]
    c = _t((rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))).astype(np.complex64))
    shape = (2,) + tuple(reversed(nm))
    f = _t((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64))

    def run(rocfft):
        if rocfft:
            monkeypatch.setenv("NUFFTX_ROCFFT", "1")
        else:
            monkeypatch.delenv("NUFFTX_ROCFFT", raising=False)
        p1 = Plan.transform(1, 3, nm, 2, 1e-6, isign, sigma, F32, modeord, torch.device(DEV))
        p2 = Plan.transform(2, 3, nm, 2, 1e-6, -isign, sigma, F32, modeord, torch.device(DEV))
        for d in range(3):
            assert p1.grid_size(d) == 2 ** int(np.log2(p1.grid_size(d)))
        p1.setpts(xs)
        p2.setpts(xs)
        o1 = torch.empty(shape, dtype=torch.complex64, device=DEV)
        o2 = torch.empty((2, M), dtype=torch.complex64, device=DEV)
        p1.execute(c, o1)
        p2.execute(o2, f)
        torch.cuda.synchronize()
        refs = None if sigma == 2.0 else _fft_stage_refs(p1, p2, c, f, nm, isign, sigma, modeord)
        return (_np(o1), _np(o2)), refs

    fused, refs = run(False)
    rocfft, _ = run(True)
    _check_fused_vs_rocfft(fused, rocfft, refs or (None, None), sigma)


@pytest.mark.parametrize(
    "nm,sigma",
    [((32, 32), 2.0), ((63, 31), 2.0), ((256, 128), 2.0), ((512, 512), 2.0), ((2048, 1024), 2.0),
     ((204, 51), 1.25), ((409, 102), 1.25)],
)
@pytest.mark.parametrize("isign", [1, -1])
@pytest.mark.parametrize("modeord", [0, 1])
def test_fused_fft_matches_rocfft(rng, monkeypatch, nm, sigma, isign, modeord):
    """Plans of 2D FP32 power-of-two grids (64 .. 4096 cells a side here) run FFT +
    deconvolution as two fused Stockham passes in both directions (default); the same plans
    built with NUFFTX_ROCFFT take rocFFT + the separate deconvolution kernels.  Type 1 and
    type 2, two transforms per call, upsampfac 2 and 1.25 (kept modes over more than half of
    each axis); relative l2 < 2e-6 (FP32 FFT rounding)."""
    from nufftax_amd._plan import F32, Plan

    M = 20000
    x = _t(rng.uniform(-np.pi, np.pi, M).astype(np.float32))
    y = _t(rng.uniform(-np.pi, np.pi, M).astype(np.float32))
    c = _t((rng.standard_normal((2, M)) + 1j * rng.standard_normal((2, M))).astype(np.complex64))
    f = _t((rng.standard_normal((2, nm[1], nm[0])) + 1j * rng.standard_normal((2, nm[1], nm[0]))).astype(np.complex64))

    def run(rocfft):
        if rocfft:
            monkeypatch.setenv("NUFFTX_ROCFFT", "1")
        else:
            monkeypatch.delenv("NUFFTX_ROCFFT", raising=False)
        p1 = Plan.transform(1, 2, nm, 2, 1e-6, isign, sigma, F32, modeord, torch.device(DEV))
        p2 = Plan.transform(2, 2, nm, 2, 1e-6, -isign, sigma, F32, modeord, torch.device(DEV))
        for d in range(2):
            assert p1.grid_size(d) == 2 ** int(np.log2(p1.grid_size(d)))
        p1.setpts([x, y])
        p2.setpts([x, y])
        o1 = torch.empty((2, nm[1], nm[0]), dtype=torch.complex64, device=DEV)
        o2 = torch.empty((2, M), dtype=torch.complex64, device=DEV)
        p1.execute(c, o1)
        p2.execute(o2, f)
        torch.cuda.synchronize()
        refs = None if sigma == 2.0 else _fft_stage_refs(p1, p2, c, f, nm, isign, sigma, modeord)
        return (_np(o1), _np(o2)), refs

    fused, refs = run(False)
    rocfft, _ = run(True)
    _check_fused_vs_rocfft(fused, rocfft, refs or (None, None), sigma)
