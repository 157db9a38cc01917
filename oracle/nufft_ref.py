"""CPU restatement of nufftax's spread -> FFT -> deconvolve pipeline.

TEST INFRASTRUCTURE ONLY.  This module is the oracle that the HIP path is
checked against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline -- never as part of the product path.

It restates, step by step and in NumPy, the reference's pure-JAX algorithm
(GragasLab/nufftax).  Every function cites the reference file:line it follows.
Arithmetic runs in the dtype of the inputs (float64/complex128 by default), so
the same code also serves as the complex64 "reference-error proxy" when fed
float32/complex64 inputs: fold, weights and FFT then run in single precision as
the reference's do.  Accumulation of the scatter uses ``np.bincount`` (float64
accumulators), which only changes summation order/precision of the reference's
``jax.ops.segment_sum``.

Conventions (all from the reference):
  * coordinates are wrapped to [-pi, pi) and folded to xs in [0, nf)
    (``nufftax/core/spread.py:78-97``); grid index j <-> x = 2*pi*j/nf - pi;
  * the kernel support of a point starts at i0 = ceil(xs - ns/2)
    (``spread.py:216``); z = (i0 + k) - xs;
  * grids are stored (n_trans, nf_d, ..., nf_1): x is the fastest axis
    (``spread.py:466``, ``spread.py:632-634``);
  * mode k (CMCL order, k in [-(n//2), (n-1)//2]) of a type-1 output lives at
    fine-grid index k mod nf and carries the factor (-1)^k / phihat[|k|]
    (``nufftax/core/deconvolve.py:27-73``);
  * isign > 0 uses the unnormalised inverse FFT (``ifft * N``), isign < 0 the
    forward FFT (``nufftax/transforms/nufft1.py:81``, ``nufft2.py:74``).

Functions whose parity is pinned only indirectly say so in their docstring;
every function here is exercised by ``tests/test_oracle.py`` against closed
forms, the direct NUDFT, adjoint identities or finite differences.
"""

from __future__ import annotations

import math
from typing import NamedTuple, Sequence

import numpy as np


# ---------------------------------------------------------------------------
# Kernel parameters and grid sizes
# ---------------------------------------------------------------------------


class KernelParams(NamedTuple):
    """``nufftax/core/kernel.py:19-25``."""

    nspread: int
    beta: float
    c: float
    upsampfac: float


def compute_kernel_params(tol: float, upsampfac: float = 2.0, max_nspread: int = 16) -> KernelParams:
    """``nufftax/core/kernel.py:165-206``: ns from tol, rounded up to even; beta formula."""
    log_tol = -math.log10(max(tol, 1e-16))
    ns = int(math.ceil(log_tol + 1))
    ns = max(2, min(ns, max_nspread))
    if ns % 2 == 1:
        ns += 1
    sigma = upsampfac
    inner = (ns**2) / (sigma**2) * (sigma - 0.5) ** 2 - 0.8
    beta = math.pi * math.sqrt(inner) if inner > 0 else 2.3 * ns
    return KernelParams(ns, beta, 4.0 / (ns**2), upsampfac)


def next_smooth_int(n: int) -> int:
    """``nufftax/utils/grid.py:6-33``: smallest 2,3,5-smooth integer >= n."""
    if n <= 1:
        return 1

    def smooth(v):
        for p in (2, 3, 5):
            while v % p == 0:
                v //= p
        return v == 1

    while not smooth(n):
        n += 1
    return n


def compute_grid_size(n_modes: int, upsampfac: float, nspread: int) -> int:
    """``nufftax/utils/grid.py:36-58``."""
    return next_smooth_int(max(int(math.ceil(upsampfac * n_modes)), 2 * nspread))


def next_smooth_even(n: int) -> int:
    """``nufftax/transforms/nufft3.py:35-47``."""
    if n <= 1:
        return 2
    if n % 2 == 1:
        n += 1
    cand = next_smooth_int(n)
    while cand % 2 != 0:
        cand = next_smooth_int(cand + 1)
    return cand


def compute_type3_grid_size(x_extent: float, s_extent: float, eps: float = 1e-6, upsampfac: float = 2.0) -> int:
    """``nufftax/transforms/nufft3.py:50-104`` (extent form; arrays reduce to (max-min)/2)."""
    x_arr, s_arr = np.asarray(x_extent), np.asarray(s_extent)
    xe = float((x_arr.max() - x_arr.min()) / 2) if x_arr.ndim > 0 else float(x_arr)
    se = float((s_arr.max() - s_arr.min()) / 2) if s_arr.ndim > 0 else float(s_arr)
    ns = compute_kernel_params(eps, upsampfac).nspread
    xe, se = max(xe, 1e-10), max(se, 1e-10)
    return next_smooth_even(int(np.ceil(2.0 * upsampfac * se * xe / np.pi + ns)))


# ---------------------------------------------------------------------------
# ES kernel and its Fourier series
# ---------------------------------------------------------------------------


def es_kernel(z: np.ndarray, beta: float, c: float) -> np.ndarray:
    """``nufftax/core/kernel.py:79-111``: exp(beta*(sqrt(1-c z^2)-1)), 0 outside support."""
    z = np.asarray(z)
    dt = z.dtype
    beta, c, one, zero = dt.type(beta), dt.type(c), dt.type(1.0), dt.type(0.0)
    arg = one - c * z * z
    res = np.exp(beta * (np.sqrt(np.maximum(arg, zero)) - one))
    return np.where(arg >= zero, res, zero).astype(dt)


def es_kernel_with_derivative(z: np.ndarray, beta: float, c: float):
    """``nufftax/core/kernel.py:114-162``: (phi, dphi/dz) with sqrt argument clamped at 1e-14."""
    z = np.asarray(z)
    dt = z.dtype
    beta, c, one, zero, tiny = dt.type(beta), dt.type(c), dt.type(1.0), dt.type(0.0), dt.type(1e-14)
    arg = one - c * z * z
    s = np.sqrt(np.maximum(arg, tiny))
    phi = np.where(arg >= zero, np.exp(beta * (s - one)), zero)
    dphi = np.where(arg > tiny, -beta * c * z / s * phi, zero)
    return phi.astype(dt), dphi.astype(dt)


def _quadrature(ns: int, beta: float, c: float, dtype):
    """Trapezoid nodes/weights shared by ``kernel.py:238-256`` and ``nufft3.py:195-206``."""
    j2 = ns / 2.0
    n_quad = max(int(4 + 3.0 * j2), 20)
    z = np.linspace(-j2, j2, n_quad, dtype=dtype)
    dz = z[1] - z[0]
    w = np.full(n_quad, dz, dtype=dtype)
    w[0] = dz / 2
    w[-1] = dz / 2
    return z, w * es_kernel(z, beta, c)


def kernel_fourier_series(nf: int, ns: int, beta: float, c: float, dtype=np.float64) -> np.ndarray:
    """``nufftax/core/kernel.py:209-260``: phihat[k], k = 0..nf//2, trapezoid rule."""
    z, wphi = _quadrature(ns, beta, c, dtype)
    k = np.arange(nf // 2 + 1, dtype=dtype)
    phase = dtype(2.0 * np.pi) * np.outer(k, z) / dtype(nf)
    return np.sum(wphi[None, :] * np.cos(phase), axis=1).astype(dtype)


def kernel_ft_at_point(k: np.ndarray, ns: int, beta: float, c: float, dtype=np.float64) -> np.ndarray:
    """``nufftax/transforms/nufft3.py:176-213``: phihat at arbitrary frequencies k."""
    z, wphi = _quadrature(ns, beta, c, dtype)
    k = np.atleast_1d(np.asarray(k)).astype(dtype)
    return np.sum(wphi[None, :] * np.cos(np.outer(k, z)), axis=1)


# ---------------------------------------------------------------------------
# Fold / weights
# ---------------------------------------------------------------------------


def wrap_pi(x: np.ndarray) -> np.ndarray:
    """Driver wrap ``mod(x+pi, 2pi) - pi`` (``nufftax/transforms/nufft1.py:71``)."""
    x = np.asarray(x)
    return (np.mod(x + x.dtype.type(np.pi), x.dtype.type(2.0 * np.pi)) - x.dtype.type(np.pi)).astype(x.dtype)


def fold_rescale(x: np.ndarray, n: int) -> np.ndarray:
    """``nufftax/core/spread.py:78-97``: [-pi,pi) -> [0, n), periodic."""
    x = np.asarray(x)
    inv_2pi = x.dtype.type(1.0 / (2.0 * np.pi))
    r = x * inv_2pi + x.dtype.type(0.5)
    r = r - np.floor(r)
    return r * x.dtype.type(n)


def kernel_weights_1d(xs: np.ndarray, nf: int, kp: KernelParams, derivative: bool = False):
    """``nufftax/core/spread.py:187-235`` (values) and ``:238-277`` (derivative).

    Returns wrapped indices (M, ns), weights (M, ns) and, with ``derivative``,
    dweights = -dphi/dz * nf/(2 pi) (the derivative w.r.t. the unscaled x).
    """
    ns = kp.nspread
    i0 = np.ceil(xs - ns / 2.0).astype(np.int64)
    idx = i0[:, None] + np.arange(ns)[None, :]
    z = idx.astype(xs.dtype) - xs[:, None]
    if not derivative:
        return idx % nf, es_kernel(z, kp.beta, kp.c)
    w, dw = es_kernel_with_derivative(z, kp.beta, kp.c)
    return idx % nf, w, -dw * xs.dtype.type(nf / (2.0 * np.pi))


def _flat_indices(idx_per_dim: Sequence[np.ndarray], nfs: Sequence[int]) -> np.ndarray:
    """Linear fine-grid index, x fastest: ``spread.py:471`` (2D), ``spread.py:632-634`` (3D)."""
    d = len(nfs)
    M, ns = idx_per_dim[0].shape
    out = np.zeros((M,) + (ns,) * d, dtype=np.int64)
    stride = 1
    for ax in range(d):  # ax 0 = x (fastest)
        shape = [M] + [1] * d
        shape[d - ax] = ns
        out = out + idx_per_dim[ax].reshape(shape) * stride
        stride *= nfs[ax]
    return out.reshape(M, -1)


def _outer_weights(w_per_dim: Sequence[np.ndarray]) -> np.ndarray:
    """Tensor-product weights laid out like ``_flat_indices`` (z outermost, x innermost)."""
    d = len(w_per_dim)
    M, ns = w_per_dim[0].shape
    out = np.ones((M,) + (ns,) * d, dtype=w_per_dim[0].dtype)
    for ax in range(d):
        shape = [M] + [1] * d
        shape[d - ax] = ns
        out = out * w_per_dim[ax].reshape(shape)
    return out.reshape(M, -1)


# ---------------------------------------------------------------------------
# Spread / interp (materialised, like the pure-JAX path)
# ---------------------------------------------------------------------------


def _as_batched(a: np.ndarray, core_ndim: int):
    a = np.asarray(a)
    if a.ndim == core_ndim:
        return a[None], False
    return a, True


def spread(coords: Sequence[np.ndarray], c: np.ndarray, nfs: Sequence[int], kp: KernelParams, chunk: int = 1 << 15):
    """Type-1 spread, ``spread.py:285-339`` (1D), ``:435-494`` (2D), ``:596-658`` (3D).

    coords: (x, y, z)[:d] in any real range (wrapped by the fold); c: (M,) or
    (n_trans, M).  Returns the fine grid (n_trans, nf_d, ..., nf_1) (batch axis
    dropped for unbatched c).  The scatter (``segment_sum``) is done with
    ``np.bincount`` in chunks of points; float64 accumulation.
    """
    d = len(nfs)
    cb, batched = _as_batched(c, 1)
    n_trans, M = cb.shape
    rdt = np.asarray(coords[0]).dtype
    ntot = int(np.prod(nfs))
    acc_re = np.zeros((n_trans, ntot))
    acc_im = np.zeros((n_trans, ntot))
    for s in range(0, M, chunk):
        e = min(M, s + chunk)
        idxs, ws = [], []
        for ax in range(d):
            xs = fold_rescale(np.asarray(coords[ax][s:e]), nfs[ax])
            i, w = kernel_weights_1d(xs, nfs[ax], kp)
            idxs.append(i)
            ws.append(w)
        flat = _flat_indices(idxs, nfs).ravel()
        w = _outer_weights(ws)
        for t in range(n_trans):
            contrib = cb[t, s:e, None] * w  # in c's dtype, like the reference
            acc_re[t] += np.bincount(flat, weights=contrib.real.ravel(), minlength=ntot)
            acc_im[t] += np.bincount(flat, weights=contrib.imag.ravel(), minlength=ntot)
    fw = (acc_re + 1j * acc_im).astype(cb.dtype).reshape((n_trans,) + tuple(reversed(nfs)))
    return fw if batched else fw[0]


def interp(coords: Sequence[np.ndarray], fw: np.ndarray, kp: KernelParams, chunk: int = 1 << 15):
    """Type-2 interp, ``spread.py:342-381`` (1D), ``:497-540`` (2D), ``:661-709`` (3D).

    fw: (n_trans, nf_d, ..., nf_1) or unbatched.  Returns (n_trans, M).
    """
    d = len(coords)
    fb, batched = _as_batched(fw, d)
    nfs = list(reversed(fb.shape[1:]))
    n_trans = fb.shape[0]
    flat_grid = fb.reshape(n_trans, -1)
    M = len(coords[0])
    out = np.zeros((n_trans, M), dtype=fb.dtype)
    for s in range(0, M, chunk):
        e = min(M, s + chunk)
        idxs, ws = [], []
        for ax in range(d):
            xs = fold_rescale(np.asarray(coords[ax][s:e]), nfs[ax])
            i, w = kernel_weights_1d(xs, nfs[ax], kp)
            idxs.append(i)
            ws.append(w)
        flat = _flat_indices(idxs, nfs)
        w = _outer_weights(ws)
        for t in range(n_trans):
            out[t, s:e] = np.sum(flat_grid[t][flat] * w, axis=1)
    return out if batched else out[0]


def spread_grad_coords(coords, c, g, kp: KernelParams):
    """Coordinate gradient of spread, ``spread.py:820-858`` (1D), ``:997-1032`` (2D), ``:1168-1212`` (3D).

    g is the cotangent on the fine grid in the reference's (JAX) convention;
    returns [d_x, d_y, d_z][:d] with d_ax[j] = sum_t Re(conj(c_tj) * sum g * dW/d ax).
    """
    return _grad_coords(coords, c, g, kp, conj_on="pointwise")


def interp_grad_coords(coords, fw, g, kp: KernelParams):
    """Coordinate gradient of interp, ``spread.py:911-944`` (1D), ``:1080-1114`` (2D), ``:1261-1305`` (3D).

    Returns d_ax[j] = sum_t Re(conj(g_tj) * sum fw * dW/d ax).
    """
    return _grad_coords(coords, g, fw, kp, conj_on="pointwise")


def _grad_coords(coords, pointwise, grid, kp, conj_on):
    d = len(coords)
    gb, _ = _as_batched(grid, d)
    pb, _ = _as_batched(pointwise, 1)
    nfs = list(reversed(gb.shape[1:]))
    flat_grid = gb.reshape(gb.shape[0], -1)
    idxs, ws, dws = [], [], []
    for ax in range(d):
        xs = fold_rescale(np.asarray(coords[ax]), nfs[ax])
        i, w, dw = kernel_weights_1d(xs, nfs[ax], kp, derivative=True)
        idxs.append(i)
        ws.append(w)
        dws.append(dw)
    flat = _flat_indices(idxs, nfs)
    outs = []
    for ax in range(d):
        wl = [dws[a] if a == ax else ws[a] for a in range(d)]
        wprod = _outer_weights(wl)
        acc = np.zeros(len(coords[0]))
        for t in range(gb.shape[0]):
            s = np.sum(flat_grid[t][flat] * wprod, axis=1)
            acc += np.real(np.conj(pb[t]) * s)
        outs.append(acc)
    return outs


# ---------------------------------------------------------------------------
# Deconvolution
# ---------------------------------------------------------------------------


def deconv_factors(phihat: np.ndarray, n_modes: int, modeord: int = 0) -> np.ndarray:
    """``nufftax/core/deconvolve.py:27-56``: (-1)^k / phihat[|k|] in output mode order."""
    kmin, kmax = -(n_modes // 2), (n_modes - 1) // 2
    if modeord == 0:
        k = np.arange(kmin, kmax + 1)
    else:
        k = np.concatenate([np.arange(kmax + 1), np.arange(kmin, 0)])
    fac = 1.0 / phihat[np.abs(k)]
    return np.where(k % 2 == 0, 1.0, -1.0).astype(phihat.dtype) * fac


def extraction_indices(nf: int, n_modes: int, modeord: int = 0) -> np.ndarray:
    """``nufftax/core/deconvolve.py:59-73``: fine-grid index of each kept mode."""
    kmin, kmax = -(n_modes // 2), (n_modes - 1) // 2
    pos, neg = np.arange(kmax + 1), np.arange(nf + kmin, nf)
    return np.concatenate([neg, pos]) if modeord == 0 else np.concatenate([pos, neg])


def deconvolve_shuffle(fw_hat: np.ndarray, phihats: Sequence[np.ndarray], n_modes: Sequence[int], modeord: int = 0):
    """Type-1 mode extraction, ``deconvolve.py:81-111`` (1D), ``:174-224`` (2D), ``:312-366`` (3D).

    phihats / n_modes are per dimension in (x, y, z) order; fw_hat is
    (..., nf_d, ..., nf_1).  Output (..., n_d, ..., n_1).
    """
    d = len(n_modes)
    out = fw_hat
    for ax in range(d):
        axis = fw_hat.ndim - 1 - ax
        nf = fw_hat.shape[axis]
        idx = extraction_indices(nf, n_modes[ax], modeord)
        fac = deconv_factors(phihats[ax], n_modes[ax], modeord)
        shape = [1] * fw_hat.ndim
        shape[axis] = n_modes[ax]
        out = np.take(out, idx, axis=axis) * fac.reshape(shape)
    return out.astype(fw_hat.dtype)


def deconvolve_pad(f: np.ndarray, phihats: Sequence[np.ndarray], nfs: Sequence[int], modeord: int = 0):
    """Type-2 pad, ``deconvolve.py:114-166`` (1D), ``:227-304`` (2D), ``:369-468`` (3D)."""
    d = len(nfs)
    lead = f.shape[: f.ndim - d]
    out = np.zeros(lead + tuple(reversed(nfs)), dtype=f.dtype)
    n_modes = [f.shape[f.ndim - 1 - ax] for ax in range(d)]
    fac = np.ones((1,) * f.ndim, dtype=f.real.dtype)
    idx = []
    for ax in range(d):
        shape = [1] * f.ndim
        shape[f.ndim - 1 - ax] = n_modes[ax]
        fac = fac * deconv_factors(phihats[ax], n_modes[ax], modeord).reshape(shape)
        idx.append(extraction_indices(nfs[ax], n_modes[ax], modeord))
    grids = np.ix_(*[idx[ax] for ax in reversed(range(d))])
    out[(Ellipsis,) + grids] = f * fac
    return out


# ---------------------------------------------------------------------------
# FFT with the reference's sign mapping
# ---------------------------------------------------------------------------


def fft_isign(a: np.ndarray, d: int, isign: int) -> np.ndarray:
    """isign>0: ifftn * N (``nufft1.py:81,154,232-235``); isign<0: fftn (``nufft2.py:74,140,211-214``)."""
    axes = tuple(range(a.ndim - d, a.ndim))
    if isign > 0:
        n = int(np.prod([a.shape[ax] for ax in axes]))
        return (np.fft.ifftn(a, axes=axes) * n).astype(a.dtype)
    return np.fft.fftn(a, axes=axes).astype(a.dtype)


# ---------------------------------------------------------------------------
# Transforms
# ---------------------------------------------------------------------------


def _n_modes_tuple(n_modes, d):
    if isinstance(n_modes, (int, np.integer)):
        return (int(n_modes),) * d
    return tuple(int(v) for v in n_modes)


def nufft_type1(coords, c, n_modes, eps=1e-6, isign=1, upsampfac=2.0, modeord=0):
    """Type 1, ``nufftax/transforms/nufft1.py:21-243``.

    coords: [x] / [x, y] / [x, y, z]; n_modes: int or (n1, n2[, n3]).
    Output (..., n_d, ..., n_1) in c's dtype.
    """
    d = len(coords)
    nm = _n_modes_tuple(n_modes, d)
    kp = compute_kernel_params(eps, upsampfac)
    nfs = [compute_grid_size(n, upsampfac, kp.nspread) for n in nm]
    rdt = np.asarray(c).real.dtype
    phihats = [kernel_fourier_series(nf, kp.nspread, kp.beta, kp.c, dtype=rdt.type) for nf in nfs]
    xn = [wrap_pi(np.asarray(x)) for x in coords]
    fw = spread(xn, c, nfs, kp)
    fw_hat = fft_isign(fw, d, isign)
    return deconvolve_shuffle(fw_hat, phihats, nm, modeord)


def nufft_type2(coords, f, eps=1e-6, isign=-1, upsampfac=2.0, modeord=0):
    """Type 2, ``nufftax/transforms/nufft2.py:21-227``.  f: (..., n_d, ..., n_1) -> (..., M)."""
    d = len(coords)
    f = np.asarray(f)
    nm = [f.shape[f.ndim - 1 - ax] for ax in range(d)]
    kp = compute_kernel_params(eps, upsampfac)
    nfs = [compute_grid_size(n, upsampfac, kp.nspread) for n in nm]
    rdt = f.real.dtype
    phihats = [kernel_fourier_series(nf, kp.nspread, kp.beta, kp.c, dtype=rdt.type) for nf in nfs]
    fw_hat = deconvolve_pad(f, phihats, nfs, modeord)
    fw = fft_isign(fw_hat, d, isign)
    xn = [wrap_pi(np.asarray(x)) for x in coords]
    return interp(xn, fw, kp)


def nufft_type3(coords, c, targets, n_modes, eps=1e-6, isign=1, upsampfac=2.0):
    """Type 3, ``nufftax/transforms/nufft3.py:216-515`` (d = len(coords)).

    n_modes are the fine-grid sizes (nf1, ...); the inner type 2 runs on the
    spread grid as CMCL "modes" (``nufft3.py:292,387,494-502``).
    """
    d = len(coords)
    nfs = _n_modes_tuple(n_modes, d)
    cb, batched = _as_batched(c, 1)
    kp = compute_kernel_params(eps, upsampfac)
    rdt = cb.real.dtype
    xs_n, ss_n, s_resc = [], [], []
    phase_pre = np.zeros(len(coords[0]), dtype=rdt)
    phase_post = np.zeros(len(targets[0]), dtype=rdt)
    for ax in range(d):
        x = np.asarray(coords[ax])
        s = np.asarray(targets[ax])
        C = (x.min() + x.max()) / 2.0
        D = (s.min() + s.max()) / 2.0
        S = max((s.max() - s.min()) / 2.0, 1e-10)
        nf = nfs[ax]
        h = 2.0 * np.pi / nf
        gamma = nf / (2.0 * upsampfac * S)
        xs_n.append(wrap_pi(((x - C) / gamma).astype(rdt)))
        phase_pre = phase_pre + D * x
        sr = (h * gamma * (s - D)).astype(rdt)
        s_resc.append(sr)
        ss_n.append(wrap_pi(sr))
        phase_post = phase_post + (s - D) * C
    cdt = cb.dtype
    c_ph = cb * np.exp(1j * isign * phase_pre).astype(cdt)[None, :]
    fw = spread(xs_n, c_ph, nfs, kp)
    f = nufft_type2(ss_n, fw, eps=eps, isign=isign, upsampfac=upsampfac)
    phi = np.ones(len(targets[0]), dtype=rdt)
    for ax in range(d):
        phi = phi * kernel_ft_at_point(s_resc[ax], kp.nspread, kp.beta, kp.c, dtype=rdt.type)
    out = (f * (np.exp(1j * isign * phase_post).astype(cdt) / phi)[None, :]).astype(cdt)
    return out if batched else out[0]


# ---------------------------------------------------------------------------
# Direct sums (FP64 truth)
# ---------------------------------------------------------------------------


def mode_grid(n: int, modeord: int = 0) -> np.ndarray:
    """Mode numbers in output order (CMCL for modeord=0)."""
    kmin, kmax = -(n // 2), (n - 1) // 2
    if modeord == 0:
        return np.arange(kmin, kmax + 1)
    return np.concatenate([np.arange(kmax + 1), np.arange(kmin, 0)])


def direct_type1(coords, c, n_modes, isign=1, mode_subset=None, chunk=1 << 14):
    """f[k] = sum_j c_j exp(i isign k.x_j) in FP64 (``tests/conftest.py:154-173`` pattern).

    mode_subset: optional (K, d) integer array of modes (k1, k2, k3); then the
    result is (..., K).  Otherwise the full CMCL cube (..., n_d, ..., n_1).
    """
    d = len(coords)
    nm = _n_modes_tuple(n_modes, d)
    X = np.stack([np.asarray(x, dtype=np.float64) for x in coords], axis=1)  # (M, d)
    cb, batched = _as_batched(np.asarray(c).astype(np.complex128), 1)
    if mode_subset is None:
        grids = np.meshgrid(*[mode_grid(n) for n in reversed(nm)], indexing="ij")
        K = np.stack([g.ravel() for g in reversed(grids)], axis=1)  # x fastest
        out_shape = tuple(reversed(nm))
    else:
        K = np.asarray(mode_subset, dtype=np.int64).reshape(-1, d)
        out_shape = (K.shape[0],)
    out = np.zeros((cb.shape[0], K.shape[0]), dtype=np.complex128)
    for s in range(0, K.shape[0], chunk):
        ph = np.exp(1j * isign * (K[s : s + chunk].astype(np.float64) @ X.T))  # (k, M)
        out[:, s : s + chunk] = cb @ ph.T
    out = out.reshape((cb.shape[0],) + out_shape)
    return out if batched else out[0]


def direct_type2(coords, f, isign=-1, point_subset=None, chunk=256):
    """c_j = sum_k f_k exp(i isign k.x_j) in FP64 (``tests/conftest.py:176-198`` pattern)."""
    d = len(coords)
    f = np.asarray(f).astype(np.complex128)
    nm = [f.shape[f.ndim - 1 - ax] for ax in range(d)]
    fb = f.reshape((-1,) + tuple(reversed(nm)))
    grids = np.meshgrid(*[mode_grid(n) for n in reversed(nm)], indexing="ij")
    K = np.stack([g.ravel() for g in reversed(grids)], axis=1).astype(np.float64)
    X = np.stack([np.asarray(x, dtype=np.float64) for x in coords], axis=1)
    if point_subset is not None:
        X = X[np.asarray(point_subset)]
    fl = fb.reshape(fb.shape[0], -1)
    out = np.zeros((fb.shape[0], X.shape[0]), dtype=np.complex128)
    for s in range(0, X.shape[0], chunk):
        ph = np.exp(1j * isign * (X[s : s + chunk] @ K.T))  # (m, K)
        out[:, s : s + chunk] = fl @ ph.T
    return out.reshape(f.shape[: f.ndim - d] + (X.shape[0],))


def direct_type3(coords, c, targets, isign=1, chunk=256):
    """f_k = sum_j c_j exp(i isign s_k.x_j) in FP64 (``tests/test_nufft3.py:25-75`` pattern)."""
    X = np.stack([np.asarray(x, dtype=np.float64) for x in coords], axis=1)
    S = np.stack([np.asarray(s, dtype=np.float64) for s in targets], axis=1)
    cb, batched = _as_batched(np.asarray(c).astype(np.complex128), 1)
    out = np.zeros((cb.shape[0], S.shape[0]), dtype=np.complex128)
    for s in range(0, S.shape[0], chunk):
        ph = np.exp(1j * isign * (S[s : s + chunk] @ X.T))
        out[:, s : s + chunk] = cb @ ph.T
    return out if batched else out[0]


def rel_l2(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# ---------------------------------------------------------------------------
# Bin-sort oracle (integer; no reference counterpart)
# ---------------------------------------------------------------------------


def bin_keys(coords, nfs, bin_sizes) -> np.ndarray:
    """Bin index of each point: floor(xs_d / bin_d) flattened with x fastest.

    xs is the FP64 fold ``((x/2pi + 1/2) mod 1) * nf`` (``spread.py:78-97``
    evaluated in float64), with values that round up to nf mapped back to 0.
    """
    key = np.zeros(len(coords[0]), dtype=np.int64)
    stride = 1
    for ax in range(len(nfs)):
        x = np.asarray(coords[ax], dtype=np.float64)
        r = x * (1.0 / (2.0 * np.pi)) + 0.5
        r = r - np.floor(r)
        xs = r * nfs[ax]
        xs = np.where(xs >= nfs[ax], xs - nfs[ax], xs)
        nb = -(-nfs[ax] // bin_sizes[ax])
        b = np.floor(xs).astype(np.int64) // bin_sizes[ax]
        key += b * stride
        stride *= nb
    return key


def bin_sort(keys: np.ndarray, n_bins: int):
    """Stable counting sort contract: (perm, offsets) with perm == argsort(keys, stable)."""
    perm = np.argsort(keys, kind="stable").astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(np.bincount(keys, minlength=n_bins))]).astype(np.int64)
    return perm, offsets

