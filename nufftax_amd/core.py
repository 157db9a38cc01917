"""Stage-level operators: spreading, interpolation, deconvolution, kernel helpers.

Mirrors ``nufftax.core`` (nufftax/core/__init__.py): ``spread_{1,2,3}d`` and
``interp_{1,2,3}d`` with the custom VJP of nufftax/core/spread.py:766-1305,
``deconvolve_{shuffle,pad}_{1,2,3}d`` (nufftax/core/deconvolve.py:81-468),
``fold_rescale`` (spread.py:78-97) and the ES kernel helpers
(nufftax/core/kernel.py:79-260).

Spreading, interpolation and the coordinate-gradient gathers run in the nufftx
HIP library (``nufftx_spread`` / ``nufftx_interp`` / ``nufftx_gather_grad``),
deconvolution in ``nufftx_deconvolve``; the kernel Fourier series is the
library's FP64 host routine.  ``es_kernel`` / ``es_kernel_with_derivative`` and
``fold_rescale`` are closed-form elementwise helpers evaluated with torch
arithmetic on the caller's tensors.

Grid layout follows the reference: (nf2, nf1) in 2D and (nf3, nf2, nf1) in 3D,
x fastest, with an optional leading n_trans axis; strengths / point values are
(M,) or (n_trans, M).  complex64 runs in FP32 (points folded in FP64),
complex128 in FP64.

Gradients use PyTorch's complex convention (dL/dRe + i dL/dIm).  With it the
reference's formulas hold as written: the adjoint of spreading in c is
interpolation and vice versa; the coordinate gradient is
``dx_j = sum_t Re(conj(P_tj) sum_k G_t[idx_jk] dw_jk/dx)`` with (P, G) = (c, grid
cotangent) for spread and (point cotangent, fw) for interp
(spread.py:820-858, 911-944).

``Kernel`` accepts the built-in exponential-of-semicircle kernel only: the HIP
kernels evaluate it in registers, and an arbitrary Python ``phi`` cannot be
compiled into them.
"""

from __future__ import annotations

import ctypes
import math
from typing import Callable, NamedTuple

import numpy as np
import torch

from . import _lib
from ._lib import F32, F64, check, lib, ptr
from ._plan import _stream_handle, grid_plan
from .params import KernelParams, compute_kernel_params
from .transforms import _run

__all__ = [
    "es_kernel", "es_kernel_with_derivative", "compute_kernel_params", "kernel_fourier_series", "KernelParams",
    "Kernel", "es_kernel_spec",
    "spread_1d", "spread_2d", "spread_3d", "interp_1d", "interp_2d", "interp_3d", "fold_rescale",
    "deconvolve_shuffle_1d", "deconvolve_pad_1d", "deconvolve_shuffle_2d", "deconvolve_pad_2d",
    "deconvolve_shuffle_3d", "deconvolve_pad_3d",
]

TWO_PI = 2.0 * math.pi


# ---------------------------------------------------------------- kernel helpers (kernel.py:28-260)


class Kernel(NamedTuple):
    """Spreading kernel (nufftax/core/kernel.py:28-66).

    Only kernels built by :func:`es_kernel_spec` are accepted by the spread /
    interp operators (``params`` carries their ES parameters); ``phi`` and
    ``phi_and_dphi`` evaluate the same function with torch arithmetic.
    """

    nspread: int
    phi: Callable
    phi_and_dphi: Callable | None = None
    params: KernelParams | None = None

    def value(self, z):
        return self.phi(z)

    def value_and_grad(self, z):
        if self.phi_and_dphi is None:
            raise NotImplementedError("kernel has no derivative")
        return self.phi_and_dphi(z)


def es_kernel_spec(kernel_params: KernelParams) -> Kernel:
    """The built-in ES kernel as a :class:`Kernel` (kernel.py:69-76)."""
    beta, c = kernel_params.beta, kernel_params.c
    return Kernel(
        nspread=kernel_params.nspread,
        phi=lambda z: es_kernel(z, beta, c),
        phi_and_dphi=lambda z: es_kernel_with_derivative(z, beta, c),
        params=kernel_params,
    )


def _es_params(kernel) -> KernelParams:
    if isinstance(kernel, Kernel):
        if kernel.params is None:
            raise NotImplementedError(
                "custom kernels are not supported by the HIP spread/interp kernels; use es_kernel_spec(params)")
        return kernel.params
    if isinstance(kernel, KernelParams):
        return kernel
    if hasattr(kernel, "nspread") and hasattr(kernel, "beta") and hasattr(kernel, "c"):
        return KernelParams(int(kernel.nspread), float(kernel.beta), float(kernel.c),
                            float(getattr(kernel, "upsampfac", 2.0)))
    raise TypeError(f"expected KernelParams or Kernel, got {type(kernel).__name__}")


def _real_tensor(z):
    z = torch.as_tensor(z)
    return z if z.is_floating_point() else z.to(torch.float32)


def es_kernel(z, beta: float, c: float):
    """phi(z) = exp(beta (sqrt(1 - c z^2) - 1)) where 1 - c z^2 >= 0, else 0, in z's dtype
    (kernel.py:79-111)."""
    z = _real_tensor(z)
    arg = 1.0 - c * z * z
    val = torch.exp(beta * (torch.sqrt(torch.clamp(arg, min=0.0)) - 1.0))
    return torch.where(arg >= 0, val, torch.zeros_like(val))


def es_kernel_with_derivative(z, beta: float, c: float):
    """(phi, dphi/dz), dphi = -beta c z / sqrt(arg) phi with arg = 1 - c z^2 clamped at 1e-14;
    phi is zero where arg < 0 and dphi where arg <= 1e-14 (kernel.py:114-162)."""
    z = _real_tensor(z)
    arg = 1.0 - c * z * z
    sq = torch.sqrt(torch.clamp(arg, min=1e-14))
    phi = torch.where(arg >= 0, torch.exp(beta * (sq - 1.0)), torch.zeros_like(z))
    dphi = torch.where(arg > 1e-14, -beta * c * z / sq * phi, torch.zeros_like(z))
    return phi, dphi


def kernel_fourier_series(nf: int, nspread: int, beta: float, c: float, dtype=torch.float32):
    """phihat[k], k = 0 .. nf//2, by the trapezoid rule on max(int(4 + 1.5 ns), 20) nodes
    over [-ns/2, ns/2] (kernel.py:209-260), computed in FP64 by the library."""
    n = nf // 2 + 1
    out = np.zeros(n, dtype=np.float64)
    check(lib().nufftx_kernel_fourier_series(int(nf), int(nspread), float(beta), float(c),
                                             out.ctypes.data_as(ctypes.c_void_p)),
          "nufftx_kernel_fourier_series")
    return torch.from_numpy(out).to(dtype)


def fold_rescale(x, n: int):
    """((x / 2 pi + 1/2) mod 1) * n, evaluated in FP64 and returned in x's dtype (spread.py:78-97)."""
    x = torch.as_tensor(x)
    r = x.to(torch.float64) / TWO_PI + 0.5
    r = (r - torch.floor(r)) * n
    return r.to(x.dtype if x.is_floating_point() else torch.float64)


# ---------------------------------------------------------------- spread / interp (spread.py:766-1305)


def _device_of(tensors):
    for t in tensors:
        if isinstance(t, torch.Tensor) and t.is_cuda:
            return t.device
    if not torch.cuda.is_available():
        raise _lib.NufftxError("nufftax_amd needs an AMD GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _prec_of(cdtype):
    if cdtype == torch.complex128:
        return F64, torch.complex128, torch.float64
    if cdtype == torch.complex64:
        return F32, torch.complex64, torch.float32
    raise TypeError(f"expected a complex64 or complex128 tensor, got {cdtype}")


def _as_cplx(t):
    if t.is_complex():
        return t
    return t.to(torch.complex128 if t.dtype == torch.float64 else torch.complex64)


class _GridCtx(NamedTuple):
    dim: int
    nf: tuple          # (nf1, nf2, nf3)[:dim]
    params: KernelParams


def _plan_for(ctx: _GridCtx, n_trans, prec, dev):
    kp = ctx.params
    return grid_plan(ctx.dim, ctx.nf, kp.nspread, kp.beta, kp.c, n_trans, prec, dev)


def _norm_coords(coords, rdt, dev):
    M = coords[0].shape[-1]
    out = []
    for x in coords:
        x = x.detach().to(device=dev, dtype=rdt)
        if x.ndim != 1 or x.shape[0] != M:
            raise ValueError(f"coordinate arrays must have shape ({M},), got {tuple(x.shape)}")
        out.append(x.contiguous())
    return out, M


def _raw_spread(ctx: _GridCtx, coords, c):
    """c (n_trans, M) -> fw (n_trans, nf[d-1], ..., nf[0])."""
    prec, cdt, rdt = _prec_of(c.dtype)
    dev = c.device
    xs, M = _norm_coords(coords, rdt, dev)
    n_trans = c.shape[0]
    if c.shape[1] != M:
        raise ValueError(f"strengths must have {M} points, got {c.shape[1]}")
    fw = torch.empty((n_trans,) + tuple(reversed(ctx.nf)), dtype=cdt, device=dev)
    if n_trans == 0:
        return fw
    plan = _plan_for(ctx, n_trans, prec, dev)
    with plan.lock:
        plan.setpts(xs)
        plan.spread(c.contiguous(), fw)
    return fw


def _raw_interp(ctx: _GridCtx, coords, fw):
    """fw (n_trans, nf...) -> c (n_trans, M)."""
    prec, cdt, rdt = _prec_of(fw.dtype)
    dev = fw.device
    xs, M = _norm_coords(coords, rdt, dev)
    n_trans = fw.shape[0]
    c = torch.empty((n_trans, M), dtype=cdt, device=dev)
    if n_trans == 0 or M == 0:
        return c.zero_()
    plan = _plan_for(ctx, n_trans, prec, dev)
    with plan.lock:
        plan.setpts(xs)
        plan.interp(fw.contiguous(), c)
    return c


def _raw_coord_grad(ctx: _GridCtx, coords, G, P):
    """grad_ax[j] = sum_t Re(conj(P_tj) sum_k G_t[idx_jk] dw_jk/d ax) for every axis."""
    prec, cdt, rdt = _prec_of(G.dtype)
    dev = G.device
    xs, M = _norm_coords(coords, rdt, dev)
    n_trans = G.shape[0]
    grads = [torch.empty(M, dtype=rdt, device=dev) for _ in range(ctx.dim)]
    if n_trans == 0 or M == 0:
        return [g.zero_() for g in grads]
    plan = _plan_for(ctx, n_trans, prec, dev)
    with plan.lock:
        plan.setpts(xs)
        plan.gather_grad(G.contiguous(), P.to(cdt).contiguous(), None, grads)
    return grads


class _Spread(torch.autograd.Function):
    """fw = spread(x, c); VJP: dc = interp(g), dx = Re(conj(c) sum g dw) (spread.py:793-817, 977-994, 1148-1165)."""

    @staticmethod
    def forward(gctx: _GridCtx, c2, *coords):
        return _raw_spread(gctx, coords, c2)

    @staticmethod
    def setup_context(ctx, inputs, output):
        gctx, c2, *coords = inputs
        ctx.gctx = gctx
        ctx.save_for_backward(c2, *coords)

    @staticmethod
    def backward(ctx, g):
        c2, *coords = ctx.saved_tensors
        gctx = ctx.gctx
        g = g.to(c2.dtype).contiguous()
        dc = _raw_interp(gctx, coords, g) if ctx.needs_input_grad[1] else None
        dco = [None] * gctx.dim
        if any(ctx.needs_input_grad[2:]):
            gr = _raw_coord_grad(gctx, coords, g, c2)
            dco = [gr[d].to(coords[d].dtype) if ctx.needs_input_grad[2 + d] else None for d in range(gctx.dim)]
        return (None, dc, *dco)


class _Interp(torch.autograd.Function):
    """c = interp(x, fw); VJP: dfw = spread(g), dx = Re(conj(g) sum fw dw) (spread.py:886-944, 1060-1114, 1241-1305)."""

    @staticmethod
    def forward(gctx: _GridCtx, fw2, *coords):
        return _raw_interp(gctx, coords, fw2)

    @staticmethod
    def setup_context(ctx, inputs, output):
        gctx, fw2, *coords = inputs
        ctx.gctx = gctx
        ctx.save_for_backward(fw2, *coords)

    @staticmethod
    def backward(ctx, g):
        fw2, *coords = ctx.saved_tensors
        gctx = ctx.gctx
        g = g.to(fw2.dtype).contiguous()
        dfw = _raw_spread(gctx, coords, g) if ctx.needs_input_grad[1] else None
        dco = [None] * gctx.dim
        if any(ctx.needs_input_grad[2:]):
            gr = _raw_coord_grad(gctx, coords, fw2, g)
            dco = [gr[d].to(coords[d].dtype) if ctx.needs_input_grad[2 + d] else None for d in range(gctx.dim)]
        return (None, dfw, *dco)


def _to_dev(a, dev):
    t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))
    return t.to(dev)


def _spread(dim, coords, c, nf, kernel_params):
    was_numpy = isinstance(c, np.ndarray)
    dev = _device_of([c, *coords])
    on_cpu = isinstance(c, torch.Tensor) and not c.is_cuda
    c = _as_cplx(_to_dev(c, dev))
    coords = [_to_dev(x, dev) for x in coords]
    batched = c.ndim == 2
    c2 = c if batched else c[None, :]
    gctx = _GridCtx(dim, tuple(int(v) for v in nf), _es_params(kernel_params))
    fw = _run(_Spread, gctx, c2, *coords)
    if not batched:
        fw = fw[0]
    if was_numpy:
        return fw.detach().cpu().numpy()
    return fw.cpu() if on_cpu else fw


def _interp(dim, coords, fw, kernel_params):
    was_numpy = isinstance(fw, np.ndarray)
    dev = _device_of([fw, *coords])
    on_cpu = isinstance(fw, torch.Tensor) and not fw.is_cuda
    fw = _as_cplx(_to_dev(fw, dev))
    coords = [_to_dev(x, dev) for x in coords]
    batched = fw.ndim == dim + 1
    if fw.ndim not in (dim, dim + 1):
        raise ValueError(f"grid must have {dim} or {dim + 1} dimensions, got shape {tuple(fw.shape)}")
    fw2 = fw if batched else fw[None]
    nf = tuple(reversed(fw2.shape[1:]))
    gctx = _GridCtx(dim, nf, _es_params(kernel_params))
    c = _run(_Interp, gctx, fw2, *coords)
    if not batched:
        c = c[0]
    if was_numpy:
        return c.detach().cpu().numpy()
    return c.cpu() if on_cpu else c


def spread_1d(x, c, nf: int, kernel_params):
    """fw[k] = sum_j c[j] phi(k - fold(x_j)); (M,) or (n_trans, M) -> (nf,) or (n_trans, nf) (spread.py:766-791)."""
    return _spread(1, [x], c, (nf,), kernel_params)


def spread_2d(x, y, c, nf1: int, nf2: int, kernel_params):
    """fw[k2, k1] = sum_j c[j] phi_x phi_y; output (nf2, nf1) or (n_trans, nf2, nf1) (spread.py:952-975)."""
    return _spread(2, [x, y], c, (nf1, nf2), kernel_params)


def spread_3d(x, y, z, c, nf1: int, nf2: int, nf3: int, kernel_params):
    """Output (nf3, nf2, nf1) or (n_trans, nf3, nf2, nf1) (spread.py:1122-1146)."""
    return _spread(3, [x, y, z], c, (nf1, nf2, nf3), kernel_params)


def interp_1d(x, fw, nf: int, kernel_params):
    """c[j] = sum_k fw[k] phi(k - fold(x_j)); fw (nf,) or (n_trans, nf) (spread.py:861-884)."""
    return _interp(1, [x], fw, kernel_params)


def interp_2d(x, y, fw, nf1: int, nf2: int, kernel_params):
    """fw (nf2, nf1) or (n_trans, nf2, nf1) -> (M,) or (n_trans, M) (spread.py:1035-1058)."""
    return _interp(2, [x, y], fw, kernel_params)


def interp_3d(x, y, z, fw, nf1: int, nf2: int, nf3: int, kernel_params):
    """fw (nf3, nf2, nf1) or (n_trans, nf3, nf2, nf1) (spread.py:1215-1239)."""
    return _interp(3, [x, y, z], fw, kernel_params)


# ---------------------------------------------------------------- deconvolution (deconvolve.py:27-468)


def _deconvolve(direction, arr, phihats, dim, n_modes, nf, modeord):
    """direction 0: shuffle (fw_hat grid -> modes), 1: pad (modes -> fw_hat grid)."""
    was_numpy = isinstance(arr, np.ndarray)
    dev = _device_of([arr])
    on_cpu = isinstance(arr, torch.Tensor) and not arr.is_cuda
    a = _as_cplx(_to_dev(arr, dev))
    prec, cdt, rdt = _prec_of(a.dtype)
    inner = tuple(reversed(nf)) if direction == 0 else tuple(reversed(n_modes))
    outer = tuple(reversed(n_modes)) if direction == 0 else tuple(reversed(nf))
    if tuple(a.shape[a.ndim - dim:]) != inner or a.ndim not in (dim, dim + 1):
        raise ValueError(f"expected trailing shape {inner}, got {tuple(a.shape)}")
    lead = a.shape[: a.ndim - dim]
    n_trans = int(math.prod(lead)) if lead else 1
    for d in range(dim):
        if n_modes[d] > nf[d] or n_modes[d] < 1:
            raise ValueError(f"n_modes {n_modes} must be in [1, nf] for nf {nf}")
    ph = []
    for d in range(dim):
        p = _to_dev(phihats[d], dev).to(rdt).contiguous()
        if p.ndim != 1 or p.shape[0] < n_modes[d] // 2 + 1:
            raise ValueError(f"phihat for axis {d} must have at least {n_modes[d] // 2 + 1} entries")
        ph.append(p)
    src = a.reshape((n_trans,) + inner).contiguous()
    out = torch.empty((n_trans,) + outer, dtype=cdt, device=dev)
    nm = (ctypes.c_int64 * 3)(*(list(n_modes) + [1] * (3 - dim)))
    nfa = (ctypes.c_int64 * 3)(*(list(nf) + [1] * (3 - dim)))
    phs = ph + [None] * (3 - dim)
    with torch.cuda.device(dev):
        check(lib().nufftx_deconvolve(direction, dim, nm, nfa, n_trans, ptr(phs[0]), ptr(phs[1]), ptr(phs[2]),
                                      int(modeord), prec, ctypes.c_void_p(_stream_handle(dev)), ptr(src), ptr(out)),
              "nufftx_deconvolve")
    out = out.reshape(lead + outer)
    if was_numpy:
        return out.cpu().numpy()
    return out.cpu() if on_cpu else out


def deconvolve_shuffle_1d(fw_hat, phihat, n_modes: int, modeord: int = 0):
    """f[m] = fw_hat[idx(k_m)] (-1)^k / phihat[|k|] (deconvolve.py:81-111)."""
    nf = _as_shape(fw_hat)[-1]
    return _deconvolve(0, fw_hat, [phihat], 1, (int(n_modes),), (nf,), modeord)


def deconvolve_pad_1d(f, phihat, nf: int, modeord: int = 0):
    """Zero grid (nf,) with the deconvolved modes in FFT order (deconvolve.py:114-166)."""
    n = _as_shape(f)[-1]
    return _deconvolve(1, f, [phihat], 1, (n,), (int(nf),), modeord)


def deconvolve_shuffle_2d(fw_hat, phihat1, phihat2, n_modes1: int, n_modes2: int, modeord: int = 0):
    """fw_hat (.., nf2, nf1) -> f (.., n_modes2, n_modes1) (deconvolve.py:174-224)."""
    s = _as_shape(fw_hat)
    return _deconvolve(0, fw_hat, [phihat1, phihat2], 2, (int(n_modes1), int(n_modes2)), (s[-1], s[-2]), modeord)


def deconvolve_pad_2d(f, phihat1, phihat2, nf1: int, nf2: int, modeord: int = 0):
    """f (.., n2, n1) -> fw_hat (.., nf2, nf1) (deconvolve.py:227-304)."""
    s = _as_shape(f)
    return _deconvolve(1, f, [phihat1, phihat2], 2, (s[-1], s[-2]), (int(nf1), int(nf2)), modeord)


def deconvolve_shuffle_3d(fw_hat, phihat1, phihat2, phihat3, n_modes1: int, n_modes2: int, n_modes3: int,
                          modeord: int = 0):
    """fw_hat (.., nf3, nf2, nf1) -> f (.., n3, n2, n1) (deconvolve.py:312-366)."""
    s = _as_shape(fw_hat)
    return _deconvolve(0, fw_hat, [phihat1, phihat2, phihat3], 3, (int(n_modes1), int(n_modes2), int(n_modes3)),
                       (s[-1], s[-2], s[-3]), modeord)


def deconvolve_pad_3d(f, phihat1, phihat2, phihat3, nf1: int, nf2: int, nf3: int, modeord: int = 0):
    """f (.., n3, n2, n1) -> fw_hat (.., nf3, nf2, nf1) (deconvolve.py:369-468)."""
    s = _as_shape(f)
    return _deconvolve(1, f, [phihat1, phihat2, phihat3], 3, (s[-1], s[-2], s[-3]), (int(nf1), int(nf2), int(nf3)),
                       modeord)


def _as_shape(a):
    return tuple(a.shape)
