"""Kernel parameters and grid sizes (host side).

``compute_kernel_params`` / ``compute_grid_size`` call the C ABI (one source of
truth with the kernels); the type-3 grid helpers follow
nufftax/transforms/nufft3.py:35-173.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from . import _lib


class KernelParams(NamedTuple):
    """``nufftax/core/kernel.py:19-25``."""

    nspread: int
    beta: float
    c: float
    upsampfac: float


def compute_kernel_params(tol: float, upsampfac: float = 2.0, max_nspread: int = 16) -> KernelParams:
    """``nufftax/core/kernel.py:165-206`` (max_nspread is fixed at 16 by the kernels)."""
    if max_nspread != 16:
        raise ValueError("the HIP kernels support nspread up to 16; max_nspread must be 16")
    ns, beta, c = _lib.kernel_params(tol, upsampfac)
    return KernelParams(ns, beta, c, float(upsampfac))


def next_smooth_int(n: int) -> int:
    """``nufftax/utils/grid.py:6-33``."""
    if n <= 1:
        return 1
    while True:
        v = n
        for p in (2, 3, 5):
            while v % p == 0:
                v //= p
        if v == 1:
            return n
        n += 1


def compute_grid_size(n_modes: int, upsampfac: float, nspread: int) -> int:
    """``nufftax/utils/grid.py:36-58``."""
    return _lib.grid_size(n_modes, upsampfac, nspread)


def _next_smooth_even(n: int) -> int:
    """``nufftax/transforms/nufft3.py:35-47``."""
    if n <= 1:
        return 2
    if n % 2 == 1:
        n += 1
    cand = next_smooth_int(n)
    while cand % 2 != 0:
        cand = next_smooth_int(cand + 1)
    return cand


def _extent(a) -> float:
    arr = np.asarray(a.detach().cpu() if hasattr(a, "detach") else a)
    if arr.ndim > 0:
        return float((np.max(arr) - np.min(arr)) / 2)
    return float(arr)


def compute_type3_grid_size(x_or_x_extent, s_or_s_extent, eps: float = 1e-6, upsampfac: float = 2.0) -> int:
    """``nufftax/transforms/nufft3.py:50-104``: arrays or half-widths."""
    xe = max(_extent(x_or_x_extent), 1e-10)
    se = max(_extent(s_or_s_extent), 1e-10)
    ns = compute_kernel_params(eps, upsampfac).nspread
    return _next_smooth_even(int(np.ceil(2.0 * upsampfac * se * xe / np.pi + ns)))


def compute_type3_grid_sizes_2d(x_extent, y_extent, s_extent, t_extent, eps=1e-6, upsampfac=2.0):
    """``nufftax/transforms/nufft3.py:107-136``."""
    ns = compute_kernel_params(eps, upsampfac).nspread
    xe, ye, se, te = (max(float(v), 1e-10) for v in (x_extent, y_extent, s_extent, t_extent))
    return (
        _next_smooth_even(int(np.ceil(2.0 * upsampfac * se * xe / np.pi + ns))),
        _next_smooth_even(int(np.ceil(2.0 * upsampfac * te * ye / np.pi + ns))),
    )


def compute_type3_grid_sizes_3d(x_extent, y_extent, z_extent, s_extent, t_extent, u_extent, eps=1e-6, upsampfac=2.0):
    """``nufftax/transforms/nufft3.py:139-173``."""
    ns = compute_kernel_params(eps, upsampfac).nspread
    xe, ye, ze, se, te, ue = (max(float(v), 1e-10) for v in (x_extent, y_extent, z_extent, s_extent, t_extent,
                                                              u_extent))
    return (
        _next_smooth_even(int(np.ceil(2.0 * upsampfac * se * xe / np.pi + ns))),
        _next_smooth_even(int(np.ceil(2.0 * upsampfac * te * ye / np.pi + ns))),
        _next_smooth_even(int(np.ceil(2.0 * upsampfac * ue * ze / np.pi + ns))),
    )
