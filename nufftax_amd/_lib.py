"""ctypes binding of the nufftx C ABI (``include/nufftx.h``).

PyTorch is imported first so that the HIP runtime and rocFFT that torch ships are
the ones the library binds to (same sonames: libamdhip64.so.7, librocfft.so.0),
giving one HIP runtime per process.  The library is never replaced by a CPU
fallback: if it is missing or fails to load, every call raises.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import torch  # noqa: F401  (binds the process to torch's HIP runtime before the library loads)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnufftx.so")
CSRC = os.path.join(_HERE, "csrc")

OK = 0
F32, F64 = 0, 1

# name -> (restype, argtypes)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
SIGNATURES = {
    "nufftx_version": (ctypes.c_int, []),
    "nufftx_last_error": (ctypes.c_char_p, []),
    "nufftx_kernel_params": (
        ctypes.c_int,
        [ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
         ctypes.POINTER(ctypes.c_double)],
    ),
    "nufftx_grid_size": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_double, ctypes.c_int]),
    "nufftx_makeplan": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_int, _c_int64_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
         ctypes.c_int, ctypes.c_int, _vp, ctypes.POINTER(_vp)],
    ),
    "nufftx_plan_grid_size": (ctypes.c_int64, [_vp, ctypes.c_int, ctypes.c_int]),
    "nufftx_set_stream": (ctypes.c_int, [_vp, _vp]),
    "nufftx_setpts": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp, _vp]),
    "nufftx_execute": (ctypes.c_int, [_vp, _vp, _vp]),
    "nufftx_execute_sign": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "nufftx_execute_many": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "nufftx_vjp": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nufftx_destroy": (ctypes.c_int, [_vp]),
    "nufftx_makeplan_grid": (
        ctypes.c_int,
        [ctypes.c_int, _c_int64_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int64, ctypes.c_int, _vp,
         ctypes.POINTER(_vp)],
    ),
    "nufftx_spread": (ctypes.c_int, [_vp, _vp, _vp]),
    "nufftx_spread_add": (ctypes.c_int, [_vp, _vp, _vp]),
    "nufftx_interp": (ctypes.c_int, [_vp, _vp, _vp]),
    "nufftx_gather_grad": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nufftx_get_sort_info": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    ),
    "nufftx_copy_sort": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "nufftx_kernel_fourier_series": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                    _vp]),
    "nufftx_deconvolve": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_int64_p, _c_int64_p, ctypes.c_int64, _vp, _vp, _vp,
                                         ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]),
    "nufftx_plan_acquire": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_int, _c_int64_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
         ctypes.c_int, ctypes.c_int, _vp, ctypes.POINTER(_vp)],
    ),
    "nufftx_plan_acquire_grid": (
        ctypes.c_int,
        [ctypes.c_int, _c_int64_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int64, ctypes.c_int, _vp,
         ctypes.POINTER(_vp)],
    ),
    "nufftx_plan_release": (ctypes.c_int, [_vp]),
    "nufftx_plan_cache_clear": (ctypes.c_int, []),
    "nufftx_plan_cache_stats": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int64)] * 4),
    "nufftx_device_allocations": (ctypes.c_int64, []),
    "nufftx_spread_noalloc": (ctypes.c_int, [ctypes.c_int, _c_int64_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp]),
}

# nufftx_spread_noalloc limits (include/nufftx.h)
NOALLOC_SCRATCH_BYTES = 16 << 20
NOALLOC_MAX_POINTS = 1 << 18

_lib = None
_lock = threading.Lock()


class NufftxError(RuntimeError):
    pass


def build(jobs: int = 8, verbose: bool = False) -> str:
    """Compile libnufftx.so for gfx950 with hipcc (make -B in csrc/: every source, every time)."""
    cmd = ["make", "-B", "-C", CSRC, f"-j{jobs}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise NufftxError(f"building libnufftx.so failed:\n{res.stdout[-4000:]}\n{res.stderr[-4000:]}")
    if verbose:
        print(res.stdout[-2000:])
    return LIB_PATH


def lib() -> ctypes.CDLL:
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NufftxError(
                    f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                    "or `make -C nufftax_amd/csrc`"
                )
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().nufftx_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = "") -> None:
    if status != OK:
        raise NufftxError(f"{what} failed (status {status}): {last_error()}")


def ptr(t) -> ctypes.c_void_p:
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def kernel_params(eps: float, upsampfac: float = 2.0):
    ns = ctypes.c_int()
    beta = ctypes.c_double()
    c = ctypes.c_double()
    check(lib().nufftx_kernel_params(float(eps), float(upsampfac), ctypes.byref(ns), ctypes.byref(beta),
                                     ctypes.byref(c)), "nufftx_kernel_params")
    return ns.value, beta.value, c.value


def grid_size(n_modes: int, upsampfac: float, nspread: int) -> int:
    v = lib().nufftx_grid_size(int(n_modes), float(upsampfac), int(nspread))
    if v < 0:
        raise NufftxError("nufftx_grid_size: invalid arguments")
    return int(v)
