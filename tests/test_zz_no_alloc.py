"""nufftx_spread_noalloc (no device allocation, working memory in a fixed device array)
against the allocating path it mirrors, nufftx_makeplan_grid + nufftx_setpts + nufftx_spread,
through the C ABI: two calls in a row per case, the second with a smaller input (the array
keeps the first call's contents), rel l2 < 1e-6 in FP32 and < 1e-12 in FP64 (the same kernels
run on the same sorted points; only the order of the grid's global atomic adds differs)."""

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


def _rel(a, b):
    return float(torch.linalg.vector_norm((a - b).to(torch.complex128)) / torch.linalg.vector_norm(b.to(torch.complex128)))


# (dim, fine grid, eps): 2D FP32 at ns <= 8 takes the matrix-core window spread, 3D FP32 the
# LDS bank-spread kernel, 1D the tiled kernel, FP64 the FP64 tiled kernels
CASES = [(2, (256, 192), 1e-6, "f32"), (2, (128, 128), 1e-4, "f32"), (3, (48, 40, 32), 1e-6, "f32"),
         (1, (4000,), 1e-6, "f32"), (2, (96, 80), 1e-9, "f64")]


@pytest.mark.parametrize("dim,nf,eps,prec", CASES)
def test_spread_noalloc_matches_plan_spread(dim, nf, eps, prec):
    from nufftax_amd import _lib
    from nufftax_amd._plan import F32, F64, grid_plan

    L = _lib.lib()
    kp = R.compute_kernel_params(eps)
    pr = F32 if prec == "f32" else F64
    rdt, cdt = (torch.float32, torch.complex64) if prec == "f32" else (torch.float64, torch.complex128)
    rng = np.random.default_rng(7)
    nfa = (ctypes.c_int64 * 3)(*(list(nf) + [1] * (3 - len(nf))))
    for M in (150_000, 20_000):  # second call smaller: stale contents must not leak into it
        centres = rng.uniform(-np.pi, np.pi, (4, dim))
        pts = centres[rng.integers(0, 4, M)] + rng.normal(0, 0.4, (M, dim))
        xs = [torch.as_tensor(pts[:, d], device=DEV).to(rdt).contiguous() for d in range(dim)]
        c = torch.as_tensor(rng.standard_normal(M) + 1j * rng.standard_normal(M), device=DEV).to(cdt)
        ref = torch.empty(tuple(reversed(nf)), dtype=cdt, device=DEV)
        plan = grid_plan(dim, nf, kp.nspread, kp.beta, kp.c, 1, pr, torch.device(DEV))
        with plan.lock:
            plan.setpts(xs)
            plan.spread(c, ref)
        out = torch.full(tuple(reversed(nf)), float("nan"), dtype=cdt, device=DEV)  # fw is overwritten
        ptr = [x.data_ptr() for x in xs] + [None] * (3 - dim)
        stream = torch.cuda.current_stream().cuda_stream
        rc = L.nufftx_spread_noalloc(dim, nfa, kp.nspread, kp.beta, kp.c, pr, M, ptr[0], ptr[1], ptr[2],
                                     c.data_ptr(), out.data_ptr(), ctypes.c_void_p(stream))
        assert rc == 0, L.nufftx_last_error()
        torch.cuda.synchronize()
        err = _rel(out, ref)
        assert err < (1e-6 if prec == "f32" else 1e-12), (dim, nf, eps, M, err)


def test_spread_noalloc_rejects_sizes_above_its_limit():
    """Host-side checks only: nothing is launched."""
    from nufftax_amd import _lib

    L = _lib.lib()
    nfa = (ctypes.c_int64 * 3)(64, 64, 1)
    big = _lib.NOALLOC_MAX_POINTS + 1
    rc = L.nufftx_spread_noalloc(2, nfa, 8, 18.6, 1.0 / 16, 0, big, None, None, None, None, None, None)
    assert rc == 1 and b"NOALLOC_MAX_POINTS" in L.nufftx_last_error()
