"""Serving loop on the GPU: results equal the direct calls, grouping, admission.

Tolerance: relative l2 1e-6 (complex64 at eps 1e-6; the spread flush adds tile
overlaps with float atomics, so repeated runs may differ in the last bits).
"""

import numpy as np
import pytest
import torch

from oracle import nufft_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _rel(a, b):
    return R.rel_l2(a.detach().cpu().numpy(), b.detach().cpu().numpy())


def test_mixed_stream_matches_direct_calls():
    from nufftax_amd import serving as S

    dev = torch.device("cuda:0")
    reqs = S.mixed_requests(16, seed=5, device=dev)
    with S.NufftServer(dev, n_streams=3) as srv:
        futs = [srv.submit(k, c, d, n_modes=nm if k == 1 else None, eps=1e-6) for k, c, d, nm in reqs]
        outs = [f.result(timeout=120) for f in futs]
    for (k, c, d, nm), o in zip(reqs, outs):
        ref = S._direct_call(k, c, d, nm)
        assert o.shape == ref.shape and o.dtype == ref.dtype
        assert _rel(o, ref) < 1e-6
    st = S.latency_stats([f.request for f in futs])
    assert st["n"] == 16 and st["p99_ms"] >= st["p50_ms"] > 0


def test_shared_points_are_grouped_and_equal():
    from nufftax_amd import serving as S
    import nufftax_amd as N

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    M = 50000
    x, y = ((torch.rand(M, device=dev, generator=g) * 2 - 1) * np.pi for _ in range(2))
    cs = [torch.randn(M, dtype=torch.complex64, device=dev, generator=g) for _ in range(6)]
    with S.NufftServer(dev, n_streams=2) as srv:
        futs = [srv.submit(1, (x, y), c, n_modes=(64, 48)) for c in cs]
        outs = [f.result(timeout=60) for f in futs]
    for c, o in zip(cs, outs):
        assert _rel(o, N.nufft2d1(x, y, c, (64, 48), eps=1e-6)) < 1e-6


def test_shared_launches_across_point_sets_match_direct_calls():
    """Uneven small requests (different points, M and grid sizes) through shared spreading /
    interpolation launches equal the same requests called one by one."""
    from nufftax_amd import serving as S

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    reqs = []
    for i in range(24):
        kind = 1 + (i % 2)
        nm = [(32, 32), (64, 48), (96, 64), (128, 128)][(i // 2) % 4]
        M = 3000 + 997 * i
        pts = tuple((torch.rand(M, device=dev, generator=g) * 2 - 1) * np.pi for _ in range(2))
        if kind == 1:
            data = torch.randn(M, dtype=torch.complex64, device=dev, generator=g)
        else:
            data = torch.randn(tuple(reversed(nm)), dtype=torch.complex64, device=dev, generator=g)
        reqs.append((kind, pts, data, nm))
    with S.NufftServer(dev, n_streams=2, max_units=8) as srv:
        srv.batch_min_queue = 1
        futs = [srv.submit(k, c, d, n_modes=nm if k == 1 else None, eps=1e-6) for k, c, d, nm in reqs]
        outs = [f.result(timeout=120) for f in futs]
        shared = srv.shared_launches
    for (k, c, d, nm), o in zip(reqs, outs):
        ref = S._direct_call(k, c, d, nm)
        assert o.shape == ref.shape and _rel(o, ref) < 1e-6
    assert shared >= 1


def test_admission_budget():
    from nufftax_amd import serving as S

    dev = torch.device("cuda:0")
    reqs = S.mixed_requests(6, seed=9, device=dev)
    need = max(S.request_bytes(k, len(c), c[0].shape[0], nm, 1e-6) for k, c, d, nm in reqs)
    with S.NufftServer(dev, n_streams=2, hbm_budget_bytes=need) as srv:
        futs = [srv.submit(k, c, d, n_modes=nm if k == 1 else None) for k, c, d, nm in reqs]
        for f in futs:
            f.result(timeout=120)
        k, c, d, nm = reqs[0]
    with S.NufftServer(dev, n_streams=1, hbm_budget_bytes=1024) as srv:
        with pytest.raises(ValueError):
            srv.submit(k, c, d, n_modes=nm if k == 1 else None)
