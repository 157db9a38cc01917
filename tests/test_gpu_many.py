"""Several plans in one nufftx_execute_many call (shared spreading / interpolation launch).

Each plan's result must equal the same plan executed alone, to relative l2 1e-6 (complex64:
the fine-grid flush adds overlapping tiles with float atomics in launch-dependent order).
"""

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


def _case(rng, dim, nm, M, kind):
    x = [torch.from_numpy(rng.uniform(-np.pi, np.pi, M).astype(np.float32)).to(DEV) for _ in range(dim)]
    if kind == 1:
        d = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    else:
        shp = tuple(reversed(nm))
        d = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(np.complex64)
    return x, torch.from_numpy(d).to(DEV)


def _run(kind, dim, cases, eps, isigns, many):
    from nufftax_amd._plan import F32, Plan

    plans, ins, outs = [], [], []
    for (x, d, nm), isign in zip(cases, isigns):
        p = Plan.transform(kind, dim, nm, 1, eps, isign, 2.0, F32, 0, torch.device(DEV))
        p.setpts(x)
        o = torch.empty(tuple(reversed(nm)) if kind == 1 else (x[0].numel(),), dtype=torch.complex64, device=DEV)
        plans.append(p)
        ins.append(d)
        outs.append(o)
    if many:
        if kind == 1:
            Plan.execute_many(plans, ins, outs)
        else:
            Plan.execute_many(plans, outs, ins)
    else:
        for p, d, o in zip(plans, ins, outs):
            if kind == 1:
                p.execute(d, o)
            else:
                p.execute(o, d)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("dim", [2, 3])
def test_execute_many_matches_one_by_one(rng, kind, dim):
    """uneven M and grid sizes, mixed signs, one shared launch"""
    shapes = [(64, 48), (128, 128), (33, 20), (256, 64), (64, 64)] if dim == 2 else \
        [(16, 16, 16), (32, 24, 16), (20, 20, 20)]
    Ms = [5000, 40000, 777, 120000, 1][: len(shapes)]
    cases = []
    for nm, M in zip(shapes, Ms):
        x, d = _case(rng, dim, nm, M, kind)
        cases.append((x, d, nm))
    isigns = [1 if (i % 2 == 0) else -1 for i in range(len(cases))]
    a = _run(kind, dim, cases, 1e-6, isigns, many=False)
    b = _run(kind, dim, cases, 1e-6, isigns, many=True)
    for u, v in zip(a, b):
        assert R.rel_l2(v, u) < 1e-6 if np.linalg.norm(u) > 0 else np.linalg.norm(v) == 0


def test_execute_many_falls_back_for_mixed_kernel_widths(rng):
    """plans of different eps cannot share a launch: they run one by one, same results"""
    cases = [(*_case(rng, 2, (64, 64), 3000, 1), (64, 64)), (*_case(rng, 2, (32, 32), 2000, 1), (32, 32))]
    from nufftax_amd._plan import F32, Plan

    plans, outs = [], []
    for (x, d, nm), eps in zip(cases, (1e-6, 1e-3)):
        p = Plan.transform(1, 2, nm, 1, eps, 1, 2.0, F32, 0, torch.device(DEV))
        p.setpts(x)
        plans.append(p)
        outs.append(torch.empty(tuple(reversed(nm)), dtype=torch.complex64, device=DEV))
    Plan.execute_many(plans, [c[1] for c in cases], outs)
    torch.cuda.synchronize()
    import nufftax_amd as N

    for (x, d, nm), eps, o in zip(cases, (1e-6, 1e-3), outs):
        ref = N.nufft2d1(x[0], x[1], d, nm, eps=eps)
        assert R.rel_l2(o.cpu().numpy(), ref.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("kind", [1, 2])
def test_execute_many_n_trans_and_fp64_plans(rng, kind):
    """two transforms per plan in the shared launch (FP32), and FP64 plans (plan-by-plan
    path): each equals the plans executed alone"""
    from nufftax_amd._plan import F32, F64, Plan

    for prec, rdt, cdt, tol in ((F32, np.float32, np.complex64, 1e-6), (F64, np.float64, np.complex128, 1e-12)):
        shapes, Ms = [(48, 40), (128, 96), (64, 64)], [7000, 30000, 1500]
        plans, ins, alone = [], [], []
        for nm, M in zip(shapes, Ms):
            x = [torch.from_numpy(rng.uniform(-np.pi, np.pi, M).astype(rdt)).to(DEV) for _ in range(2)]
            shp = (2, M) if kind == 1 else (2,) + tuple(reversed(nm))
            d = torch.from_numpy((rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(cdt)).to(DEV)
            p = Plan.transform(kind, 2, nm, 2, 1e-6 if prec == F32 else 1e-12, 1, 2.0, prec, 0, torch.device(DEV))
            p.setpts(x)
            oshape = (2,) + tuple(reversed(nm)) if kind == 1 else (2, M)
            o = torch.empty(oshape, dtype=d.dtype, device=DEV)
            if kind == 1:
                p.execute(d, o)
            else:
                p.execute(o, d)
            torch.cuda.synchronize()
            alone.append(o.cpu().numpy())
            plans.append(p)
            ins.append(d)
        outs = [torch.empty_like(torch.from_numpy(a)).to(DEV) for a in alone]
        if kind == 1:
            Plan.execute_many(plans, ins, outs)
        else:
            Plan.execute_many(plans, outs, ins)
        torch.cuda.synchronize()
        for a, o in zip(alone, outs):
            assert R.rel_l2(o.cpu().numpy(), a) < tol
