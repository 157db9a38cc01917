"""Benchmark: NU points/s of 2D type-1 + type-2 NUFFT on one MI355X.

Workload (the 2D radial configuration): M = 512 spokes x 7,812 samples = 3,999,744
points, angles U[0, pi), radius U[-pi, pi), Gaussian jitter 1e-3, wrapped;
complex64 strengths / modes N(0,1); 512 x 512 modes (1024 x 1024 fine grid),
eps = 1e-6 (ns = 8), upsampfac 2.

One step = nufft2d1 (bin-sort + spread + FFT + deconvolve) followed by
nufft2d2 (bin-sort + deconvolve/pad + FFT + interp) on the same points, inputs
resident in HBM.  value = 2 * M / step time (NU points per second over both
transforms).  The roofline object is for the dominant kernel (spread or
interp), timed with HIP events on the plan stream.

Secondary figures on the same JSON line (not part of value): config_3d (3D
clustered type 1 / type 2, complex64, eps 1e-6 and 1e-4) and config_extra (the
2D gradient round trip, 3D complex128 at eps 1e-12, 3D type 3), ms per call.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes moved per launch of the dominant kernels on this workload, from two
# rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch) scaled by what
# the same counters read for a 256 MiB copy (FETCH_SIZE x2.0, WRITE_SIZE x1.0);
# the summary with the commands is profiles/nufft2d_radial_4M_hbm_traffic.md.
TRAFFIC_BYTES = {
    "k_spread_banked<float,2,8,false,16,false>": 499800.5 * 1024 * 2.0 + 78768.4 * 1024,  # 1.11 GB
    "k_gather_wide<float,2,8,true>": 51023.6 * 1024 * 2.0 + 123608.3 * 1024,           # 0.23 GB
}


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D secondary figures")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the gradient round trip, 3D complex128 and 3D type-3 figures")
    ap.add_argument("--cpu-sample", type=int, default=1_000_000, help="points in the CPU-baseline sample")
    ap.add_argument("--cpu-runs", type=int, default=7, help="timed CPU-baseline runs (median reported)")
    ap.add_argument("--profile-only", action="store_true", help="run a few steps, print nothing but a status line")
    return ap.parse_args()


def direct_subset_type1(x, y, c, modes, isign, torch, dev):
    """FP64 direct sum for a subset of (k1, k2) modes, on the GPU."""
    xd = torch.as_tensor(x, dtype=torch.float64, device=dev)
    yd = torch.as_tensor(y, dtype=torch.float64, device=dev)
    cd = torch.as_tensor(c, dtype=torch.complex128, device=dev)
    out = []
    for k1, k2 in modes:
        ph = isign * (k1 * xd + k2 * yd)
        out.append(torch.sum(cd * torch.polar(torch.ones_like(ph), ph)))
    return torch.stack(out).cpu().numpy()


def direct_subset_type2(x, y, f, idx, isign, torch, dev):
    n2, n1 = f.shape
    k1 = torch.arange(-(n1 // 2), (n1 + 1) // 2, device=dev, dtype=torch.float64)
    k2 = torch.arange(-(n2 // 2), (n2 + 1) // 2, device=dev, dtype=torch.float64)
    fd = torch.as_tensor(f, dtype=torch.complex128, device=dev)
    out = []
    for j in idx:
        ph = isign * (k2[:, None] * float(y[j]) + k1[None, :] * float(x[j]))
        out.append(torch.sum(fd * torch.polar(torch.ones_like(ph), ph)))
    return torch.stack(out).cpu().numpy()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cpu_baseline(x, y, c, f, nm, eps, sample, runs):
    """The oracle (NumPy restatement of the reference's pure-JAX algorithm, complex64 like
    the reference) on the first `sample` points of the same workload, 1 thread."""
    from oracle import nufft_ref as R

    xs, ys, cs = x[:sample], y[:sample], c[:sample]
    R.nufft_type1([xs[:1000], ys[:1000]], cs[:1000], nm, eps=eps)  # warm
    times = []
    for _ in range(max(1, runs)):
        t0 = time.perf_counter()
        R.nufft_type1([xs, ys], cs, nm, eps=eps, isign=1)
        R.nufft_type2([xs, ys], f, eps=eps, isign=-1)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {
        "value": 2 * sample / t,
        "unit": "NU pts/s",
        "cores": 1,
        "kind": "port",
        "sample": f"nufft2d1 + nufft2d2 on the first {sample} radial points (512x512 modes, complex64), "
                  f"median of {len(times)} runs, {t:.2f} s each",
    }


def bench_extra(torch, dev):
    """The remaining measured workloads, through the public transforms (plans cached, points
    sorted in every call), ms per call (mean of 5 after one untimed call):
    * 2D round trip with gradients: loss = sum |nufft2d1(x, y, nufft2d2(x, y, f)) - g|^2 on 1M
      uniform points, 256^2 modes, complex64, eps 1e-6; forward + backward w.r.t. f, x, y;
    * 3D complex128, eps 1e-12: 500k clustered points, 64^3 modes, type 1 and type 2;
    * 3D type 3: 200k sources, 200k targets in [-64, 64]^3, complex128, eps 1e-9."""
    import nufftax_amd as N
    from nufftax_amd import workloads

    def timeit(fn, n=5):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    out = {}
    (x, y), _ = workloads.uniform_nd(seed=0, M=1_000_000, dim=2)
    rng = np.random.default_rng(2)
    f0 = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))
    g = torch.from_numpy((rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))).astype(np.complex64)).to(dev)
    xs = torch.from_numpy(x.astype(np.float32)).to(dev).requires_grad_(True)
    ys = torch.from_numpy(y.astype(np.float32)).to(dev).requires_grad_(True)
    f = torch.from_numpy(f0.astype(np.complex64)).to(dev).requires_grad_(True)

    def roundtrip():
        for t in (xs, ys, f):
            t.grad = None
        o = N.nufft2d1(xs, ys, N.nufft2d2(xs, ys, f, eps=1e-6), (256, 256), eps=1e-6)
        torch.sum(torch.abs(o - g) ** 2).backward()

    out["roundtrip_2d_grad_1M_256sq_c64"] = {"forward_backward_ms": timeit(roundtrip)}

    coords, c = workloads.gaussian_mixture_3d(seed=1, M=500_000)
    xd = [torch.from_numpy(v).to(dev) for v in coords]
    cd = torch.from_numpy(c).to(dev)
    fd = torch.from_numpy((rng.standard_normal((64, 64, 64)) + 1j * rng.standard_normal((64, 64, 64)))).to(dev)
    out["clusters_3d_500k_64cube_c128_eps1e-12"] = {
        "nufft3d1_ms": timeit(lambda: N.nufft3d1(*xd, cd, (64, 64, 64), eps=1e-12)),
        "nufft3d2_ms": timeit(lambda: N.nufft3d2(*xd, fd, eps=1e-12)),
    }

    src, tgt, c3 = workloads.type3_3d(seed=0)
    nf = tuple(N.compute_type3_grid_size(src[d], tgt[d], eps=1e-9) for d in range(3))
    sd = [torch.from_numpy(v).to(dev) for v in src]
    td = [torch.from_numpy(v).to(dev) for v in tgt]
    c3d = torch.from_numpy(c3).to(dev)
    out["type3_3d_200k_c128_eps1e-9"] = {"fine_grid": list(nf),
                                         "nufft3d3_ms": timeit(lambda: N.nufft3d3(*sd, c3d, *td, nf, eps=1e-9))}
    return out


def bench_3d(args, torch, dev):
    """3D type 1 + type 2 on 8M points in 16 Gaussian clusters, 128^3 modes, complex64, eps
    1e-6 and 1e-4: ms per transform (points sorted in every call) and NU pts/s."""
    from nufftax_amd import workloads
    from nufftax_amd._lib import lib, ptr, check
    from nufftax_amd._plan import F32, transform_plan

    coords, c = workloads.gaussian_mixture_3d(seed=0)
    M = coords[0].size
    xs = [torch.from_numpy(x.astype(np.float32)).to(dev) for x in coords]
    cd = torch.from_numpy(c.astype(np.complex64)).to(dev)
    nm = (128, 128, 128)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    fd = torch.randn(nm, dtype=torch.complex64, device=dev, generator=g)
    fout = torch.empty(nm, dtype=torch.complex64, device=dev)
    cout = torch.empty(M, dtype=torch.complex64, device=dev)
    L = lib()
    out = {"workload": "3D Gaussian clusters type-1 + type-2", "M": M, "n_modes": list(nm)}
    for eps in (1e-6, 1e-4):
        p1 = transform_plan(1, 3, nm, 1, eps, 1, 2.0, F32, dev)
        p2 = transform_plan(2, 3, nm, 1, eps, -1, 2.0, F32, dev)

        def run(p, a, b):
            p._sync_stream()
            check(L.nufftx_setpts(p.h, M, ptr(xs[0]), ptr(xs[1]), ptr(xs[2]), 0, None, None, None), "setpts")
            check(L.nufftx_execute(p.h, ptr(a), ptr(b)), "execute")

        def timeit(fn, n=10):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(n):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / n

        t1 = timeit(lambda: run(p1, cd, fout))
        t2 = timeit(lambda: run(p2, cout, fd))
        out[f"eps_{eps:g}"] = {"nufft3d1_ms": t1 * 1e3, "nufft3d2_ms": t2 * 1e3, "value": 2 * M / (t1 + t2),
                               "unit": "NU pts/s"}
    return out


def main():
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU runs are not part of this "
                                   "build (--gpus must be 1)"}))
        sys.exit(2)
    import torch

    from nufftax_amd import workloads
    from nufftax_amd._lib import lib, ptr, check
    from nufftax_amd._plan import F32, Plan, transform_plan, grid_plan

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    nm = (512, 512)
    eps = 1e-6
    x, y, c, f = workloads.radial_2d(seed=0, n_modes=nm)
    M = x.size
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    c64 = c.astype(np.complex64)
    f64 = f.astype(np.complex64)
    xd = torch.from_numpy(x32).to(dev)
    yd = torch.from_numpy(y32).to(dev)
    cd = torch.from_numpy(c64).to(dev)
    fd = torch.from_numpy(f64).to(dev)
    fout = torch.empty((nm[1], nm[0]), dtype=torch.complex64, device=dev)
    cout = torch.empty(M, dtype=torch.complex64, device=dev)

    p1 = transform_plan(1, 2, nm, 1, eps, 1, 2.0, F32, dev)
    p2 = transform_plan(2, 2, nm, 1, eps, -1, 2.0, F32, dev)
    L = lib()

    def setpts(p):
        p._sync_stream()
        check(L.nufftx_setpts(p.h, M, ptr(xd), ptr(yd), None, 0, None, None, None), "setpts")

    def step():
        setpts(p1)
        check(L.nufftx_execute(p1.h, ptr(cd), ptr(fout)), "execute t1")
        setpts(p2)
        check(L.nufftx_execute(p2.h, ptr(cout), ptr(fd)), "execute t2")

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    if args.profile_only:
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        print("profile-only run done", flush=True)
        return

    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps

    # per-transform and resident-points breakdown
    def timeit(fn, n=10):
        fn()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / n

    t_t1 = timeit(lambda: (setpts(p1), check(L.nufftx_execute(p1.h, ptr(cd), ptr(fout)), "t1")))
    t_t2 = timeit(lambda: (setpts(p2), check(L.nufftx_execute(p2.h, ptr(cout), ptr(fd)), "t2")))
    setpts(p1)
    setpts(p2)
    t_t1_res = timeit(lambda: check(L.nufftx_execute(p1.h, ptr(cd), ptr(fout)), "t1"))
    t_t2_res = timeit(lambda: check(L.nufftx_execute(p2.h, ptr(cout), ptr(fd)), "t2"))
    t_sort = timeit(lambda: setpts(p1))

    # dominant-kernel timing with HIP events on the plan stream: the spread stage of the
    # type-1 plan and the interpolation stage of the type-2 plan (nufftx_spread_add /
    # nufftx_interp run exactly the kernels of nufftx_execute on the plan's sorted points)
    from nufftax_amd._lib import kernel_params
    ns, beta, kc = kernel_params(eps, 2.0)
    nf = (p1.grid_size(0), p1.grid_size(1))
    setpts(p1)
    setpts(p2)
    grid = torch.zeros((nf[1], nf[0]), dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream(dev)
    reps = 20

    def ev_time(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    t_spread = ev_time(lambda: check(L.nufftx_spread_add(p1.h, ptr(cd), ptr(grid)), "spread"))
    t_interp = ev_time(lambda: check(L.nufftx_interp(p2.h, ptr(grid), ptr(cout)), "interp"))
    nfc = nf[0] * nf[1]
    bytes_spread = M * (4 + 8 + 8) + 8 * nfc   # perm, x/y, c gathered; grid written once
    bytes_interp = M * (4 + 8 + 8) + 8 * nfc   # perm, x/y, c written; grid read once
    # 2D FP32 ns <= 8 runs on the matrix cores when NUFFTX_MFMA is set, else on the LDS-tiled kernels
    mfma = os.environ.get("NUFFTX_MFMA") is not None and os.environ.get("NUFFTX_NO_MFMA") is None
    if t_spread >= t_interp:
        kname = "k_spread_mfma2d<8>" if mfma else "k_spread_banked<float,2,8,false,16,false>"
        kt, kb = t_spread, bytes_spread
    else:
        kname = "k_interp_mfma2d<8>" if mfma else "k_gather_wide<float,2,8,true>"
        kt, kb = t_interp, bytes_interp
    achieved = kb / kt / 1e9
    traffic = None if mfma else TRAFFIC_BYTES.get(kname)

    # accuracy against the FP64 direct NUDFT on subsets
    rng = np.random.default_rng(1)
    modes = [(int(a), int(b)) for a, b in zip(rng.integers(-256, 256, 64), rng.integers(-256, 256, 64))]
    setpts(p1)
    check(L.nufftx_execute(p1.h, ptr(cd), ptr(fout)), "t1")
    fo = fout.cpu().numpy()
    got1 = np.array([fo[k2 + 256, k1 + 256] for k1, k2 in modes])
    ref1 = direct_subset_type1(x32, y32, c64, modes, 1, torch, dev)
    idx = rng.integers(0, M, 64)
    setpts(p2)
    check(L.nufftx_execute(p2.h, ptr(cout), ptr(fd)), "t2")
    got2 = cout.cpu().numpy()[idx]
    ref2 = direct_subset_type2(x32, y32, f64, idx, -1, torch, dev)

    res = {
        "metric": "NU pts/s (2D type 1 + type 2, eps 1e-6)",
        "value": 2 * M / t_step,
        "unit": "NU pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "2D radial type-1 + type-2", "M": M, "n_modes": list(nm), "fine_grid": list(nf),
                   "eps": eps, "nspread": ns, "upsampfac": 2.0},
        "breakdown_ms": {"nufft2d1": t_t1 * 1e3, "nufft2d2": t_t2 * 1e3, "nufft2d1_resident_points": t_t1_res * 1e3,
                         "nufft2d2_resident_points": t_t2_res * 1e3, "bin_sort": t_sort * 1e3,
                         "spread_kernel": t_spread * 1e3, "interp_kernel": t_interp * 1e3},
        "accuracy": {"nufft2d1_rel_l2_vs_direct_64_modes": rel_l2(got1, ref1),
                     "nufft2d2_rel_l2_vs_direct_64_points": rel_l2(got2, ref2)},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes": kb,
                     "kernel_ms": kt * 1e3},
    }
    # whole-transform roofline: minimal HBM bytes of each transform (sort, spread/interp, FFT,
    # deconvolution) over its measured time
    nmc = nm[0] * nm[1]
    b1 = M * (2 * 2 * 4 + 8) + M * 8 + 3 * 8 * nfc + 2 * 8 * nmc
    b2 = M * (2 * 2 * 4 + 8) + M * 8 + 4 * 8 * nfc + 8 * nmc
    res["transform_roofline"] = {
        "nufft2d1": {"bytes": b1, "ms": t_t1 * 1e3, "frac": b1 / t_t1 / (HBM_PEAK_GBS * 1e9)},
        "nufft2d2": {"bytes": b2, "ms": t_t2 * 1e3, "frac": b2 / t_t2 / (HBM_PEAK_GBS * 1e9)},
    }
    if not args.no_3d:
        res["config_3d"] = bench_3d(args, torch, dev)
    if not args.no_extra:
        res["config_extra"] = bench_extra(torch, dev)
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(x32, y32, c64, f64, nm, eps, args.cpu_sample, args.cpu_runs)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
