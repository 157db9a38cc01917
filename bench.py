"""Benchmark: NU points/s of 2D type-1 + type-2 NUFFT on one MI355X.

Workload (the 2D radial configuration): M = 512 spokes x 7,812 samples = 3,999,744
points, angles U[0, pi), radius U[-pi, pi), Gaussian jitter 1e-3, wrapped;
complex64 strengths / modes N(0,1); 512 x 512 modes (1024 x 1024 fine grid),
eps = 1e-6 (ns = 8), upsampfac 2.

One step = nufft2d1 (bin-sort + spread + FFT + deconvolve) followed by
nufft2d2 (bin-sort + deconvolve/pad + FFT + interp) on the same points, inputs
resident in HBM.  value = 2 * M / step time (NU points per second over both
transforms).  The roofline object is for the dominant stage, the spreading of the
type-1 transform (its two launches: the pass that packs the points in sorted order and
the matrix-core window spread), timed with HIP events on the plan stream.

Every reported figure carries its accuracy: relative l2 error against an FP64 direct
NUDFT (evaluated in FP64 on the GPU by plain torch arithmetic) on seeded subsets of
modes / points, next to the reference-error proxy where one is computed (the complex64
restatement of the reference's algorithm on the same full input).

Secondary figures on the same JSON line (not part of value): config_1d (1D type 1),
config_3d (3D clustered type 1 / type 2, complex64, eps 1e-6 and 1e-4), config_extra (2D
gradient round trip, 3D complex128 at eps 1e-12, 3D type 3), serving (the mixed request
stream through the server against per-call dispatch, synchronous and asynchronous), and
cpu_baseline (the reference algorithm on the host cores, run before the GPU is touched).
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
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "nufft2d_radial_4M_traffic.json")


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D secondary figures")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 1D, gradient round trip, 3D complex128 and 3D type-3 figures")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving stream")
    ap.add_argument("--cpu-cores", type=int, default=0, help="CPU-baseline worker processes (0: host cores, <= 16)")
    ap.add_argument("--cpu-runs", type=int, default=2, help="timed CPU-baseline runs (median reported)")
    ap.add_argument("--serving-requests", type=int, default=200)
    ap.add_argument("--profile-only", action="store_true", help="run a few steps, print nothing but a status line")
    return ap.parse_args()


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


# ---------------------------------------------------------------- CPU baseline (no GPU use)

_W = {}  # worker state (fork inherits it)


def _w_spread(bounds):
    from oracle import nufft_ref as R

    s, e = bounds
    kp, nfs = _W["kp"], _W["nfs"]
    return R.spread([R.wrap_pi(_W["x"][s:e]), R.wrap_pi(_W["y"][s:e])], _W["c"][s:e], nfs, kp)


def _w_interp(bounds):
    from oracle import nufft_ref as R

    s, e = bounds
    return R.interp([R.wrap_pi(_W["x"][s:e]), R.wrap_pi(_W["y"][s:e])], _W["fw"], _W["kp"])


def cpu_baseline(x, y, c, f, nm, eps, cores, runs, modes, idx):
    """The oracle -- the NumPy restatement of the reference's pure-JAX algorithm
    (materialised weights, segment-sum scatter, gather-reduce, same FFT and deconvolution),
    complex64 like the reference -- on the full workload, the points split over `cores`
    worker processes (spreading and interpolation are linear in the points: only the
    summation order changes).  Also returns its outputs at the accuracy subsets: the
    reference-error proxy."""
    import multiprocessing as mp

    from oracle import nufft_ref as R

    kp = R.compute_kernel_params(eps)
    nfs = [R.compute_grid_size(n, 2.0, kp.nspread) for n in nm]
    phihats = [R.kernel_fourier_series(nf, kp.nspread, kp.beta, kp.c, dtype=np.float32) for nf in nfs]
    M = x.size
    # the interpolated grid depends on f only: the workers get it through fork (the parent
    # recomputes it inside every timed run)
    _W.update(x=x, y=y, c=c, kp=kp, nfs=nfs, fw=R.fft_isign(R.deconvolve_pad(f, phihats, nfs), 2, -1))
    cuts = np.linspace(0, M, 4 * cores + 1).astype(np.int64)
    chunks = [(int(a), int(b)) for a, b in zip(cuts[:-1], cuts[1:])]
    ctx = mp.get_context("fork")  # the GPU is not initialised yet in this process
    times, F, C = [], None, None
    with ctx.Pool(cores) as pool:
        pool.map(_w_spread, [(0, 1000)] * cores)  # warm the workers
        for _ in range(max(1, runs)):
            t0 = time.perf_counter()
            fw = np.zeros(tuple(reversed(nfs)), dtype=np.complex128)
            for part in pool.imap_unordered(_w_spread, chunks):
                fw += part
            F = R.deconvolve_shuffle(R.fft_isign(fw.astype(np.complex64), 2, 1), phihats, nm)
            fwi = R.fft_isign(R.deconvolve_pad(f, phihats, nfs), 2, -1)
            assert np.array_equal(fwi, _W["fw"])
            C = np.concatenate(pool.map(_w_interp, chunks))
            times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    got1 = np.array([F[k2 + nm[1] // 2, k1 + nm[0] // 2] for k1, k2 in modes])
    return {
        "value": 2 * M / t,
        "unit": "NU pts/s",
        "cores": cores,
        "kind": "port",
        "sample": f"the full workload: nufft2d1 + nufft2d2 on all {M} radial points (512x512 modes, complex64), "
                  f"points split over {cores} processes, median of {len(times)} runs, {t:.2f} s each",
    }, got1, C[idx]


# ---------------------------------------------------------------- FP64 direct sums on the GPU


def direct_modes(torch, dev, coords, c, modes, isign, chunk=1 << 22):
    """sum_j c_j exp(i isign k . x_j) for mode tuples k (FP64)."""
    xs = [torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev) for x in coords]
    cd = torch.as_tensor(np.asarray(c, dtype=np.complex128), device=dev)
    out = torch.zeros(len(modes), dtype=torch.complex128, device=dev)
    for s in range(0, cd.shape[0], chunk):
        e = min(cd.shape[0], s + chunk)
        for i, k in enumerate(modes):
            ph = sum(float(kd) * x[s:e] for kd, x in zip(k, xs)) * isign
            out[i] += torch.sum(cd[s:e] * torch.polar(torch.ones_like(ph), ph))
    return out.cpu().numpy()


def direct_points(torch, dev, coords, f, idx, isign):
    """sum_k f_k exp(i isign k . x_j) at points idx (f is (n_d, .., n_1), CMCL), FP64."""
    dim = len(coords)
    fd = torch.as_tensor(np.asarray(f, dtype=np.complex128), device=dev) if isinstance(f, np.ndarray) else \
        f.to(torch.complex128)
    shape = tuple(fd.shape)
    ks = [torch.arange(-(n // 2), (n + 1) // 2, device=dev, dtype=torch.float64) for n in reversed(shape)]
    out = []
    for j in idx:
        ph = torch.zeros(shape, dtype=torch.float64, device=dev)
        for d in range(dim):
            sh = [1] * dim
            sh[dim - 1 - d] = shape[dim - 1 - d]
            ph = ph + ks[d].reshape(sh) * float(coords[d][j])
        out.append(torch.sum(fd * torch.polar(torch.ones_like(ph), ph * isign)))
    return torch.stack(out).cpu().numpy()


def pick(out, modes, nm):
    return np.array([out[tuple(k[d] + nm[d] // 2 for d in reversed(range(len(nm))))] for k in modes])


def rand_modes(rng, nm, n=64):
    return [tuple(int(rng.integers(-(m // 2), (m + 1) // 2)) for m in nm) for _ in range(n)]


def timeit(torch, fn, n=5):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / n


# ---------------------------------------------------------------- secondary workloads


def bench_1d(torch, dev):
    """1D type 1: M = 100,000 points U[-pi, pi), 1024 modes, complex64, eps 1e-6."""
    import nufftax_amd as N
    from nufftax_amd import workloads

    x, c = workloads.uniform_1d_type1(seed=0)
    x32, c64 = x.astype(np.float32), c.astype(np.complex64)
    xd, cd = torch.from_numpy(x32).to(dev), torch.from_numpy(c64).to(dev)
    t = timeit(torch, lambda: N.nufft1d1(xd, cd, 1024, eps=1e-6), 100)  # ~50 us a call: 100 calls
    out = N.nufft1d1(xd, cd, 1024, eps=1e-6).cpu().numpy()
    modes = [(k,) for k in range(-512, -448)]  # output indices 0..63
    truth = direct_modes(torch, dev, [x32.astype(np.float64)], c64, modes, 1)
    return {"workload": "1D uniform type 1", "M": x.size, "n_modes": 1024, "eps": 1e-6,
            "nufft1d1_ms": t * 1e3, "value": x.size / t, "unit": "NU pts/s",
            "accuracy": {"nufft1d1_rel_l2_vs_direct_first_64_modes": rel_l2(out[:64], truth)}}


def bench_3d(torch, dev):
    """3D type 1 + type 2 on 8M points in 16 Gaussian clusters, 128^3 modes, complex64, eps
    1e-6 and 1e-4: ms per transform (points sorted in every call), NU pts/s, accuracy."""
    from nufftax_amd import workloads
    from nufftax_amd._lib import check, lib, ptr
    from nufftax_amd._plan import F32, transform_plan

    coords, c = workloads.gaussian_mixture_3d(seed=0)
    M = coords[0].size
    c32 = [x.astype(np.float32) for x in coords]
    xs = [torch.from_numpy(x).to(dev) for x in c32]
    c64 = c.astype(np.complex64)
    cd = torch.from_numpy(c64).to(dev)
    nm = (128, 128, 128)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    fd = torch.randn(nm, dtype=torch.complex64, device=dev, generator=g)
    fout = torch.empty(nm, dtype=torch.complex64, device=dev)
    cout = torch.empty(M, dtype=torch.complex64, device=dev)
    L = lib()
    rng = np.random.default_rng(3)
    modes = rand_modes(rng, nm)
    idx = rng.integers(0, M, 64)
    c64d = [x.astype(np.float64) for x in c32]
    truth1 = direct_modes(torch, dev, c64d, c64, modes, 1)
    truth2 = direct_points(torch, dev, c64d, fd, idx, -1)
    out = {"workload": "3D Gaussian clusters type-1 + type-2", "M": M, "n_modes": list(nm)}
    for eps in (1e-6, 1e-4):
        p1 = transform_plan(1, 3, nm, 1, eps, 1, 2.0, F32, dev)
        p2 = transform_plan(2, 3, nm, 1, eps, -1, 2.0, F32, dev)

        def run(p, a, b):
            p._sync_stream()
            check(L.nufftx_setpts(p.h, M, ptr(xs[0]), ptr(xs[1]), ptr(xs[2]), 0, None, None, None), "setpts")
            check(L.nufftx_execute(p.h, ptr(a), ptr(b)), "execute")

        t1 = timeit(torch, lambda: run(p1, cd, fout), 10)
        t2 = timeit(torch, lambda: run(p2, cout, fd), 10)
        out[f"eps_{eps:g}"] = {
            "nufft3d1_ms": t1 * 1e3, "nufft3d2_ms": t2 * 1e3, "value": 2 * M / (t1 + t2), "unit": "NU pts/s",
            "accuracy": {"nufft3d1_rel_l2_vs_direct_64_modes": rel_l2(pick(fout.cpu().numpy(), modes, nm), truth1),
                         "nufft3d2_rel_l2_vs_direct_64_points": rel_l2(cout.cpu().numpy()[idx], truth2)}}
    return out


def bench_extra(torch, dev):
    """The remaining measured workloads, through the public transforms (points sorted in
    every call), ms per call (mean of 5 after one untimed call), each with its accuracy:
    * 2D round trip with gradients: loss = sum |nufft2d1(x, y, nufft2d2(x, y, f)) - g|^2 on 1M
      uniform points, 256^2 modes, complex64, eps 1e-6; forward + backward w.r.t. f, x, y;
      accuracy = rel l2 of each gradient against the same round trip in complex128 at eps 1e-12;
    * 3D complex128, eps 1e-12: 500k clustered points, 64^3 modes, type 1 and type 2;
    * 3D type 3: 200k sources, 200k targets in [-64, 64]^3, complex128, eps 1e-9."""
    import nufftax_amd as N
    from nufftax_amd import workloads

    out = {}
    (x, y), _ = workloads.uniform_nd(seed=0, M=1_000_000, dim=2)
    rng = np.random.default_rng(2)
    f0 = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))
    g0 = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))

    def leaves(cdt, rdt):
        return (torch.from_numpy(x).to(dev, rdt).requires_grad_(True),
                torch.from_numpy(y).to(dev, rdt).requires_grad_(True),
                torch.from_numpy(f0).to(dev, cdt).requires_grad_(True), torch.from_numpy(g0).to(dev, cdt))

    def roundtrip(ls, eps):
        xs, ys, f, g = ls
        for t in (xs, ys, f):
            t.grad = None
        o = N.nufft2d1(xs, ys, N.nufft2d2(xs, ys, f, eps=eps), (256, 256), eps=eps)
        torch.sum(torch.abs(o - g) ** 2).backward()
        return [t.grad for t in (f, xs, ys)]

    l32 = leaves(torch.complex64, torch.float32)
    t_rt = timeit(torch, lambda: roundtrip(l32, 1e-6))
    g32 = [v.detach().cpu().numpy() for v in roundtrip(l32, 1e-6)]
    g64 = [v.detach().cpu().numpy() for v in roundtrip(leaves(torch.complex128, torch.float64), 1e-12)]
    out["roundtrip_2d_grad_1M_256sq_c64"] = {
        "forward_backward_ms": t_rt * 1e3,
        "accuracy": {f"grad_{n}_rel_l2_vs_complex128": rel_l2(a, b) for n, a, b in zip(("f", "x", "y"), g32, g64)}}

    coords, c = workloads.gaussian_mixture_3d(seed=1, M=500_000)
    xd = [torch.from_numpy(v).to(dev) for v in coords]
    cd = torch.from_numpy(c).to(dev)
    f3 = rng.standard_normal((64, 64, 64)) + 1j * rng.standard_normal((64, 64, 64))
    fd = torch.from_numpy(f3).to(dev)
    modes = rand_modes(rng, (64, 64, 64))
    idx = rng.integers(0, coords[0].size, 64)
    o1 = N.nufft3d1(*xd, cd, (64, 64, 64), eps=1e-12).cpu().numpy()
    o2 = N.nufft3d2(*xd, fd, eps=1e-12).cpu().numpy()
    out["clusters_3d_500k_64cube_c128_eps1e-12"] = {
        "nufft3d1_ms": timeit(torch, lambda: N.nufft3d1(*xd, cd, (64, 64, 64), eps=1e-12)) * 1e3,
        "nufft3d2_ms": timeit(torch, lambda: N.nufft3d2(*xd, fd, eps=1e-12)) * 1e3,
        "accuracy": {
            "nufft3d1_rel_l2_vs_direct_64_modes": rel_l2(pick(o1, modes, (64, 64, 64)),
                                                         direct_modes(torch, dev, coords, c, modes, 1)),
            "nufft3d2_rel_l2_vs_direct_64_points": rel_l2(o2[idx], direct_points(torch, dev, coords, f3, idx, -1))}}

    src, tgt, c3 = workloads.type3_3d(seed=0)
    nf = tuple(N.compute_type3_grid_size(src[d], tgt[d], eps=1e-9) for d in range(3))
    sd = [torch.from_numpy(v).to(dev) for v in src]
    td = [torch.from_numpy(v).to(dev) for v in tgt]
    c3d = torch.from_numpy(c3).to(dev)
    o3 = N.nufft3d3(*sd, c3d, *td, nf, eps=1e-9).cpu().numpy()
    jt = rng.integers(0, tgt[0].size, 64)
    truth3 = direct_modes(torch, dev, src, c3, [tuple(float(s[j]) for s in tgt) for j in jt], 1)
    out["type3_3d_200k_c128_eps1e-9"] = {
        "fine_grid": list(nf), "nufft3d3_ms": timeit(torch, lambda: N.nufft3d3(*sd, c3d, *td, nf, eps=1e-9)) * 1e3,
        "accuracy": {"nufft3d3_rel_l2_vs_direct_64_targets": rel_l2(o3[jt], truth3)}}
    return out


def bench_serving(torch, dev, n_requests):
    """The mixed request stream (70% 2D, 30% 3D, type 1 / type 2, uneven M and grids,
    complex64, eps 1e-6; nufftax_amd.serving.mixed_requests) replayed three ways: through
    the server (4 streams), per-call dispatch synchronised after every call, and per-call
    dispatch issued back to back on one stream (asynchronous, completion from a per-call
    event).  Saturation: every request arrives at once; fixed rate: Poisson arrivals."""
    from nufftax_amd import serving as S

    reqs = S.mixed_requests(n_requests, 0, dev)
    seen = set()
    for kind, coords, data, nm in reqs:  # plan caches of the per-call paths
        if (kind, len(coords), nm) not in seen:
            seen.add((kind, len(coords), nm))
            S._direct_call(kind, coords, data, nm)
    torch.cuda.synchronize()
    svc = S.service_times(reqs)
    out = {"workload": "mixed 2D/3D type 1/2 stream, complex64, eps 1e-6", "requests": n_requests,
           "service_alone_ms": {k: svc[k] for k in ("p50_ms", "p99_ms", "mean_ms")}}
    keys = ("throughput_rps", "p50_ms", "p99_ms", "mean_ms")
    for label, qps in (("saturation", 0.0), ("qps_1000", 1000.0)):
        res = {}
        for mode in ("server", "per-call", "per-call-async"):
            st = S.run_stream(reqs, qps, 0, mode)
            res[mode.replace("-", "_")] = {k: st[k] for k in keys}
        res["server_vs_per_call_async_throughput"] = res["server"]["throughput_rps"] / \
            res["per_call_async"]["throughput_rps"]
        out[label] = res
    return out


# ---------------------------------------------------------------- main


def main():
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU runs are not part of this "
                                   "build (--gpus must be 1)"}))
        sys.exit(2)
    from nufftax_amd import workloads

    nm = (512, 512)
    eps = 1e-6
    x, y, c, f = workloads.radial_2d(seed=0, n_modes=nm)
    M = x.size
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    c64 = c.astype(np.complex64)
    f64 = f.astype(np.complex64)
    rng = np.random.default_rng(1)
    modes = rand_modes(rng, nm)
    idx = rng.integers(0, M, 64)

    cpu = None
    if not args.no_cpu_baseline and not args.profile_only:
        cores = args.cpu_cores or min(len(os.sched_getaffinity(0)), 16)
        cpu = cpu_baseline(x32, y32, c64, f64, nm, eps, cores, args.cpu_runs, modes, idx)

    import torch

    from nufftax_amd._lib import check, lib, ptr
    from nufftax_amd._plan import F32, transform_plan

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
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

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps

    # per-transform and resident-points breakdown
    t_t1 = timeit(torch, lambda: (setpts(p1), check(L.nufftx_execute(p1.h, ptr(cd), ptr(fout)), "t1")), 10)
    t_t2 = timeit(torch, lambda: (setpts(p2), check(L.nufftx_execute(p2.h, ptr(cout), ptr(fd)), "t2")), 10)
    setpts(p1)
    setpts(p2)
    t_t1_res = timeit(torch, lambda: check(L.nufftx_execute(p1.h, ptr(cd), ptr(fout)), "t1"), 10)
    t_t2_res = timeit(torch, lambda: check(L.nufftx_execute(p2.h, ptr(cout), ptr(fd)), "t2"), 10)
    t_sort = timeit(torch, lambda: setpts(p1), 10)

    # dominant stage timed with HIP events on the plan stream: the spreading of the type-1
    # plan (nufftx_spread_add runs exactly the launches of nufftx_execute's spreading, on the
    # plan's sorted points) and the interpolation of the type-2 plan
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
    win = os.environ.get("NUFFTX_SPREAD_WIN", "1") != "0"
    if win:
        # pack: coordinates, key, slot, strength read (4 + 4 + 4 + 4 + 8 B), 16-B record written;
        # window spread: the 16-B records read, the grid written once
        f32win = os.environ.get("NUFFTX_WIN_F32", "0") == "1"
        poly = os.environ.get("NUFFTX_ES_POLY", "1") != "0"  # polynomial kernel weights (ns = 8)
        kname = "k_pack2d<float> + " + ("k_spread_win2d<8, true>" if f32win else
                                        "k_spread_win2dh<8, %s>" % ("true" if poly else "false"))
        bytes_spread = M * (24 + 16) + M * 16 + 8 * nfc
    else:
        kname = "k_absmax + k_spread_banked<float, 2, 8, false, 16, false>"
        bytes_spread = M * 8 + M * (4 + 8 + 8) + 8 * nfc
    traffic = None
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as fh:
            tj = json.load(fh)
        ks = [k.strip() for k in kname.split("+")]
        if all(k in tj.get("kernels", {}) for k in ks):
            fac_f, fac_w = tj["calibration"]["fetch_factor"], tj["calibration"]["write_factor"]
            traffic = sum((tj["kernels"][k]["fetch_kib"] * fac_f + tj["kernels"][k]["write_kib"] * fac_w) * 1024
                          for k in ks)
    achieved = bytes_spread / t_spread / 1e9

    # accuracy against the FP64 direct NUDFT on subsets (and the reference-error proxy)
    setpts(p1)
    check(L.nufftx_execute(p1.h, ptr(cd), ptr(fout)), "t1")
    got1 = pick(fout.cpu().numpy(), modes, nm)
    coords64 = [x32.astype(np.float64), y32.astype(np.float64)]
    ref1 = direct_modes(torch, dev, coords64, c64, modes, 1)
    setpts(p2)
    check(L.nufftx_execute(p2.h, ptr(cout), ptr(fd)), "t2")
    got2 = cout.cpu().numpy()[idx]
    ref2 = direct_points(torch, dev, coords64, f64, idx, -1)
    accuracy = {"nufft2d1_rel_l2_vs_direct_64_modes": rel_l2(got1, ref1),
                "nufft2d2_rel_l2_vs_direct_64_points": rel_l2(got2, ref2)}
    if cpu is not None:
        accuracy["reference_error_proxy"] = {
            "nufft2d1": rel_l2(cpu[1], ref1), "nufft2d2": rel_l2(cpu[2], ref2),
            "what": "complex64 restatement of the reference algorithm on the full input, same subsets"}

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
                         "spread_stage": t_spread * 1e3, "interp_kernel": t_interp * 1e3},
        "accuracy": accuracy,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes": bytes_spread,
                     "kernel_ms": t_spread * 1e3},
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
    if not args.no_extra:
        res["config_1d"] = bench_1d(torch, dev)
    if not args.no_3d:
        res["config_3d"] = bench_3d(torch, dev)
    if not args.no_extra:
        res["config_extra"] = bench_extra(torch, dev)
    if not args.no_serving:
        res["serving"] = bench_serving(torch, dev, args.serving_requests)
    if cpu is not None:
        res["cpu_baseline"] = cpu[0]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
