"""Serving loop: many independent NUFFT requests on one GPU.

There is no counterpart in the reference (its nearest relative is the n_trans
batching of ``nufftax/core/spread.py:100-173`` and the batcher fast path of
``nufftax/transforms/primitives.py:41-42``).  Design:

* requests are type-1 or type-2 transforms (1D/2D/3D, complex64/complex128) whose
  inputs already live on the GPU; ``submit`` returns a ``concurrent.futures.Future``;
* a dispatcher thread admits requests against an HBM budget (points, strengths,
  fine grid and output of every request in flight) without blocking on one that does not
  fit while later ones do (those go ahead of it, a bounded number of times), and batches
  what is queued:
  requests with the same plan signature that share their point set become one
  n_trans unit; with ``max_units`` > 1 (default 8) and at least ``batch_min_queue`` requests
  queued, up to ``max_units`` units of one kind, dimension, tolerance, sign and precision with
  distinct grid sizes (any points, M up to ``batch_max_points``, eps below 1e-4) run through
  ``nufftx_execute_many``, which gives them one shared spreading (type 1) or
  interpolation (type 2) launch; each batch is issued on one of ``n_streams`` HIP
  streams (round robin),
  each with its own pool of plans, so that batches overlap on the GPU;
* a completion thread waits on each launch's HIP event in issue order per
  stream, resolves the futures and releases the budget;
* every request records its arrival, start and completion time (host clock),
  from which ``latency_stats`` reports p50 / p99; ``future.request`` is the
  ``ServeRequest``.

The launch path is ``nufftx_setpts`` per unit + ``nufftx_execute`` (one unit, the same
kernels as a direct call) or ``nufftx_execute_many`` (shared launches: the LDS-atomic
spread and the tiled interpolation over all units, whose results equal the same plans
called one by one to FP32 rounding -- relative l2 within 1e-6 --, not bit for bit: a 2D
plan called alone spreads with the matrix-core window kernel, and tiles sum in a different
fixed point).
"""

from __future__ import annotations

import math
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._plan import F32, F64, Plan
from .params import compute_grid_size, compute_kernel_params

__all__ = ["NufftServer", "ServeRequest", "latency_stats", "request_bytes"]


@dataclass
class ServeRequest:
    kind: int                    # 1 or 2
    coords: tuple                # (x, y, z)[:dim] device tensors, float32 / float64
    data: torch.Tensor           # type 1: c (n_trans?, M); type 2: f (n_trans?, n_d, .., n_1)
    n_modes: tuple               # (n1, .., nd)
    eps: float
    isign: int
    future: Future = field(default_factory=Future)
    ready: object = None         # HIP event recorded on the submitter's stream
    consumer: object = None      # the submitter's stream, which reads the result
    stream_index: int | None = None  # pin to one of the server's streams (plan warm-up)
    nbytes: int = 0
    t_submit: float = 0.0
    t_start: float = 0.0         # taken by the dispatcher
    t_issued: float = 0.0        # all of its launches queued on the stream
    t_done: float = 0.0
    skipped: int = 0             # times a later request was admitted ahead of this one
    _sig: tuple = None           # signature(), computed once

    @property
    def dim(self):
        return len(self.coords)

    def signature(self):
        if self._sig is None:
            prec = F64 if (self.data.dtype == torch.complex128 or self.eps < 1e-7) else F32
            self._sig = (self.kind, self.dim, tuple(self.n_modes), float(self.eps), int(self.isign), prec)
        return self._sig

    def batch_key(self):
        """Requests with equal keys may share a spreading / interpolation launch."""
        kind, dim, _, eps, isign, prec = self.signature()
        return (kind, dim, eps, isign, prec)


def request_bytes(kind, dim, M, n_modes, eps, complex_bytes=8, n_trans=1):
    """HBM held by one request in flight: points, strengths, output, plan fine grid and sort state."""
    kp = compute_kernel_params(eps)
    nf = math.prod(compute_grid_size(n, 2.0, kp.nspread) for n in n_modes)
    nm = math.prod(n_modes)
    real = complex_bytes // 2
    return int(M * (dim * real + 4) + n_trans * (M + nm) * complex_bytes + n_trans * nf * complex_bytes)


class NufftServer:
    """Asynchronous NUFFT executor over ``n_streams`` HIP streams with HBM admission control."""

    def __init__(self, device="cuda:0", n_streams: int = 4, hbm_budget_bytes: int = 32 << 30, max_group: int = 64,
                 max_units: int = 8, batch_max_points: int = 200_000):
        if not torch.cuda.is_available():
            raise _lib.NufftxError("NufftServer needs an AMD GPU")
        self.device = torch.device(device)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(n_streams)]  # default priority
        self.budget = int(hbm_budget_bytes)
        self.max_group = int(max_group)
        # point sets per shared launch (library limit 8).  Default 1: on the mixed stream the
        # server is bound by its dispatcher thread's issue rate, and shared launches (fewer
        # kernels, one stream per batch) measured 5% below separate ones at saturation
        self.max_units = max(1, min(int(max_units), 8))
        self.shared_launches = 0
        # only requests this small share launches with other point sets: their kernels are
        # launch-latency bound; larger ones fill the GPU alone and keep the streams' overlap
        self.batch_max_points = int(batch_max_points)
        # queued requests looked at when forming one group: the scan runs on the dispatcher
        # thread, whose issue rate bounds the throughput of small requests
        self.scan_depth = 16
        self.batch_min_queue = 4  # queued requests before point sets share launches
        # admission does not block on a request that does not fit the budget while later ones
        # do: those are admitted ahead of it, at most this many times (then it waits first)
        self.max_skips = 64
        self.skipped_admissions = 0
        self._in_use = 0
        self._cv = threading.Condition()
        self._q: "queue.Queue[ServeRequest | None]" = queue.Queue()
        self._done_q: "queue.Queue" = queue.Queue()
        self._next_stream = 0
        self._plans = {}  # (signature, n_trans, stream) -> [Plan, ...], kept for the server's lifetime
        self.plans_built = 0
        self._closed = False
        self._dispatcher = threading.Thread(target=self._dispatch_loop, name="nufftx-dispatch", daemon=True)
        self._completer = threading.Thread(target=self._complete_loop, name="nufftx-complete", daemon=True)
        self._dispatcher.start()
        self._completer.start()

    # ------------------------------------------------------------------ API
    def submit(self, kind, coords, data, n_modes=None, eps=1e-6, isign=None, stream_index=None) -> Future:
        """Queue a type-1 (kind=1, data = strengths) or type-2 (kind=2, data = modes) transform.
        stream_index pins the request to one server stream (default: round robin)."""
        if self._closed:
            raise RuntimeError("server is closed")
        coords = tuple(coords)
        dim = len(coords)
        if kind == 1:
            if n_modes is None:
                raise ValueError("type 1 needs n_modes")
            n_modes = (int(n_modes),) * dim if isinstance(n_modes, int) else tuple(int(v) for v in n_modes)
            isign = 1 if isign is None else isign
        elif kind == 2:
            n_modes = tuple(reversed(data.shape[data.ndim - dim:]))
            isign = -1 if isign is None else isign
        else:
            raise ValueError("kind must be 1 or 2")
        for t in coords + (data,):
            if not (isinstance(t, torch.Tensor) and t.device == self.device):
                raise ValueError("serving inputs must be tensors on the server's device")
        M = coords[0].shape[-1]
        lead = data.shape[:-1] if kind == 1 else data.shape[: data.ndim - dim]
        n_trans = int(math.prod(lead)) if lead else 1
        req = ServeRequest(kind, coords, data, n_modes, float(eps), 1 if isign > 0 else -1)
        req.nbytes = request_bytes(kind, dim, M, n_modes, eps, data.element_size(), n_trans)
        req.stream_index = None if stream_index is None else int(stream_index) % len(self.streams)
        if req.nbytes > self.budget:
            raise ValueError(f"request needs {req.nbytes} B of HBM, over the server budget of {self.budget} B")
        req.consumer = torch.cuda.current_stream(self.device)
        req.ready = torch.cuda.Event()
        req.ready.record(req.consumer)
        req.future.request = req
        req.t_submit = time.perf_counter()
        self._q.put(req)
        return req.future

    def close(self):
        if self._closed:
            return
        self._closed = True
        self._q.put(None)
        self._dispatcher.join()
        self._done_q.put(None)
        self._completer.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ internals
    def _next_group(self, pending: list):
        """The next group to issue, its budget already taken: the group led by the oldest
        pending request if it fits the budget now, else the first group led by a later request
        (among the next ``scan_depth``) that fits, unless the oldest has been passed over
        ``max_skips`` times; waits for completions only when nothing fits.  Returns (units, the
        requests left pending, in arrival order, bytes taken)."""
        with self._cv:
            while True:
                head = pending[0]
                last = 1 if head.skipped >= self.max_skips else min(len(pending), self.scan_depth + 1)
                for start in range(last):
                    if start and pending[start].nbytes + self._in_use > self.budget:
                        continue
                    units, rest = self._take_group(pending[start:])
                    nbytes = sum(g.nbytes for u in units for g in u)
                    if self._in_use + nbytes <= self.budget:
                        self._in_use += nbytes
                        if start:
                            self.skipped_admissions += 1
                            for r in pending[:start]:
                                r.skipped += 1
                        return units, pending[:start] + rest, nbytes
                self._cv.wait()

    def _release(self, nbytes):
        with self._cv:
            self._in_use -= nbytes
            self._cv.notify_all()

    def _take_group(self, pending: list):
        """pending[0] plus the requests after it (at most ``scan_depth`` looked at) that can run
        with it: units of requests that share their point tensors and plan signature (one
        n_trans launch each), and up to ``max_units`` such units of one kind, dimension,
        tolerance, sign and precision (one shared spreading / interpolation launch).  Returns
        (units, the requests left pending, in arrival order)."""
        first = pending[0]
        units = [[first]]
        held = []
        total = first.nbytes
        n = 1
        # other point sets join only under load: with a short queue the shared launch only
        # delays the first request (measured: higher p50 at 1000 requests/s, no throughput gain)
        cross = len(pending) >= self.batch_min_queue and self._batchable(first)
        for idx in range(1, len(pending)):
            r = pending[idx]
            if n >= self.max_group or idx > self.scan_depth:
                held.extend(pending[idx:])
                break
            if r.stream_index != first.stream_index or total + r.nbytes > self.budget or \
                    r.batch_key() != first.batch_key():
                held.append(r)
                continue
            unit = next((u for u in units if r.signature() == u[0].signature() and r.data.shape == u[0].data.shape
                         and all(a is b for a, b in zip(r.coords, u[0].coords))), None)
            if unit is not None:
                unit.append(r)
            elif cross and len(units) < self.max_units and self._batchable(r) and \
                    all(r.signature() != u[0].signature() for u in units):
                # one unit per plan signature, so that every unit of a batch runs on the stream's
                # pooled plan of its signature (no plan is built for a second unit of one shape)
                units.append([r])
            else:
                held.append(r)
                continue
            total += r.nbytes
            n += 1
        return units, held

    def _batchable(self, r: ServeRequest) -> bool:
        # small point sets only, and tolerances below 1e-4: there a plan called alone also
        # spreads into 64-bit tiles (above it, into packed ones, which shared launches do not use)
        return r.coords[0].shape[-1] <= self.batch_max_points and r.eps < 1e-4

    def _dispatch_loop(self):
        pending: list = []  # queued requests in arrival order
        stopping = False
        while True:
            if not pending and not stopping:
                pending.append(self._q.get())
            while True:  # everything queued so far joins the pending list
                try:
                    pending.append(self._q.get_nowait())
                except queue.Empty:
                    break
            if None in pending:  # close(): finish what was submitted before it
                stopping = True
                pending = [p for p in pending if p is not None]
            if not pending:
                break
            units, pending, nbytes = self._next_group(pending)
            r = units[0][0]
            group = [g for u in units for g in u]
            if r.stream_index is not None:
                stream = self.streams[r.stream_index]
            else:
                stream = self.streams[self._next_stream]
                self._next_stream = (self._next_stream + 1) % len(self.streams)
            t0 = time.perf_counter()
            try:
                with torch.cuda.stream(stream):
                    for g in group:
                        stream.wait_event(g.ready)
                    outs = self._launch(units)
                    ev = torch.cuda.Event()
                    ev.record(stream)
            except Exception as e:  # noqa: BLE001 - reported through the futures
                self._release(nbytes)
                for g in group:
                    g.future.set_exception(e)
                continue
            t1 = time.perf_counter()
            for g in group:
                g.t_start, g.t_issued = t0, t1
            self._done_q.put((ev, group, outs, nbytes))

    def _plan(self, sig, n_trans, k):
        """The k-th pooled plan of this signature, n_trans and the current stream."""
        kind, dim, n_modes, eps, isign, prec = sig
        key = (sig, n_trans, torch.cuda.current_stream(self.device).cuda_stream)
        pool = self._plans.setdefault(key, [])
        while len(pool) <= k:
            pool.append(Plan.transform(kind, dim, n_modes, n_trans, eps, isign, 2.0, prec, 0, self.device))
            self.plans_built += 1
        return pool[k]

    def _launch(self, units):
        """Issue the units of one batch on the current stream; returns the outputs in request order."""
        kind = units[0][0].kind
        plans, srcs, dsts, parts = [], [], [], []
        used = {}
        for unit in units:
            first = unit[0]
            sig = first.signature()
            _, dim, n_modes, _, _, prec = sig
            cdt = torch.complex128 if prec == F64 else torch.complex64
            rdt = torch.float64 if prec == F64 else torch.float32
            coords = [x if (x.dtype == rdt and x.is_contiguous()) else x.to(rdt).contiguous() for x in first.coords]
            M = coords[0].shape[0]
            if kind == 1:
                datas = [g.data.reshape(-1, M) for g in unit]
            else:
                datas = [g.data.reshape((-1,) + tuple(reversed(n_modes))) for g in unit]
            sizes = [d.shape[0] for d in datas]
            src = torch.cat(datas, 0).to(cdt).contiguous() if len(datas) > 1 else datas[0].to(cdt).contiguous()
            n_trans = src.shape[0]
            k = used.get((sig, n_trans), 0)
            used[(sig, n_trans)] = k + 1
            plan = self._plan(sig, n_trans, k)
            plan.setpts(coords)
            if kind == 1:
                out = torch.empty((n_trans,) + tuple(reversed(n_modes)), dtype=cdt, device=self.device)
            else:
                out = torch.empty((n_trans, M), dtype=cdt, device=self.device)
            plans.append(plan)
            srcs.append(src)
            dsts.append(out)
            parts.append((unit, sizes, n_modes, M, dim))
        if len(plans) == 1:
            if kind == 1:
                plans[0].execute(srcs[0], dsts[0])
            else:
                plans[0].execute(dsts[0], srcs[0])
        elif kind == 1:
            self.shared_launches += 1
            Plan.execute_many(plans, srcs, dsts)
        else:
            self.shared_launches += 1
            Plan.execute_many(plans, dsts, srcs)
        outs = []
        for out, (unit, sizes, n_modes, M, dim) in zip(dsts, parts):
            off = 0
            for g, n in zip(unit, sizes):
                o = out[off:off + n]
                off += n
                lead = g.data.shape[:-1] if kind == 1 else g.data.shape[: g.data.ndim - dim]
                tail = tuple(reversed(n_modes)) if kind == 1 else (M,)
                outs.append(o.reshape(tuple(lead) + tail).to(g.data.dtype))
        return outs

    def _complete_loop(self):
        while True:
            item = self._done_q.get()
            if item is None:
                break
            ev, group, outs, nbytes = item
            ev.synchronize()
            t = time.perf_counter()
            self._release(nbytes)
            for g, o in zip(group, outs):
                g.t_done = t
                # the result was allocated on a server stream: tell the caching allocator that
                # the submitter's stream uses it, so that its memory is not handed to a later
                # batch while work the submitter queued on the result may still be reading it
                if g.consumer is not None:
                    o.record_stream(g.consumer)
                g.future.set_result(o)


def latency_stats(requests) -> dict:
    """p50 / p99 / mean latency (submit -> done) in ms, requests per second over the span, and
    the mean of each phase: queued (submit -> taken), issue (taken -> launches queued) and
    device (launches queued -> completion seen on the host)."""
    lat = np.array([(r.t_done - r.t_submit) * 1e3 for r in requests])
    span = max(r.t_done for r in requests) - min(r.t_submit for r in requests)
    mean = lambda v: float(np.mean(v)) * 1e3  # noqa: E731
    return {"n": len(lat), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "throughput_rps": len(lat) / span if span > 0 else float("inf"),
            "phase_mean_ms": {"queued": mean([r.t_start - r.t_submit for r in requests]),
                              "issue": mean([r.t_issued - r.t_start for r in requests]),
                              "device": mean([r.t_done - r.t_issued for r in requests])}}


# ---------------------------------------------------------------------- stream harness


def mixed_requests(n: int, seed: int = 0, device="cuda:0"):
    """The mixed serving workload: 70% 2D (n_modes 128^2 / 256^2 / 512^2, M log-uniform in
    [1e4, 1e6]) and 30% 3D (32^3 / 64^3 / 128^3, M log-uniform in [1e5, 4e6]); type 1 or 2 with
    probability 1/2; complex64, eps 1e-6.  Host draws (numpy default_rng(seed)), in order per
    request: dimension, modes, log M, kind, then a torch generator seed for the device data."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        two_d = rng.random() < 0.7
        if two_d:
            n1 = int(rng.choice([128, 256, 512]))
            nm, dim = (n1, n1), 2
            M = int(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        else:
            n1 = int(rng.choice([32, 64, 128]))
            nm, dim = (n1, n1, n1), 3
            M = int(np.exp(rng.uniform(np.log(1e5), np.log(4e6))))
        kind = 1 if rng.random() < 0.5 else 2
        g = torch.Generator(device=device)
        g.manual_seed(int(rng.integers(0, 2**31 - 1)))
        coords = tuple((torch.rand(M, device=device, generator=g) * 2 - 1) * math.pi for _ in range(dim))
        if kind == 1:
            data = torch.randn(M, dtype=torch.complex64, device=device, generator=g)
        else:
            data = torch.randn(tuple(reversed(nm)), dtype=torch.complex64, device=device, generator=g)
        out.append((kind, coords, data, nm))
    return out


def _direct_call(kind, coords, data, nm, eps=1e-6):
    from . import transforms as T

    dim = len(coords)
    if kind == 1:
        fn = [T.nufft1d1, T.nufft2d1, T.nufft3d1][dim - 1]
        return fn(*coords, data, nm if dim > 1 else nm[0], eps=eps)
    fn = [T.nufft1d2, T.nufft2d2, T.nufft3d2][dim - 1]
    return fn(*coords, data, eps=eps)


def run_stream(reqs, qps: float, seed: int = 0, mode: str = "server", n_streams: int = 4, max_units: int = 8):
    """Replay reqs with Poisson arrivals at qps (qps <= 0: every request arrives at once, which
    measures the saturated throughput).  mode "server": NufftServer (max_units point sets per
    shared launch; 1 turns cross-point batching off); mode "per-call": the same transforms called
    one at a time (each synchronised) in arrival order; mode "per-call-async": the same calls
    issued back to back on one stream without synchronising, each followed by an event whose
    completion (seen by a waiting thread, in issue order) is the request's completion time."""
    rng = np.random.default_rng(seed + 1)
    if qps > 0:
        arrivals = np.cumsum(rng.exponential(1.0 / qps, len(reqs)))
    else:
        arrivals = np.zeros(len(reqs))
    dev = reqs[0][1][0].device
    torch.cuda.synchronize(dev)
    records = []
    if mode == "server":
        with NufftServer(dev, n_streams=n_streams, max_units=max_units) as srv:
            # build every stream's plans for the distinct shapes, then replay the stream once
            # untimed (all at once, the largest batches) so that the plan pools are filled
            firsts = {}
            for r in reqs:
                firsts.setdefault((r[0], len(r[1]), r[3]), r)
            for si in range(n_streams):
                warm = [srv.submit(k, c, d, n_modes=nm if k == 1 else None, eps=1e-6, stream_index=si)
                        for k, c, d, nm in firsts.values()]
                for f in warm:
                    f.result()
            warm = [srv.submit(k, c, d, n_modes=nm if k == 1 else None, eps=1e-6) for k, c, d, nm in reqs]
            for f in warm:
                f.result()
            torch.cuda.synchronize(dev)
            built0 = srv.plans_built
            t0 = time.perf_counter()
            futs = []
            for (kind, coords, data, nm), ta in zip(reqs, arrivals):
                now = time.perf_counter() - t0
                if ta > now:
                    time.sleep(ta - now)
                fut = srv.submit(kind, coords, data, n_modes=nm if kind == 1 else None, eps=1e-6)
                fut.request.t_submit = t0 + ta  # latency counts from the scheduled arrival
                futs.append(fut)
            for f in futs:
                f.result()
            records = [f.request for f in futs]
            built = srv.plans_built - built0
    elif mode == "per-call-async":
        done_q: "queue.Queue" = queue.Queue()

        def complete():
            while True:
                item = done_q.get()
                if item is None:
                    break
                ev, r, out = item
                ev.synchronize()
                r.t_done = time.perf_counter()
                del out

        waiter = threading.Thread(target=complete, daemon=True)
        waiter.start()
        t0 = time.perf_counter()
        for (kind, coords, data, nm), ta in zip(reqs, arrivals):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            r = ServeRequest(kind, coords, data, nm, 1e-6, 1 if kind == 1 else -1)
            r.t_submit = t0 + ta
            r.t_start = time.perf_counter()
            out = _direct_call(kind, coords, data, nm)
            ev = torch.cuda.Event()
            ev.record(torch.cuda.current_stream(dev))
            r.t_issued = time.perf_counter()
            done_q.put((ev, r, out))
            records.append(r)
        done_q.put(None)
        waiter.join()
    else:
        t0 = time.perf_counter()
        for (kind, coords, data, nm), ta in zip(reqs, arrivals):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            r = ServeRequest(kind, coords, data, nm, 1e-6, 1 if kind == 1 else -1)
            r.t_submit = t0 + ta
            r.t_start = time.perf_counter()
            _direct_call(kind, coords, data, nm)
            r.t_issued = time.perf_counter()
            torch.cuda.synchronize(dev)
            r.t_done = time.perf_counter()
            records.append(r)
    st = latency_stats(records)
    if mode == "server":
        st["plans_built_in_replay"] = built
    return st


def service_times(reqs) -> dict:
    """Each request alone on an idle GPU (synchronised before and after): the service-time
    distribution in ms, overall and per (type, dimension, n_modes) class."""
    dev = reqs[0][1][0].device
    per_class = {}
    all_ms = []
    for kind, coords, data, nm in reqs:
        torch.cuda.synchronize(dev)
        t = time.perf_counter()
        _direct_call(kind, coords, data, nm)
        torch.cuda.synchronize(dev)
        ms = (time.perf_counter() - t) * 1e3
        all_ms.append(ms)
        per_class.setdefault(f"type{kind} {len(coords)}D {nm[0]}^{len(nm)}", []).append(ms)
    a = np.array(all_ms)
    return {"n": len(a), "p50_ms": float(np.percentile(a, 50)), "p99_ms": float(np.percentile(a, 99)),
            "mean_ms": float(a.mean()),
            "per_class_mean_ms": {k: round(float(np.mean(v)), 4) for k, v in sorted(per_class.items())}}


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="p50/p99 latency of the mixed NUFFT serving stream")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--qps", type=float, nargs="+", default=[100.0],
                    help="one stream replay per rate; 0 = all requests at once (saturated throughput)")
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--max-units", type=int, default=8,
                    help="point sets per shared spreading / interpolation launch in the server (1..8)")
    ap.add_argument("--compare-batched", type=int, default=0, metavar="UNITS",
                    help="also replay through a server with this many point sets per shared launch")
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    reqs = mixed_requests(args.requests, args.seed, dev)
    # warm the plan caches of both paths with one pass over the distinct shapes
    seen = set()
    for kind, coords, data, nm in reqs:
        key = (kind, len(coords), nm)
        if key not in seen:
            seen.add(key)
            _direct_call(kind, coords, data, nm)
    torch.cuda.synchronize(dev)
    svc = service_times(reqs)
    print(json.dumps({"workload": "mixed 2D/3D type 1/2 stream, complex64, eps 1e-6",
                      "requests": args.requests, "service": svc}), flush=True)
    for qps in args.qps:
        res = {"qps": qps, "streams": args.streams,
               "server": run_stream(reqs, qps, args.seed, "server", args.streams, args.max_units),
               "per_call": run_stream(reqs, qps, args.seed, "per-call"),
               "per_call_async": run_stream(reqs, qps, args.seed, "per-call-async")}
        if args.compare_batched:
            res[f"server_units{args.compare_batched}"] = run_stream(reqs, qps, args.seed, "server", args.streams,
                                                                   args.compare_batched)
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
