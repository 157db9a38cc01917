"""Python handle around a nufftx plan, plus a small per-stream plan cache.

A plan owns its device buffers (sort state, fine grid, rocFFT plans).  The
public transforms sort their points on every call, so a result depends only on
the arguments of that call (the reference's transforms are pure functions,
nufftax/transforms/nufft1.py:92-162).  Keeping a sort resident across calls is
an explicit opt-in (``setpts(..., reuse=True)``, used by callers that own the
coordinate tensors, such as the serving loop's registered point sets).  The
plan keeps the last coordinate tensors alive because execute re-reads
coordinates through the sorted permutation.

Threads: the cache is locked, and a plan is used by one thread at a time
(``with plan.lock:`` around setpts + execute); plans are per stream, so work of
one plan stays in stream order.
"""

from __future__ import annotations

import ctypes
import threading
from collections import OrderedDict

import torch

from . import _lib
from ._lib import F32, F64, check, lib, ptr


def _stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _tensor_key(t):
    if t is None:
        return None
    return (t.data_ptr(), t._version, tuple(t.shape), t.dtype)


class Plan:
    """One nufftx plan (type 1/2/3 transform, or type 0 grid operator)."""

    def __init__(self, handle: ctypes.c_void_p, kind: int, dim: int, prec: int, n_trans: int, device: torch.device):
        self.h = handle
        self.kind = kind
        self.dim = dim
        self.prec = prec
        self.n_trans = n_trans
        self.device = device
        self._pts_key = None
        self._pts_refs = None
        self.M = 0
        self.N = 0
        self.lock = threading.RLock()

    # ---------------------------------------------------------------- creation
    @classmethod
    def transform(cls, kind, dim, n_modes, n_trans, eps, isign, upsampfac, prec, modeord, device):
        nm = (ctypes.c_int64 * 3)(*(list(n_modes) + [1] * (3 - len(n_modes))))
        h = ctypes.c_void_p()
        stream = _stream_handle(device)
        with torch.cuda.device(device):
            check(
                lib().nufftx_makeplan(kind, dim, nm, int(isign), int(n_trans), float(eps), float(upsampfac), prec,
                                      int(modeord), ctypes.c_void_p(stream), ctypes.byref(h)),
                "nufftx_makeplan",
            )
        return cls(h, kind, dim, prec, n_trans, device)

    @classmethod
    def grid_op(cls, dim, nf, nspread, beta, kc, n_trans, prec, device):
        nfa = (ctypes.c_int64 * 3)(*(list(nf) + [1] * (3 - len(nf))))
        h = ctypes.c_void_p()
        stream = _stream_handle(device)
        with torch.cuda.device(device):
            check(
                lib().nufftx_makeplan_grid(dim, nfa, int(nspread), float(beta), float(kc), int(n_trans), prec,
                                           ctypes.c_void_p(stream), ctypes.byref(h)),
                "nufftx_makeplan_grid",
            )
        return cls(h, 0, dim, prec, n_trans, device)

    def grid_size(self, d: int, inner: bool = False) -> int:
        return int(lib().nufftx_plan_grid_size(self.h, d, 1 if inner else 0))

    # ---------------------------------------------------------------- state
    def _sync_stream(self):
        check(lib().nufftx_set_stream(self.h, ctypes.c_void_p(_stream_handle(self.device))), "nufftx_set_stream")

    def setpts(self, coords, targets=None, reuse=False):
        """Fold and bin-sort the points.  reuse=True skips the sort when the same tensors (same
        storage, in-place version, shape and dtype) were the last ones sorted: only for callers
        that guarantee the coordinates were not written behind the version counter since."""
        coords = list(coords)
        targets = list(targets) if targets is not None else []
        key = (tuple(_tensor_key(t) for t in coords), tuple(_tensor_key(t) for t in targets))
        self._sync_stream()
        if reuse and key == self._pts_key:
            return
        M = coords[0].numel()
        N = targets[0].numel() if targets else 0
        xs = coords + [None] * (3 - len(coords))
        ss = targets + [None] * (3 - len(targets))
        self._pts_key = None
        check(
            lib().nufftx_setpts(self.h, M, ptr(xs[0]), ptr(xs[1]), ptr(xs[2]), N, ptr(ss[0]), ptr(ss[1]),
                                ptr(ss[2])),
            "nufftx_setpts",
        )
        self._pts_key = key
        self._pts_refs = (coords, targets)
        self.M, self.N = M, N

    def execute(self, c, f, isign=None):
        self._sync_stream()
        if isign is None:
            check(lib().nufftx_execute(self.h, ptr(c), ptr(f)), "nufftx_execute")
        else:
            check(lib().nufftx_execute_sign(self.h, ptr(c), ptr(f), int(isign)), "nufftx_execute_sign")

    @staticmethod
    def execute_many(plans, cs, fs):
        """Run plans[i] on (cs[i], fs[i]) as execute() does, sharing the spreading /
        interpolation launch across the plans where the library can (nufftx_execute_many)."""
        n = len(plans)
        if n == 0:
            return
        for p in plans:
            p._sync_stream()
        arr = ctypes.c_void_p * n
        check(lib().nufftx_execute_many(n, arr(*[p.h for p in plans]), arr(*[ptr(c) for c in cs]),
                                        arr(*[ptr(f) for f in fs])), "nufftx_execute_many")

    def vjp(self, src, g, grad_src, grads):
        self._sync_stream()
        gs = list(grads) + [None] * (3 - len(grads))
        check(lib().nufftx_vjp(self.h, ptr(src), ptr(g), ptr(grad_src), ptr(gs[0]), ptr(gs[1]), ptr(gs[2])),
              "nufftx_vjp")

    def spread(self, c, fw):
        self._sync_stream()
        check(lib().nufftx_spread(self.h, ptr(c), ptr(fw)), "nufftx_spread")

    def spread_add(self, c, fw):
        self._sync_stream()
        check(lib().nufftx_spread_add(self.h, ptr(c), ptr(fw)), "nufftx_spread_add")

    def interp(self, fw, c):
        self._sync_stream()
        check(lib().nufftx_interp(self.h, ptr(fw), ptr(c)), "nufftx_interp")

    def gather_grad(self, G, P, gval, grads):
        self._sync_stream()
        gs = list(grads) + [None] * (3 - len(grads))
        check(lib().nufftx_gather_grad(self.h, ptr(G), ptr(P), ptr(gval), ptr(gs[0]), ptr(gs[1]), ptr(gs[2])),
              "nufftx_gather_grad")

    def sort_state(self, inner: bool = False):
        """(perm, bin_start, nbins, bin_size, bins_per_dim) of the last setpts (device tensors)."""
        nb = ctypes.c_int()
        bs = (ctypes.c_int * 3)()
        nbd = (ctypes.c_int * 3)()
        check(lib().nufftx_get_sort_info(self.h, 1 if inner else 0, ctypes.byref(nb), bs, nbd), "nufftx_get_sort_info")
        n = self.N if inner else self.M
        perm = torch.empty(n, dtype=torch.int32, device=self.device)
        bst = torch.empty(nb.value + 1, dtype=torch.int32, device=self.device)
        self._sync_stream()
        check(lib().nufftx_copy_sort(self.h, 1 if inner else 0, ptr(perm), ptr(bst)), "nufftx_copy_sort")
        return perm, bst, nb.value, tuple(bs), tuple(nbd)

    def destroy(self):
        if self.h is not None and self.h.value:
            try:
                torch.cuda.synchronize(self.device)
            except Exception:
                pass
            lib().nufftx_destroy(self.h)
        self.h = None
        self._pts_refs = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PlanCache:
    """LRU cache of plans keyed by their static parameters, device and stream."""

    def __init__(self, capacity: int = 24):
        self.capacity = capacity
        self._plans: OrderedDict = OrderedDict()
        self._lock = threading.Lock()

    def get(self, key, factory):
        with self._lock:
            p = self._plans.get(key)
            if p is not None:
                self._plans.move_to_end(key)
                return p
            p = factory()
            self._plans[key] = p
            evicted = []
            while len(self._plans) > self.capacity:
                evicted.append(self._plans.popitem(last=False)[1])
        for old in evicted:
            with old.lock:
                old.destroy()
        return p

    def clear(self):
        with self._lock:
            plans = list(self._plans.values())
            self._plans.clear()
        for p in plans:
            with p.lock:
                p.destroy()


CACHE = PlanCache()


def transform_plan(kind, dim, n_modes, n_trans, eps, isign, upsampfac, prec, device, modeord=0) -> Plan:
    device = torch.device(device)
    stream = _stream_handle(device)
    key = ("t", kind, dim, tuple(int(v) for v in n_modes), int(n_trans), float(eps), int(isign), float(upsampfac),
           prec, modeord, device.index, stream)
    return CACHE.get(key, lambda: Plan.transform(kind, dim, n_modes, n_trans, eps, isign, upsampfac, prec, modeord,
                                                 device))


def grid_plan(dim, nf, nspread, beta, kc, n_trans, prec, device) -> Plan:
    device = torch.device(device)
    stream = _stream_handle(device)
    key = ("g", dim, tuple(int(v) for v in nf), int(nspread), float(beta), float(kc), int(n_trans), prec,
           device.index, stream)
    return CACHE.get(key, lambda: Plan.grid_op(dim, nf, nspread, beta, kc, n_trans, prec, device))


__all__ = ["Plan", "PlanCache", "CACHE", "transform_plan", "grid_plan", "F32", "F64"]
