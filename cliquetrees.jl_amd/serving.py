"""Serving loop: a stream of factor + solve requests on one GPU.

Each request carries a sparse symmetric matrix (a value update of some
pattern) and right-hand sides.  The loop

* analyses every pattern once (ordering, symbolic, device plan) and reuses it
  for every later request with that pattern;
* keeps factor objects (panels, update matrices, workspaces) resident on the
  device between requests and admits work only while the resident factors fit
  an HBM budget, evicting idle ones least recently used first;
* batches continuously: when it dispatches, it takes every request of one
  pattern that has arrived by then (up to ``max_batch``, rounded down to a
  power of two) and factors them as one forest, so the persistent kernels walk
  the same elimination-tree levels of all of them in one launch;
* overlaps dispatches on ``streams`` HIP streams, so requests of other
  patterns (and later arrivals) run while earlier work is on the GPU.

Each request's result is what it gets alone: a forest of k copies has k
independent trees, factored with the same fronts and the same arithmetic.
"""
from __future__ import annotations

import ctypes
import time
from collections import OrderedDict, deque
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _native
from .factorization import (MF_CHOLESKY, MF_FP64, MF_LDLT, MF_REG_NONE, STAT_NAMES, DeviceFactor,
                            _check, device_plan)
from .symbolic import flatindices, mmd, symbolic


@dataclass
class Request:
    K: sp.csc_matrix
    B: np.ndarray               # n x nrhs (float64)
    arrival: float = 0.0        # seconds, relative to the start of run()
    tag: object = None
    X: np.ndarray | None = None
    relres: float = float("nan")    # max_j ||B_j - K X_j|| / ||B_j||, from the device
    info: int = 0
    start: float = 0.0
    done: float = 0.0
    batch: int = 0

    @property
    def latency(self):
        return self.done - self.arrival


def _pattern_key(K):
    return (K.shape[0], hash(K.indptr.tobytes()), hash(K.indices.tobytes()))


class _Pattern:
    def __init__(self, K, order, diag):
        self.n = K.shape[0]
        self.K = K
        self.order = order
        self.diag = diag
        self.variants = {}      # k -> (P, S, flat, plan)

    def variant(self, k):
        v = self.variants.get(k)
        if v is None:
            Kk = sp.block_diag([self.K] * k, format="csc")
            Kk.sort_indices()
            order = np.concatenate([self.order + c * self.n for c in range(k)])
            P, S = symbolic(Kk, alg=order)
            flat = flatindices(S, Kk)
            plan = device_plan(S)
            st = plan.stats()
            v = (P, S, flat, plan, st, Kk)
            self.variants[k] = v
        return v


def _slot_bytes(st, n, nnz, precision, nrhs=1):
    """Device bytes of one resident factor object of a plan for ``nrhs``
    right-hand sides: panels, U, W, reference arrays, per-column vectors and
    counters, the pattern's scatter targets, the registered matrix, the
    solve and refinement workspaces (update vectors, permuted right-hand
    sides, staging, B copy / X / residual and norms) and the slot's own value
    and right-hand-side buffers."""
    es = 4 if precision != MF_FP64 else 8
    fixed = (es * (st["panel_elems"] + st["u_elems"] + st["w_elems"])
             + 8 * (st["ndz"] + st["nlz"]) + 64 * n + 16 * st["internal_fronts"])
    pattern = 16 * nnz + 8 * (n + 1) + 16 * nnz + 8 * nnz     # targets, CSC of A, values
    per_rhs = es * st["mv_elems1"] + es * n + 8 * n + 3 * 8 * n + 2 * 8 * n + 32
    return int(fixed + pattern + nrhs * per_rhs)


class _Slot:
    """A resident factor object for (pattern, k) with its own stream and buffers."""

    def __init__(self, pat, k, precision, torch):
        P, S, flat, plan, st, Kk = pat.variant(k)
        self.pat, self.k = pat, k
        self.f = DeviceFactor(plan, precision)
        self.perm = np.ascontiguousarray(P.perm, dtype=np.int64)
        lib = _native.hip()
        _check(lib.mf_factor_set_pattern(self.f.handle, len(flat), flat.ctypes.data_as(_native.i64p)))
        # A itself (same nonzero order as the pattern, so that every value
        # refresh updates it too) for the FP64 residuals of the refinement
        colptr = np.ascontiguousarray(Kk.indptr, dtype=np.int64) + 1
        rowval = np.ascontiguousarray(Kk.indices, dtype=np.int64) + 1
        vals = np.ascontiguousarray(Kk.data, dtype=np.float64)
        _check(lib.mf_factor_set_matrix(self.f.handle, Kk.shape[0], colptr.ctypes.data_as(_native.i64p),
                                        rowval.ctypes.data_as(_native.i64p),
                                        vals.ctypes.data_as(ctypes.c_void_p), 0))
        self.seg = np.arange(k + 1, dtype=np.int64) * pat.n
        self.stream = torch.cuda.Stream()
        _check(lib.mf_factor_set_stream(self.f.handle, ctypes.c_void_p(self.stream.cuda_stream)))
        n = S.n
        self.nnz = len(flat)
        self.vals_h = torch.empty(self.nnz, dtype=torch.float64, pin_memory=True)
        self.vals_d = torch.empty(self.nnz, dtype=torch.float64, device="cuda")
        self.n = n
        self.nrhs = 0
        self.B_h = self.B_d = self.X_d = self.nrm_h = self.nrm_d = None
        self.st = st
        self.precision = precision
        self.bytes = _slot_bytes(st, n, self.nnz, precision)
        self.busy = False
        self.event = None
        self.batch = []
        self.last_used = 0.0

    def ensure_rhs(self, nrhs, torch):
        """Right-hand-side buffers for nrhs columns; returns the change of the
        slot's device bytes (the workspaces grow with nrhs)."""
        old = self.bytes
        if self.nrhs != nrhs:
            self.B_h = torch.empty((nrhs, self.n), dtype=torch.float64, pin_memory=True)
            self.B_d = torch.empty((nrhs, self.n), dtype=torch.float64, device="cuda")
            self.X_d = torch.empty((nrhs, self.n), dtype=torch.float64, device="cuda")
            self.nrm_h = torch.empty(2 * self.k * nrhs, dtype=torch.float64, pin_memory=True)
            self.nrm_d = torch.empty(2 * self.k * nrhs, dtype=torch.float64, device="cuda")
            self.nrhs = nrhs
            self.bytes = max(self.bytes, _slot_bytes(self.st, self.n, self.nnz, self.precision, nrhs))
        return self.bytes - old

    def measured_bytes(self):
        """What the factor object holds on the device (mf_factor_bytes) plus
        the slot's own buffers."""
        b = ctypes.c_int64(0)
        _check(_native.hip().mf_factor_bytes(self.f.handle, ctypes.byref(b)))
        own = 8 * self.nnz + (2 * 8 * self.n + 16 * self.k) * self.nrhs
        return int(b.value) + own

    def release(self):
        self.f = None


class ServingLoop:
    """Continuous-batching factor + solve server (see the module docstring).

    Each dispatch solves with ``refine_steps`` FP64 corrections
    X += F \\ (B - K X) (mf_refine_steps, no host synchronization) and
    reports every request's final relative residual in ``Request.relres``.

    diag:       MF_LDLT (quasi-definite KKT systems, default signs) or MF_CHOLESKY
    ordering:   "mmd" or a callable K -> 1-based order
    hbm_budget: bytes of device memory the resident factor objects may use
    """

    def __init__(self, diag=MF_LDLT, precision=MF_FP64, ordering="mmd", hbm_budget=64 << 30,
                 max_batch=8, streams=4, refine_steps=1):
        import torch

        self.torch = torch
        self.diag, self.precision = diag, precision
        self.ordering = ordering
        self.budget = int(hbm_budget)
        self.max_batch = max(1, int(max_batch))
        self.refine_steps = max(0, int(refine_steps))
        self.nstreams = max(1, int(streams))
        self.patterns = {}
        self.slots = OrderedDict()    # (pattern key, k, serial) -> _Slot, LRU order
        self.resident = 0
        self._serial = 0
        self.analysis_s = 0.0

    # -- pattern analysis (once per pattern) --
    def pattern(self, K):
        key = _pattern_key(K)
        p = self.patterns.get(key)
        if p is None:
            t = time.perf_counter()
            order = mmd(K) if self.ordering == "mmd" else np.asarray(self.ordering(K), np.int64)
            p = _Pattern(K, np.asarray(order, np.int64), self.diag)
            p.key = key
            self.patterns[key] = p
            self.analysis_s += time.perf_counter() - t
        return p

    def prepare(self, K, ks=(1,)):
        """Analyse a pattern (and its k-copy forests) ahead of the stream."""
        p = self.pattern(K)
        t = time.perf_counter()
        for k in ks:
            p.variant(k)
        self.analysis_s += time.perf_counter() - t
        return p

    # -- device memory admission --
    def _slot_for(self, pat, k, nrhs=1):
        for key, s in self.slots.items():
            if key[0] == pat.key and key[1] == k and not s.busy:
                self.slots.move_to_end(key)
                return s
        P, S, flat, plan, st, _ = pat.variant(k)
        need = _slot_bytes(st, S.n, len(flat), self.precision, nrhs)
        if need > self.budget:
            raise MemoryError(f"one factor of this pattern needs {need} bytes > budget {self.budget}")
        # evict idle slots (least recently used first) until the new one fits
        while self.resident + need > self.budget:
            victim = next((key for key, v in self.slots.items() if not v.busy), None)
            if victim is None:
                return None        # wait for a running dispatch to finish
            self.resident -= self.slots[victim].bytes
            self.slots.pop(victim).release()
        s = _Slot(pat, k, self.precision, self.torch)
        self._serial += 1
        self.slots[(pat.key, k, self._serial)] = s
        self.resident += s.bytes
        return s

    # -- one dispatch: refresh + factorize + solve of a batch on a slot's stream --
    def _dispatch(self, slot, batch, now):
        torch = self.torch
        lib = _native.hip()
        nrhs = batch[0].B.shape[1]
        if any(r.B.shape != (slot.pat.n, nrhs) for r in batch):
            raise ValueError("requests of one pattern must have the same number of right-hand sides")
        self.resident += slot.ensure_rhs(nrhs, torch)
        vals = np.concatenate([np.ascontiguousarray(r.K.data, dtype=np.float64) for r in batch])
        slot.vals_h.numpy()[:] = vals
        Bh = slot.B_h.numpy()
        n1 = slot.pat.n
        for c, r in enumerate(batch):
            Bh[:, c * n1:(c + 1) * n1] = r.B.T
        with torch.cuda.stream(slot.stream):
            slot.vals_d.copy_(slot.vals_h, non_blocking=True)
            slot.B_d.copy_(slot.B_h, non_blocking=True)
        _check(lib.mf_factor_refresh(slot.f.handle, ctypes.c_void_p(slot.vals_d.data_ptr()), 1))
        _check(lib.mf_factorize_async(slot.f.handle, self.diag, None, MF_REG_NONE, 0.0, 0.0))
        _check(lib.mf_refine_steps(slot.f.handle, slot.perm.ctypes.data_as(_native.i64p),
                                   ctypes.c_void_p(slot.B_d.data_ptr()), slot.n,
                                   ctypes.c_void_p(slot.X_d.data_ptr()), slot.n, nrhs,
                                   self.refine_steps, slot.k, slot.seg.ctypes.data_as(_native.i64p),
                                   ctypes.c_void_p(slot.nrm_d.data_ptr())))
        with torch.cuda.stream(slot.stream):
            slot.B_h.copy_(slot.X_d, non_blocking=True)
            slot.nrm_h.copy_(slot.nrm_d, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        slot.busy, slot.event, slot.batch = True, ev, batch
        for r in batch:
            r.start = now
            r.batch = len(batch)

    def _complete(self, slot, now):
        info = ctypes.c_int64(0)
        _check(_native.hip().mf_factor_info(slot.f.handle, ctypes.byref(info)))
        X = slot.B_h.numpy()
        nrm = slot.nrm_h.numpy()
        k, nrhs = slot.k, slot.nrhs
        b2, r2 = nrm[:k * nrhs].reshape(nrhs, k), nrm[k * nrhs:].reshape(nrhs, k)
        n1 = slot.pat.n
        for c, r in enumerate(slot.batch):
            r.X = np.array(X[:, c * n1:(c + 1) * n1].T)
            r.relres = float(np.max(np.sqrt(r2[:, c] / np.where(b2[:, c] > 0, b2[:, c], 1.0))))
            g = int(info.value)
            r.info = 0
            if g:
                v = int(slot.perm[g - 1]) - 1
                if c * n1 <= v < (c + 1) * n1:
                    r.info = v - c * n1 + 1
            r.done = now
        slot.busy, slot.event, slot.batch = False, None, []
        slot.last_used = now
        # account what the factor object really holds once its workspaces exist
        real = slot.measured_bytes()
        if real > slot.bytes:
            self.resident += real - slot.bytes
            slot.bytes = real

    def run(self, requests, realtime=True, batching=True):
        """Serve requests (sorted by arrival).  realtime: requests become visible
        at their arrival time; otherwise all are visible at once.  batching=False
        serves one request at a time on one stream (the baseline)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            self.pattern(r.K)
        pending = deque(reqs)
        queue = OrderedDict()          # pattern key -> deque of arrived requests
        inflight = []
        nstreams = self.nstreams if batching else 1
        maxb = self.max_batch if batching else 1
        t0 = time.perf_counter()
        while pending or any(queue.values()) or inflight:
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                r = pending.popleft()
                queue.setdefault(_pattern_key(r.K), deque()).append(r)
            for slot in list(inflight):
                if slot.event.query():
                    self._complete(slot, time.perf_counter() - t0)
                    inflight.remove(slot)
            while len(inflight) < nstreams:
                cand = [(q[0].arrival, key) for key, q in queue.items() if q]
                if not cand:
                    break
                _, key = min(cand)
                q = queue[key]
                take = 1
                while take * 2 <= min(len(q), maxb):
                    take *= 2
                pat = self.patterns[key]
                slot = self._slot_for(pat, take, q[0].B.shape[1])
                if slot is None:
                    break
                batch = [q.popleft() for _ in range(take)]
                self._dispatch(slot, batch, time.perf_counter() - t0)
                inflight.append(slot)
            if inflight and not any(s.event.query() for s in inflight):
                time.sleep(0.0002)
            elif not inflight and pending and realtime:
                time.sleep(max(0.0, min(0.001, pending[0].arrival - (time.perf_counter() - t0))))
        self.torch.cuda.synchronize()
        return reqs


def poisson_arrivals(count, rate, seed):
    """Arrival times (s) of a Poisson process with the given rate (1/s)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


def latency_stats(reqs):
    lat = np.array([r.latency for r in reqs]) * 1e3
    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "requests": len(reqs),
            "solves_per_s": float(sum(r.B.shape[1] for r in reqs) / span) if span > 0 else None,
            "makespan_s": float(span)}


__all__ = ["Request", "ServingLoop", "poisson_arrivals", "latency_stats", "MF_LDLT",
           "MF_CHOLESKY", "STAT_NAMES"]
