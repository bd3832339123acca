"""Autotuner of the tile-DAG planner parameters, per pattern and precision.

The shape of the big-front schedule is set by seven planner parameters
(``TUNING_NAMES``: panel steps per trailing-update tile, parent rows per
extend-add task, the front width from which update-matrix tiles are
128 x 128, the workgroups the ticket order is simulated with, and the
explicit zeros a chain merge may add, the task kinds run at raised wave
priority, and whether the diagonal-tile task also runs the next panel solve
and look-ahead update).  Their best values depend on the
pattern (front sizes, tree shape) and on the precision of the fronts (FP32
fronts run three workgroups per CU, FP64 two).

``tune`` measures the device-resident value refresh + factorization of one
pattern for candidate parameter sets on the GPU (coordinate search: one
parameter at a time over its candidate values, the others fixed at the best
so far, repeated while it improves) and returns the fastest.  ``lookup``
returns a tuned set from the table ``tuning.json`` next to this file (keyed
by the pattern's symbolic signature and the precision), or None, which
means the planner's defaults.  ``save`` adds a result to that table.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import time

import numpy as np

from . import _native
from .factorization import (MF_CHOLESKY, MF_FP32, MF_FP64, TUNING_NAMES, DeviceFactor, DevicePlan,
                            _check)
from .symbolic import flatindices

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

DEFAULTS = {"upd_group": 8, "ea_rows": 16384, "wide_k": 128, "sim_workers": 512, "zero_ppm": 50000,
            "prio": 0, "fuse": 0, "wg_per_cu": 0}
CANDIDATES = {
    "upd_group": [4, 8, 12, 16],
    "ea_rows": [8192, 16384, 32768],
    "wide_k": [128, 256, 512],
    "sim_workers": [512, 768, 1024],
    "zero_ppm": [20000, 50000, 100000],
    "prio": [0, 4, 4 | 8 | 64],      # diagonal tiles; + panel solves and the look-ahead update
    "fuse": [0, 1],
}


def signature(S) -> str:
    """Digest of the symbolic factorization (the supernode partition and the
    separators): equal for every value update of one pattern."""
    h = hashlib.sha1()
    for a in (S.resptr, S.sepptr, S.septgt):
        h.update(np.ascontiguousarray(a, dtype=np.int64).tobytes())
    return h.hexdigest()[:20]


def _key(S, precision):
    return f"{signature(S)}|{'fp32' if precision == MF_FP32 else 'fp64'}"


def _vector(d):
    return [int(d[k]) for k in TUNING_NAMES]


def lookup(S, precision=MF_FP64, table=TABLE):
    """The tuned parameters of this pattern and precision, or None."""
    if not os.path.exists(table):
        return None
    with open(table) as f:
        entry = json.load(f).get(_key(S, precision))
    return None if entry is None else entry["tuning"]


def save(S, precision, result, desc="", table=TABLE):
    """Record a tune() result in the table."""
    data = {}
    if os.path.exists(table):
        with open(table) as f:
            data = json.load(f)
    data[_key(S, precision)] = {"tuning": _vector(result["best"]), "names": TUNING_NAMES,
                                "ms": round(result["best_ms"], 4),
                                "default_ms": round(result["default_ms"], 4),
                                "n": int(S.n), "desc": desc}
    with open(table, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def time_factorization(S, A, precision, tuning, reps=3, diag=MF_CHOLESKY, torch=None):
    """Median milliseconds of refresh + factorization (values already on the
    device) with a plan built from ``tuning`` (dict or None)."""
    if torch is None:
        import torch
    lib = _native.hip()
    plan = DevicePlan(S, None if tuning is None else _vector(tuning))
    fac = DeviceFactor(plan, precision)
    h = fac.handle
    flat = flatindices(S, A)
    _check(lib.mf_factor_set_pattern(h, len(flat), flat.ctypes.data_as(_native.i64p)))
    stream = torch.cuda.Stream()
    _check(lib.mf_factor_set_stream(h, ctypes.c_void_p(stream.cuda_stream)))
    nz = torch.tensor(np.ascontiguousarray(A.data), dtype=torch.float64, device="cuda")

    def step():
        _check(lib.mf_factor_refresh(h, ctypes.c_void_p(nz.data_ptr()), 1))
        _check(lib.mf_factorize_async(h, diag, None, 0, 0.0, 0.0))

    step()
    stream.synchronize()
    info = ctypes.c_int64(0)
    _check(lib.mf_factor_info(h, ctypes.byref(info)))
    if info.value != 0:
        raise RuntimeError(f"factorization failed while tuning (info {info.value})")
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        step()
        b.record(stream)
        stream.synchronize()
        times.append(a.elapsed_time(b))
    del fac, plan
    return float(np.median(times))


def tune(S, A, precision=MF_FP64, candidates=None, reps=3, rounds=2, diag=MF_CHOLESKY, log=None):
    """Coordinate search of the planner parameters for one pattern; returns
    {"best": dict, "best_ms", "default_ms", "trials": [(dict, ms), ...],
    "seconds"}."""
    cand = dict(CANDIDATES if candidates is None else candidates)
    start = dict(DEFAULTS)
    if precision == MF_FP32:
        start["sim_workers"] = 768
    t0 = time.perf_counter()
    trials = []
    seen = {}

    def measure(d):
        k = tuple(_vector(d))
        if k not in seen:
            ms = time_factorization(S, A, precision, d, reps=reps, diag=diag)
            seen[k] = ms
            trials.append((dict(d), ms))
            if log:
                log(f"tune {dict(d)}: {ms:.3f} ms")
        return seen[k]

    default_ms = measure(DEFAULTS)
    best, best_ms = dict(start), measure(start)
    if default_ms < best_ms:
        best, best_ms = dict(DEFAULTS), default_ms
    for _ in range(rounds):
        improved = False
        for name, values in cand.items():
            for v in values:
                if v == best[name]:
                    continue
                d = dict(best)
                d[name] = v
                ms = measure(d)
                if ms < best_ms * 0.995:
                    best, best_ms, improved = d, ms, True
        if not improved:
            break
    return {"best": best, "best_ms": best_ms, "default_ms": default_ms, "trials": trials,
            "seconds": time.perf_counter() - t0}


__all__ = ["tune", "lookup", "save", "signature", "time_factorization", "DEFAULTS", "CANDIDATES",
           "TABLE", "MF_FP32", "MF_FP64"]
