"""Benchmark of the multifrontal numeric engine on one MI355X.

Default workload (lap7_64): plain FP64 sparse Cholesky of the 3-D 7-point
weighted Laplacian on a 64^3 grid (n = 262,144; edge weights i.i.d.
U[0.5, 1.5], seed 2, plus 1e-3 I), geometric nested-dissection order, 16
right-hand sides N(0,1) (seed 2).  One step = assembly of A's values into the
factor storage (setflatindex! of every nonzero, device-resident values) +
numeric factorization + a 16-RHS solve, all on the GPU.

  value        = reference factor flops / step time, TFLOP/s (whole step:
                 assembly + factorization + solve), with the flop count of
                 the reference's own supernode structure:
                 sum_j nn^3/3 + na nn^2 + na^2 nn.
  roofline     = the dominant kernel (k_dag, the persistent tile-DAG of the
                 big fronts, FP64 MFMA) timed with HIP events on the
                 engine's stream over the timed region; its algorithmic work
                 is the reference flop count of the fronts it factors (the
                 reference's own fronts, not the engine's merged ones).
  cpu_baseline = the CPU oracle (restatement of the reference's serial front
                 loop, dense kernels in host OpenBLAS on all host threads) on
                 a bounded sample of the same factorization (a subtree with
                 1/8 of the flops), median of the repetitions; the whole
                 factorization + solve measured by --workload cpu_reference
                 on the same host is attached from profiles/.
  workloads    = the other configurations measured in the same run
                 (lap7_64_mixed, st27_48, st27_48_mixed, batch4096,
                 kkt_serving), fewer steps each; --headline-only skips them.

Other workloads (--workload):
  lap7_64_mixed  the lap7_64 matrix with FP32 fronts + FP64 iterative
                 refinement to 1e-10, against the plain FP64 step on it
  st27_48_mixed  3-D 27-point stencil 48^3 (seed 3), FP32 fronts + FP64
                 refinement to 1e-10, error vs the FP64 solution, plain FP64
                 step alongside
  st27_48        the same 27-point matrix in plain FP64
  batch4096      4,096 independent grid Laplacians, n ~ U{50..2000} (seed 4),
                 half 2-D / half 3-D, factored and solved as one forest (FP64
                 and FP32 + refinement); value = systems per second
  kkt_serving    Poisson stream of KKT LDLt requests, patterns reused over 20
                 value updates, continuous batching; p50/p99 latency, solves/s
                 against one-at-a-time serving
  lap7_64_selinv selected inversion (selinv!) of the lap7_64 FP64 factor
  lap7_64_complete
                 completion (complete!) of that selected inverse through the
                 drop-in mf_complete_impl (host arrays, PCIe inside the call);
                 not among the default run's workloads (its first call builds
                 the completion's forest plans, about 20 s)
  copy           a 1 GiB device copy (calibration of the HBM traffic counters)
  cpu_reference  the CPU reference path (oracle + host OpenBLAS) on the whole
                 lap7_64 factorization and 16-RHS solve, median / p99 of
                 at least 5 repetitions (no GPU)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
FP64_PEAK = 78.6   # TFLOP/s, MI355X dense FP64 matrix
FP32_PEAK = 157.3  # TFLOP/s, MI355X dense FP32 matrix
HBM_PEAK = 8000.0  # GB/s
METRIC = "factor TFLOP/s + solves/s at <=1e-10 rel residual vs plain FP64; p50/p99 ms"


def log(msg):
    """Progress line on stderr (the JSON result alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def traffic_from_profile(kernel, tag):
    """HBM bytes per launch of ``kernel`` from the committed counter summary
    ``profiles/<kernel>_<tag>_hbm_traffic.md`` (FETCH_SIZE and WRITE_SIZE
    passes of rocprofv3, each scaled by what the same counter reads for a
    1 GiB device copy); None when no such summary exists."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{tag}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def load_package():
    if "cliquetrees_amd" in sys.modules:
        return sys.modules["cliquetrees_amd"]
    pkg = os.path.join(ROOT, "cliquetrees.jl_amd")
    spec = importlib.util.spec_from_file_location("cliquetrees_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cliquetrees_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(ct, name):
    if name in ("lap7_64", "lap7_64_mixed"):
        A = ct.problems.laplacian3d_7pt(64, 64, 64, seed=2)
        mp = name == "lap7_64_mixed"
        # with FP32 fronts, one correction meets 1e-10 when the solves run in
        # FP64 arithmetic (worst column 8.8e-11; 1.5e-10 with FP32 solves,
        # which then need a second correction)
        return dict(A=A, order=ct.grid_nd(64, 64, 64), nrhs=16, seed=2, tag=name,
                    precision=ct.MF_FP32 if mp else ct.MF_FP64, refine=mp, solve_fp64=mp,
                    desc="3D 7-point Laplacian 64^3, w~U[0.5,1.5] + 1e-3 I, geometric ND, 16 RHS, "
                         + ("FP32 fronts + FP64 iterative refinement to 1e-10 (FP64-arithmetic "
                            "solves on the widened FP32 factor)" if mp
                            else "plain FP64 Cholesky"))
    if name in ("st27_48", "st27_48_mixed"):
        A = ct.problems.stencil3d_27pt(48, 48, 48, seed=3)
        mp = name == "st27_48_mixed"
        return dict(A=A, order=ct.grid_nd(48, 48, 48), nrhs=16, seed=3, tag=name,
                    precision=ct.MF_FP32 if mp else ct.MF_FP64, refine=mp,
                    desc="3D 27-point stencil 48^3, w~U[0.5,1.5], diagonally dominant, geometric ND, "
                         "16 RHS, " + ("FP32 fronts + FP64 iterative refinement to 1e-10" if mp
                                       else "plain FP64 Cholesky"))
    raise SystemExit(f"unknown workload {name}")


def host_info():
    """Host CPU of the run: model name, logical CPUs, and the BLAS threads the
    oracle uses (all of them)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = ncpu
    try:
        from threadpoolctl import threadpool_info
        threads = max(int(i.get("num_threads", 1)) for i in threadpool_info()) or ncpu
    except Exception:
        pass
    return model, ncpu, threads


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.use_host_blas(True)
    return oracle


def cpu_baseline(ct, A, P, S, frac=0.125, min_s=10.0, min_reps=3):
    """The oracle (restated serial front loop of chol_impl!, its dense kernels
    and the nn = 1 rank-1 updates routed to host OpenBLAS with all host
    threads, as the reference does for Float64) on a bounded sample of the
    same factorization: the largest elimination subtree holding at most
    ``frac`` of the reference flops (for a nested-dissection order a sub-box
    with the whole problem's mix of front sizes), factored until ``min_s``
    seconds and ``min_reps`` repetitions; median and p99 of the repetitions.
    The whole factorization and solve are timed by ``--workload cpu_reference``."""
    oracle = _oracle()
    Dv, Lv = oracle.assemble(S, P, A)
    nn = S.nn().astype(np.float64)
    na = S.na().astype(np.float64)
    fl = nn ** 3 / 3 + na * nn ** 2 + na ** 2 * nn
    sub = fl.copy()
    for j in range(S.nfr):
        p = S.pnt[j] - 1
        if p >= 0:
            sub[p] += sub[j]
    total = float(fl.sum())
    ok = np.nonzero(sub <= frac * total)[0]
    root = int(ok[np.argmax(sub[ok])]) + 1
    model, ncpu, threads = host_info()
    times, flops, info = [], 0.0, 0
    while sum(times) < min_s or len(times) < min_reps:
        dt, flops, info = oracle.chol_subtree(S, Dv, Lv, root)
        times.append(dt)
        if len(times) >= 50:
            break
    med = float(np.median(times))
    return {"value": flops / med / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "host": model, "nproc": ncpu,
            "median_s": round(med, 4), "p99_s": round(float(np.percentile(times, 99)), 4),
            "sample": f"subtree of front {root} ({flops / total:.1%} of the factor flops, "
                      f"{flops:.3e} flops) factored {len(times)} times ({sum(times):.1f} s), "
                      f"info={info}: oracle.c serial front loop + OpenBLAS dpotrf/dtrsm/dsyrk/"
                      f"dgemm/dsyr with {threads} threads; median per factorization"}


def run_cpu_reference(args, ct):
    """The CPU reference path on the whole default workload: the oracle's
    factorization of the 64^3 Laplacian (restated chol_impl! + host OpenBLAS
    on all host threads) and its 16-RHS solve (div_impl! + permutations),
    1 warm-up then ``--steps`` repetitions (at least 5), median and p99 of
    each.  Symbolic analysis and assembly are excluded."""
    W = workload(ct, "lap7_64")
    A = W["A"]
    P, S = ct.symbolic(A, alg=W["order"])
    oracle = _oracle()
    Dv, Lv = oracle.assemble(S, P, A)
    B0 = np.random.default_rng(W["seed"]).standard_normal((A.shape[0], W["nrhs"]))
    model, ncpu, threads = host_info()
    reps = max(5, args.steps)
    tf, ts, res = [], [], None
    for k in range(reps + 1):
        t0 = time.perf_counter()
        D, L, d, info = oracle.chol_impl(S, Dv, Lv)
        t1 = time.perf_counter()
        X = oracle.solve(S, P, D, L, d, B0)
        t2 = time.perf_counter()
        log(f"cpu_reference rep {k}: factor {t1 - t0:.2f} s, solve {t2 - t1:.2f} s, info {info}")
        if k > 0:
            tf.append(t1 - t0)
            ts.append(t2 - t1)
        res = float(np.linalg.norm(A @ X - B0) / np.linalg.norm(B0))
    flops = S.factor_flops()
    mf, ms = float(np.median(tf)), float(np.median(ts))
    return {"metric": "CPU reference path: factor TFLOP/s and 16-RHS solve time",
            "value": round(flops / mf / 1e12, 5), "unit": "TFLOP/s", "n_gpus": 0,
            "steps": reps, "warmup": 1, "ms_per_step": round((mf + ms) * 1e3, 1),
            "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": W["desc"], "n": int(A.shape[0]), "ref_flops": flops},
            "factor_median_s": round(mf, 3), "factor_p99_s": round(float(np.percentile(tf, 99)), 3),
            "solve_median_s": round(ms, 3), "solve_p99_s": round(float(np.percentile(ts, 99)), 3),
            "rel_residual": res, "cores": threads, "nproc": ncpu, "host": model,
            "kind": "port",
            "path": "oracle.c (restated chol_impl!/div_impl! serial front loops) + OpenBLAS "
                    "dpotrf/dtrsm/dsyrk/dgemm/dsyr through scipy"}


class Engine:
    """One factor object of a workload, its stream and device buffers."""

    def __init__(self, ct, torch, A, P, S, precision, nrhs, B0, stream, solve_fp64=False):
        from cliquetrees_amd import _native
        self.ct, self.torch, self.lib, self._native = ct, torch, _native.hip(), _native
        self.A, self.n, self.nrhs = A, A.shape[0], nrhs
        # planner parameters tuned for this pattern and precision (autotune.py's
        # table), else the defaults
        self.plan = ct.device_plan(S, ct.autotune.lookup(S, precision))
        self.stats = self.plan.stats()
        self.fac = ct.DeviceFactor(self.plan, precision)
        self.h = self.fac.handle
        if precision == ct.MF_FP32 and solve_fp64:
            # FP32 factorization, FP64-arithmetic solves on the widened factor:
            # the refinement's corrections contract faster than with FP32 solves
            self.c(self.lib.mf_factor_set_solve_fp64(self.h, 1))
        self.precision = precision
        flat = ct.flatindices(S, A)
        self.c(self.lib.mf_factor_set_pattern(self.h, len(flat), flat.ctypes.data_as(_native.i64p)))
        self.stream = stream
        self.c(self.lib.mf_factor_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream)))
        dev = torch.device("cuda:0")
        self.nz = torch.tensor(np.ascontiguousarray(A.data), dtype=torch.float64, device=dev)
        self.B0 = torch.tensor(np.asfortranarray(B0).T.copy(), dtype=torch.float64, device=dev)
        self.X = self.B0.clone()
        self.perm = np.ascontiguousarray(P.perm, dtype=np.int64)
        self.iters = []
        self.relres = []

    def c(self, rc):
        self.ct.factorization._check(rc)

    def register_matrix(self):
        A = self.A
        colptr = np.ascontiguousarray(A.indptr, dtype=np.int64) + 1
        rowval = np.ascontiguousarray(A.indices, dtype=np.int64) + 1
        self.c(self.lib.mf_factor_set_matrix(self.h, self.n, colptr.ctypes.data_as(self._native.i64p),
                                             rowval.ctypes.data_as(self._native.i64p),
                                             ctypes.c_void_p(self.nz.data_ptr()), 1))

    def factor(self):
        self.c(self.lib.mf_factor_refresh(self.h, ctypes.c_void_p(self.nz.data_ptr()), 1))
        self.c(self.lib.mf_factorize_async(self.h, self.ct.MF_CHOLESKY, None, 0, 0.0, 0.0))

    def step_plain(self):
        self.factor()
        self.X.copy_(self.B0)
        self.c(self.lib.mf_solve(self.h, self.perm.ctypes.data_as(self._native.i64p),
                                 ctypes.c_void_p(self.X.data_ptr()), self.nrhs, self.n, 1))

    def step_refine(self, tol=1e-10, maxit=20, segments=None):
        """Factorization + refinement until every row segment (stacked
        system) meets tol."""
        self.factor()
        it, rr = ctypes.c_int(0), ctypes.c_double(0.0)
        seg = np.ascontiguousarray([0, self.n] if segments is None else segments, dtype=np.int64)
        self.c(self.lib.mf_refine_segmented(self.h, self.perm.ctypes.data_as(self._native.i64p),
                                            ctypes.c_void_p(self.B0.data_ptr()), self.n,
                                            ctypes.c_void_p(self.X.data_ptr()), self.n, self.nrhs, 1,
                                            tol, maxit, len(seg) - 1,
                                            seg.ctypes.data_as(self._native.i64p),
                                            ctypes.byref(it), ctypes.byref(rr)))
        self.iters.append(it.value)
        self.relres.append(rr.value)

    def info(self):
        info = ctypes.c_int64(0)
        self.c(self.lib.mf_factor_info(self.h, ctypes.byref(info)))
        return int(info.value)

    def solution(self):
        return self.X.T.contiguous().cpu().numpy()


def time_steps(torch, stream, step, steps, warmup):
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k in range(steps):
        step()
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = np.array([evs[k].elapsed_time(evs[k + 1]) for k in range(steps)])
    return wall * 1e3 / steps, step_ms


def read_timing(ct, lib, _native, h):
    nkc = len(ct.KC_NAMES)
    times = np.zeros(nkc)
    cnts = np.zeros(nkc, dtype=np.int64)
    work = np.zeros(nkc)
    ct.factorization._check(lib.mf_timing_read(h, times.ctypes.data_as(_native.f64p),
                                               cnts.ctypes.data_as(_native.i64p),
                                               work.ctypes.data_as(_native.f64p), 1))
    return times, cnts, work


def run_single(args, ct, torch, name):
    from cliquetrees_amd import _native
    W = workload(ct, name)
    A = W["A"]
    n = A.shape[0]
    t0 = time.perf_counter()
    P, S = ct.symbolic(A, alg=W["order"])
    t_sym = time.perf_counter() - t0
    log(f"{name}: n={n}, symbolic {t_sym:.2f} s")
    dev = torch.device("cuda:0")
    # a dedicated (non-null, default-priority) stream shared by torch and the engine
    stream = torch.cuda.Stream(dev)
    torch.cuda.set_stream(stream)
    rng = np.random.default_rng(W["seed"])
    B0 = rng.standard_normal((n, W["nrhs"]))
    E = Engine(ct, torch, A, P, S, W["precision"], W["nrhs"], B0, stream, W.get("solve_fp64", False))
    lib, h = E.lib, E.h
    step = E.step_refine if W["refine"] else E.step_plain
    if W["refine"]:
        E.register_matrix()

    log("device plan and factor object ready")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"{args.warmup} warmup steps done")
    if E.info() != 0:
        raise SystemExit(f"factorization failed, info={E.info()}")
    E.iters.clear()
    E.relres.clear()
    E.c(lib.mf_timing(h, 1))
    if args.dag_profile:
        E.c(lib.mf_dag_profile(h, 1))
    read_timing(ct, lib, _native, h)
    ms_per_step, step_ms = time_steps(torch, stream, step, args.steps, 0)
    log(f"{args.steps} timed steps: {ms_per_step:.3f} ms/step")
    times, cnts, work = read_timing(ct, lib, _native, h)
    E.c(lib.mf_timing(h, 0))
    if args.dag_profile:
        prof = np.zeros(21)
        E.c(lib.mf_dag_profile_read(h, prof.ctypes.data_as(_native.f64p), 1))
        for k, kname in enumerate(["ea", "small", "potrf", "trsm", "upd", "syrk"]):
            w, b, c = prof[3 * k:3 * k + 3] / args.steps
            print(f"dag {kname:6s} tasks/step {c:9.0f}  wait {w / 1e3:9.2f} ms  work {b / 1e3:9.2f} ms"
                  f"  work/task {b / max(c, 1):8.2f} us", file=sys.stderr)
        if args.trace_out:
            nt = int(E.stats["dag_tasks"])
            tasks = np.zeros(6 * nt, dtype=np.int32)
            tr = np.zeros(2 * nt, dtype=np.uint64)
            E.c(E.lib.mf_plan_dag(E.plan.handle, tasks.ctypes.data_as(ctypes.c_void_p), len(tasks)))
            E.c(E.lib.mf_dag_trace_read(h, tr.ctypes.data_as(ctypes.c_void_p), len(tr)))
            ng = int(E.stats["gemm_tasks"])
            gm = np.zeros(4 * max(ng, 1), dtype=np.int32)
            E.c(E.lib.mf_plan_gemm(E.plan.handle, gm.ctypes.data_as(ctypes.c_void_p), len(gm)))
            np.savez_compressed(args.trace_out, tasks=tasks.reshape(-1, 6), trace=tr.reshape(-1, 2),
                                gemm=gm.reshape(-1, 4))
        rel, acq, tick = prof[18:21] / args.steps
        print(f"dag sync: release {rel / 1e3:.2f} ms  acquire {acq / 1e3:.2f} ms  ticket {tick / 1e3:.2f} ms"
              f" (summed over workgroups, per step)", file=sys.stderr)
    info = E.info()
    X = E.solution()
    res = np.linalg.norm(A @ X - B0) / np.linalg.norm(B0)

    stats = E.stats
    ref_flops = float(stats["ref_flops"])
    kc = {k: i for i, k in enumerate(ct.KC_NAMES)}
    g = int(np.argmax(times))
    dom_ms = times[g] / max(cnts[g], 1)
    dom_flops_per_launch = work[g] / max(cnts[g], 1)
    achieved = dom_flops_per_launch / (dom_ms * 1e-3) / 1e12 if dom_ms > 0 else 0.0
    fp32 = W["precision"] == ct.MF_FP32
    peak = FP32_PEAK if fp32 else FP64_PEAK
    solve_ms = times[kc["solve"]] / max(args.steps, 1)
    factor_ms = sum(times[kc[c]] for c in ["small", "ea", "diag", "gemm", "copyout", "dag"]) / args.steps
    out = {
        "metric": METRIC,
        "value": round(ref_flops / (ms_per_step * 1e-3) / 1e12, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": W["desc"], "n": n, "nnz": int(A.nnz), "nrhs": W["nrhs"],
                   "fronts": stats["nfr"], "internal_fronts": stats["internal_fronts"],
                   "max_front": stats["max_nj"], "ref_flops": ref_flops,
                   "engine_flops": float(stats["engine_flops"]), "symbolic_s": round(t_sym, 3)},
        "p50_ms": round(float(np.percentile(step_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(step_ms, 99)), 4),
        "solves_per_s": round(W["nrhs"] * 1e3 / ms_per_step, 2),
        "solve_ms_per_step": round(float(solve_ms), 4),
        "factor_ms_per_step": round(float(factor_ms), 4),
        "factor_tflops": round(ref_flops / (factor_ms * 1e-3) / 1e12, 4) if factor_ms > 0 else None,
        "rel_residual": float(res),
        "info": info,
        "kernel_ms_per_step": {k: round(times[i] / args.steps, 4) for i, k in enumerate(ct.KC_NAMES)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4),
                     "traffic": traffic_from_profile("k_" + ct.KC_NAMES[g], W["tag"]),
                     "kernel": "k_" + ct.KC_NAMES[g], "launches": int(cnts[g]),
                     "flops_per_launch": float(dom_flops_per_launch),
                     "avg_launch_ms": round(float(dom_ms), 5)},
    }
    if W["refine"]:
        out["refinement"] = {"iters_mean": float(np.mean(E.iters)), "iters_max": int(np.max(E.iters)),
                             "relres_max": float(np.max(E.relres))}
        # the plain FP64 step on the same matrix, and the error against its solution
        log("plain FP64 step on the same matrix")
        fac64 = Engine(ct, torch, A, P, S, ct.MF_FP64, W["nrhs"], B0, stream)
        ms64, _ = time_steps(torch, stream, fac64.step_plain, max(3, args.steps // 2), 1)
        X64 = fac64.solution()
        out["plain_fp64"] = {"ms_per_step": round(ms64, 4),
                             "value": round(ref_flops / (ms64 * 1e-3) / 1e12, 4),
                             "rel_residual": float(np.linalg.norm(A @ X64 - B0) / np.linalg.norm(B0))}
        out["error_vs_fp64"] = float(np.linalg.norm(X - X64) / np.linalg.norm(X64))
        out["speedup_vs_plain_fp64"] = round(ms64 / ms_per_step, 3)
    if not args.no_cpu and not fp32:
        log("CPU baseline (oracle + host BLAS on a subtree sample)")
        out["cpu_baseline"] = cpu_baseline(ct, A, P, S)
        full = cpu_reference_from_profile(W["tag"])
        if full:
            out["cpu_baseline"]["whole_factor_and_solve"] = full
        log("CPU baseline done")
    return out


def cpu_reference_from_profile(tag):
    """The whole-workload CPU reference measurement committed under profiles/
    (``bench.py --workload cpu_reference`` on the MI355X machine's host), or None."""
    path = os.path.join(ROOT, "profiles", f"cpu_reference_{tag}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("json:"):
            d = json.loads(line.split(":", 1)[1])
            keep = ["factor_median_s", "factor_p99_s", "solve_median_s", "solve_p99_s", "value",
                    "unit", "cores", "nproc", "host", "steps"]
            out = {k: d[k] for k in keep if k in d}
            out["source"] = os.path.relpath(path, ROOT)
            return out
    return None


def compact(d):
    """The figures of a sub-workload's line that the headline line carries."""
    keep = ["value", "unit", "ms_per_step", "p50_ms", "p99_ms", "dtype", "speedup_vs_plain_fp64",
            "error_vs_fp64", "rel_residual", "refinement", "plain_fp64", "factor_ms_per_step",
            "solve_ms_per_step", "roofline", "fp64", "fp32_ir", "one_at_a_time",
            "worst_rel_residual", "max_rel_diff_batched_vs_alone", "kernel_ms_per_step", "mode",
            "root_column_vs_solve", "metric"]
    out = {k: d[k] for k in keep if k in d}
    out["workload"] = d["config"]["workload"]
    return out


def run_batched(args, ct, torch):
    """4,096 uneven systems as one forest (block-diagonal matrix, per-system
    orders offset into the block); one step = value refresh + factorization
    + solve of every system, device-resident, FP64 and FP32 + refinement."""
    specs = ct.problems.batch_specs(4096, 4)
    mats = [ct.problems.grid_laplacian_random(d, 4000 + k) for k, d in enumerate(specs)]
    orders = [ct.grid_nd(*(d if len(d) == 3 else (d[0], d[1], 1))) for d in specs]
    log("batch4096: 4096 systems generated")
    t0 = time.perf_counter()
    # the forest as BatchedFactorization builds it
    sizes = np.array([m.shape[0] for m in mats], dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    import scipy.sparse as sp
    A = sp.block_diag(mats, format="csc")
    A.sort_indices()
    order = np.concatenate([o + off for o, off in zip(orders, offsets[:-1])])
    P, S = ct.symbolic(A, alg=order)
    t_setup = time.perf_counter() - t0
    log(f"batch4096: forest analysed ({t_setup:.1f} s, n = {A.shape[0]})")
    rng = np.random.default_rng(4)
    B = rng.standard_normal((A.shape[0], 1))
    stream = torch.cuda.Stream(torch.device("cuda:0"))
    torch.cuda.set_stream(stream)
    res = {}
    st = None
    for label, prec in [("fp64", ct.MF_FP64), ("fp32_ir", ct.MF_FP32)]:
        E = Engine(ct, torch, A, P, S, prec, 1, B, stream)
        st = E.stats
        if prec == ct.MF_FP32:
            E.register_matrix()
            step = lambda: E.step_refine(segments=offsets)  # noqa: E731
        else:
            step = E.step_plain
        ms, step_ms = time_steps(torch, stream, step, args.steps, args.warmup)
        if E.info() != 0:
            raise SystemExit(f"batch4096 {label}: factorization failed, info={E.info()}")
        X = E.solution()[:, 0]
        worst = max(np.linalg.norm(m @ X[o:o + m.shape[0]] - B[o:o + m.shape[0], 0])
                    / np.linalg.norm(B[o:o + m.shape[0], 0]) for m, o in zip(mats, offsets[:-1]))
        log(f"batch4096 {label}: {ms:.2f} ms per batch, worst residual {worst:.2e}")
        res[label] = {"ms_per_batch": round(float(ms), 3),
                      "p50_ms": round(float(np.percentile(step_ms, 50)), 3),
                      "p99_ms": round(float(np.percentile(step_ms, 99)), 3),
                      "systems_per_s": round(4096 / (ms * 1e-3), 1),
                      "worst_rel_residual": float(worst)}
        if prec == ct.MF_FP32:
            res[label]["refine_iters_max"] = int(max(E.iters))
        del E
    best = max(res.values(), key=lambda r: r["systems_per_s"])
    return {
        "metric": "systems/s (factor + solve of every system) at <=1e-10 rel residual",
        "value": best["systems_per_s"], "unit": "systems/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_batch"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 | f32+f64", "data": "synthetic",
        "config": {"workload": "4096 grid Laplacians n~U{50..2000}, half 2-D/half 3-D, "
                               "w~U[0.5,1.5]+1e-3 I, geometric ND, one forest, 1 RHS each",
                   "n_total": int(A.shape[0]), "internal_fronts": st["internal_fronts"],
                   "ref_flops": float(st["ref_flops"]), "setup_s": round(t_setup, 2)},
        "fp64": res["fp64"], "fp32_ir": res["fp32_ir"],
    }


def run_serving(args, ct, torch):
    """Poisson stream of KKT LDLt requests with pattern reuse."""
    srv = ct.serving
    rng = np.random.default_rng(5)
    npat, updates = 6, 20
    # n_H log-uniform; capped at 3e4: random couplings make the fill grow
    # roughly quadratically (n_H = 2.7e4 already has a 1e4-wide front)
    nhs = np.exp(rng.uniform(np.log(1e3), np.log(3e4), size=npat)).astype(int)
    pats = []
    for p, nh in enumerate(nhs):
        K0, nh_, m = ct.problems.kkt_matrix(ct.problems.near_cubic_dims(int(nh), 3), seed=500 + p)
        pats.append((K0, nh_))
    reqs = []
    for p, (K0, nh_) in enumerate(pats):
        for u in range(updates):
            K = ct.problems.kkt_value_update(K0, nh_, seed=10000 * p + u)
            reqs.append(srv.Request(K=K, B=rng.standard_normal((K.shape[0], 1)), tag=(p, u)))
    order = rng.permutation(len(reqs))
    reqs = [reqs[i] for i in order]

    def fresh(rs, t):
        return [srv.Request(K=r.K, B=r.B, arrival=float(a), tag=r.tag) for r, a in zip(rs, t)]

    batched = getattr(args, "serving_mix", False)
    loop = srv.ServingLoop(diag=ct.MF_LDLT, max_batch=8, streams=4, hbm_budget=96 << 30)
    for K0, _ in pats:
        loop.prepare(K0, ks=(1, 2, 4, 8) if batched else (1,))
    log(f"kkt_serving: {npat} patterns analysed ({loop.analysis_s:.1f} s)")
    # warm every variant once
    loop.run(fresh(reqs[:len(pats) * 2], np.zeros(len(pats) * 2)), realtime=False, batching=batched)
    # one-at-a-time capacity: all requests visible, one stream, no batching
    one = loop.run(fresh(reqs, np.zeros(len(reqs))), realtime=False, batching=False)
    s_one = srv.latency_stats(one)
    log(f"kkt_serving: one-at-a-time pass {s_one['makespan_s']:.2f} s")
    rate = 1.5 * len(reqs) / s_one["makespan_s"]      # offered load above one-at-a-time capacity
    t_arr = srv.poisson_arrivals(len(reqs), rate, seed=5)
    seq = loop.run(fresh(reqs, t_arr), realtime=True, batching=False)
    log("kkt_serving: one-at-a-time Poisson pass done")
    if batched:
        # continuous batching on four streams under the Poisson arrivals
        mix = loop.run(fresh(reqs, t_arr), realtime=True, batching=True)
        log("kkt_serving: batched Poisson pass done")
        mode = "continuous batching, 4 streams, Poisson arrivals"
    else:
        # the default line reports the one-request-at-a-time server under the
        # Poisson arrivals: the batched run on four streams once stopped on the
        # engine's device watchdog and is kept out until that is understood
        mix = seq
        mode = "one request at a time, 1 stream, Poisson arrivals"
    s_mix, s_seq = srv.latency_stats(mix), srv.latency_stats(seq)
    worst = max(np.linalg.norm(r.K @ r.X - r.B) / np.linalg.norm(r.B) for r in mix)
    worst_dev = max(r.relres for r in mix)
    diff = max(np.linalg.norm(a.X - b.X) / np.linalg.norm(b.X) for a, b in zip(mix, seq))
    return {
        "metric": "serving solves/s at <=1e-10 rel residual; p50/p99 ms",
        "value": round(s_mix["solves_per_s"], 2), "unit": "solves/s", "n_gpus": 1,
        "steps": len(reqs), "warmup": len(pats) * 2, "ms_per_step": round(s_mix["mean_ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"KKT [H A'; A -1e-6 I] LDLt, {npat} patterns x {updates} updates, "
                               f"n_H log-uniform in [1e3, 3e4], MMD, Poisson arrivals at "
                               f"{rate:.1f}/s, 1 RHS",
                   "n_H": [int(x) for x in nhs], "analysis_s": round(loop.analysis_s, 2)},
        "p50_ms": round(s_mix["p50_ms"], 3), "p99_ms": round(s_mix["p99_ms"], 3), "mode": mode,
        "one_at_a_time": {"p50_ms": round(s_seq["p50_ms"], 3), "p99_ms": round(s_seq["p99_ms"], 3),
                          "solves_per_s": round(s_seq["solves_per_s"], 2),
                          "capacity_solves_per_s": round(s_one["solves_per_s"], 2)},
        "worst_rel_residual": float(worst), "worst_rel_residual_device": float(worst_dev),
        "refine_steps": loop.refine_steps, "max_rel_diff_batched_vs_alone": float(diff),
    }


def selinv_ref_flops(S):
    """Flops of selinv_impl! on the reference's own fronts (selinv.jl:122-260):
    trsm na nn^2, symm 2 na^2 nn, trrk na nn^2, trtri nn^3 / 3, syrk nn^3 / 3."""
    nn = S.nn().astype(np.float64)
    na = S.na().astype(np.float64)
    return float((2 * na ** 2 * nn + 2 * na * nn ** 2 + 2.0 / 3.0 * nn ** 3).sum())


def run_selinv(args, ct, torch, name="lap7_64"):
    """Selected inversion (selinv!) of the workload's FP64 Cholesky factor on
    the device: the factor is computed once, then each step is one inversion
    (mf_selinv) on the factor object's stream, timed by HIP events around the
    launch.  Checked at size by comparing one column of the selected inverse
    with a solve against the unit vector."""
    from cliquetrees_amd import _native
    W = workload(ct, name)
    A = W["A"]
    P, S = ct.symbolic(A, alg=W["order"])
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    lib = _native.hip()
    h = F._f.handle
    chk = ct.factorization._check
    for _ in range(args.warmup):
        chk(lib.mf_selinv(h))
    chk(lib.mf_factor_sync(h))
    chk(lib.mf_timing(h, 1))
    read_timing(ct, lib, _native, h)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        chk(lib.mf_selinv(h))
    chk(lib.mf_factor_sync(h))
    wall = (time.perf_counter() - t0) * 1e3 / args.steps
    times, cnts, work = read_timing(ct, lib, _native, h)
    chk(lib.mf_timing(h, 0))
    kc = ct.KC_NAMES.index("selinv")
    ms = times[kc] / max(1, cnts[kc])
    ct.selinv_(F)
    j = S.nfr - 1
    r0, nn = int(S.resptr[j] - 1), int(S.resptr[j + 1] - S.resptr[j])
    e = np.zeros(S.n)
    e[P.perm[r0] - 1] = 1.0
    x = ct.solve(F, e)[P.perm - 1]
    col = F.Dval[S.Dptr[j] - 1: S.Dptr[j] - 1 + nn]
    err = float(np.linalg.norm(col - x[r0:r0 + nn]) / np.linalg.norm(x[r0:r0 + nn]))
    flops = selinv_ref_flops(S)
    log(f"selinv {name}: {ms:.3f} ms per inversion, root column vs solve {err:.2e}")
    return {"metric": "selected inversion TFLOP/s (reference selinv_impl! flops)",
            "value": round(flops / (ms * 1e-3) / 1e12, 3), "unit": "TFLOP/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
            "wall_ms_per_step": round(wall, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": W["desc"].split(", 16 RHS")[0] + ", selinv! of its Cholesky factor",
                       "ref_flops": flops, "engine_flops": float(work[kc] / max(1, cnts[kc]))},
            "root_column_vs_solve": err}


def complete_ref_flops(S):
    """Flops of complete_impl! on the reference's own fronts
    (complete.jl:140-260): potrf of Sigma(sep, sep) na^3 / 3, two trsm with
    it 2 na^2 nn, syrk na nn^2 / 2, potrf of the Schur complement nn^3 / 3,
    trsm na nn^2, trtri nn^3 / 3 multiply-adds, two flops each."""
    nn = S.nn().astype(np.float64)
    na = S.na().astype(np.float64)
    return float(2 * (na ** 3 / 3 + 2 * na ** 2 * nn + 1.5 * na * nn ** 2 + 2.0 / 3.0 * nn ** 3).sum())


def run_complete(args, ct, torch, name="lap7_64"):
    """Completion (complete!) of the selected inverse of the workload's FP64
    Cholesky factor through the drop-in mf_complete_impl: host arrays in and
    out (the PCIe transfers are inside the timed call), the first call
    (forest plans, X schedules and the device arena) untimed.  Checked at
    size: the result is the factor again."""
    W = workload(ct, name)
    A = W["A"]
    P, S = ct.symbolic(A, alg=W["order"])
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    D0, L0 = F.Dval, F.Lval
    ct.selinv_(F)
    SD, SL = F.Dval, F.Lval
    t0 = time.perf_counter()
    ct.complete_impl(S, SD, SL)
    first = time.perf_counter() - t0
    log(f"complete {name}: first call {first:.2f} s")
    for _ in range(max(0, args.warmup - 1)):
        ct.complete_impl(S, SD, SL)
    ts = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        CD, CL, info = ct.complete_impl(S, SD, SL)
        ts.append((time.perf_counter() - t0) * 1e3)
        log(f"complete {name}: {ts[-1]:.1f} ms")
    err = max(float(np.abs(CD - D0).max() / np.abs(D0).max()), float(np.abs(CL - L0).max() / np.abs(L0).max()))
    ms = float(np.median(ts))
    flops = complete_ref_flops(S)
    return {"metric": "completion TFLOP/s (reference complete_impl! flops, host arrays in and out)",
            "value": round(flops / (ms * 1e-3) / 1e12, 3), "unit": "TFLOP/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 1),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": W["desc"].split(", 16 RHS")[0] + ", complete! of its selected inverse",
                       "ref_flops": flops, "first_call_s": round(first, 2)},
            "info": int(info), "max_rel_diff_vs_factor": err}


def run_copy(args, ct, torch):
    """A plain device copy of 1 GiB (read 1 GiB, write 1 GiB): the reference
    kernel whose FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic
    counters of the engine's kernels."""
    n = (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    ms, _ = time_steps(torch, stream, lambda: dst.copy_(src), args.steps, args.warmup)
    return {"metric": "device copy GB/s (read + write)", "value": round(2 * (1 << 30) / (ms * 1e-3) / 1e9, 1),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "torch copy_ of 1 GiB float64 (copy-kernel calibration)",
                       "bytes_read": 1 << 30, "bytes_written": 1 << 30}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="lap7_64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="lap7_64 alone (without the other workloads in the same line)")
    ap.add_argument("--dag-profile", action="store_true",
                    help="per task kind wait/work time of the tile-DAG kernel (stderr)")
    ap.add_argument("--serving-mix", action="store_true",
                    help="kkt_serving: continuous batching under the Poisson arrivals")
    ap.add_argument("--trace-out", default=None,
                    help="with --dag-profile: save the per-ticket start/end times (npz)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine runs "
              f"on a single MI355X and has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    ct = load_package()
    if args.workload == "cpu_reference":
        print(json.dumps(run_cpu_reference(args, ct)))
        return
    import torch
    if args.workload == "copy":
        out = run_copy(args, ct, torch)
    elif args.workload == "batch4096":
        out = run_batched(args, ct, torch)
    elif args.workload == "kkt_serving":
        out = run_serving(args, ct, torch)
    elif args.workload == "lap7_64_selinv":
        out = run_selinv(args, ct, torch, "lap7_64")
    elif args.workload == "lap7_64_complete":
        out = run_complete(args, ct, torch, "lap7_64")
    else:
        out = run_single(args, ct, torch, args.workload)
        if args.workload == "lap7_64" and not args.headline_only:
            # the other configurations, measured in the same run (fewer steps)
            sub = argparse.Namespace(**vars(args))
            sub.steps, sub.warmup, sub.no_cpu = max(3, args.steps // 2), 1, True
            sub.dag_profile, sub.trace_out = False, None
            out["workloads"] = {}
            for name in ["lap7_64_mixed", "st27_48", "st27_48_mixed", "batch4096", "kkt_serving",
                         "lap7_64_selinv"]:
                t0 = time.perf_counter()
                log(f"sub-workload {name}")
                try:
                    if name == "batch4096":
                        r = run_batched(sub, ct, torch)
                    elif name == "kkt_serving":
                        r = run_serving(sub, ct, torch)
                    elif name == "lap7_64_selinv":
                        r = run_selinv(sub, ct, torch, "lap7_64")
                    else:
                        r = run_single(sub, ct, torch, name)
                    out["workloads"][name] = compact(r)
                except Exception as e:  # the headline line is still printed
                    log(f"sub-workload {name} failed: {e}")
                    out["workloads"][name] = {"error": str(e)}
                out["workloads"][name]["wall_s"] = round(time.perf_counter() - t0, 1)
                torch.cuda.empty_cache()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
