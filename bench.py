"""Benchmark of the multifrontal numeric engine on one MI355X.

Workload (default): plain FP64 sparse Cholesky of the 3-D 7-point weighted
Laplacian on a 64^3 grid (n = 262,144; edge weights i.i.d. U[0.5, 1.5],
seed 2, plus 1e-3 I), geometric nested-dissection order, 16 right-hand sides
N(0,1) (seed 2).  One step = assembly of A's values into the factor storage
(setflatindex! of every nonzero, device-resident values) + numeric
factorization + a 16-RHS solve, all on the GPU.

  value        = reference factor flops / step time, TFLOP/s (whole step:
                 assembly + factorization + solve), with the flop count of
                 the reference's own supernode structure:
                 sum_j nn^3/3 + na nn^2 + na^2 nn.
  roofline     = the dominant kernel (k_dag, the persistent tile-DAG of the
                 big fronts, FP64 MFMA) timed with HIP events on the
                 engine's stream over the timed region; its algorithmic work
                 is the engine flop count of the fronts it factors.
  cpu_baseline = the CPU oracle (restatement of the reference's serial front
                 loop) with host OpenBLAS on the same problem, timed on the
                 host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg2|cfg3]
"""
from __future__ import annotations

import argparse
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
    if name == "cfg2":
        A = ct.problems.laplacian3d_7pt(64, 64, 64, seed=2)
        return dict(A=A, order=ct.grid_nd(64, 64, 64), nrhs=16, seed=2, precision=ct.MF_FP64,
                    desc="3D 7-point Laplacian 64^3, w~U[0.5,1.5] + 1e-3 I, geometric ND, "
                         "16 RHS, plain FP64 Cholesky")
    if name == "cfg3":
        A = ct.problems.stencil3d_27pt(48, 48, 48, seed=3)
        return dict(A=A, order=ct.grid_nd(48, 48, 48), nrhs=16, seed=3, precision=ct.MF_FP64,
                    desc="3D 27-point stencil 48^3, diagonally dominant, geometric ND, 16 RHS, FP64")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(ct, A, P, S, frac=0.125):
    """The oracle (restated serial front loop, its dense kernels routed to host
    OpenBLAS as the reference does for Float64) on a bounded sample of the
    same factorization: the largest elimination subtree holding at most
    ``frac`` of the reference flops (a sub-box of the nested dissection)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.use_host_blas(True)
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
    dt, flops, info = oracle.chol_subtree(S, Dv, Lv, root)
    cores = os.cpu_count() or 1
    try:
        from threadpoolctl import threadpool_info
        cores = max(int(i.get("num_threads", 1)) for i in threadpool_info()) or cores
    except Exception:
        pass
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sample": f"subtree of front {root} ({flops / total:.1%} of the factor flops, "
                      f"{flops:.3e} flops, one run {dt:.2f} s, info={info}): oracle.c serial front "
                      f"loop + OpenBLAS dpotrf/dtrsm/dsyrk/dgemm"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dag-profile", action="store_true",
                    help="per task kind wait/work time of the tile-DAG kernel (stderr)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine runs "
              f"on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import torch
    import ctypes
    ct = load_package()
    from cliquetrees_amd import _native
    lib = _native.hip()

    W = workload(ct, args.workload)
    A = W["A"]
    n = A.shape[0]
    t0 = time.perf_counter()
    P, S = ct.symbolic(A, alg=W["order"])
    t_sym = time.perf_counter() - t0
    plan = ct.device_plan(S)
    stats = plan.stats()
    fac = ct.DeviceFactor(plan, W["precision"])
    h = fac.handle
    flat = ct.flatindices(S, A)
    _c = lambda rc: ct.factorization._check(rc)
    _c(lib.mf_factor_set_pattern(h, len(flat), flat.ctypes.data_as(_native.i64p)))

    dev = torch.device("cuda:0")
    # a dedicated (non-null, default-priority) stream shared by torch and the engine
    stream = torch.cuda.Stream(dev)
    torch.cuda.set_stream(stream)
    _c(lib.mf_factor_set_stream(h, ctypes.c_void_p(stream.cuda_stream)))
    nz = torch.tensor(np.ascontiguousarray(A.data), dtype=torch.float64, device=dev)
    rng = np.random.default_rng(W["seed"])
    B0 = rng.standard_normal((n, W["nrhs"]))
    Bdev0 = torch.tensor(np.asfortranarray(B0).T.copy(), dtype=torch.float64, device=dev)  # (nrhs, n)
    Bdev = Bdev0.clone()
    perm = np.ascontiguousarray(P.perm, dtype=np.int64)
    perm_p = perm.ctypes.data_as(_native.i64p)
    nrhs = W["nrhs"]

    def step():
        _c(lib.mf_factor_refresh(h, ctypes.c_void_p(nz.data_ptr()), 1))
        _c(lib.mf_factorize_async(h, ct.MF_CHOLESKY, None, 0, 0.0, 0.0))
        Bdev.copy_(Bdev0)
        _c(lib.mf_solve(h, perm_p, ctypes.c_void_p(Bdev.data_ptr()), nrhs, n, 1))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    info = ctypes.c_int64(0)
    _c(lib.mf_factor_info(h, ctypes.byref(info)))
    if info.value != 0:
        raise SystemExit(f"factorization failed, info={info.value}")

    _c(lib.mf_timing(h, 1))
    if args.dag_profile:
        _c(lib.mf_dag_profile(h, 1))
    nkc = len(ct.KC_NAMES)
    times = np.zeros(nkc)
    cnts = np.zeros(nkc, dtype=np.int64)
    work = np.zeros(nkc)
    _c(lib.mf_timing_read(h, times.ctypes.data_as(_native.f64p), cnts.ctypes.data_as(_native.i64p),
                          work.ctypes.data_as(_native.f64p), 1))
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k in range(args.steps):
        step()
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = np.array([evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps)])
    _c(lib.mf_timing_read(h, times.ctypes.data_as(_native.f64p), cnts.ctypes.data_as(_native.i64p),
                          work.ctypes.data_as(_native.f64p), 1))
    _c(lib.mf_timing(h, 0))
    if args.dag_profile:
        prof = np.zeros(18)
        _c(lib.mf_dag_profile_read(h, prof.ctypes.data_as(_native.f64p), 1))
        for k, name in enumerate(["ea", "small", "potrf", "trsm", "upd", "syrk"]):
            w, b, c = prof[3 * k:3 * k + 3] / args.steps
            print(f"dag {name:6s} tasks/step {c:9.0f}  wait {w / 1e3:9.2f} ms  work {b / 1e3:9.2f} ms"
                  f"  work/task {b / max(c, 1):8.2f} us", file=sys.stderr)
    _c(lib.mf_factor_info(h, ctypes.byref(info)))

    # accuracy of the last step
    X = Bdev.T.contiguous().cpu().numpy()
    res = np.linalg.norm(A @ X - B0) / np.linalg.norm(B0)

    ms_per_step = wall * 1e3 / args.steps
    ref_flops = float(stats["ref_flops"])
    kc = {name: i for i, name in enumerate(ct.KC_NAMES)}
    # dominant kernel: the one with the largest share of the timed region
    g = int(np.argmax(times))
    dom_ms = times[g] / max(cnts[g], 1)
    dom_flops_per_launch = work[g] / max(cnts[g], 1)
    achieved = dom_flops_per_launch / (dom_ms * 1e-3) / 1e12 if dom_ms > 0 else 0.0
    peak = FP64_PEAK if W["precision"] == ct.MF_FP64 else FP32_PEAK
    solve_ms = times[kc["solve"]] / max(cnts[kc["solve"]], 1)
    factor_ms = sum(times[kc[c]] for c in ["small", "ea", "diag", "gemm", "copyout", "dag"]) / args.steps
    out = {
        "metric": "factor TFLOP/s + solves/s at <=1e-10 rel residual vs plain FP64; p50/p99 ms",
        "value": round(ref_flops / (ms_per_step * 1e-3) / 1e12, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if W["precision"] == ct.MF_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": W["desc"], "n": n, "nnz": int(A.nnz), "nrhs": nrhs,
                   "fronts": stats["nfr"], "internal_fronts": stats["internal_fronts"],
                   "max_front": stats["max_nj"], "ref_flops": ref_flops,
                   "engine_flops": float(stats["engine_flops"]), "symbolic_s": round(t_sym, 3)},
        "p50_ms": round(float(np.percentile(step_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(step_ms, 99)), 4),
        "solves_per_s": round(nrhs / (solve_ms * 1e-3), 2) if solve_ms > 0 else None,
        "factor_tflops": round(ref_flops / (factor_ms * 1e-3) / 1e12, 4) if factor_ms > 0 else None,
        "rel_residual": float(res),
        "info": int(info.value),
        "kernel_ms_per_step": {name: round(times[i] / args.steps, 4) for i, name in enumerate(ct.KC_NAMES)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": None,
                     "kernel": "k_" + ct.KC_NAMES[g], "launches": int(cnts[g]),
                     "flops_per_launch": float(dom_flops_per_launch),
                     "avg_launch_ms": round(float(dom_ms), 5)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ct, A, P, S)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
