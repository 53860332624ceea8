"""Micro-benchmark of the drop-in solve over the reference's benchmark matrix.

The reference's GPU benchmark (benchmarks/cuda/my_bm.cu:1-74, Google
Benchmark) times one pentadsolver_gpsv_batch call per iteration for

    N = 32, 64, 128, 256 (range multiplier 2), FP32 and FP64,
    ndims 1 (solvedim 0), ndims 2 (solvedim 0, 1), ndims 3 (solvedim 0, 1, 2),

on an N^ndims grid, with the scratch buffer allocated outside the timed loop
and the right-hand side restored outside the timed region after every call
(benchmarks/cuda/my_bm.cu:18-45).  This harness runs the same matrix through
this library's C ABI: every case is timed call by call with HIP events on
the handle's stream (warm-up calls first, the restore of x between calls not
timed), reported as the median time per call, rows per second and the
algorithmic HBM rate (7 words per row).  Each case's first solve is also
checked against the FP64 CPU oracle (forward error), when the oracle is
importable (it is test infrastructure; the timing never uses it).

Inputs are this library's seeded diagonally dominant recipe
(workloads.dominant_torch, ratio U(2, 3)) rather than the reference's Mesh
random fill (include/util/mesh.hpp:106-120), which is not diagonally
dominant and whose per-thread seeding depends on the thread count.

Usage:  python microbench.py [--reps 20] [--max-n 256]
Prints a table and, with --json, one JSON object per case.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import statistics
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "pentadiagonal-solver_amd")


def _pkg():
    if "pentadsolver_amd" in sys.modules:
        return sys.modules["pentadsolver_amd"]
    spec = importlib.util.spec_from_file_location(
        "pentadsolver_amd", os.path.join(PKG, "__init__.py"), submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pentadsolver_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def matrix(max_n: int = 256):
    """(N, ndims, solvedim) cases of benchmarks/cuda/my_bm.cu:55-71."""
    out = []
    n = 32
    while n <= max_n:
        out.append((n, 1, 0))
        out += [(n, 2, sd) for sd in (0, 1)]
        out += [(n, 3, sd) for sd in (0, 1, 2)]
        n *= 2
    return out


def run_case(P, W, torch, n, ndims, sd, dtype, reps, warmup, check, handle):
    dims = [n] * ndims
    ds, dl, d, du, dw, rhs = W.dominant_torch(dims, seed=1000 + 10 * ndims + sd, dtype=dtype)
    x = rhs.clone()
    nbytes = P.gpsv_batch_buffer_size_ext(ds, dl, d, du, dw, x, dims, sd, handle)
    buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda") if nbytes else None
    err = None
    P.gpsv_batch(ds, dl, d, du, dw, x, dims, sd, buf, handle)
    torch.cuda.synchronize()
    if check is not None:
        import numpy as np
        a = [t.double().cpu().numpy() for t in (ds, dl, d, du, dw)]
        ref = check.solve(*a, rhs.double().cpu().numpy(), dims, sd)
        got = x.double().cpu().numpy()
        err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    stream = torch.cuda.current_stream()
    times = []
    for it in range(warmup + reps):
        x.copy_(rhs)  # restore, outside the timed region
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        P.gpsv_batch(ds, dl, d, du, dw, x, dims, sd, buf, handle, check=False)
        e1.record(stream)
        e1.synchronize()
        if it >= warmup:
            times.append(e0.elapsed_time(e1))
    ms = statistics.median(times)
    rows = n ** ndims
    esz = 8 if dtype == torch.float64 else 4
    return {
        "N": n, "ndims": ndims, "solvedim": sd,
        "dtype": "f64" if esz == 8 else "f32",
        "kernel": handle.kernel_config(sd if ndims == 3 or sd == 0 else 2),
        "ms_per_call": ms, "grows": rows / (ms * 1e-3) / 1e9,
        "gbs": 7 * esz * rows / (ms * 1e-3) / 1e9,
        "scratch_bytes": nbytes, "max_rel_err_vs_fp64_oracle": err,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--max-n", type=int, default=256)
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--json", action="store_true")
    args = ap.parse_args(argv)
    import torch
    P = _pkg()
    W = importlib.import_module("pentadsolver_amd.workloads")
    check = None
    if not args.no_check:
        try:
            if ROOT not in sys.path:
                sys.path.insert(0, ROOT)
            from oracle import oracle as check  # checker only
        except ImportError:
            check = None
    if not torch.cuda.is_available():
        raise SystemExit("microbench needs a GPU")
    rows = []
    with P.Handle() as h:
        for dtype in (torch.float32, torch.float64):
            for n, nd, sd in matrix(args.max_n):
                r = run_case(P, W, torch, n, nd, sd, dtype, args.reps, args.warmup, check, h)
                rows.append(r)
                if args.json:
                    print(json.dumps(r), flush=True)
                else:
                    e = r["max_rel_err_vs_fp64_oracle"]
                    print(f"{r['dtype']} N={n:4d} ndims={nd} solvedim={sd}  "
                          f"{r['ms_per_call'] * 1e3:9.1f} us  {r['grows']:8.3f} GRows/s  "
                          f"{r['gbs']:8.1f} GB/s  {r['kernel']:<36s} "
                          f"err {'-' if e is None else f'{e:.1e}'}", flush=True)
    return rows


if __name__ == "__main__":
    main()
