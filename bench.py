"""Benchmark: three-direction ADI step (x, then y, then z, in place) on one MI355X.

Prints ONE JSON line.  A "step" is one x->y->z sweep over the whole grid
(inputs resident in HBM); rows = unknowns solved, summed over directions.
The timed region brackets exactly --steps steps with synchronisation on
both sides; the per-direction kernel durations come from HIP events
recorded on the stream the kernels run on (the handle's stream is torch's
current stream).  After timing, every direction is re-solved from a fresh
right-hand side and a sample of lines is checked against the CPU oracle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 512]
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
PKG = os.path.join(ROOT, "pentadiagonal-solver_amd")
sys.path.insert(0, ROOT)

METRIC = "GRows/s + % HBM roofline, FP32+refinement vs plain FP64 at 1e-12 rel err"
HBM_PEAK_GBS = 8000.0


def import_pkg():
    spec = importlib.util.spec_from_file_location(
        "pentadsolver_amd", os.path.join(PKG, "__init__.py"),
        submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pentadsolver_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sample_lines(dims, sd, count, seed):
    """Flat indices (count, len) of `count` random lines along sd."""
    rng = np.random.default_rng(seed)
    strides = [1]
    for v in dims[:-1]:
        strides.append(strides[-1] * v)
    out = []
    for _ in range(count):
        pos = [int(rng.integers(0, v)) for v in dims]
        pos[sd] = 0
        base = sum(p * s for p, s in zip(pos, strides))
        out.append(base + strides[sd] * np.arange(dims[sd]))
    return np.array(out)


def run_mode(P, torch, dims, dtype, steps, warmup, seed):
    """Time the ADI step and each direction; return a dict of results."""
    from pentadsolver_amd import workloads as W
    n = int(np.prod(dims))
    arrs = W.dominant_torch(dims, seed, dtype=dtype)
    ds, dl, d, du, dw, b = arrs
    x = b.clone()
    stream = torch.cuda.current_stream()
    h = P.Handle(stream=stream)
    h.freeze_env(True)
    bufs = []
    for sd in range(3):
        nb = P.gpsv_batch_buffer_size_ext(ds, dl, d, du, dw, x, dims, sd, h)
        bufs.append(torch.empty(max(nb, 1), dtype=torch.uint8, device="cuda") if nb else None)

    def step():
        for sd in range(3):
            P.gpsv_batch(ds, dl, d, du, dw, x, dims, sd, bufs[sd], h, check=False)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    ev_ms = ev0.elapsed_time(ev1) / steps
    status = h.last_status()
    if status != 0:
        raise RuntimeError(f"solve failed, status {status}")

    # per-direction kernel durations (pass B)
    per_dir = []
    for sd in range(3):
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        for a, z in evs:
            a.record(stream)
            P.gpsv_batch(ds, dl, d, du, dw, x, dims, sd, bufs[sd], h, check=False)
            z.record(stream)
        torch.cuda.synchronize()
        per_dir.append(sum(a.elapsed_time(z) for a, z in evs) / steps)
    kernels = [h.kernel_config(k) for k in range(3)]

    # sampled-line parity at full size, each direction from a fresh RHS
    from oracle import oracle as O
    err = 0.0
    for sd in range(3):
        x.copy_(b)
        P.gpsv_batch(ds, dl, d, du, dw, x, dims, sd, bufs[sd], h)
        torch.cuda.synchronize()
        lines = sample_lines(dims, sd, 32, 100 + sd)
        idx = torch.from_numpy(lines.reshape(-1)).cuda()
        got = x[idx].double().cpu().numpy().reshape(lines.shape)
        coeffs = [a[idx].double().cpu().numpy().reshape(lines.shape) for a in (ds, dl, d, du, dw, b)]
        for j in range(lines.shape[0]):
            line = [np.ascontiguousarray(c[j]) for c in coeffs]
            ref = O.solve_ld(*line, [dims[sd]], 0)
            e = float(np.max(np.abs(got[j] - ref)) / np.max(np.abs(ref)))
            err = max(err, e)
    h.close()
    del arrs, ds, dl, d, du, dw, b, x, bufs
    torch.cuda.empty_cache()
    return {"ms_per_step": ev_ms, "wall_ms_per_step": wall_ms, "dir_ms": per_dir,
            "kernels": kernels, "max_rel_err": err, "elem_bytes": torch.finfo(dtype).bits // 8}


def cpu_baseline(dims, seed, iters=2):
    """The reference's own OpenMP+AVX2 CPU solver (oracle/_ref, built from its
    sources) timed on this host, FP64, method of apps/adi_cpu.cpp:49-103."""
    from oracle import oracle as O
    from pentadsolver_amd import workloads as W
    kind = "reference" if O.ref_available("avx") else "port"
    arrs = W.dominant_np(dims, seed)
    x = arrs[5].copy()
    solve = (lambda sd: O.ref_gpsv_batch(*arrs[:5], x, dims, sd)) if kind == "reference" \
        else (lambda sd: O.gpsv_batch(*arrs[:5], x, dims, sd))
    for sd in range(3):  # warm-up, faults in every page
        solve(sd)
    t = np.zeros(3)
    for _ in range(iters):
        for sd in range(3):
            t0 = time.perf_counter()
            solve(sd)
            t[sd] += time.perf_counter() - t0
    t /= iters
    rows = 3 * int(np.prod(dims))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": rows / t.sum() / 1e9, "unit": "GRows/s", "cores": threads, "kind": kind,
            "sample": f"FP64 ADI step x/y/z on {dims[0]}x{dims[1]}x{dims[2]}, {iters} timed "
                      f"steps after 1 warm-up, ms per direction "
                      f"{t[0]*1e3:.1f}/{t[1]*1e3:.1f}/{t[2]*1e3:.1f}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the per-GPU path would shard whole lines across GPUs, see DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    import torch
    P = import_pkg()
    P.load_library()
    torch.cuda.set_device(0)
    dims = [args.size] * 3
    n = int(np.prod(dims))
    rows = 3 * n

    log(f"bench: fp64 ADI step {dims}")
    r64 = run_mode(P, torch, dims, torch.float64, args.steps, args.warmup, seed=3)
    log(f"bench: fp64 {r64['ms_per_step']:.3f} ms/step dirs {r64['dir_ms']} err {r64['max_rel_err']:.2e}")
    log(f"bench: fp32 ADI step {dims}")
    r32 = run_mode(P, torch, dims, torch.float32, args.steps, args.warmup, seed=3)
    log(f"bench: fp32 {r32['ms_per_step']:.3f} ms/step dirs {r32['dir_ms']} err {r32['max_rel_err']:.2e}")

    head = r64
    bpr = 7 * head["elem_bytes"]  # 5 coefficients + RHS read, solution written
    dom = int(np.argmax(head["dir_ms"]))
    achieved = bpr * n / (head["dir_ms"][dom] * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": rows / (head["ms_per_step"] * 1e-3) / 1e9,
        "unit": "GRows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ADI step x->y->z, {dims[0]}^3 grid, "
                               f"{n // dims[0]} systems x {dims[0]} per direction, "
                               "diagonally dominant seeded coefficients",
                   "grid": dims, "mode": "plain FP64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"dir {dom}: {head['kernels'][dom]}",
                     "bytes_per_row": bpr},
        "max_rel_err_vs_oracle": head["max_rel_err"],
        "per_direction_ms": head["dir_ms"],
        "kernels": head["kernels"],
        "wall_ms_per_step": head["wall_ms_per_step"],
        "fp32": {"value": rows / (r32["ms_per_step"] * 1e-3) / 1e9,
                 "ms_per_step": r32["ms_per_step"], "per_direction_ms": r32["dir_ms"],
                 "kernels": r32["kernels"], "max_rel_err_vs_fp64_oracle": r32["max_rel_err"]},
    }
    if not args.no_cpu:
        log("bench: cpu baseline")
        result["cpu_baseline"] = cpu_baseline(dims, seed=3)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
