"""Benchmark: the three-direction ADI step (x, then y, then z) on one MI355X.

Prints ONE JSON line (diagnostics go to stderr).

Workload: a 512^3 grid, 262144 systems of length 512 per direction,
diagonally dominant seeded coefficients (off-diagonals U(-1,1), main
diagonal = U(2,3) * sum|off|, RHS U(-1,1)), the same five coefficient arrays
for every direction.  A "step" is one x->y->z sweep over the whole grid with
the field and the coefficients resident in HBM; rows = unknowns solved,
summed over the three directions.

Headline ("value"): FP32 bulk solve + FP64 iterative refinement (coefficients
stored in FP32, field in FP64).  Beside it: plain FP64 and plain FP32 on the
same driver, and the in-place drop-in per-direction solves (FP64, and the
mixed-precision entry point: FP32 coefficients, FP64 field).  Each mode
captures --steps steps as one hipGraph (after a --warmup-step graph) and the
timed region brackets exactly one replay of it (--steps steps) with a
synchronisation on both sides.  Per-direction kernel
durations come from HIP events recorded on the handle's stream (torch's
current stream), and the dominant kernel gives the roofline.  After timing,
one step is re-run from the initial field and compared on the full grid with
the CPU oracle: FP64 against the FP64 oracle, FP32+refinement against the
FP64 oracle of the same FP32-stored coefficients widened to FP64.

Beside the headline line's own fields, "weak_set" repeats the refinement
step on the weakly dominant set (c = 1.05 sum|off|), and "workloads" holds
the other configurations, each timed and checked against the oracle:
256^3 x 50 iterations (FP64, FP32, and the constant-coefficient shared-fact
variant; FP64 also against the reference's own CPU solver), 8192 systems of
log-uniform length in [32, 8192] in one launch vs one call per system, and
the 384x256x640 hyperdiffusion run (100 steps as one hipGraph, FP32+FP64
refinement to --c5-tol vs FP64, the first --c5-cpu-steps steps also on the
CPU; and the same operators as structured coefficients, generated in the
kernels from one scale per line, FP64 and FP32), and the 2-D cavity
consumer's pressure-correction step (1024 x 1024 and 1024 x 8192, the
reference's assembled arrays vs the fused solve; at 1024 x 1024 also the
reference's CPU pipeline).
--no-extra skips them.

Usage: python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--size 512]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

# The CPU baseline's OpenMP placement (method of the reference's CPU
# measurements): set before any library that links an OpenMP runtime loads.
os.environ.setdefault("OMP_PROC_BIND", "spread")
os.environ.setdefault("OMP_PLACES", "cores")
# CPUs this job may use, read before an OpenMP runtime binds the main thread
try:
    AFFINITY_CPUS = len(os.sched_getaffinity(0))
except AttributeError:
    AFFINITY_CPUS = None

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "pentadiagonal-solver_amd")
sys.path.insert(0, ROOT)

METRIC = "GRows/s + % HBM roofline, FP32+refinement vs plain FP64 at 1e-12 rel err"
HBM_PEAK_GBS = 8000.0
BYTES_PER_ROW = {"ir": 36, "fp64": 56, "fp32": 28}
# HBM bytes per launch measured with rocprofv3 (FETCH_SIZE x 2 + WRITE_SIZE, in
# KiB per dispatch; the factor 2 calibrated on the 6-read/1-write probe of
# known size), keyed by (kernel configuration, line length).  Kept under
# profiles/ with the commands that produced them.
MEASURED_TRAFFIC_BYTES = {
    ("ir thomas-pcr rows=8 lanes=64", 512): (3_670_240 + 1_048_579) * 1024,
    ("ir interleaved thomas-pcr rows=8 lanes=64", 512): (3_670_252 + 1_048_619) * 1024,
    ("thomas-pcr rows=8 lanes=64", 512): (6_294_100 + 1_048_576) * 1024,
    ("thomas-pcr rows=16 lanes=32", 512): (3_150_078 + 524_289) * 1024,
}


def import_pkg():
    if "pentadsolver_amd" in sys.modules:
        return sys.modules["pentadsolver_amd"]
    spec = importlib.util.spec_from_file_location(
        "pentadsolver_amd", os.path.join(PKG, "__init__.py"),
        submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pentadsolver_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_adi(P, torch, mode_name, coeffs, b, dims, steps, warmup, h, max_refine, tol,
             warmup_ms=1500.0):
    """Graph-replay timing of the ADI step for one mode; returns a dict.
    The warm-up graph (`warmup` steps) is replayed until at least
    `warmup_ms` of wall time has passed (the reference's warm-up rule,
    apps/adi_cuda.cu:403-435)."""
    mode = {"fp64": P.ADI_FP64, "fp32": P.ADI_FP32, "ir": P.ADI_IR}[mode_name]
    n = int(np.prod(dims))
    stream = torch.cuda.current_stream()
    x = b.clone()
    adi = P.Adi(mode, dims, [coeffs], handle=h, max_refine=max_refine, tol=tol)
    # warm-up graph of `warmup` steps, then the timed graph of `steps` steps
    warm_wall = 0.0
    warm_replays = 0
    if warmup > 0:
        adi.capture(x, warmup)
        tw = time.perf_counter()
        while True:
            adi.replay(1)
            torch.cuda.synchronize()
            warm_replays += 1
            warm_wall = (time.perf_counter() - tw) * 1e3
            if warm_wall >= warmup_ms:
                break
    x.copy_(b)
    adi.capture(x, steps)
    torch.cuda.synchronize()
    if mode_name == "ir":
        adi.refine_steps()  # reset the counter
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    adi.replay(1)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ms = ev0.elapsed_time(ev1) / steps
    refine = adi.refine_steps() if mode_name == "ir" else None
    x.copy_(b)
    dir_ms = adi.direction_ms(x, reps=10)
    # one step from the initial field for the accuracy check
    x.copy_(b)
    adi.step(x, 1)
    torch.cuda.synchronize()
    if mode_name == "ir":
        adi.refine_steps()  # reset, then count the checked step alone
        x.copy_(b)
        adi.step(x, 1)
        refine_checked = adi.refine_steps()
    else:
        refine_checked = None
    result = x.double().cpu().numpy()
    cfg = [adi.config(k) for k in range(len(dims))]
    adi.close()
    del x
    bpr = BYTES_PER_ROW[mode_name]
    dom = int(np.argmax(dir_ms))
    # the dominant kernel's mean launch duration inside the timed replay: the
    # replay's time per step shared out over the directions in proportion to
    # their event-timed durations (when every direction runs the same kernel
    # configuration this is ms_per_step / ndim, the mean over all its launches)
    same = all(c == cfg[dom] for c in cfg)
    n_same = sum(1 for c in cfg if c == cfg[dom])
    kernel_ms = (ms / n_same) if same else ms * dir_ms[dom] / sum(dir_ms)
    return {"ms_per_step": ms, "wall_ms_per_step": wall, "dir_ms": dir_ms, "dominant": dom,
            "kernel_ms_timed": kernel_ms, "kernel_launches_per_step": n_same,
            "achieved_GBs": bpr * n / (kernel_ms * 1e-3) / 1e9, "bytes_per_row": bpr,
            "step_GBs": bpr * n * len(dims) / (ms * 1e-3) / 1e9,
            "grows": len(dims) * n / (ms * 1e-3) / 1e9, "kernels": cfg,
            "refine_steps_timed_max": refine, "refine_steps_checked": refine_checked,
            "warmup_wall_ms": warm_wall, "warmup_steps": warm_replays * warmup,
            "one_step": result}


def dropin_direction_ms(P, torch, coeffs, b, dims, h, reps=3, ir=False):
    """In-place per-direction solves through pentadsolver_D_gpsv_batch, or
    (ir) the mixed-precision pentadsolver_DS_gpsv_batch_ir (FP32 coefficients,
    FP64 field)."""
    x = b.clone()
    stream = torch.cuda.current_stream()
    out = []
    for sd in range(len(dims)):
        if ir:
            nb = P.gpsv_batch_ir_buffer_size_ext(*coeffs, x, dims, sd, h)
        else:
            nb = P.gpsv_batch_buffer_size_ext(*coeffs, x, dims, sd, h)
        buf = torch.empty(nb, dtype=torch.uint8, device="cuda") if nb else None

        def call(check=True):
            if ir:
                P.gpsv_batch_ir(*coeffs, x, dims, sd, buffer=buf, handle=h, check=check)
            else:
                P.gpsv_batch(*coeffs, x, dims, sd, buf, h, check=check)

        call()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            call(check=False)
        e1.record(stream)
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1) / reps)
        del buf
    return out, [h.kernel_config(k) for k in range(len(dims))]


def oracle_step(coeffs_np, b_np, dims):
    from oracle import oracle as O
    x = b_np.copy()
    for sd in range(len(dims)):
        O.gpsv_batch(*coeffs_np, x, dims, sd)
    return x


def cpu_info():
    """lscpu model, sockets, cores and threads, plus the CPUs this process
    may run on (the GPU box gives a job a share of the host)."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets",
                "Core(s) per socket": "cores_per_socket", "Thread(s) per core": "threads_per_core",
                "CPU(s)": "cpus", "NUMA node(s)": "numa_nodes"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        info["lscpu"] = "unavailable"
    info["affinity_cpus"] = AFFINITY_CPUS
    info["OMP_NUM_THREADS"] = os.environ.get("OMP_NUM_THREADS")
    info["OMP_PROC_BIND"] = os.environ.get("OMP_PROC_BIND")
    info["OMP_PLACES"] = os.environ.get("OMP_PLACES")
    return info


def cpu_adi_steps(O, kind, coeffs_np, b_np, dims, warm, iters):
    """ms per direction of the reference CPU solver (or the oracle port) on
    x/y/z in place: `warm` untimed steps, then the mean over `iters` timed
    steps, timed per direction with a monotonic clock (apps/adi_cpu.cpp:49-103)."""
    x = b_np.copy()

    def solve(sd):
        if kind == "reference":
            O.ref_gpsv_batch(*coeffs_np, x, dims, sd)
        else:
            O.gpsv_batch(*coeffs_np, x, dims, sd)

    for _ in range(warm):
        for sd in range(3):
            solve(sd)
    t = np.zeros(3)
    for _ in range(iters):
        for sd in range(3):
            t0 = time.perf_counter()
            solve(sd)
            t[sd] += time.perf_counter() - t0
    return t / iters * 1e3


def cpu_baseline(coeffs_np, b_np, dims, iters=5, warm=2):
    """The reference's own OpenMP+AVX2 CPU solver (oracle/_ref, built from its
    sources) timed on this host, method of apps/adi_cpu.cpp:49-103: 2 warm-up
    steps, then per-direction times averaged over the timed steps.  The value
    is FP64 on the full grid with every thread this job may use; beside it an
    FP32 row on the same grid and a thread sweep (1, 2, 4, ... max) of the
    FP64 step on the 256^3 grid (a bounded sample: one thread takes about
    16x longer).  OMP_PROC_BIND=spread, OMP_PLACES=cores."""
    from oracle import oracle as O
    kind = "reference" if O.ref_available("avx") else "port"
    info = cpu_info()
    env_threads = os.environ.get("OMP_NUM_THREADS")
    max_threads = int(env_threads) if env_threads else int(info.get("affinity_cpus") or 1)
    if kind == "reference":
        O.ref_set_threads(max_threads)
        info["omp_proc_bind_in_force"] = O.ref_proc_bind()
    rows = 3 * int(np.prod(dims))
    t64 = cpu_adi_steps(O, kind, coeffs_np, b_np, dims, warm, iters)
    out = {"value": float(rows / (t64.sum() * 1e-3) / 1e9), "unit": "GRows/s",
           "cores": max_threads, "kind": kind,
           "sample": f"FP64 ADI step x/y/z on {dims[0]}x{dims[1]}x{dims[2]} (same seeded "
                     f"coefficients), {iters} timed steps after {warm} warm-up steps, ms per "
                     f"direction {t64[0]:.1f}/{t64[1]:.1f}/{t64[2]:.1f}, {max_threads} OpenMP "
                     f"threads, OMP_PROC_BIND={info['OMP_PROC_BIND']} "
                     f"OMP_PLACES={info['OMP_PLACES']}",
           "host": info,
           "vis_GBs_fp64": [7 * 8 * rows / 3 / (t * 1e-3) / 1e9 for t in t64]}
    if kind == "reference":
        c32 = [a.astype(np.float32) for a in coeffs_np]
        t32 = cpu_adi_steps(O, kind, c32, b_np.astype(np.float32), dims, warm, iters)
        out["fp32"] = {"grows": float(rows / (t32.sum() * 1e-3) / 1e9),
                       "ms_per_direction": t32.tolist(),
                       "vis_GBs": [7 * 4 * rows / 3 / (t * 1e-3) / 1e9 for t in t32]}
        del c32
        # thread sweep on the 256^3 corner of the same field
        q = [min(256, v) for v in dims]
        sub = tuple(slice(0, v) for v in reversed(q))
        shape = tuple(reversed(dims))
        cs = [np.ascontiguousarray(a.reshape(shape)[sub]).reshape(-1) for a in coeffs_np]
        bs = np.ascontiguousarray(b_np.reshape(shape)[sub]).reshape(-1)
        sweep = []
        nt = 1
        while True:
            O.ref_set_threads(nt)
            t = cpu_adi_steps(O, kind, cs, bs, q, 1, 2)
            sweep.append({"threads": nt, "ms_per_step": float(t.sum()),
                          "grows": float(3 * int(np.prod(q)) / (t.sum() * 1e-3) / 1e9)})
            if nt >= max_threads:
                break
            nt = min(2 * nt, max_threads)
        O.ref_set_threads(max_threads)
        out["thread_sweep_fp64"] = {"grid": q, "steps": "1 warm-up + 2 timed", "points": sweep}
    return out


# ---------------------------------------------------------------------------
# the other configurations: reported beside the headline, each with its check
# ---------------------------------------------------------------------------
def graph_ms_per_step(torch, adi, x, b, steps):
    """Capture `steps` ADI steps on x as one hipGraph, replay once to warm,
    then time one replay from the initial field (ms per step)."""
    stream = torch.cuda.current_stream()
    x.copy_(b)
    adi.capture(x, steps)
    adi.replay(1)
    torch.cuda.synchronize()
    x.copy_(b)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    adi.replay(1)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def max_rel(a, ref):
    return float(np.max(np.abs(np.asarray(a, np.float64) - ref)) / np.max(np.abs(ref)))


def weak_set(P, torch, W, h, size, steps, max_refine, tol, check):
    """Precision config, weakly dominant set: c = 1.05 * sum|off| (seed 4)."""
    dims = [size] * 3
    n = size ** 3
    ds, dl, d, du, dw, b = W.dominant_torch(dims, 4, ratio=(1.05, 1.05), dtype=torch.float64)
    c32 = tuple(a.float() for a in (ds, dl, d, du, dw))
    del ds, dl, d, du, dw
    x = b.clone()
    with P.Adi(P.ADI_IR, dims, [c32], handle=h, max_refine=max_refine, tol=tol) as adi:
        ms = graph_ms_per_step(torch, adi, x, b, steps)
        adi.refine_steps()
        x.copy_(b)
        adi.step(x, 1)
        steps_taken = adi.refine_steps()
        cfg = [adi.config(k) for k in range(3)]
    out = {"ms_per_step": ms, "grows": 3 * n / (ms * 1e-3) / 1e9, "kernels": cfg,
           "refine_steps_one_step": steps_taken}
    if check:
        cw = [a.double().cpu().numpy() for a in c32]
        ref = oracle_step(cw, b.cpu().numpy(), dims)
        out["max_rel_err_vs_fp64_oracle"] = max_rel(x.cpu().numpy(), ref)
    return out


def workload_c2(P, torch, W, h, check, iters=50):
    """256^3 drop-in config: 50 ADI iterations, FP64 and FP32, variable
    coefficients (seed 2) and the constant-coefficient (shared-fact) variant;
    FP64 compared with the reference's own CPU solver (oracle/_ref) and the
    oracle after one step.  The ADI driver (hipGraph replay) and, beside it,
    the step as three in-place drop-in calls (pentadsolver_{D,S}_gpsv_batch)."""
    from oracle import oracle as O
    dims = [256] * 3
    n = 256 ** 3
    ds, dl, d, du, dw, b = W.dominant_torch(dims, 2, dtype=torch.float64)
    c64 = (ds, dl, d, du, dw)
    c32 = tuple(a.float() for a in c64)
    res = {"grid": dims, "iterations": iters}
    one = {}
    for mode, cs, x0 in (("fp64", c64, b), ("fp32", c32, b.float())):
        m = P.ADI_FP64 if mode == "fp64" else P.ADI_FP32
        x = x0.clone()
        with P.Adi(m, dims, [cs], handle=h) as adi:
            ms = graph_ms_per_step(torch, adi, x, x0, iters)
            x.copy_(x0)
            adi.step(x, 1)
            cfg = [adi.config(k) for k in range(3)]
        one[mode] = x.double().cpu().numpy()
        res[mode] = {"ms_per_step": ms, "grows": 3 * n / (ms * 1e-3) / 1e9, "kernels": cfg}
        # the same step through the drop-in entry point, in place, x -> y -> z
        d_ms, d_k = dropin_direction_ms(P, torch, cs, x0, dims, h)
        res[mode]["dropin"] = {"ms_per_step": sum(d_ms), "per_direction_ms": d_ms,
                               "grows": 3 * n / (sum(d_ms) * 1e-3) / 1e9, "kernels": d_k}
    # constant-coefficient variant: one profile per direction, replicated
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    prof_sets = []
    for ax in range(3):
        off = [torch.empty(256, dtype=torch.float64, device="cuda").uniform_(-1, 1, generator=g)
               for _ in range(4)]
        r = torch.empty(256, dtype=torch.float64, device="cuda").uniform_(2, 3, generator=g)
        dd = r * sum(o.abs() for o in off)
        view = [1, 1, 1]
        view[2 - ax] = 256
        prof_sets.append([a.reshape(view).expand(256, 256, 256).contiguous().reshape(-1)
                          for a in (off[0], off[1], dd, off[2], off[3])])
    for mode, dt, m in (("shared_fp64", torch.float64, P.ADI_SHARED_FP64),
                        ("shared_fp32", torch.float32, P.ADI_SHARED_FP32)):
        sets = [[a.to(dt) for a in cs] for cs in prof_sets]
        x0 = b.to(dt)
        x = x0.clone()
        with P.Adi(m, dims, sets, handle=h) as adi:
            ms = graph_ms_per_step(torch, adi, x, x0, iters)
            x.copy_(x0)
            adi.step(x, 1)
            cfg = [adi.config(k) for k in range(3)]
        one[mode] = x.double().cpu().numpy()
        res[mode] = {"ms_per_step": ms, "grows": 3 * n / (ms * 1e-3) / 1e9, "kernels": cfg,
                     "bytes_per_row": 16 if dt == torch.float64 else 8}
    if check:
        b_np = b.cpu().numpy()
        c_np = [a.cpu().numpy() for a in c64]
        ref = oracle_step(c_np, b_np, dims)
        res["fp64"]["max_rel_err_vs_fp64_oracle"] = max_rel(one["fp64"], ref)
        if O.ref_available("avx"):
            xr = b_np.copy()
            for sd in range(3):
                O.ref_gpsv_batch(*c_np, xr, dims, sd)
            res["fp64"]["max_rel_err_vs_reference_cpu"] = max_rel(one["fp64"], xr)
        c32_np = [a.double().cpu().numpy() for a in c32]
        ref32 = oracle_step(c32_np, b.float().double().cpu().numpy(), dims)
        res["fp32"]["max_rel_err_vs_fp64_oracle"] = max_rel(one["fp32"], ref32)
        for mode, dt in (("shared_fp64", torch.float64), ("shared_fp32", torch.float32)):
            x = b.to(dt).double().cpu().numpy()
            for sd in range(3):
                cs = [a.to(dt).double().cpu().numpy() for a in prof_sets[sd]]
                O.gpsv_batch(*cs, x, dims, sd)
            res[mode]["max_rel_err_vs_fp64_oracle"] = max_rel(one[mode], x)
    return res


def workload_c4(P, torch, W, h, check, reps=10):
    """Batching config: 8192 systems, lengths log-uniform in [32, 8192]
    (seed 5), dominance ratio U(1.2, 3) (seed 6); one persistent launch per
    solve vs one solve call per system (FP64).  Before each mode is timed the
    plan's lanes per system are autotuned per length bucket on the same
    arrays (the untuned schedule's time is reported beside)."""
    from oracle import oracle as O
    off = W.varlen_offsets(8192, 5)
    n = int(off[-1])
    rng = np.random.default_rng(6)
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(1.2, 3.0, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    c64 = [torch.from_numpy(a).cuda() for a in (ds, dl, d, du, dw)]
    c32 = [t.float() for t in c64]
    bt = torch.from_numpy(b).cuda()
    stream = torch.cuda.current_stream()
    res = {"systems": 8192, "rows": n}
    with P.Varlen(off, handle=h) as vl:
        res["work_items"] = vl.items()
        def timed(cs, x, steps):
            vl.solve(*cs, x, steps=steps)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                vl.solve(*cs, x, steps=steps)
            e1.record(stream)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        # "ir": FP32 coefficients with an FP64 x, FP32 solve + FP64 refinement
        # (varlen_set_refine(plan, 1)); "fp64_widened": the same coefficients
        # solved once in FP64 after widening on load (the plan's default)
        for mode, cs, x0 in (("fp64", c64, bt), ("fp32", c32, bt.float()), ("ir", c32, bt),
                             ("fp64_widened", c32, bt)):
            vl.set_refine(mode == "ir")
            steps = torch.zeros(1, dtype=torch.int32, device="cuda") if mode == "ir" else None
            x = x0.clone()
            vl.set_spread([0] * 10)
            ms_default = timed(cs, x, steps)
            x.copy_(x0)
            vl.autotune(*cs, x, reps=3)
            spread = vl.spread()
            ms = timed(cs, x, steps)
            x.copy_(x0)
            vl.solve(*cs, x, steps=steps)
            r = {"ms": ms, "grows": n / (ms * 1e-3) / 1e9, "spread_per_bucket": spread,
                 "ms_untuned": ms_default, "kernel": h.kernel_config(0)}
            if steps is not None:
                steps.zero_()
                x.copy_(x0)
                vl.solve(*cs, x, steps=steps)
                r["refine_steps_max"] = int(steps.item())
            if check:
                if mode == "fp64":
                    ref = O.varlen(ds, dl, d, du, dw, b, off)
                else:
                    w = [a.cpu().numpy().astype(np.float64) for a in c32]
                    rhs = b.astype(np.float32).astype(np.float64) if mode == "fp32" else b
                    ref = O.varlen(*w, rhs, off)
                r["max_rel_err_vs_fp64_oracle"] = max_rel(x.double().cpu().numpy(), ref)
            res[mode] = r
    # one solve call per system (FP64, in place, through the drop-in entry point)
    x = bt.clone()
    lens = np.diff(off)
    scratch = torch.empty(3 * int(lens.max()) * 8, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(len(lens)):
        lo, hi = int(off[k]), int(off[k + 1])
        P.gpsv_batch(*[t[lo:hi] for t in c64], x[lo:hi], [hi - lo], 0, scratch, h, check=False)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    res["per_system_fp64"] = {"ms": ms, "grows": n / (ms * 1e-3) / 1e9,
                              "note": "8192 gpsv_batch calls (1-D, one system each), wall clock"}
    res["speedup_one_launch_vs_per_system_fp64"] = ms / res["fp64"]["ms"]
    return res


def workload_cavity(P, torch, W, h, check, reps=20, cpu_reps=2):
    """The 2-D cavity consumer's pressure-correction step (reference
    apps/2D_cavity_cpu.cpp:195-219): assembly of the systems
    (apps/preproc.hpp:248-525) + x solve + y solve of the field, rows of
    nx_pad = nx.  Grids 1024 x 1024 (the consumer's N_MAX) and 1024 x 8192;
    rows per step = 2 nx ny.  Two ways on the GPU: the reference's sequence
    (assembly kernel writing the right-hand side and both sets of five
    diagonals, then the padded batch solves: 224 B per point) and the fused
    solve (coefficients generated inside the solve kernels: 56 B per point).
    At 1024 x 1024 also the reference's own CPU pipeline (its compiled
    assembly and AVX solver, oracle/_ref) on the same inputs."""
    from oracle import oracle as O
    stream = torch.cuda.current_stream()
    res = {"bytes_per_point": {"assembled": 224, "fused": 56}}

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    for nx, ny in ((1024, 1024), (1024, 8192)):
        up, vp, uv, aU, dx, dy = W.cavity_torch(nx, ny, seed=11)
        dims = pads = [nx, ny]
        n = nx * ny
        arr = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(11)]
        x, xset, yset = arr[0], arr[1:6], arr[6:11]
        bufs = []
        for sd, cs in ((0, xset), (1, yset)):
            nb = P.gpsv_batch_padded_buffer_size(*cs, x, dims, pads, sd, h)
            bufs.append(torch.empty(nb, dtype=torch.uint8, device="cuda") if nb else None)

        def assembled():
            P.cavity_pressure_coef(nx, ny, nx, aU, dx, dy, up, vp, uv, x, xset, yset, handle=h)
            P.gpsv_batch_padded(*xset, x, dims, pads, 0, bufs[0], h)
            P.gpsv_batch_padded(*yset, x, dims, pads, 1, bufs[1], h)

        xf = torch.empty(n, dtype=torch.float64, device="cuda")

        def fused():
            P.cavity_pressure_solve(nx, ny, nx, aU, dx, dy, up, vp, uv, xf, handle=h)

        ms_a = timed(assembled, reps)
        cfg_a = [h.kernel_config(0), h.kernel_config(2)]
        ms_f = timed(fused, reps)
        cfg_f = [h.kernel_config(0), h.kernel_config(2)]
        rows = 2 * n
        r = {"grid": [nx, ny],
             "assembled": {"ms_per_step": ms_a, "grows": rows / (ms_a * 1e-3) / 1e9,
                           "gbs": 224 * n / (ms_a * 1e-3) / 1e9, "kernels": cfg_a},
             "fused": {"ms_per_step": ms_f, "grows": rows / (ms_f * 1e-3) / 1e9,
                       "gbs": 56 * n / (ms_f * 1e-3) / 1e9, "kernels": cfg_f},
             "fused_speedup": ms_a / ms_f,
             "max_rel_diff_fused_vs_assembled": max_rel(xf.cpu().numpy(), x.cpu().numpy())}
        if check and nx * ny <= 1 << 20 and O.ref_cavity_available() and O.ref_available("avx"):
            npin = [t.cpu().numpy() for t in (up, vp, uv)]
            t0 = time.perf_counter()
            for _ in range(cpu_reps):
                ro = O.ref_cavity_pressure_coef(nx, ny, nx, aU, dx, dy, *npin)
                xr = ro["x"]
                O.ref_gpsv_batch(*[ro[k] for k in ("s_x", "l_x", "d_x", "u_x", "w_x")], xr,
                                 dims, 0)
                O.ref_gpsv_batch(*[ro[k] for k in ("s_y", "l_y", "d_y", "u_y", "w_y")], xr,
                                 dims, 1)
            ms_c = (time.perf_counter() - t0) * 1e3 / cpu_reps
            r["reference_cpu"] = {"ms_per_step": ms_c, "grows": rows / (ms_c * 1e-3) / 1e9,
                                  "threads": os.environ.get("OMP_NUM_THREADS", "all"),
                                  "note": "compiled assembly (inputs converted to its nested "
                                          "vectors inside the timed region) + AVX solver x, y"}
            r["max_rel_err_fused_vs_reference_cpu"] = max_rel(xf.cpu().numpy(), xr)
        res[f"{nx}x{ny}"] = r
        del arr, x, xset, yset, xf, up, vp, uv, bufs
        torch.cuda.empty_cache()
    return res


def workload_c5(P, torch, W, h, check, steps=100, cpu_steps=2, max_refine=5, tol=1e-13):
    """Hyperdiffusion config: 384x256x640, I + nu dt/h^4 [1,-4,6,-4,1] per
    direction with nu ~ U(0.5, 2) per grid line (seed 7), N(0,1) initial
    field (seed 8), 100 steps as one hipGraph; FP32+FP64 refinement vs FP64."""
    dims = [384, 256, 640]
    n = int(np.prod(dims))
    sets64 = W.hyperdiffusion_torch(dims, 7)
    sets32 = [[a.float() for a in cs] for cs in sets64]
    b = W.gaussian_field_torch(dims, 8)
    res = {"grid": dims, "steps": steps, "dt_over_h4": W.HYPERDIFFUSION_DT_OVER_H4}
    out = {}
    for mode, sets, m in (("fp64", sets64, P.ADI_FP64), ("ir", sets32, P.ADI_IR)):
        x = b.clone()
        with P.Adi(m, dims, sets, handle=h, max_refine=max_refine, tol=tol) as adi:
            ms = graph_ms_per_step(torch, adi, x, b, steps)
            refine = adi.refine_steps() if mode == "ir" else None
            x.copy_(b)
            adi.capture(x, steps)
            adi.replay(1)
            torch.cuda.synchronize()
            cfg = [adi.config(k) for k in range(3)]
        out[mode] = x.clone()
        res[mode] = {"ms_per_step": ms, "grows": 3 * n / (ms * 1e-3) / 1e9, "kernels": cfg}
        if refine is not None:
            res[mode]["refine_steps_max"] = refine
    # structured coefficients: the same operators generated in the kernels
    # from one scale per line (16 B per row FP64, 8 B FP32, against 56 / 36)
    # ("stencil_ir": FP32 solves refined against the FP64 operator, so its
    # answer is the FP64 operator's, not the FP32-stored one's)
    scales = W.hyperdiffusion_scales_torch(dims, 7)
    for mode, sc in (("stencil_fp64", scales), ("stencil_fp32", [t.float() for t in scales]),
                     ("stencil_ir", scales)):
        x = b.float() if mode == "stencil_fp32" else b.clone()
        with P.AdiStencil(dims, sc, W.HYPERDIFFUSION_WEIGHTS, 1.0, handle=h,
                          refine=(mode == "stencil_ir")) as adi:
            if mode == "stencil_ir":
                adi.set_refine(max_refine, tol)
            ms = graph_ms_per_step(torch, adi, x, x.clone(), steps)
            if mode == "stencil_ir":
                refine = adi.refine_steps()
            x.copy_(b)
            adi.capture(x, steps)
            adi.replay(1)
            torch.cuda.synchronize()
            cfg = [adi.config(k) for k in range(3)]
        res[mode] = {"ms_per_step": ms, "grows": 3 * n / (ms * 1e-3) / 1e9, "kernels": cfg,
                     "bytes_per_row": 8 if mode == "stencil_fp32" else 16,
                     "max_rel_err_vs_assembled_fp64_after_steps": float(
                         ((x.double() - out["fp64"]).abs().max() /
                          out["fp64"].abs().max()).item())}
        if mode == "stencil_ir":
            res[mode]["refine_steps_max"] = refine
            out["stencil_ir"] = x.clone()
        del x
    del scales
    res["stencil_fp64_speedup_vs_fp64"] = res["fp64"]["ms_per_step"] / res["stencil_fp64"]["ms_per_step"]
    res["stencil_ir_speedup_vs_stencil_fp64"] = (res["stencil_fp64"]["ms_per_step"] /
                                                 res["stencil_ir"]["ms_per_step"])
    # the refinement run solves the FP32-stored operator: its solver error after
    # all steps is measured against an FP64 run of that same (widened) operator
    xw = b.clone()
    with P.Adi(P.ADI_FP64, dims, [[a.double() for a in cs] for cs in sets32], handle=h) as adi:
        adi.capture(xw, steps)
        adi.replay(1)
        torch.cuda.synchronize()
    res["ir_max_rel_err_vs_fp64_same_operator_after_steps"] = float(
        ((out["ir"] - xw).abs().max() / xw.abs().max()).item())
    # and the FP32 storage of the operator itself moves the answer by:
    res["fp32_operator_vs_fp64_operator_after_steps"] = float(
        ((xw - out["fp64"]).abs().max() / out["fp64"].abs().max()).item())
    del xw
    res["energy_ratio_fp64"] = float((out["fp64"].norm() / b.norm()).item())
    res["ir_speedup_vs_fp64"] = res["fp64"]["ms_per_step"] / res["ir"]["ms_per_step"]
    if check and cpu_steps > 0:
        # FP64 CPU run (the oracle, all host threads) of the first cpu_steps steps
        from oracle import oracle as O
        x = b.clone()
        with P.Adi(P.ADI_FP64, dims, sets64, handle=h) as adi:
            adi.step(x, cpu_steps)
        with P.Adi(P.ADI_IR, dims, sets32, handle=h, max_refine=max_refine, tol=tol) as adi:
            xi = b.clone()
            adi.step(xi, cpu_steps)
        torch.cuda.synchronize()
        xc = b.cpu().numpy()
        c64 = [[a.cpu().numpy() for a in cs] for cs in sets64]
        t0 = time.perf_counter()
        for _ in range(cpu_steps):
            for sd in range(3):
                O.gpsv_batch(*c64[sd], xc, dims, sd)
        cpu_s = time.perf_counter() - t0
        del c64
        res["fp64_max_rel_err_vs_cpu_fp64"] = max_rel(x.cpu().numpy(), xc)
        # the structured refinement run against the same CPU run of the FP64
        # operator (its residual uses that operator)
        scales = W.hyperdiffusion_scales_torch(dims, 7)
        with P.AdiStencil(dims, scales, W.HYPERDIFFUSION_WEIGHTS, 1.0, handle=h,
                          refine=True) as adi:
            adi.set_refine(max_refine, tol)
            xs = b.clone()
            adi.step(xs, cpu_steps)
            torch.cuda.synchronize()
        res["stencil_ir_max_rel_err_vs_cpu_fp64"] = max_rel(xs.cpu().numpy(), xc)
        del scales, xs
        c32w = [[a.double().cpu().numpy() for a in cs] for cs in sets32]
        xw = b.cpu().numpy()
        for _ in range(cpu_steps):
            for sd in range(3):
                O.gpsv_batch(*c32w[sd], xw, dims, sd)
        res["ir_max_rel_err_vs_cpu_fp64"] = max_rel(xi.cpu().numpy(), xw)
        res["cpu_check_steps"] = cpu_steps
        res["cpu_fp64_oracle_grows"] = 3 * n * cpu_steps / cpu_s / 1e9
    return res


def workload_dist(P, torch, W, h, check, size=512, nranks=8, reps=5):
    """The distributed line solve's one-GPU share: the z lines of a size^3 FP64
    field split into nranks slabs (slab k = rank k's size x size x size/nranks
    arrays, a view of the field).  Per rank: phase 1 (forward elimination,
    reads 48 B per row) and phase 3 (separator rows + the interior solve,
    56 B per row), timed on every slab and reported as the mean per rank;
    phase 2 (the reduced block system of all nranks ranks, which on several
    GPUs would be spread over them) timed for the whole line set.  The whole
    solve is checked against the oracle's z solve of the field."""
    dims = [size, size, size]
    ds, dl, d, du, dw, b = W.dominant_torch(dims, 3, dtype=torch.float64)
    c = (ds, dl, d, du, dw)
    dist = P.DistSlabs(dims, 2, nranks, torch.float64, handle=h)
    stream = torch.cuda.current_stream()
    x = b.clone()
    dist.solve(*c, x)  # warm-up, and the result checked below
    torch.cuda.synchronize()
    out = {"field": dims, "ranks": nranks, "slab": dist.local,
           "lines_per_rank": dist.n_sys, "rows_per_rank": dist.n_sys * dist.m}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    t_f = t_r = t_b = 0.0
    xs = b.clone()
    for _ in range(reps):
        xs.copy_(b)
        ev[0].record(stream)
        for k in range(nranks):
            dist.forward(*c, xs, k)
        ev[1].record(stream)
        dist.reduced()
        ev[2].record(stream)
        for k in range(nranks):
            dist.backward(*c, xs, k)
        ev[3].record(stream)
        torch.cuda.synchronize()
        t_f += ev[0].elapsed_time(ev[1])
        t_r += ev[1].elapsed_time(ev[2])
        t_b += ev[2].elapsed_time(ev[3])
    t_f, t_r, t_b = t_f / reps / nranks, t_r / reps, t_b / reps / nranks
    rows = dist.n_sys * dist.m
    out.update({
        "forward_ms_per_rank": t_f, "forward_GBs": 48 * rows / (t_f * 1e-3) / 1e9,
        "backward_ms_per_rank": t_b, "backward_GBs": 56 * rows / (t_b * 1e-3) / 1e9,
        "reduced_ms_all_ranks": t_r,
        "rank_ms": t_f + t_b, "rank_grows": rows / ((t_f + t_b) * 1e-3) / 1e9,
        "interior_kernel": h.kernel_config(2)})
    # the same z lines solved whole on one GPU (the drop-in call), for scale
    x1 = b.clone()
    P.solve(*c, x1, dims, 2, handle=h)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        P.gpsv_batch(*c, x1, dims, 2, None, h, check=False)
    e1.record(stream)
    torch.cuda.synchronize()
    out["whole_line_z_ms_one_gpu"] = e0.elapsed_time(e1) / reps
    if check:
        from oracle import oracle as O
        ref = O.solve(*[a.cpu().numpy() for a in c], b.cpu().numpy(), dims, 2)
        out["max_rel_err_vs_fp64_oracle"] = max_rel(x.cpu().numpy(), ref)
    log(f"bench: dist {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--warmup-ms", type=float, default=1500.0,
                    help="replay the warm-up graph until this much wall time has passed")
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--max-refine", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-13)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other configurations (weak set, 256^3, varlen, hyperdiffusion)")
    ap.add_argument("--c5-tol", type=float, default=1e-15,
                    help="refinement tolerance of the hyperdiffusion run: its 100 steps "
                         "compound the per-solve error on a field that decays 30x, so each "
                         "solve is refined to the FP64 limit (1e-13 leaves 2.5e-12 after "
                         "100 steps, 1e-15 leaves 1.2e-14 and costs 5 percent more time)")
    ap.add_argument("--c5-cpu-steps", type=int, default=2,
                    help="hyperdiffusion steps re-run on the CPU oracle for the error check")
    ap.add_argument("--probe", action="store_true",
                    help="also launch the 6-read/1-write bandwidth probe once per precision "
                         "over the same field size (calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); lines are "
              "independent within a direction, so a multi-GPU run would shard whole lines "
              "(see DESIGN.md, sharding)", file=sys.stderr)
        sys.exit(2)

    import torch
    P = import_pkg()
    from pentadsolver_amd import workloads as W
    P.load_library()
    torch.cuda.set_device(0)
    dims = [args.size] * 3
    n = int(np.prod(dims))

    log(f"bench: generating {dims} coefficients (seed 3)")
    ds, dl, d, du, dw, b = W.dominant_torch(dims, 3, dtype=torch.float64)
    c64 = (ds, dl, d, du, dw)
    c32 = tuple(a.float() for a in c64)
    b32 = b.float()
    h = P.Handle(stream=torch.cuda.current_stream())

    res = {}
    for mode, coeffs, rhs in (("ir", c32, b), ("fp64", c64, b), ("fp32", c32, b32)):
        log(f"bench: {mode} ADI step")
        res[mode] = time_adi(P, torch, mode, coeffs, rhs, dims, args.steps, args.warmup, h,
                             args.max_refine, args.tol, args.warmup_ms)
        r = res[mode]
        log(f"bench: {mode} {r['ms_per_step']:.3f} ms/step {r['grows']:.1f} GRows/s dirs "
            f"{['%.3f' % v for v in r['dir_ms']]} kernels {r['kernels']} refine {r['refine_steps_timed_max']}")
    if args.probe:
        log("bench: bandwidth probe (6 reads + 1 write per element)")
        for arrs, out in ((c64 + (b,), torch.empty_like(b)), (c32 + (b32,), torch.empty_like(b32))):
            P.probe_stream(list(arrs), out, handle=h)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(torch.cuda.current_stream())
            P.probe_stream(list(arrs), out, handle=h)
            e1.record(torch.cuda.current_stream())
            torch.cuda.synchronize()
            gbs = 7 * out.numel() * out.element_size() / (e0.elapsed_time(e1) * 1e-3) / 1e9
            log(f"bench: probe {out.dtype}: {gbs:.0f} GB/s")
            del out
        # the refinement solve's byte mix as 7, 3 and 2 streams (best of 10)
        probe = {}
        xo = torch.empty_like(b)
        for streams in (7, 3, 2):
            if streams == 7:
                cf = list(c32)
            else:
                cf = [torch.zeros((5 if streams == 3 else 7) * n, dtype=torch.float32,
                                  device="cuda")]
            best = 1e30
            for _ in range(10):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(torch.cuda.current_stream())
                P.probe_ir_mix(streams, cf, b, xo, handle=h)
                e1.record(torch.cuda.current_stream())
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1))
            probe[f"{streams}_streams"] = {"ms": best, "GBs": 36 * n / (best * 1e-3) / 1e9}
            del cf
        # the same bytes with the headline kernel's per-lane load shape:
        # coalesced writes at full occupancy, and the kernel's
        # workgroup structure with its rotated tile store, with 0..150 rounds
        # of 8 FP64 FMAs per pass standing in for the solve (the kernel
        # itself issues about 850 VALU instructions per pass)
        cf = torch.zeros(5 * n, dtype=torch.float32, device="cuda")
        layouts = [(0, 0)] + ([(1, r) for r in (0, 25, 50, 100, 150)] if dims == [512] * 3 else [])
        for layout, rounds in layouts:
            best = 1e30
            for _ in range(10):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(torch.cuda.current_stream())
                P.probe_ir_layout(layout, rounds, cf, b, xo, handle=h)
                e1.record(torch.cuda.current_stream())
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1))
            key = "lanes_coalesced_writes" if layout == 0 else f"kernel_structure_rotated_valu{8 * rounds}"
            probe[key] = {"ms": best, "GBs": 36 * n / (best * 1e-3) / 1e9}
        del cf
        # the FP64 strided solve's I/O without the solve, y and z lines of
        # 512 rows: whole lines per workgroup (the drop-in kernel's
        # structure) and lines shared by 2 or 4 workgroups
        strided = {}
        if dims == [512] * 3:
            xo.zero_()
            for sd, split in ((1, 1), (2, 1), (2, 2), (2, 4)):
                best = 1e30
                for _ in range(10):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(torch.cuda.current_stream())
                    P.probe_strided_io(split, list(c64), xo, dims, sd, handle=h)
                    e1.record(torch.cuda.current_stream())
                    torch.cuda.synchronize()
                    best = min(best, e0.elapsed_time(e1))
                strided[f"{'yz'[sd - 1]}_split{split}"] = {
                    "ms": best, "GBs": 56 * n / (best * 1e-3) / 1e9}
        del xo
        log(f"bench: refinement byte-mix probe {probe}; strided I/O probe {strided}")
    log("bench: drop-in in-place FP64 direction solves")
    dropin_ms, dropin_k = dropin_direction_ms(P, torch, c64, b, dims, h)
    log("bench: drop-in in-place mixed-precision direction solves")
    dropin_ir_ms, dropin_ir_k = dropin_direction_ms(P, torch, c32, b, dims, h, ir=True)
    log("bench: drop-in FP64 direction solves with PENTA_AUTOTUNE=1 (tuned on first use)")
    os.environ["PENTA_AUTOTUNE"] = "1"
    try:
        with P.Handle() as ht:
            tuned_ms, tuned_k = dropin_direction_ms(P, torch, c64, b, dims, ht)
    finally:
        del os.environ["PENTA_AUTOTUNE"]

    check = {}
    c64_np = c32_np = b_np = None
    if not args.no_check or not args.no_cpu:
        c64_np = [a.cpu().numpy() for a in c64]
        b_np = b.cpu().numpy()
    if not args.no_check:
        from oracle import oracle as O  # noqa: F401  (checker only)
        log("bench: oracle FP64 step")
        ref64 = oracle_step(c64_np, b_np, dims)
        scale = np.max(np.abs(ref64))
        check["fp64_max_rel_err"] = float(np.max(np.abs(res["fp64"]["one_step"] - ref64)) / scale)
        del ref64
        log("bench: oracle FP64 step of the FP32-stored coefficients")
        c32_np = [a.cpu().numpy().astype(np.float64) for a in c32]
        ref_w = oracle_step(c32_np, b_np, dims)
        scale = np.max(np.abs(ref_w))
        check["ir_max_rel_err"] = float(np.max(np.abs(res["ir"]["one_step"] - ref_w)) / scale)
        ref32 = oracle_step(c32_np, b32.double().cpu().numpy(), dims)
        check["fp32_max_rel_err"] = float(np.max(np.abs(res["fp32"]["one_step"] - ref32)) /
                                          np.max(np.abs(ref32)))
        del ref_w, ref32, c32_np
        log(f"bench: errors {check}")
    for r in res.values():
        r.pop("one_step", None)

    head = res["ir"]
    dom = head["dominant"]
    result = {
        "metric": METRIC,
        "value": head["grows"],
        "unit": "GRows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 solve + f64 residual/update",
        "data": "synthetic",
        "config": {"workload": f"ADI step x->y->z on a {dims[0]}x{dims[1]}x{dims[2]} grid, "
                               f"{n // dims[0]} systems of length {dims[0]} per direction, "
                               "FP32-stored diagonally dominant coefficients (ratio U(2,3)), "
                               "FP64 field, iterative refinement to tol "
                               f"{args.tol:g} (max {args.max_refine} steps), hipGraph replay",
                   "grid": dims, "mode": "FP32 + FP64 refinement"},
        "roofline": {"bound": "hbm", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC_BYTES.get((head["kernels"][dom], dims[dom])),
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE)",
                     "algorithmic_bytes": head["bytes_per_row"] * n,
                     "kernel": f"direction {dom}: {head['kernels'][dom]}",
                     "bytes_per_row": head["bytes_per_row"],
                     "rows_per_launch": n, "kernel_ms": head["kernel_ms_timed"],
                     "kernel_ms_how": "mean launch duration inside the timed graph replay "
                                      "(HIP events around the replay, time per step / "
                                      f"{head['kernel_launches_per_step']} launches of this "
                                      "kernel per step)",
                     "kernel_ms_eager_events": head["dir_ms"][dom]},
        "max_rel_err_vs_fp64_oracle": check.get("ir_max_rel_err"),
        "refine_steps": {"timed_max": head["refine_steps_timed_max"],
                         "one_step_from_initial_field": head["refine_steps_checked"]},
        "per_direction_ms": head["dir_ms"],
        "kernels": head["kernels"],
        "step_GBs": head["step_GBs"],
        "wall_ms_per_step": head["wall_ms_per_step"],
        "fp64": {"value": res["fp64"]["grows"], "ms_per_step": res["fp64"]["ms_per_step"],
                 "per_direction_ms": res["fp64"]["dir_ms"],
                 "roofline_frac": res["fp64"]["achieved_GBs"] / HBM_PEAK_GBS,
                 "kernels": res["fp64"]["kernels"],
                 "max_rel_err_vs_fp64_oracle": check.get("fp64_max_rel_err")},
        "fp32": {"value": res["fp32"]["grows"], "ms_per_step": res["fp32"]["ms_per_step"],
                 "per_direction_ms": res["fp32"]["dir_ms"],
                 "roofline_frac": res["fp32"]["achieved_GBs"] / HBM_PEAK_GBS,
                 "kernels": res["fp32"]["kernels"],
                 "max_rel_err_vs_fp64_oracle": check.get("fp32_max_rel_err")},
        "ir_speedup_vs_fp64": res["fp64"]["ms_per_step"] / head["ms_per_step"],
        "dropin_fp64_inplace": {"per_direction_ms": dropin_ms, "kernels": dropin_k,
                                "grows": 3 * n / (sum(dropin_ms) * 1e-3) / 1e9},
        "dropin_ir_inplace": {"per_direction_ms": dropin_ir_ms, "kernels": dropin_ir_k,
                              "grows": 3 * n / (sum(dropin_ir_ms) * 1e-3) / 1e9},
        "dropin_fp64_autotuned": {"per_direction_ms": tuned_ms, "kernels": tuned_k,
                                  "grows": 3 * n / (sum(tuned_ms) * 1e-3) / 1e9,
                                  "how": "fresh handle, PENTA_AUTOTUNE=1: the first solve of "
                                         "each direction times every configuration on the "
                                         "same arrays (untimed), later solves use the winner"},
    }
    if args.probe:
        result["probe_ir_byte_mix"] = probe
        result["probe_strided_io"] = strided
    if not args.no_cpu:
        log("bench: cpu baseline (reference CPU solver)")
        result["cpu_baseline"] = cpu_baseline(c64_np, b_np, dims)
    del c64_np, b_np, c64, c32, b32, ds, dl, d, du, dw, b
    torch.cuda.empty_cache()
    if not args.no_extra:
        chk = not args.no_check
        log("bench: weak set (c = 1.05 sum|off|)")
        result["weak_set"] = weak_set(P, torch, W, h, args.size, args.steps, args.max_refine,
                                      args.tol, chk)
        torch.cuda.empty_cache()
        log("bench: 256^3, 50 iterations, FP64 / FP32 / shared-fact")
        wl = {"c2_256cubed": workload_c2(P, torch, W, h, chk)}
        torch.cuda.empty_cache()
        log("bench: 8192 variable-length systems")
        wl["c4_varlen"] = workload_c4(P, torch, W, h, chk)
        torch.cuda.empty_cache()
        log("bench: hyperdiffusion 384x256x640, 100 steps")
        wl["c5_hyperdiffusion"] = workload_c5(P, torch, W, h, chk, cpu_steps=args.c5_cpu_steps,
                                              max_refine=args.max_refine,
                                              tol=args.c5_tol)
        log("bench: distributed z solve, one GPU's share (8 slabs of 512x512x64)")
        wl["dist_z_slabs"] = workload_dist(P, torch, W, h, chk, size=args.size)
        torch.cuda.empty_cache()
        log("bench: 2-D cavity consumer pressure step, assembled vs fused")
        wl["cavity_pressure_step"] = workload_cavity(P, torch, W, h, chk)
        result["workloads"] = wl
    h.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
