"""Autotune sweep: time every kernel configuration per direction and report
algorithmic GB/s against the HBM peak and the measured probe bandwidth.

    python pentadiagonal-solver_amd/tune.py --size 512 --mode f64 f32 ir

Bytes per row: f64 56 (5 coefficients + RHS in, solution out, 8 B each),
f32 28, ir 36 (coefficients FP32, RHS/solution FP64).
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK = 8000.0


def _pkg():
    if "pentadsolver_amd" in sys.modules:
        return sys.modules["pentadsolver_amd"]
    spec = importlib.util.spec_from_file_location(
        "pentadsolver_amd", os.path.join(HERE, "__init__.py"), submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pentadsolver_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def probe(P, torch, n, dtype, reps=5):
    ins = [torch.ones(n, dtype=dtype, device="cuda") for _ in range(6)]
    out = torch.empty(n, dtype=dtype, device="cuda")
    P.probe_stream(ins, out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        P.probe_stream(ins, out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del ins, out
    torch.cuda.empty_cache()
    return 7 * n * torch.finfo(dtype).bits / 8 / (ms * 1e-3) / 1e9


def sweep(size, modes, reps=3, dims=None):
    import torch
    P = _pkg()
    from pentadsolver_amd import workloads as W
    P.load_library()
    dims = dims or [size] * 3
    n = 1
    for v in dims:
        n *= v
    results = {"dims": dims}
    h = P.Handle(stream=torch.cuda.current_stream())
    for mode in modes:
        dt = torch.float64 if mode in ("f64", "ir") else torch.float32
        ds, dl, d, du, dw, b = W.dominant_torch(dims, 3, dtype=torch.float64)
        if mode == "ir":
            coeffs = [a.float() for a in (ds, dl, d, du, dw)]
            x = b
            bpr = 36
        else:
            coeffs = [a.to(dt) for a in (ds, dl, d, du, dw)]
            x = b.to(dt)
            bpr = 7 * (8 if dt == torch.float64 else 4)
        del ds, dl, d, du, dw
        torch.cuda.empty_cache()
        results[mode] = {}
        for sd in range(len(dims)):
            rows = P.autotune(*coeffs, x, dims, sd, h, reps=reps, ir=(mode == "ir"))
            table = []
            for cfg, ms in rows:
                gbs = bpr * n / (ms * 1e-3) / 1e9
                table.append({"config": cfg, "ms": ms, "GBs": gbs, "frac": gbs / HBM_PEAK})
            table.sort(key=lambda r: r["ms"])
            results[mode][sd] = table
            print(f"[{mode} dir {sd}] best {table[0]['config']} {table[0]['ms']:.3f} ms "
                  f"{table[0]['GBs']:.0f} GB/s ({100 * table[0]['frac']:.1f}%)", flush=True)
            for r in table[1:]:
                print(f"      {r['config']:<40} {r['ms']:.3f} ms {r['GBs']:.0f} GB/s", flush=True)
        del coeffs, x
        torch.cuda.empty_cache()
    results["probe_GBs"] = {"f64": probe(P, torch, n, torch.float64),
                            "f32": probe(P, torch, n, torch.float32)}
    print(f"probe 6r/1w: f64 {results['probe_GBs']['f64']:.0f} GB/s, "
          f"f32 {results['probe_GBs']['f32']:.0f} GB/s", flush=True)
    h.close()
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--dims", type=int, nargs="*")
    ap.add_argument("--mode", nargs="*", default=["f64", "f32", "ir"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--json")
    args = ap.parse_args()
    res = sweep(args.size, args.mode, args.reps, args.dims)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
