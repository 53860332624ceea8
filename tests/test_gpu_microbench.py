"""The micro-benchmark harness (the reference's benchmark matrix,
benchmarks/cuda/my_bm.cu:55-71) runs every case through the C ABI and every
solve meets the parity bar: FP64 <= 1e-12, FP32 <= 1e-5 (reference rule,
test/catch_utils.hpp:19-25) against the FP64 oracle."""
import importlib

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_microbench_matrix_small():
    mb = importlib.import_module("microbench")
    assert len(mb.matrix(256)) == 24
    rows = mb.main(["--max-n", "64", "--reps", "2", "--warmup", "1"])
    assert len(rows) == 2 * len(mb.matrix(64))
    for r in rows:
        print(r, flush=True)
        bar = 1e-12 if r["dtype"] == "f64" else 1e-5
        assert r["max_rel_err_vs_fp64_oracle"] is not None
        assert r["max_rel_err_vs_fp64_oracle"] < bar
        assert r["ms_per_call"] > 0


def test_native_microbench_matrix():
    """lib/microbench_dropin, the same matrix run natively through the C++
    overloads (no Python in the timed loop): one line per case, every first
    solve's host-side FP64 residual |A x - b| / |b| at FP64 (<= 1e-13) or
    FP32 (<= 1e-5) rounding, positive times."""
    import os
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "..", "pentadiagonal-solver_amd", "lib",
                       "microbench_dropin")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: build it with make -C pentadiagonal-solver_amd")
    out = subprocess.run([exe, "64", "3"], capture_output=True, text=True, timeout=100)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    lines = [l for l in out.stdout.splitlines() if l.strip()]
    assert len(lines) == 2 * 2 * (1 + 2 + 3)  # precision x N in {32, 64} x (ndims, solvedim)
    for l in lines:
        kv = dict(re.findall(r"(\w+)=(\S+)", l))
        bar = 1e-13 if l.startswith("f64") else 1e-5
        assert float(kv["resid"]) < bar, l
        assert float(kv["gpu_us"]) > 0 and float(kv["launch_us"]) > 0, l
