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
