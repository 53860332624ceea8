"""A C++ caller written against the reference's interface (pentadsolver.hpp
overloads, create / buffer size / gpsv_batch / kernel_that_ran / destroy, the
x -> y -> z step of apps/adi_cuda.cu:167-176) linked against this library and
run as a child process; its output is checked against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pentadiagonal-solver_amd", "lib", "adi_dropin")


def gen(dims, seed, dtype):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(2, 3, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    return [a.astype(dtype) for a in (ds, dl, d, du, dw, b)]


@pytest.mark.parametrize("dims", [[64, 48, 40], [37, 50, 21], [256, 32, 16]])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_cpp_caller_adi_iterations(tmp_path, dims, prec):
    assert os.path.exists(EXE), "build the package first (make -C pentadiagonal-solver_amd)"
    dtype = np.float64 if prec == "f64" else np.float32
    arrs = gen(dims, sum(dims), dtype)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate(arrs).tofile(fin)
    iters = 2
    print(f"cpp caller dims={dims} {prec}", flush=True)
    r = subprocess.run([EXE, str(fin), str(fout), *map(str, dims), str(iters), prec],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)
    x = np.fromfile(fout, dtype=dtype)
    w = [a.astype(np.float64) for a in arrs]
    ref = w[5].copy()
    for _ in range(iters):
        for sd in range(3):
            ref = O.solve(*w[:5], ref, dims, sd)
    err = float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))
    assert err < (1e-12 if prec == "f64" else 5e-5), err
