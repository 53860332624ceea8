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


CAVITY_EXE = os.path.join(ROOT, "pentadiagonal-solver_amd", "lib", "cavity_dropin")


@pytest.mark.parametrize("nx,ny", [(30, 20), (64, 64), (250, 33)])
def test_cpp_caller_cavity_pressure_step(tmp_path, nx, ny):
    """The 2-D cavity consumer's pressure step from C++ (examples/
    cavity_dropin.cpp): assembly + padded x / y solves with the reference's
    nx_pad (nx rounded up to 4), and the fused call, against the oracle's
    assembly (nx_pad = nx, embedded) solved in long double; the padding
    untouched by both."""
    import importlib.util
    assert os.path.exists(CAVITY_EXE), "build the package first (make -C pentadiagonal-solver_amd)"
    spec = importlib.util.spec_from_file_location(
        "_workloads", os.path.join(ROOT, "pentadiagonal-solver_amd", "workloads.py"))
    W = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(W)
    nx_pad = -(-nx // 4) * 4
    up, vp, uv, aU, dx, dy = W.cavity_np(nx, ny, seed=nx + ny)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([up, vp, uv]).tofile(fin)
    print(f"cpp cavity nx={nx} ny={ny} nx_pad={nx_pad}", flush=True)
    r = subprocess.run([CAVITY_EXE, str(fin), str(fout), str(nx), str(ny), str(nx_pad)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)
    out = np.fromfile(fout, dtype=np.float64).reshape(2, -1)
    o = O.cavity_pressure_coef(nx, ny, nx, aU, dx, dy, up, vp, uv)
    keys = ("s_x", "l_x", "d_x", "u_x", "w_x")
    xs = O.solve_ld(*[o[k] for k in keys], o["x"], [nx, ny], 0)
    ref = O.solve_ld(*[o[k] for k in keys], xs, [nx, ny], 1)
    mask = O.padding_mask([nx, ny], [nx_pad, ny])
    for k in range(2):
        got = O.extract_padded(out[k], [nx, ny], [nx_pad, ny])
        err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        assert err < 1e-12, (k, err)
        assert np.all(out[k][mask] == 0.0)
