"""Shared factorisation ("shared-fact") on the GPU vs the CPU oracle.

The mode factors system 0 of a direction once and applies that factorisation
to every system, as the reference's shared-fact kernels do
(src/cuda/singlenode/pentad_cuda.cu:908-1195, system 0 trusted at :962-963).
Invariant checked here (the method of verify_shared_fact_cuda.cu): with
coefficients replicated along the direction, the result equals the general
FP64 oracle solve to 1e-12 max-norm relative forward error (FP64) or 2e-5
(FP32, vs the FP64 oracle of the float32-rounded inputs).  With coefficients
that differ between systems, the result equals the oracle solve of system 0's
coefficients replicated.
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-5


def relerr(a, ref):
    return float(np.max(np.abs(np.asarray(a, np.float64) - ref)) / np.max(np.abs(ref)))


def profile(n, seed, ratio=(2.0, 3.0)):
    """One line of the five diagonals (dominance recipe of the configs)."""
    rng = np.random.default_rng(seed)
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(*ratio, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    return [ds, dl, d, du, dw]


def replicate(prof, dims, sd):
    """Full arrays (caller's layout, dims[0] fastest) whose every line along
    sd is `prof`."""
    shape = tuple(reversed(dims))          # numpy order: slowest first
    axis = len(dims) - 1 - sd
    view = [1] * len(dims)
    view[axis] = dims[sd]
    return [np.ascontiguousarray(np.broadcast_to(p.reshape(view), shape)).reshape(-1)
            for p in prof]


def run(coefs, b, dims, sd, dtype, algo="shared-fact"):
    t = [torch.from_numpy(np.ascontiguousarray(a.astype(dtype))).cuda() for a in coefs]
    x = torch.from_numpy(b.astype(dtype)).cuda()
    with P.Handle() as h:
        h.set_algo(sd, algo)
        P.solve(*t, x, dims, sd, handle=h)
        torch.cuda.synchronize()
        cfg = h.kernel_config(sd)
    return x.cpu().numpy(), cfg, P.kernel_that_ran(sd)


SHAPES = [([512, 6, 5], 0), ([256, 8, 8], 0), ([384, 4, 3], 0), ([64, 20, 3], 0),
          ([48, 16, 4], 0), ([16, 40, 2], 0), ([128, 9, 7], 0),
          ([9, 512, 7], 1), ([16, 256, 8], 1), ([12, 96, 5], 1), ([8, 640, 3], 1),
          ([9, 7, 512], 2), ([16, 8, 256], 2), ([20, 6, 64], 2), ([5, 6, 48], 2)]


@pytest.mark.parametrize("dims,sd", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_shared_fact_replicated(dims, sd, dtype):
    print(f"shared-fact dims={dims} sd={sd} {np.dtype(dtype).name}", flush=True)
    n = int(np.prod(dims))
    prof = profile(dims[sd], 11 + dims[sd])
    coefs = [a.astype(dtype) for a in replicate(prof, dims, sd)]
    b = np.random.default_rng(5).uniform(-1, 1, n)
    x, cfg, ran = run(coefs, b, dims, sd, dtype)
    assert ran == "shared-fact" and cfg.startswith("shared-fact"), cfg
    ref = O.solve(*[a.astype(np.float64) for a in coefs], b.astype(dtype).astype(np.float64),
                  dims, sd)
    err = relerr(x, ref)
    print(f"  {cfg} err={err:.2e}", flush=True)
    assert err < (FP64_TOL if dtype == np.float64 else FP32_TOL)


@pytest.mark.parametrize("sd", [0, 1, 2])
def test_shared_fact_uses_system_zero(sd):
    """Coefficients that differ per system: shared-fact solves every system
    with system 0's coefficients (the reference's documented contract)."""
    dims = [64, 32, 48]
    n = int(np.prod(dims))
    rng = np.random.default_rng(21)
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(2, 3, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    coefs = [ds, dl, d, du, dw]
    b = rng.uniform(-1, 1, n)
    x, cfg, ran = run(coefs, b, dims, sd, np.float64)
    assert ran == "shared-fact", cfg
    # system 0's line along sd, replicated
    shape = tuple(reversed(dims))
    axis = len(dims) - 1 - sd
    idx = [0, 0, 0]
    idx[axis] = slice(None)
    prof = [a.reshape(shape)[tuple(idx)] for a in coefs]
    ref = O.solve(*replicate(prof, dims, sd), b, dims, sd)
    assert relerr(x, ref) < FP64_TOL


def test_shared_fact_unsupported_length_falls_back():
    """A length the shared-fact kernels do not cover (37) is solved by the
    automatic choice, and kernel_that_ran says so."""
    dims = [37, 20, 3]
    n = int(np.prod(dims))
    coefs = replicate(profile(37, 3), dims, 0)
    b = np.random.default_rng(2).uniform(-1, 1, n)
    x, cfg, ran = run(coefs, b, dims, 0, np.float64)
    assert ran != "shared-fact"
    ref = O.solve(*coefs, b, dims, 0)
    assert relerr(x, ref) < FP64_TOL


def test_shared_fact_buffer_is_the_table():
    dims = [512, 64, 64]
    t = [torch.zeros(int(np.prod(dims)), dtype=torch.float64, device="cuda") for _ in range(6)]
    with P.Handle() as h:
        h.set_algo(0, "shared-fact")
        nb = P.gpsv_batch_buffer_size_ext(*t[:5], t[5], dims, 0, h)
        h.set_algo(0, "auto")
        nb_auto = P.gpsv_batch_buffer_size_ext(*t[:5], t[5], dims, 0, h)
    # 124 values x 64 chunks x 8 bytes for rows=8 (7*6 + 10 + 12*6 values per chunk)
    assert nb == 124 * 64 * 8
    assert nb_auto == 0
