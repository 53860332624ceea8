"""Hyperdiffusion ADI (one operator I + nu dt/h^4 [1,-4,6,-4,1] per direction,
nu per grid line) on the GPU vs the CPU oracle, at a reduced anisotropic grid
whose line lengths use the same kernel shapes as 384x256x640 (6, 4/8 and 10
rows per lane).

Bars: FP64 and FP32 + FP64 refinement within 1e-12 max-norm relative error of
the oracle after 10 steps (refinement against the oracle of the FP32-stored
coefficients widened to FP64); every FP64 step shrinks the 2-norm (the
operator is symmetric positive definite with eigenvalues >= 1); the 10-step
hipGraph replay equals 10 eager steps bit for bit.
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from pentadsolver_amd import workloads as W
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DIMS = [48, 32, 80]
STEPS = 10


def oracle_run(sets, b, steps, dims=DIMS):
    x = b.copy()
    for _ in range(steps):
        for sd in range(3):
            x = O.solve(*sets[sd], x, dims, sd)
    return x


def relerr(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


@pytest.mark.parametrize("mode", ["fp64", "ir"])
def test_hyperdiffusion_matches_oracle(mode):
    print(f"hyperdiffusion {DIMS} {mode}", flush=True)
    sets = W.hyperdiffusion_np(DIMS, 7)
    b = np.random.default_rng(8).standard_normal(int(np.prod(DIMS)))
    if mode == "fp64":
        tsets = [[torch.from_numpy(a).cuda() for a in cs] for cs in sets]
        ref = oracle_run(sets, b, STEPS)
        m = P.ADI_FP64
    else:
        s32 = [[a.astype(np.float32) for a in cs] for cs in sets]
        tsets = [[torch.from_numpy(a).cuda() for a in cs] for cs in s32]
        ref = oracle_run([[a.astype(np.float64) for a in cs] for cs in s32], b, STEPS)
        m = P.ADI_IR
    x = torch.from_numpy(b.copy()).cuda()
    with P.Adi(m, DIMS, tsets) as adi:
        adi.capture(x, STEPS)
        adi.replay(1)
        torch.cuda.synchronize()
        graph = x.cpu().numpy()
        x.copy_(torch.from_numpy(b))
        norms = [np.linalg.norm(b)]
        for _ in range(STEPS):
            adi.step(x, 1)
            torch.cuda.synchronize()
            norms.append(float(torch.linalg.norm(x).item()))
        eager = x.cpu().numpy()
        cfg = [adi.config(k) for k in range(3)]
    print(f"  kernels {cfg} err={relerr(eager, ref):.2e}", flush=True)
    assert np.array_equal(graph, eager)
    assert relerr(eager, ref) < 1e-12
    assert all(b2 < a2 for a2, b2 in zip(norms, norms[1:])), norms


@pytest.mark.parametrize("dims", [[384, 256, 12], [24, 16, 640]])
@pytest.mark.parametrize("mode", ["fp64", "ir"])
def test_hyperdiffusion_at_full_line_lengths(dims, mode):
    """The hyperdiffusion operators at the 384x256x640 grid's line lengths
    (the 64- and 32-lane kernels of that workload), two steps against the
    oracle: <= 1e-12."""
    print(f"hyperdiffusion {dims} {mode}", flush=True)
    sets = W.hyperdiffusion_np(dims, 7)
    b = np.random.default_rng(8).standard_normal(int(np.prod(dims)))
    if mode == "ir":
        sets = [[a.astype(np.float32) for a in cs] for cs in sets]
    tsets = [[torch.from_numpy(a).cuda() for a in cs] for cs in sets]
    ref = oracle_run([[a.astype(np.float64) for a in cs] for cs in sets], b, 2, dims)
    x = torch.from_numpy(b.copy()).cuda()
    with P.Adi(P.ADI_IR if mode == "ir" else P.ADI_FP64, dims, tsets) as adi:
        adi.step(x, 2)
        torch.cuda.synchronize()
        cfg = [adi.config(k) for k in range(3)]
    err = relerr(x.cpu().numpy(), ref)
    print(f"  kernels {cfg} err={err:.2e}", flush=True)
    assert err < 1e-12
