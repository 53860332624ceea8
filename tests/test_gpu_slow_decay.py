"""Systems whose separator couplings decay slowly, on the GPU vs the FP64 oracle.

The partitioned kernels truncate block PCR once every coupling of a round is
at most 2^-28 (FP32) / 2^-57 (FP64), and the refinement kernels store the
multipliers of at most three rounds, redoing any later ones inside every
solve.  Diagonally dominant systems stop after two or three rounds; these
operators, I + diag(s) [1, -4, 6, -4, 1] with s = 100 U(0.5, 2) (not
diagonally dominant, condition numbers about 3000), keep couplings above
the thresholds up to about 128 rows apart, so with 64 chunks of 8 rows they
need four rounds in FP32 and five in FP64: the refinement runs rounds after
its three stored ones, the plain solves most of theirs.

Bars: FP64 and FP32 + refinement <= 1e-12 max-norm relative forward error;
plain FP32 (s up to 4) <= 1e-5 relative to max|x|.
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def stiff(dims, seed, s0):
    """Row-wise fourth-difference operator plus identity, s = s0 * U(0.5, 2)."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    s = s0 * rng.uniform(0.5, 2.0, n)
    ds, dl, d, du, dw = s, -4.0 * s, 1.0 + 6.0 * s, -4.0 * s, s.copy()
    b = rng.standard_normal(n)
    return [ds, dl, d, du, dw], b


def relerr(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


SHAPES = [([512, 16, 8], 0), ([384, 8, 6], 0), ([16, 512, 8], 1), ([8, 16, 512], 2),
          ([256, 32, 16], 0), ([640, 4, 4], 0)]


@pytest.mark.parametrize("dims,sd", SHAPES)
def test_fp64_slow_decay(dims, sd):
    a, b = stiff(dims, 11, 100.0)
    print(f"fp64 stiff dims={dims} sd={sd}", flush=True)
    t = [torch.from_numpy(v).cuda() for v in a]
    x = torch.from_numpy(b.copy()).cuda()
    P.solve(*t, x, dims, sd)
    torch.cuda.synchronize()
    ref = O.solve(*a, b.copy(), dims, sd)
    err = relerr(x.cpu().numpy(), ref)
    print(f"  err={err:.2e} kernel={P.kernel_that_ran(sd)}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("dims,sd", SHAPES)
def test_fp32_slow_decay(dims, sd):
    a, b = stiff(dims, 12, 2.0)
    a32 = [v.astype(np.float32) for v in a]
    print(f"fp32 stiff dims={dims} sd={sd}", flush=True)
    t = [torch.from_numpy(v).cuda() for v in a32]
    x = torch.from_numpy(b.astype(np.float32)).cuda()
    P.solve(*t, x, dims, sd)
    torch.cuda.synchronize()
    ref = O.solve(*[v.astype(np.float64) for v in a32], b.astype(np.float32).astype(np.float64),
                  dims, sd)
    err = relerr(x.cpu().numpy().astype(np.float64), ref)
    print(f"  err={err:.2e} kernel={P.kernel_that_ran(sd)}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("dims,sd", SHAPES)
@pytest.mark.parametrize("algo", ["auto", "thomas-pcr"])
def test_ir_slow_decay(dims, sd, algo):
    """Mixed precision: the automatic choice (along y / z the FP64 solve of
    the widened coefficients, no refinement sweep) and the refinement
    kernels asked for by name."""
    a, b = stiff(dims, 13, 100.0)
    a32 = [v.astype(np.float32) for v in a]
    print(f"ir stiff dims={dims} sd={sd} {algo}", flush=True)
    t = [torch.from_numpy(v).cuda() for v in a32]
    x = torch.from_numpy(b.copy()).cuda()
    steps = torch.zeros(1, dtype=torch.int32, device="cuda")
    d = sd if len(dims) == 3 or sd == 0 else 2
    with P.Handle() as h:
        h.set_algo(d, algo)
        P.solve_ir(*t, x, dims, sd, max_refine=8, tol=1e-13, steps=steps, handle=h)
        torch.cuda.synchronize()
        cfg = h.kernel_config(d)
    ref = O.solve(*[v.astype(np.float64) for v in a32], b.copy(), dims, sd)
    err = relerr(x.cpu().numpy(), ref)
    print(f"  err={err:.2e} steps={int(steps.item())} kernel={cfg}", flush=True)
    assert err < 1e-12
    if cfg.startswith("ir widened"):
        assert algo == "auto" and sd != 0 and int(steps.item()) == 0
    else:
        assert 1 <= int(steps.item()) <= 8
