"""FP32 solve + FP64 iterative refinement on the GPU vs the FP64 oracle.

The oracle solves widen(A32) x = b in FP64 (the system the refinement
converges to).  Bar: max-norm relative forward error <= 1e-12 for both the
well-conditioned (dominance ratio 2-3) and the weakly dominant (1.05) sets.
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def gen32(dims, seed, ratio=(2.0, 3.0)):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    r = rng.uniform(*ratio, n) if ratio[1] > ratio[0] else np.full(n, ratio[0])
    d = r * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    a32 = [a.astype(np.float32) for a in (ds, dl, d, du, dw)]
    return a32, b


def run_ir(a32, b, dims, sd, handle=None, max_refine=5, tol=1e-13):
    t = [torch.from_numpy(a).cuda() for a in a32]
    x = torch.from_numpy(b.copy()).cuda()
    steps = torch.zeros(1, dtype=torch.int32, device="cuda")
    P.solve_ir(*t, x, dims, sd, max_refine=max_refine, tol=tol, steps=steps, handle=handle)
    torch.cuda.synchronize()
    return x.cpu().numpy(), int(steps.item())


def steps_ok(cfg, steps, lo=1, hi=5):
    """The FP64 solve of the widened coefficients (strided lines, automatic
    choice) takes no refinement sweep; the refinement kernels take 1..5."""
    return steps == 0 if cfg.startswith("ir widened") else lo <= steps <= hi


def reference(a32, b, dims, sd):
    a64 = [a.astype(np.float64) for a in a32]
    return O.solve(*a64, b.copy(), dims, sd)


def relerr(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


SHAPES = [([512, 16, 8], 0), ([16, 512, 8], 1), ([16, 8, 512], 2), ([256, 32, 16], 0),
          ([32, 256, 16], 1), ([16, 32, 256], 2), ([384, 12, 10], 0), ([10, 12, 384], 2),
          ([64, 64, 64], 0), ([64, 64, 64], 1), ([64, 64, 64], 2),
          ([37, 9, 7], 0), ([9, 37, 7], 1), ([7, 9, 37], 2)]


@pytest.mark.parametrize("dims,sd", SHAPES)
@pytest.mark.parametrize("cond", ["well", "weak"])
def test_ir_matches_fp64_oracle(dims, sd, cond):
    ratio = (2.0, 3.0) if cond == "well" else (1.05, 1.05)
    a32, b = gen32(dims, 3 if cond == "well" else 4, ratio)
    print(f"ir dims={dims} sd={sd} {cond}", flush=True)
    with P.Handle() as h:
        x, steps = run_ir(a32, b, dims, sd, handle=h)
        cfg = h.kernel_config(sd)
    err = relerr(x, reference(a32, b, dims, sd))
    print(f"  err={err:.2e} steps={steps} {cfg}", flush=True)
    assert err < TOL
    assert steps_ok(cfg, steps), (cfg, steps)
    assert cfg.startswith("ir widened") == (sd != 0)


def test_ir_fixed_steps_when_tol_zero():
    dims = [256, 16, 4]
    a32, b = gen32(dims, 5)
    x, steps = run_ir(a32, b, dims, 0, max_refine=3, tol=0.0)
    assert steps == 3
    assert relerr(x, reference(a32, b, dims, 0)) < TOL


@pytest.mark.parametrize("sd,algo", [(0, "auto"), (1, "thomas-pcr"), (2, "thomas-pcr")])
def test_ir_zero_steps_is_plain_fp32(sd, algo):
    """The refinement kernels with no sweep return the FP32 solve."""
    dims = [256, 16, 4] if sd == 0 else ([16, 256, 4] if sd == 1 else [16, 4, 256])
    a32, b = gen32(dims, 6)
    with P.Handle() as h:
        h.set_algo(sd, algo)
        x, steps = run_ir(a32, b, dims, sd, max_refine=0, handle=h)
        assert h.kernel_config(sd).startswith("ir thomas-pcr"), h.kernel_config(sd)
    err = relerr(x, reference(a32, b, dims, sd))
    assert steps == 0
    assert 1e-12 < err < 2e-5  # FP32 accuracy only, refinement needed


@pytest.mark.parametrize("sd", [1, 2])
def test_ir_strided_auto_is_one_fp64_solve(sd):
    """Along y / z the automatic choice solves the widened coefficients in
    FP64: FP64 accuracy with no refinement sweep, whatever max_refine says;
    asking for "thomas-pcr" by name gives the refinement kernel, with the
    same result to FP64 rounding."""
    dims = [16, 512, 8] if sd == 1 else [16, 8, 512]
    a32, b = gen32(dims, 9)
    ref = reference(a32, b, dims, sd)
    with P.Handle() as h:
        x0, s0 = run_ir(a32, b, dims, sd, max_refine=0, handle=h)
        cfg = h.kernel_config(sd)
        h.set_algo(sd, "thomas-pcr")
        x1, s1 = run_ir(a32, b, dims, sd, handle=h)
        cfg1 = h.kernel_config(sd)
    print(f"ir strided {cfg}: {relerr(x0, ref):.2e} | {cfg1}: {relerr(x1, ref):.2e} steps {s1}",
          flush=True)
    assert cfg.startswith("ir widened") and s0 == 0
    assert relerr(x0, ref) < TOL
    assert cfg1.startswith("ir thomas-pcr") and 1 <= s1 <= 5
    assert relerr(x1, ref) < TOL
    assert relerr(x0, x1) < 1e-13


IRX = [(4, 8), (4, 64), (6, 16), (6, 64), (8, 4), (8, 32), (8, 64), (10, 16), (10, 32), (10, 64)]


def test_ir_x_all_configs():
    with P.Handle() as h:
        for rows, lanes in IRX:
            n = rows * lanes
            dims = [n, 21, 3]
            print(f"irx rows={rows} lanes={lanes}", flush=True)
            h.set_kernel_params(0, rows, lanes)
            a32, b = gen32(dims, rows + lanes)
            x, _ = run_ir(a32, b, dims, 0, handle=h)
            assert h.kernel_config(0) == f"ir thomas-pcr rows={rows} lanes={lanes}"
            assert relerr(x, reference(a32, b, dims, 0)) < TOL


IRYZ = [(4, 16, 64), (4, 8, 256), (6, 8, 192), (8, 16, 256), (8, 8, 512), (8, 64, 64),
        (10, 8, 640), (10, 16, 160), (6, 4, 768)]


@pytest.mark.parametrize("sd", [1, 2])
def test_ir_yz_all_configs(sd):
    with P.Handle() as h:
        for rows, spw, n in IRYZ:
            dims = [24, n, 3] if sd == 1 else [24, 3, n]
            print(f"iryz sd={sd} rows={rows} spw={spw} n={n}", flush=True)
            h.set_kernel_params(sd, rows, spw)
            a32, b = gen32(dims, rows * spw)
            x, _ = run_ir(a32, b, dims, sd, handle=h)
            assert h.kernel_config(sd) == f"ir thomas-pcr rows={rows} sys_per_wg={spw}"
            assert relerr(x, reference(a32, b, dims, sd)) < TOL


def test_ir_lane_fallback_forced():
    dims = [256, 20, 6]
    a32, b = gen32(dims, 9)
    with P.Handle() as h:
        for sd in range(3):
            h.set_algo(sd, "naive")
            x, steps = run_ir(a32, b, dims, sd, handle=h)
            assert h.kernel_config(sd) == "ir naive"
            assert relerr(x, reference(a32, b, dims, sd)) < TOL


@pytest.mark.parametrize("n", [5, 37, 100, 300])
@pytest.mark.parametrize("sd", [1, 2])
def test_ir_yz_padded_lengths(n, sd):
    """y/z lengths without a whole-chunk split: the refinement kernel pads the
    last chunk with identity rows (not the one-lane fallback)."""
    dims = [24, n, 3] if sd == 1 else [24, 3, n]
    print(f"ir yz padded len={n} sd={sd}", flush=True)
    a32, b = gen32(dims, n + 7 * sd)
    with P.Handle() as h:
        x, steps = run_ir(a32, b, dims, sd, handle=h)
        cfg = h.kernel_config(sd)
        assert cfg.startswith(("ir thomas-pcr", "ir widened")), cfg
    assert relerr(x, reference(a32, b, dims, sd)) < TOL
    assert steps_ok(cfg, steps), (cfg, steps)


@pytest.mark.parametrize("n", [3, 5, 37, 100, 300, 500])
def test_ir_x_any_length(n):
    """x lengths without a whole-chunk split: coefficient spans staged through
    LDS into the refinement kernel (not the one-lane fallback)."""
    dims = [n, 13, 5]
    print(f"ir x any length len={n}", flush=True)
    a32, b = gen32(dims, 3 * n + 1)
    with P.Handle() as h:
        x, steps = run_ir(a32, b, dims, 0, handle=h)
        assert h.kernel_config(0).startswith("ir transpose"), h.kernel_config(0)
    assert relerr(x, reference(a32, b, dims, 0)) < TOL
    assert 1 <= steps <= 5


def test_ir_x_forced_transpose_on_exact_length():
    dims = [512, 9, 5]
    a32, b = gen32(dims, 77)
    with P.Handle() as h:
        h.set_algo(0, "transpose")
        x, _ = run_ir(a32, b, dims, 0, handle=h)
        assert h.kernel_config(0) == "ir transpose rows=8 lanes=64"
    assert relerr(x, reference(a32, b, dims, 0)) < TOL


@pytest.mark.parametrize("n,cfg", [(768, "ir widened rows=12 lanes=64"),
                                   (1024, "ir widened rows=8 lanes=128"),
                                   (640 + 128, "ir widened rows=12 lanes=64"),
                                   (2048, "ir widened rows=8 lanes=256"),
                                   (1000, "ir widened rows="),
                                   (5000, "ir widened rows=")])
@pytest.mark.parametrize("cond", ["well", "weak"])
def test_ir_x_long_lines_widened(n, cfg, cond):
    """x lengths no refinement kernel covers (past 64 chunks of 10 rows, or
    no power-of-two split past 512 rows): one FP64 solve of the coefficients
    widened on load, by the FP64 x kernel of whole chunks (up to 256 lanes of
    8 rows) or, for other lengths, the workgroup-wide kernel -- the system the
    refinement converges to (<= 1e-12 against the oracle of the widened
    coefficients, no refinement step), where these lengths took the
    one-lane kernel with its scratch passes before."""
    dims = [n, 7, 3]
    print(f"ir x long len={n} {cond}", flush=True)
    a32, b = gen32(dims, n + 11, (2.0, 3.0) if cond == "well" else (1.05, 1.05))
    with P.Handle() as h:
        x, steps = run_ir(a32, b, dims, 0, handle=h)
        got = h.kernel_config(0)
    print(f"  {got} err={relerr(x, reference(a32, b, dims, 0)):.2e}", flush=True)
    assert got.startswith(cfg), got
    assert relerr(x, reference(a32, b, dims, 0)) < TOL
    assert steps == 0
