"""The batched-solve autotuner on the GPU: pentadsolver_D_autotune,
pentadsolver_S_autotune and pentadsolver_DS_autotune_ir (no reference
counterpart; they replace the fixed rules and PENTA_PCR_LANES of
src/cuda/singlenode/pentad_cuda.cu:613-697).

For every entry point and direction: the right-hand side is bit for bit
unchanged by the tuning run, the next solve of that length runs the
configuration the report names fastest and matches the oracle (FP64 and
refinement <= 1e-12, FP32 2e-5), and pentadsolver_clear_tuning brings the
untuned choice back.
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def gen(dims, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(2, 3, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    return [ds, dl, d, du, dw], rng.uniform(-1, 1, n)


def relerr(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


CASES = [([512, 64, 24], 0), ([256, 40, 24], 0), ([32, 384, 16], 1), ([24, 40, 256], 2), ([192, 40, 30], 0)]


@pytest.mark.parametrize("dims,sd", CASES)
@pytest.mark.parametrize("prec", ["fp64", "fp32", "ir"])
def test_autotune_then_solve(dims, sd, prec):
    print(f"autotune {prec} dims={dims} sd={sd}", flush=True)
    coef, b = gen(dims, 40 + sd)
    if prec == "fp64":
        c, xh = coef, b
    elif prec == "fp32":
        c, xh = [a.astype(np.float32) for a in coef], b.astype(np.float32)
    else:
        c, xh = [a.astype(np.float32) for a in coef], b
    ref = O.solve(*[a.astype(np.float64) for a in c], xh.astype(np.float64), dims, sd)
    tol = 2e-5 if prec == "fp32" else 1e-12
    t = [torch.from_numpy(a).cuda() for a in c]

    def solve(h):
        x = torch.from_numpy(xh.copy()).cuda()
        if prec == "ir":
            P.solve_ir(*t, x, dims, sd, handle=h)
        else:
            P.solve(*t, x, dims, sd, handle=h)
        torch.cuda.synchronize()
        return x.cpu().numpy(), h.kernel_config(sd)

    with P.Handle() as h:
        x0, cfg0 = solve(h)
        assert relerr(x0, ref) < tol
        xt = torch.from_numpy(xh.copy()).cuda()
        res = P.autotune(*t, xt, dims, sd, h, reps=2, ir=(prec == "ir"))
        torch.cuda.synchronize()
        # the tuning run leaves the right-hand side as it was, bit for bit
        assert np.array_equal(xt.cpu().numpy(), xh)
        assert len(res) >= 1, res
        # the report prints 0.1-us resolution: any configuration at the
        # printed minimum may be the one the tuner measured fastest
        best_ms = min(ms for _, ms in res)
        best = [c for c, ms in res if ms == best_ms]
        x1, cfg1 = solve(h)
        print(f"  {len(res)} configs, best {best}; untuned {cfg0}; tuned {cfg1}; "
              f"err {relerr(x1, ref):.2e}", flush=True)
        assert cfg1 in best, (cfg1, best)
        assert relerr(x1, ref) < tol
        P.load_library().pentadsolver_clear_tuning(h.value)
        x2, cfg2 = solve(h)
        assert cfg2 == cfg0, (cfg2, cfg0)
        assert relerr(x2, ref) < tol


def test_autotune_keys_by_length_and_precision():
    """A tuned entry applies to its (precision, direction, length) only."""
    dims_a, dims_b = [512, 32, 8], [256, 32, 8]
    ca, ba = gen(dims_a, 1)
    cb, bb = gen(dims_b, 2)
    with P.Handle() as h:
        ta = [torch.from_numpy(a).cuda() for a in ca]
        tb = [torch.from_numpy(a).cuda() for a in cb]
        xb = torch.from_numpy(bb.copy()).cuda()
        P.solve(*tb, xb, dims_b, 0, handle=h)
        untuned_b = h.kernel_config(0)
        # tune length 512 only
        P.autotune(*ta, torch.from_numpy(ba.copy()).cuda(), dims_a, 0, h, reps=1)
        xb = torch.from_numpy(bb.copy()).cuda()
        P.solve(*tb, xb, dims_b, 0, handle=h)
        torch.cuda.synchronize()
        assert h.kernel_config(0) == untuned_b
        # FP32 of the same length is untouched by the FP64 tuning
        t32 = [a.float() for a in ta]
        x32 = torch.from_numpy(ba.astype(np.float32)).cuda()
        with P.Handle() as h2:
            P.solve(*t32, x32, dims_a, 0, handle=h2)
            want = h2.kernel_config(0)
        x32 = torch.from_numpy(ba.astype(np.float32)).cuda()
        P.solve(*t32, x32, dims_a, 0, handle=h)
        assert h.kernel_config(0) == want


@pytest.mark.parametrize("dims,sd", [([512, 64, 24], 0), ([32, 384, 16], 1), ([24, 40, 256], 2)])
def test_env_autotune_on_first_use(dims, sd, monkeypatch):
    """PENTA_AUTOTUNE=1: the first automatic-choice solve of a (precision,
    direction, length) tunes on the caller's arrays, solves with the winner
    and keeps it; without the variable nothing is tuned."""
    print(f"env autotune dims={dims} sd={sd}", flush=True)
    coef, b = gen(dims, 70 + sd)
    ref = O.solve(*coef, b, dims, sd)
    t = [torch.from_numpy(a).cuda() for a in coef]
    with P.Handle() as h:
        cands = {c for c, _ in P.autotune(*t, torch.from_numpy(b.copy()).cuda(), dims, sd, h,
                                          reps=1)}
        P.load_library().pentadsolver_clear_tuning(h.value)
        monkeypatch.delenv("PENTA_AUTOTUNE", raising=False)
        x = torch.from_numpy(b.copy()).cuda()
        P.solve(*t, x, dims, sd, handle=h)
        torch.cuda.synchronize()
        untuned = h.kernel_config(sd)
        monkeypatch.setenv("PENTA_AUTOTUNE", "1")
        cfgs = []
        for _ in range(2):
            x = torch.from_numpy(b.copy()).cuda()
            P.solve(*t, x, dims, sd, handle=h)
            torch.cuda.synchronize()
            cfgs.append(h.kernel_config(sd))
            assert relerr(x.cpu().numpy(), ref) < 1e-12
        print(f"  untuned {untuned}; tuned {cfgs}; candidates {sorted(cands)}", flush=True)
        assert cfgs[0] in cands and cfgs[1] == cfgs[0]
    # a fresh handle with the variable unset keeps the fixed rules
    monkeypatch.delenv("PENTA_AUTOTUNE", raising=False)
    with P.Handle() as h:
        x = torch.from_numpy(b.copy()).cuda()
        P.solve(*t, x, dims, sd, handle=h)
        torch.cuda.synchronize()
        assert h.kernel_config(sd) == untuned
