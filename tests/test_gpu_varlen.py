"""Variable-length batch on the GPU vs the CPU oracle.

Invariant: the batch equals solving every system separately.  The oracle
(po_varlen_d: Thomas per system over the packed offsets) is that per-system
solve; the GPU's own per-system solves through pentadsolver_D_gpsv_batch are
checked too.  Bars: FP64, and FP32 coefficients with an FP64 x (the FP64
solve of the coefficients widened on load, and the FP32 solve + FP64
refinement), 1e-12 max-norm relative forward error; FP32 2e-5 (vs the FP64
oracle of the float32-rounded inputs).
Lengths span 3 (the smallest the API accepts) to 8192 (the largest).
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from pentadsolver_amd import workloads as W
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def gen(offsets, seed, ratio=(1.2, 3.0)):
    rng = np.random.default_rng(seed)
    n = int(offsets[-1])
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(*ratio, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    return [ds, dl, d, du, dw], b


def relerr(a, ref):
    return float(np.max(np.abs(np.asarray(a, np.float64) - ref)) / np.max(np.abs(ref)))


def offsets_of(lens):
    off = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    return off


CASES = {
    "mixed": offsets_of([3, 4, 5, 7, 8, 16, 17, 31, 64, 100, 511, 512, 513, 1000, 2047, 4096,
                         4097, 8191, 8192, 33, 6]),
    "loguniform": W.varlen_offsets(400, 5, lo=3, hi=8192),
    "short": offsets_of(np.random.default_rng(1).integers(3, 40, 1500)),
    "one": offsets_of([8192]),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("mode", ["fp64", "fp32", "ir", "ir-refine"])
def test_varlen_matches_per_system_oracle(case, mode):
    """"ir": FP32 coefficients with an FP64 x, solved by default as one FP64
    solve of the coefficients widened on load (no refinement sweep: the
    steps counter stays 0); "ir-refine": the FP32 solve + FP64 refinement
    kernel.  Both meet the FP64 bar against the oracle on the widened
    coefficients."""
    off = CASES[case]
    print(f"varlen {case} n_sys={len(off) - 1} rows={off[-1]} {mode}", flush=True)
    coef, b = gen(off, 7)
    steps = None
    if mode == "fp64":
        tc = [torch.from_numpy(a).cuda() for a in coef]
        x = torch.from_numpy(b.copy()).cuda()
        ref = O.varlen(*coef, b, off)
        tol = 1e-12
    elif mode == "fp32":
        c32 = [a.astype(np.float32) for a in coef]
        tc = [torch.from_numpy(a).cuda() for a in c32]
        x = torch.from_numpy(b.astype(np.float32)).cuda()
        ref = O.varlen(*[a.astype(np.float64) for a in c32], b.astype(np.float32).astype(np.float64), off)
        tol = 2e-5
    else:
        c32 = [a.astype(np.float32) for a in coef]
        tc = [torch.from_numpy(a).cuda() for a in c32]
        x = torch.from_numpy(b.copy()).cuda()
        ref = O.varlen(*[a.astype(np.float64) for a in c32], b, off)
        tol = 1e-12
        steps = torch.zeros(1, dtype=torch.int32, device="cuda")
    with P.Handle() as h:
        with P.Varlen(off, handle=h) as vl:
            assert not vl.refine
            if mode == "ir-refine":
                vl.set_refine(True)
                assert vl.refine
            vl.solve(*tc, x, steps=steps)
            torch.cuda.synchronize()
            cfg = h.kernel_config(0)
    kind = {"ir": "varlen widened", "ir-refine": "varlen ir"}.get(mode, "varlen rows")
    assert cfg.startswith(kind), cfg
    err = relerr(x.cpu().numpy(), ref)
    print(f"  err={err:.2e} {cfg}" + (f" steps={int(steps.item())}" if steps is not None else ""),
          flush=True)
    assert err < tol
    if mode == "ir":
        assert int(steps.item()) == 0
    elif mode == "ir-refine":
        assert 1 <= int(steps.item()) <= 5


def test_varlen_equals_gpu_per_system_solves():
    off = W.varlen_offsets(60, 9, lo=3, hi=3000)
    coef, b = gen(off, 3)
    tc = [torch.from_numpy(a).cuda() for a in coef]
    x = torch.from_numpy(b.copy()).cuda()
    with P.Varlen(off) as vl:
        vl.solve(*tc, x)
    y = torch.from_numpy(b.copy()).cuda()
    for k in range(len(off) - 1):
        lo, hi = int(off[k]), int(off[k + 1])
        P.solve(*[t[lo:hi] for t in tc], y[lo:hi], [hi - lo], 0)
    torch.cuda.synchronize()
    assert relerr(x.cpu().numpy(), y.cpu().numpy()) < 1e-12


def test_varlen_rejects_bad_lengths():
    for lens in ([2, 10], [10, 8193]):
        with pytest.raises(P.PentadError):
            P.Varlen(offsets_of(lens))


@pytest.mark.parametrize("mode", ["fp64", "ir"])
def test_varlen_graph_replay_matches_eager(mode):
    off = W.varlen_offsets(200, 4, lo=3, hi=5000)
    coef, b = gen(off, 4)
    if mode == "ir":  # FP32 coefficients, FP64 x: the widened FP64 solve
        coef = [a.astype(np.float32) for a in coef]
    tc = [torch.from_numpy(a).cuda() for a in coef]
    x = torch.from_numpy(b.copy()).cuda()
    with P.Varlen(off) as vl:
        vl.solve(*tc, x)
        torch.cuda.synchronize()
        eager = x.cpu().numpy()
        x.copy_(torch.from_numpy(b))
        s = torch.cuda.Stream()
        with P.Handle(stream=s) as h, P.Varlen(off, handle=h) as vg:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                vg.solve(*tc, x)
            for rep in range(3):
                x.copy_(torch.from_numpy(b))
                torch.cuda.synchronize()
                g.replay()
                torch.cuda.synchronize()
                assert np.array_equal(x.cpu().numpy(), eager), rep


def run_mode(vl, coef, b, off, mode):
    """(solution as float64 numpy, oracle reference, tolerance) of one solve."""
    if mode == "fp64":
        tc = [torch.from_numpy(a).cuda() for a in coef]
        x = torch.from_numpy(b.copy()).cuda()
        ref, tol = O.varlen(*coef, b, off), 1e-12
    else:
        c32 = [a.astype(np.float32) for a in coef]
        tc = [torch.from_numpy(a).cuda() for a in c32]
        w = [a.astype(np.float64) for a in c32]
        if mode == "fp32":
            x = torch.from_numpy(b.astype(np.float32)).cuda()
            ref, tol = O.varlen(*w, b.astype(np.float32).astype(np.float64), off), 2e-5
        else:
            vl.set_refine(mode == "ir-refine")
            x = torch.from_numpy(b.copy()).cuda()
            ref, tol = O.varlen(*w, b, off), 1e-12
    vl.solve(*tc, x)
    torch.cuda.synchronize()
    return x.double().cpu().numpy(), ref, tol, tc, x


@pytest.mark.parametrize("spread", [[1] * 10, [2] * 10, [0, 1, 2, 0, 1, 2, 0, 1, 2, 0],
                                    [9] * 10])
@pytest.mark.parametrize("mode", ["fp64", "fp32", "ir", "ir-refine"])
def test_varlen_spread_same_solution(spread, mode):
    """More lanes per system than the minimum (identity-padded chunks; capped
    at 512 lanes): the same solution, to the same bars."""
    off = CASES["mixed"]
    coef, b = gen(off, 11)
    print(f"varlen spread {spread} {mode}", flush=True)
    with P.Varlen(off) as vl:
        vl.set_spread(spread)
        assert vl.spread() == spread
        got, ref, tol, _, _ = run_mode(vl, coef, b, off, mode)
    assert relerr(got, ref) < tol


@pytest.mark.parametrize("mode", ["fp64", "fp32", "ir", "ir-refine"])
def test_varlen_autotune(mode):
    """The autotuner times spreads 0..2 per length bucket on the caller's
    arrays, leaves x unchanged, keeps the fastest per bucket (or the default
    when the whole schedule is not faster with them), and the tuned schedule
    solves to the same bars."""
    off = W.varlen_offsets(600, 8, lo=3, hi=8192)
    coef, b = gen(off, 12)
    print(f"varlen autotune {mode}", flush=True)
    with P.Varlen(off) as vl:
        got, ref, tol, tc, x = run_mode(vl, coef, b, off, mode)
        x.copy_(torch.from_numpy(b.astype(np.float32) if mode == "fp32" else b.copy()).cuda())
        before = x.clone()
        rep = vl.autotune(*tc, x, reps=2)
        torch.cuda.synchronize()
        assert torch.equal(x, before)
        buckets = {bk for bk, _, _ in rep}
        assert buckets and all(0 <= bk <= 9 for bk in buckets)
        sp = vl.spread()
        # either the fastest spread of every bucket, or (when the whole
        # schedule with those was not faster) the default
        if any(sp):
            for bk in buckets:
                times = {s: ms for b2, s, ms in rep if b2 == bk}
                assert sp[bk] in times and times[sp[bk]] <= min(times.values()) + 1e-6
        print(f"  spread {sp}", flush=True)
        got2, _, _, _, _ = run_mode(vl, coef, b, off, mode)
    assert relerr(got, ref) < tol
    assert relerr(got2, ref) < tol
