"""Structured-coefficient solves (one scale per line) on the GPU vs the FP64 oracle.

The oracle is given the five diagonals assembled on the host from the same
scales and weights (ds = s*w0, dl = s*w1, d = id + s*w2, du = s*w3,
dw = s*w4, couplings past the ends of a line zeroed); the GPU kernels build
them in registers.  Bars: FP64 <= 1e-12 max-norm relative forward error,
FP32 the reference's rule (rel 1e-5 of max|x|); the ADI driver over
structured coefficients against the ADI driver over the assembled arrays
(<= 1e-12 FP64), its hipGraph replay bit-identical to eager steps.
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

HYPER = ((1.0, -4.0, 6.0, -4.0, 1.0), 1.0)
SKEW = ((0.3, -1.1, 2.9, -0.7, 0.2), 1.5)


def line_shape(dims, sd):
    """numpy shape (slowest axis first) of one value per line of axis sd."""
    shp = [int(v) for v in reversed(dims)]
    shp[len(dims) - 1 - sd] = 1
    return shp


def scales_np(dims, sd, seed, lo=0.5, hi=2.0, mult=0.5):
    rng = np.random.default_rng(seed)
    return mult * rng.uniform(lo, hi, line_shape(dims, sd))


def assemble(dims, sd, s_line, weights, idw):
    """Five diagonals (flat, standard layout) of id*I + s*w along axis sd."""
    nd = len(dims)
    shape = tuple(int(v) for v in reversed(dims))
    n = dims[sd]
    i = np.arange(n).reshape([-1 if q == nd - 1 - sd else 1 for q in range(nd)])
    s = np.broadcast_to(s_line, shape)
    w = weights
    ds = np.where(i >= 2, s * w[0], 0.0)
    dl = np.where(i >= 1, s * w[1], 0.0)
    d = idw + s * w[2]
    du = np.where(i + 1 < n, s * w[3], 0.0)
    dw = np.where(i + 2 < n, s * w[4], 0.0)
    return [np.ascontiguousarray(np.broadcast_to(a, shape)).reshape(-1) for a in (ds, dl, d, du, dw)]


def relerr(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


SHAPES = [([512, 12, 8], 0), ([384, 8, 6], 0), ([256, 16, 8], 0), ([37, 9, 7], 0),
          ([12, 512, 8], 1), ([16, 100, 8], 1), ([8, 12, 512], 2), ([7, 9, 37], 2),
          ([64, 64, 64], 0), ([64, 64, 64], 1), ([64, 64, 64], 2), ([640, 20], 0), ([20, 640], 1)]


@pytest.mark.parametrize("dims,sd", SHAPES)
@pytest.mark.parametrize("op", ["hyper", "skew"])
def test_stencil_fp64_matches_oracle(dims, sd, op):
    weights, idw = HYPER if op == "hyper" else SKEW
    s = scales_np(dims, sd, 21, mult=0.5 if op == "hyper" else 1.0)
    rng = np.random.default_rng(22)
    b = rng.standard_normal(int(np.prod(dims)))
    print(f"stencil fp64 {op} dims={dims} sd={sd}", flush=True)
    x = torch.from_numpy(b.copy()).cuda()
    P.gpsv_stencil_batch(torch.from_numpy(s.reshape(-1).copy()).cuda(), weights, idw, x, dims, sd)
    torch.cuda.synchronize()
    ref = O.solve(*assemble(dims, sd, s, weights, idw), b.copy(), dims, sd)
    err = relerr(x.cpu().numpy(), ref)
    print(f"  err={err:.2e} kernel={P.kernel_that_ran(sd if len(dims) == 3 or sd == 0 else 2)}",
          flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("dims,sd", SHAPES[:8])
def test_stencil_fp32_matches_oracle(dims, sd):
    weights, idw = HYPER
    s = scales_np(dims, sd, 23).astype(np.float32)
    rng = np.random.default_rng(24)
    b = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    print(f"stencil fp32 dims={dims} sd={sd}", flush=True)
    x = torch.from_numpy(b.copy()).cuda()
    P.gpsv_stencil_batch(torch.from_numpy(s.reshape(-1).copy()).cuda(), weights, idw, x, dims, sd)
    torch.cuda.synchronize()
    # the oracle gets the diagonals the FP32 kernels form: FP32 products of
    # the FP32 scale and weights, widened
    s32 = s.astype(np.float32)
    a32 = [a.astype(np.float32).astype(np.float64)
           for a in assemble(dims, sd, s32.astype(np.float64), weights, idw)]
    ref = O.solve(*a32, b.astype(np.float64), dims, sd)
    err = relerr(x.cpu().numpy().astype(np.float64), ref)
    print(f"  err={err:.2e}", flush=True)
    assert err < 1e-5


def test_stencil_rejects_bad_arguments():
    x = torch.zeros(64, dtype=torch.float64, device="cuda")
    s = torch.ones(1, dtype=torch.float64, device="cuda")
    with pytest.raises(P.PentadError):
        P.gpsv_stencil_batch(s, HYPER[0], 1.0, x, [64], 1)  # solvedim out of range
    with pytest.raises(TypeError):
        P.gpsv_stencil_batch(s.float(), HYPER[0], 1.0, x, [64], 0)


def adi_scales(dims, seed, dtype):
    return [torch.from_numpy(scales_np(dims, d, seed + d).reshape(-1).copy()).to(dtype).cuda()
            for d in range(len(dims))]


@pytest.mark.parametrize("dims", [[64, 32, 48], [128, 64, 32], [384, 32, 40], [96, 80]])
def test_adi_stencil_matches_assembled_adi(dims):
    weights, idw = HYPER
    sc = adi_scales(dims, 31, torch.float64)
    sets = [[torch.from_numpy(a).cuda()
             for a in assemble(dims, d, sc[d].cpu().numpy().reshape(line_shape(dims, d)),
                               weights, idw)]
            for d in range(len(dims))]
    rng = np.random.default_rng(32)
    b = torch.from_numpy(rng.standard_normal(int(np.prod(dims)))).cuda()
    print(f"adi stencil dims={dims}", flush=True)
    xs = b.clone()
    with P.AdiStencil(dims, sc, weights, idw) as adi:
        adi.step(xs, 3)
        torch.cuda.synchronize()
        cfg = [adi.config(k) for k in range(len(dims))]
    xg = b.clone()
    with P.Adi(P.ADI_FP64, dims, sets) as adi:
        adi.step(xg, 3)
        torch.cuda.synchronize()
    err = float(((xs - xg).abs().max() / xg.abs().max()).item())
    print(f"  err vs assembled ADI {err:.2e} {cfg}", flush=True)
    assert err < 1e-12
    # one step against the CPU oracle
    x1 = b.clone()
    with P.AdiStencil(dims, sc, weights, idw) as adi:
        adi.step(x1, 1)
        torch.cuda.synchronize()
    ref = b.cpu().numpy().copy()
    for d in range(len(dims)):
        ref = O.solve(*[t.cpu().numpy() for t in sets[d]], ref, dims, d)
    err1 = relerr(x1.cpu().numpy(), ref)
    print(f"  one step vs oracle {err1:.2e}", flush=True)
    assert err1 < 1e-12


def test_adi_stencil_graph_replay_bit_identical():
    dims = [128, 64, 48]
    weights, idw = HYPER
    sc = adi_scales(dims, 41, torch.float64)
    rng = np.random.default_rng(42)
    b = torch.from_numpy(rng.standard_normal(int(np.prod(dims)))).cuda()
    with P.AdiStencil(dims, sc, weights, idw) as adi:
        xe = b.clone()
        adi.step(xe, 2)
        xg = b.clone()
        adi.capture(xg, 2)
        for rep in range(3):
            xg.copy_(b)
            torch.cuda.synchronize()
            adi.replay(1)
            torch.cuda.synchronize()
            assert torch.equal(xg, xe), rep


@pytest.mark.parametrize("dims", [[128, 64, 32], [640, 32, 16]])
def test_adi_stencil_fp32(dims):
    weights, idw = HYPER
    sc64 = adi_scales(dims, 51, torch.float64)
    sc = [t.float() for t in sc64]
    rng = np.random.default_rng(52)
    b = torch.from_numpy(rng.standard_normal(int(np.prod(dims))).astype(np.float32)).cuda()
    x = b.clone()
    with P.AdiStencil(dims, sc, weights, idw) as adi:
        adi.step(x, 2)
        torch.cuda.synchronize()
    sets = [[torch.from_numpy(a.astype(np.float32)).cuda()
             for a in assemble(dims, d, sc[d].double().cpu().numpy().reshape(line_shape(dims, d)),
                               weights, idw)]
            for d in range(len(dims))]
    xg = b.clone()
    with P.Adi(P.ADI_FP32, dims, sets) as adi:
        adi.step(xg, 2)
        torch.cuda.synchronize()
    err = float(((x - xg).abs().max() / xg.abs().max()).item())
    print(f"fp32 stencil ADI vs assembled FP32 ADI {err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("dims", [[64, 32, 48], [384, 256, 12], [24, 16, 640], [96, 80]])
@pytest.mark.parametrize("wset", ["hyper", "skew"])
def test_adi_stencil_refinement_matches_fp64_operator(dims, wset):
    """Structured coefficients in FP32 + FP64 refinement (stencil mode 2):
    the residual uses the FP64 operator, so two steps match the oracle's
    FP64 solves of the FP64 operator (assembled on the host) to <= 1e-12,
    which the FP32-stored operator alone misses by about 1e-6; graph replay
    bit-identical to eager steps."""
    weights, idw = HYPER if wset == "hyper" else SKEW
    sc = adi_scales(dims, 61, torch.float64)
    rng = np.random.default_rng(62)
    b = torch.from_numpy(rng.standard_normal(int(np.prod(dims)))).cuda()
    print(f"adi stencil refinement dims={dims} {wset}", flush=True)
    x = b.clone()
    with P.AdiStencil(dims, sc, weights, idw, refine=True) as adi:
        adi.step(x, 2)
        torch.cuda.synchronize()
        steps = adi.refine_steps()
        cfg = [adi.config(k) for k in range(len(dims))]
        xg = b.clone()
        adi.capture(xg, 2)
        adi.replay(1)
        torch.cuda.synchronize()
        assert torch.equal(xg, x)
    ref = b.cpu().numpy().copy()
    for _ in range(2):
        for d in range(len(dims)):
            a = assemble(dims, d, sc[d].cpu().numpy().reshape(line_shape(dims, d)), weights, idw)
            ref = O.solve(*a, ref, dims, d)
    err = relerr(x.cpu().numpy(), ref)
    print(f"  err vs FP64 operator {err:.2e} steps {steps} {cfg}", flush=True)
    assert all(c.startswith("stencil ir thomas-pcr") for c in cfg), cfg
    assert 1 <= steps <= 5
    assert err < 1e-12


def test_adi_stencil_refinement_set_refine():
    """No correction at all leaves the FP32 solve of the FP32-rounded
    operator (error well above 1e-12); the default stopping rule reaches
    the FP64 operator."""
    dims = [128, 64, 32]
    weights, idw = HYPER
    sc = adi_scales(dims, 71, torch.float64)
    b = torch.from_numpy(np.random.default_rng(72).standard_normal(int(np.prod(dims)))).cuda()
    ref = b.cpu().numpy().copy()
    for d in range(3):
        a = assemble(dims, d, sc[d].cpu().numpy().reshape(line_shape(dims, d)), weights, idw)
        ref = O.solve(*a, ref, dims, d)
    with P.AdiStencil(dims, sc, weights, idw, refine=True) as adi:
        x = b.clone()
        adi.set_refine(0, 0.0)
        adi.step(x, 1)
        torch.cuda.synchronize()
        e0 = relerr(x.cpu().numpy(), ref)
        assert adi.refine_steps() == 0
        x = b.clone()
        adi.set_refine(5, 1e-15)
        adi.step(x, 1)
        torch.cuda.synchronize()
        e1 = relerr(x.cpu().numpy(), ref)
    print(f"no correction {e0:.2e}, refined {e1:.2e}", flush=True)
    assert 1e-10 < e0 < 1e-4
    assert e1 < 1e-12
    with pytest.raises(TypeError):
        P.AdiStencil(dims, [t.float() for t in sc], weights, idw, refine=True)
