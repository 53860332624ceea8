"""ADI driver (x, then y, then z in place, rotating layouts, hipGraph) on the
GPU against the oracle's in-place direction solves in the standard layout."""
import numpy as np
import pytest

import pentadsolver_amd as P
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def gen(dims, seed, ratio=(2.0, 3.0)):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(*ratio, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    return [ds, dl, d, du, dw], b


def oracle_steps(coef_sets, b, dims, nsteps):
    x = b.copy()
    for _ in range(nsteps):
        for sd in range(len(dims)):
            cs = coef_sets[0] if len(coef_sets) == 1 else coef_sets[sd]
            x = O.solve(*cs, x, dims, sd)
    return x


def relerr(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


CASES = [[256, 64, 48], [64, 64, 64], [48, 96, 128], [384, 16, 40], [37, 20, 11], [128, 256], [96, 64, 10],
         [512, 32, 16], [32, 512, 24]]


@pytest.mark.parametrize("dims", CASES)
@pytest.mark.parametrize("mode", ["fp64", "fp32", "ir"])
def test_adi_step_matches_oracle(dims, mode):
    coef, b = gen(dims, 5)
    print(f"adi dims={dims} mode={mode}", flush=True)
    if mode == "fp64":
        tc = [torch.from_numpy(a).cuda() for a in coef]
        x = torch.from_numpy(b.copy()).cuda()
        ref = oracle_steps([coef], b, dims, 2)
        tol = 1e-12
        m = P.ADI_FP64
    elif mode == "fp32":
        c32 = [a.astype(np.float32) for a in coef]
        tc = [torch.from_numpy(a).cuda() for a in c32]
        x = torch.from_numpy(b.astype(np.float32)).cuda()
        ref = oracle_steps([[a.astype(np.float64) for a in c32]], b.astype(np.float32).astype(np.float64), dims, 2)
        tol = 5e-5
        m = P.ADI_FP32
    else:
        c32 = [a.astype(np.float32) for a in coef]
        tc = [torch.from_numpy(a).cuda() for a in c32]
        x = torch.from_numpy(b.copy()).cuda()
        ref = oracle_steps([[a.astype(np.float64) for a in c32]], b, dims, 2)
        tol = 1e-12
        m = P.ADI_IR
    with P.Adi(m, dims, [tc]) as adi:
        adi.step(x, 2)
        torch.cuda.synchronize()
        got = x.double().cpu().numpy()
        cfg = [adi.config(k) for k in range(len(dims))]
        steps = adi.refine_steps() if mode == "ir" else None
    err = relerr(got, ref)
    print(f"  err={err:.2e} cfg={cfg} steps={steps}", flush=True)
    assert err < tol


@pytest.mark.parametrize("dims,rows", [([640, 32, 16], 10), ([32, 320, 16], 10),
                                       ([768, 16, 32], 12), ([16, 32, 768], 12)])
def test_adi_fp64_partial_lookahead_chunks(dims, rows):
    """FP64 ADI directions of 10- and 12-row chunks with 32 or 64 lanes per
    system: the pipelined kernel that loads two arrays of the next pass ahead
    and the other four after the solve."""
    coef, b = gen(dims, 21)
    print(f"adi fp64 dims={dims} rows={rows}", flush=True)
    tc = [torch.from_numpy(a).cuda() for a in coef]
    x = torch.from_numpy(b.copy()).cuda()
    with P.Adi(P.ADI_FP64, dims, [tc]) as adi:
        adi.step(x, 2)
        torch.cuda.synchronize()
        got = x.cpu().numpy()
        cfg = [adi.config(k) for k in range(len(dims))]
    err = relerr(got, oracle_steps([coef], b, dims, 2))
    print(f"  err={err:.2e} cfg={cfg}", flush=True)
    assert any(c.startswith(f"thomas-pcr rows={rows} ") for c in cfg), cfg  # rotating driver
    assert err < 1e-12


def test_adi_per_direction_coefficients_and_graph_replay():
    dims = [128, 64, 32]
    sets, b = [], None
    for k in range(3):
        coef, bb = gen(dims, 10 + k)
        sets.append(coef)
        b = bb if b is None else b
    tsets = [[torch.from_numpy(a).cuda() for a in cs] for cs in sets]
    x_eager = torch.from_numpy(b.copy()).cuda()
    x_graph = torch.from_numpy(b.copy()).cuda()
    with P.Adi(P.ADI_FP64, dims, tsets) as adi:
        adi.step(x_eager, 3)
        adi.capture(x_graph, 1)
        for _ in range(3):  # replays with eager work and syncs in between
            adi.replay(1)
            torch.cuda.synchronize()
            _ = (x_eager * 1.0).sum().item()
        torch.cuda.synchronize()
    ref = oracle_steps(sets, b, dims, 3)
    assert relerr(x_eager.cpu().numpy(), ref) < 1e-12
    # one-step graph replayed 3 times vs 3 eager steps: the eager run keeps an
    # internal layout between its steps, the same arithmetic in another order
    assert relerr(x_graph.cpu().numpy(), ref) < 1e-12


def test_adi_multistep_graph_bit_identical_to_eager():
    dims = [128, 64, 32]
    coef, b = gen(dims, 12)
    tc = [torch.from_numpy(a).cuda() for a in coef]
    xe = torch.from_numpy(b.copy()).cuda()
    xg = torch.from_numpy(b.copy()).cuda()
    with P.Adi(P.ADI_FP64, dims, [tc]) as adi:
        adi.step(xe, 4)
        adi.capture(xg, 4)
        adi.replay(1)
        torch.cuda.synchronize()
    assert torch.equal(xe, xg)
    ref = oracle_steps([coef], b, dims, 4)
    assert relerr(xe.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("dims,ratio", [([512, 64, 32], (2.0, 3.0)), ([256, 32, 40], (2.0, 3.0)),
                                        ([512, 16, 24], (1.05, 1.05)), ([384, 256, 8], (2.0, 3.0)),
                                        ([640, 12, 20], (2.0, 3.0)), ([20, 640, 24], (1.05, 1.05)),
                                        ([320, 24, 12], (2.0, 3.0))])
def test_adi_ir_interleaved_coefficients(dims, ratio):
    """Refinement directions read their coefficients interleaved (written
    when the driver permutes them): per 4 rows for 4- and 8-row chunks, per
    2 rows for 6-row chunks and for the 10-row chunks of 640-row lines; the
    others (10-row chunks of other lengths) keep five arrays.  Two steps
    against the oracle (<= 1e-12, well-conditioned and weak sets), and a
    captured graph replays to the eager result bit for bit."""
    coef, b = gen(dims, 31, ratio)
    print(f"adi ir interleaved dims={dims} ratio={ratio}", flush=True)
    c32 = [a.astype(np.float32) for a in coef]
    tc = [torch.from_numpy(a).cuda() for a in c32]
    ref = oracle_steps([[a.astype(np.float64) for a in c32]], b, dims, 2)
    xe = torch.from_numpy(b.copy()).cuda()
    xg = torch.from_numpy(b.copy()).cuda()
    with P.Adi(P.ADI_IR, dims, [tc]) as adi:
        adi.step(xe, 2)
        adi.capture(xg, 2)
        adi.replay(1)
        torch.cuda.synchronize()
        cfg = [adi.config(k) for k in range(len(dims))]
    err = relerr(xe.cpu().numpy(), ref)
    print(f"  err={err:.2e} cfg={cfg}", flush=True)
    for c, n in zip(cfg, dims):
        rows = int(c.split("rows=")[1].split()[0])
        assert ("interleaved" in c) == (rows in (4, 6, 8) or (rows == 10 and n == 640)), c
    assert any("interleaved" in c for c in cfg), cfg
    assert err < 1e-12
    assert np.array_equal(xg.cpu().numpy(), xe.cpu().numpy())


@pytest.mark.parametrize("dims", [[768, 24, 16], [16, 1024, 12], [24, 16, 768]])
def test_adi_ir_long_lines_widened(dims):
    """Refinement-mode directions whose length no refinement kernel covers
    (768 = 12 x 64, 1024 = 16 x 64 rows) keep the rotating layouts with the
    FP64 solve of the widened coefficients ("ir widened"): two steps against
    the oracle of the widened coefficients (<= 1e-12), graph replay equal to
    the eager steps bit for bit."""
    coef, b = gen(dims, 41)
    print(f"adi ir long dims={dims}", flush=True)
    c32 = [a.astype(np.float32) for a in coef]
    tc = [torch.from_numpy(a).cuda() for a in c32]
    ref = oracle_steps([[a.astype(np.float64) for a in c32]], b, dims, 2)
    xe = torch.from_numpy(b.copy()).cuda()
    xg = torch.from_numpy(b.copy()).cuda()
    with P.Adi(P.ADI_IR, dims, [tc]) as adi:
        adi.step(xe, 2)
        adi.capture(xg, 2)
        adi.replay(1)
        torch.cuda.synchronize()
        cfg = [adi.config(k) for k in range(3)]
    err = relerr(xe.cpu().numpy(), ref)
    print(f"  err={err:.2e} cfg={cfg}", flush=True)
    for k, n in enumerate(dims):
        if n >= 768:
            assert cfg[k].startswith("ir widened rows="), cfg
    assert err < 1e-12
    assert np.array_equal(xg.cpu().numpy(), xe.cpu().numpy())


@pytest.mark.parametrize("dims,mode", [([1024, 20], "fp64"), ([20, 1024], "fp64"),
                                       ([1024, 24, 16], "fp64"), ([24, 16, 1024], "fp64"),
                                       ([2048, 12], "fp64"), ([2048, 20], "fp32"),
                                       ([24, 16, 2048], "fp32"), ([1024, 24, 16], "ir"),
                                       ([2048, 12], "ir")])
def test_adi_wide_lines(dims, mode):
    """Directions of 128 or 256 chunks (1024 / 2048 rows) on the wide x
    kernel with the rotated store: several systems per pass (8-pass tiles,
    the last workgroup partly empty in the 2-D cases) or one; two steps
    against the oracle, graph replay equal to the eager steps bit for bit."""
    coef, b = gen(dims, 43)
    print(f"adi wide dims={dims} mode={mode}", flush=True)
    if mode == "fp64":
        tc = [torch.from_numpy(a).cuda() for a in coef]
        x0 = b.copy()
        ref = oracle_steps([coef], b, dims, 2)
        tol, m = 1e-12, P.ADI_FP64
    else:
        c32 = [a.astype(np.float32) for a in coef]
        tc = [torch.from_numpy(a).cuda() for a in c32]
        x0 = b.astype(np.float32) if mode == "fp32" else b.copy()
        ref = oracle_steps([[a.astype(np.float64) for a in c32]], x0.astype(np.float64), dims, 2)
        tol, m = (5e-5, P.ADI_FP32) if mode == "fp32" else (1e-12, P.ADI_IR)
    xe = torch.from_numpy(x0.copy()).cuda()
    xg = torch.from_numpy(x0.copy()).cuda()
    with P.Adi(m, dims, [tc]) as adi:
        adi.step(xe, 2)
        adi.capture(xg, 2)
        adi.replay(1)
        torch.cuda.synchronize()
        cfg = [adi.config(k) for k in range(len(dims))]
    err = relerr(xe.double().cpu().numpy(), ref)
    print(f"  err={err:.2e} cfg={cfg}", flush=True)
    for k, n in enumerate(dims):
        if n >= 1024:
            r = 16 if mode == "fp32" else 8
            assert f"rows={r} lanes={n // r}" in cfg[k], cfg
    assert err < tol
    assert np.array_equal(xg.cpu().numpy(), xe.cpu().numpy())


def test_adi_ir_graph_replay_matches_eager():
    dims = [256, 32, 64]
    coef, b = gen(dims, 7)
    tc = [torch.from_numpy(a.astype(np.float32)).cuda() for a in coef]
    xe = torch.from_numpy(b.copy()).cuda()
    xg = torch.from_numpy(b.copy()).cuda()
    with P.Adi(P.ADI_IR, dims, [tc]) as adi:
        adi.step(xe, 4)
        adi.capture(xg, 2)
        adi.replay(2)
        torch.cuda.synchronize()
        assert adi.refine_steps() >= 1
    ref = oracle_steps([[a.astype(np.float32).astype(np.float64) for a in coef]], b, dims, 4)
    assert relerr(xe.cpu().numpy(), ref) < 1e-12
    assert relerr(xg.cpu().numpy(), ref) < 1e-12


def replicate(prof, dims, sd):
    """Full arrays whose every line along sd is prof (dims[0] fastest)."""
    shape = tuple(reversed(dims))
    view = [1] * len(dims)
    view[len(dims) - 1 - sd] = dims[sd]
    return [np.ascontiguousarray(np.broadcast_to(p.reshape(view), shape)).reshape(-1)
            for p in prof]


@pytest.mark.parametrize("dims", [[64, 32, 48], [256, 64, 16], [128, 256], [48, 16, 96]])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_adi_shared_fact_constant_coefficients(dims, dtype):
    """Shared-factorisation ADI (one coefficient profile per direction,
    replicated over its lines) against the oracle's general solves of the
    replicated arrays; graph replay bit-identical to eager steps."""
    print(f"adi shared dims={dims} {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(len(dims) + dims[0])
    sets = []
    for sd in range(len(dims)):
        n = dims[sd]
        ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
        d = rng.uniform(2, 3, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
        sets.append([a.astype(dtype) for a in replicate([ds, dl, d, du, dw], dims, sd)])
    b = rng.uniform(-1, 1, int(np.prod(dims))).astype(dtype)
    tsets = [[torch.from_numpy(a).cuda() for a in cs] for cs in sets]
    mode = P.ADI_SHARED_FP64 if dtype == np.float64 else P.ADI_SHARED_FP32
    x = torch.from_numpy(b.copy()).cuda()
    with P.Adi(mode, dims, tsets) as adi:
        assert all(adi.config(k).startswith("shared-fact") for k in range(len(dims)))
        adi.step(x, 2)
        torch.cuda.synchronize()
        eager = x.cpu().numpy()
        x.copy_(torch.from_numpy(b))
        adi.capture(x, 2)
        adi.replay(1)
        torch.cuda.synchronize()
        graph = x.cpu().numpy()
    assert np.array_equal(eager, graph)
    ref = oracle_steps([[a.astype(np.float64) for a in cs] for cs in sets],
                       b.astype(np.float64), dims, 2)
    err = relerr(eager.astype(np.float64), ref)
    print(f"  err={err:.2e}", flush=True)
    assert err < (1e-12 if dtype == np.float64 else 5e-5)


# The line lengths of the 384x256x640 hyperdiffusion grid (384 = 6 rows x 64
# lanes, 256 = 8 x 32, 640 = 10 x 64) on grids the oracle finishes in
# seconds: each case runs the driver's kernels for two of the three lengths.
C5_LINE_CASES = [[384, 256, 12], [640, 12, 384], [16, 640, 256]]


@pytest.mark.parametrize("dims", C5_LINE_CASES)
@pytest.mark.parametrize("mode", ["fp64", "ir"])
@pytest.mark.parametrize("cond", ["well", "weak"])
def test_adi_c5_line_lengths(dims, mode, cond):
    """FP64 and FP32 + FP64 refinement ADI steps at the 384-, 256- and
    640-row line lengths (the 64-lane kernels, among them the 10-row-chunk
    refinement kernel), for the well-conditioned (ratio 2-3) and the weakly
    dominant (c = 1.05 sum|off|) coefficient sets, against the oracle:
    max-norm relative error <= 1e-12."""
    ratio = (2.0, 3.0) if cond == "well" else (1.05, 1.05)
    coef, b = gen(dims, 31 if cond == "well" else 32, ratio)
    print(f"adi c5 lines dims={dims} mode={mode} {cond}", flush=True)
    if mode == "fp64":
        tc = [torch.from_numpy(a).cuda() for a in coef]
        ref_coef = coef
        m = P.ADI_FP64
    else:
        c32 = [a.astype(np.float32) for a in coef]
        tc = [torch.from_numpy(a).cuda() for a in c32]
        ref_coef = [a.astype(np.float64) for a in c32]
        m = P.ADI_IR
    x = torch.from_numpy(b.copy()).cuda()
    with P.Adi(m, dims, [tc]) as adi:
        adi.step(x, 1)
        torch.cuda.synchronize()
        got = x.cpu().numpy()
        cfg = [adi.config(k) for k in range(3)]
        steps = adi.refine_steps() if mode == "ir" else None
    err = relerr(got, oracle_steps([ref_coef], b, dims, 1))
    print(f"  err={err:.2e} cfg={cfg} steps={steps}", flush=True)
    want = {384: "rows=6 lanes=64", 256: "rows=8 lanes=32", 640: "rows=10 lanes=64"}
    for k, n in enumerate(dims):
        if n in want:
            assert cfg[k].endswith(want[n]), (n, cfg[k])
            assert cfg[k].startswith("ir " if mode == "ir" else "thomas-pcr"), cfg[k]
    if mode == "ir":
        assert 1 <= steps <= 5, steps
    assert err < 1e-12


@pytest.mark.parametrize("dims", [[640, 12], [20, 640], [640, 24]])
def test_adi_ir_640_rows_partial_tile(dims):
    """Refinement ADI steps whose 640-row direction has a system count that
    is not a multiple of the 16-system tile (the unpadded-tile kernel's last
    workgroup holds clamped systems it must not write), against the oracle
    of the widened FP32 coefficients: max-norm relative error <= 1e-12.
    (The rotating driver needs every length to split into a power of two of
    4- to 10-row chunks, so 3-D grids always give multiples of 16 systems;
    2-D grids do not.)"""
    coef, b = gen(dims, 41)
    print(f"adi ir 640 partial dims={dims}", flush=True)
    c32 = [a.astype(np.float32) for a in coef]
    tc = [torch.from_numpy(a).cuda() for a in c32]
    x = torch.from_numpy(b.copy()).cuda()
    with P.Adi(P.ADI_IR, dims, [tc]) as adi:
        adi.step(x, 2)
        torch.cuda.synchronize()
        got = x.cpu().numpy()
        cfg = [adi.config(k) for k in range(len(dims))]
    err = relerr(got, oracle_steps([[a.astype(np.float64) for a in c32]], b, dims, 2))
    print(f"  err={err:.2e} cfg={cfg}", flush=True)
    assert cfg[dims.index(640)] == "ir interleaved thomas-pcr rows=10 lanes=64", cfg
    assert err < 1e-12
