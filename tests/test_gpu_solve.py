"""GPU parity of the batched solve against the CPU oracle (MI355X only).

Bars: FP64 within 1e-12 max-norm relative forward error of the FP64 oracle;
FP32 within 2e-5 of the FP64 oracle solving the same float32-rounded inputs
(about 100 u32 for these well-conditioned systems) and the reference fixture
rule (rel 1e-5 or abs 1e-6) on the shipped fixtures.  Every call goes through
the C ABI (libpentadsolver.so via ctypes).
"""
import os

import numpy as np
import pytest

import pentadsolver_amd as P
from conftest import FIXTURES, FIXTURE_NAMES, GENERATED_3D
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-5


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def relerr(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


def gen(dims, seed, dtype=np.float64, ratio=(2.0, 3.0)):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(*ratio, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    return [a.astype(dtype) for a in (ds, dl, d, du, dw, b)]


def gpu_solve(arrs, dims, sd, handle=None):
    t = [dev(a) for a in arrs]
    P.solve(*t[:5], t[5], dims, sd, handle=handle)
    torch.cuda.synchronize()
    return t[5].cpu().numpy()


def oracle_solve(arrs, dims, sd):
    a64 = [a.astype(np.float64) for a in arrs]
    return O.solve(*a64, dims, sd)


def check(arrs, dims, sd, handle=None):
    x = gpu_solve(arrs, dims, sd, handle)
    ref = oracle_solve(arrs, dims, sd)
    tol = FP64_TOL if arrs[0].dtype == np.float64 else FP32_TOL
    err = relerr(x, ref)
    assert np.all(np.isfinite(x))
    assert err < tol, (dims, sd, err)
    return err


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", FIXTURE_NAMES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("algo", ["auto", "naive", "thomas-pcr"])
def test_reference_fixtures_on_gpu(name, dtype, algo):
    fx = O.read_fixture(os.path.join(FIXTURES, name), dtype)
    with P.Handle() as h:
        for k in range(3):
            h.set_algo(k, algo)
        arrs = [fx[k] for k in ("ds", "dl", "d", "du", "dw", "x")]
        x = gpu_solve(arrs, fx["dims"], fx["solvedim"], h)
    assert O.allclose_reference_rule(fx["u"], x, dtype), name


@pytest.mark.parametrize("name", GENERATED_3D)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("algo", ["auto", "naive", "thomas-pcr", "shared-fact"])
def test_generated_3d_fixtures_on_gpu(name, dtype, algo, generated_fixtures):
    """The reference's missing 32^3 fixtures, rewritten by its generator's
    algorithm (tests/golden/gen_fixtures.py), every solve dimension.
    Shared-fact solves every system with system 0's coefficients (the
    reference's contract), so it is held to the oracle of that system."""
    fx = O.read_fixture(os.path.join(generated_fixtures, name), dtype)
    print(f"generated fixture {name} {np.dtype(dtype).name} {algo}", flush=True)
    arrs = [fx[k] for k in ("ds", "dl", "d", "du", "dw", "x")]
    with P.Handle() as h:
        for k in range(3):
            h.set_algo(k, algo)
        x = gpu_solve(arrs, fx["dims"], fx["solvedim"], h)
        cfg = h.kernel_config(fx["solvedim"])
    if algo == "shared-fact":
        ln = fx["dims"][fx["solvedim"]]
        sysc = []
        for a in arrs[:5]:
            v = np.moveaxis(a.reshape(list(reversed(fx["dims"]))), 2 - fx["solvedim"], -1)
            line0 = v.reshape(-1, ln)[0]
            sysc.append(np.ascontiguousarray(
                np.moveaxis(np.broadcast_to(line0, v.shape), -1, 2 - fx["solvedim"])).reshape(-1))
        ref = O.solve(*[c.astype(np.float64) for c in sysc], arrs[5].astype(np.float64),
                      fx["dims"], fx["solvedim"])
        assert relerr(x, ref) < (1e-12 if dtype == np.float64 else 2e-5), cfg
    else:
        assert O.allclose_reference_rule(fx["u"], x, dtype), (name, cfg)


SHAPES = [
    ([256, 64, 9], 0), ([256, 64, 9], 1), ([64, 256, 9], 1), ([9, 64, 256], 2),
    ([512, 40], 0), ([40, 512], 1), ([64, 64, 64], 0), ([64, 64, 64], 1),
    ([64, 64, 64], 2), ([128, 24, 10], 0), ([24, 128, 10], 1), ([10, 24, 128], 2),
    ([37, 11, 5], 0), ([11, 37, 5], 1), ([5, 11, 37], 2),  # lengths with no partition
]


@pytest.mark.parametrize("dims,sd", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_auto_dispatch_matches_oracle(dims, sd, dtype):
    print(f"case dims={dims} sd={sd} dtype={np.dtype(dtype).name}", flush=True)
    check(gen(dims, 17 + sd, dtype), dims, sd)


@pytest.mark.parametrize("dims,sd", [([256, 64, 9], 1), ([64, 64, 64], 2), ([200, 30], 0)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_naive_kernel_matches_oracle(dims, sd, dtype):
    with P.Handle() as h:
        h.set_algo(sd, "naive")
        check(gen(dims, 5, dtype), dims, sd, h)
        assert P.kernel_that_ran(0 if sd == 0 else (2 if sd == len(dims) - 1 else 1)) == "naive"


# every partitioned instantiation, x: (rows, lanes) with rows*lanes = len
X_CONFIGS = {
    np.float64: [(4, 16), (4, 64), (6, 32), (8, 8), (8, 32), (8, 64), (10, 32), (12, 16), (12, 32),
                 (12, 64), (16, 16), (16, 32)],
    np.float32: [(4, 16), (4, 64), (6, 32), (8, 8), (8, 32), (8, 64), (10, 32), (10, 64),
                 (12, 16), (16, 16), (16, 32)],
}


@pytest.mark.parametrize("n", [640, 320, 160, 384, 96])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_x_lengths_with_6_or_10_row_chunks(n, dtype):
    """Lengths 10 * 2^k and 6 * 2^k take the partitioned kernel in both
    precisions (not the LDS-staged one)."""
    dims = [n, 21, 4]
    print(f"x auto len={n} {np.dtype(dtype).name}", flush=True)
    with P.Handle() as h:
        check(gen(dims, n + 1, dtype), dims, 0, h)
        cfg = h.kernel_config(0)
        assert cfg.startswith("thomas-pcr"), cfg


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_partition_x_all_configs(dtype):
    with P.Handle() as h:
        for rows, lanes in X_CONFIGS[dtype]:
            n = rows * lanes
            dims = [n, 37, 3]
            print(f"x rows={rows} lanes={lanes} len={n} {np.dtype(dtype).name}", flush=True)
            h.set_kernel_params(0, rows, lanes)
            check(gen(dims, rows * 100 + lanes, dtype), dims, 0, h)
            assert h.kernel_config(0) == f"thomas-pcr rows={rows} lanes={lanes}"


YZ_CONFIGS = [(4, 16, 256), (4, 64, 64), (6, 16, 192), (8, 16, 256), (8, 32, 128),
              (8, 64, 64), (12, 16, 384), (16, 16, 512), (16, 32, 256), (12, 64, 96)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("sd", [1, 2])
def test_partition_yz_all_configs(dtype, sd):
    with P.Handle() as h:
        for rows, spw, n in YZ_CONFIGS:
            P_ = n // rows
            from_ok = (spw * P_) <= (1024 if (dtype == np.float32 and rows <= 12) or rows <= 4 else 512)
            if not from_ok:
                continue
            dims = [40, n, 3] if sd == 1 else [40, 3, n]
            print(f"yz sd={sd} rows={rows} spw={spw} len={n} {np.dtype(dtype).name}", flush=True)
            h.set_kernel_params(sd, rows, spw)
            check(gen(dims, rows * 7 + spw, dtype), dims, sd, h)
            assert h.kernel_config(sd) == f"thomas-pcr rows={rows} sys_per_wg={spw}"


def test_weakly_dominant_fp64():
    dims = [256, 32, 32]
    arrs = gen(dims, 4, np.float64, ratio=(1.05, 1.05))
    for sd in range(3):
        check(arrs, dims, sd)


def test_outofrange_couplings_ignored():
    dims = [256, 16, 16]
    arrs = gen(dims, 8)
    idx = np.arange(int(np.prod(dims))).reshape(16, 16, 256)
    for sd in range(3):
        lines = np.moveaxis(idx, 2 - sd, -1).reshape(-1, dims[sd])
        a2 = [a.copy() for a in arrs]
        a2[0][lines[:, :2]] = np.nan
        a2[1][lines[:, 0]] = np.inf
        a2[3][lines[:, -1]] = np.nan
        a2[4][lines[:, -2:]] = -np.inf
        x = gpu_solve(a2, dims, sd)
        ref = oracle_solve(arrs, dims, sd)
        assert relerr(x, ref) < FP64_TOL, sd


def test_adi_sweep_in_place_256():
    """x, then y, then z in place on one field with the same coefficient arrays
    (reference apps/adi_cuda.cu:167-176), 2 iterations, vs the oracle."""
    dims = [256, 64, 48]
    arrs = gen(dims, 2)
    t = [dev(a) for a in arrs]
    ref = arrs[5].copy()
    for _ in range(2):
        for sd in range(3):
            P.solve(*t[:5], t[5], dims, sd)
            ref = O.solve(*arrs[:5], ref, dims, sd)
    torch.cuda.synchronize()
    assert relerr(t[5].cpu().numpy(), ref) < FP64_TOL


def test_buffer_contract_and_status():
    dims = [37, 8, 8]
    arrs = gen(dims, 1)
    t = [dev(a) for a in arrs]
    with P.Handle() as h:
        assert P.gpsv_batch_buffer_size_ext(*t, dims, 0, h) == 0  # LDS-staged kernel
        h.set_algo(0, "naive")
        need = P.gpsv_batch_buffer_size_ext(*t, dims, 0, h)
        assert need == 3 * 37 * 64 * 8
        with pytest.raises(P.PentadError):
            P.gpsv_batch(*t, dims, 0, buffer=None, handle=h)
        with pytest.raises(P.PentadError):
            P.gpsv_batch(*t, [37, 8, 8], 3, handle=h)


@pytest.mark.parametrize("n", [3, 5, 7, 37, 100, 300, 511, 700, 1000])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_transpose_x_any_length(n, dtype):
    """x lengths with no whole-chunk partition: the LDS-staged kernel
    (rows past the end are identity rows) is the automatic choice."""
    dims = [n, 13, 7]
    print(f"transpose x len={n} {np.dtype(dtype).name}", flush=True)
    with P.Handle() as h:
        check(gen(dims, n, dtype), dims, 0, h)
        assert h.kernel_config(0).startswith("transpose"), h.kernel_config(0)
        assert P.kernel_that_ran(0) == "transpose"


@pytest.mark.parametrize("n", [64, 512, 1024])
def test_transpose_x_forced(n):
    dims = [n, 20, 3]
    with P.Handle() as h:
        h.set_algo(0, "transpose")
        check(gen(dims, 3 * n, np.float64), dims, 0, h)
        assert h.kernel_config(0).startswith("transpose")


@pytest.mark.parametrize("n", [5, 37, 100, 200, 300])
@pytest.mark.parametrize("sd", [1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_partition_yz_padded_lengths(n, sd, dtype):
    """y/z lengths that do not split into whole chunks: the partitioned
    kernel pads the last chunk with identity rows."""
    dims = [40, n, 3] if sd == 1 else [40, 3, n]
    print(f"yz padded len={n} sd={sd} {np.dtype(dtype).name}", flush=True)
    with P.Handle() as h:
        check(gen(dims, n + sd, dtype), dims, sd, h)
        assert h.kernel_config(sd).startswith("thomas-pcr"), h.kernel_config(sd)


@pytest.mark.parametrize("n", [1024, 1000, 2048, 8192])
@pytest.mark.parametrize("sd", [1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_partition_yz_long_lines(n, sd, dtype):
    """Lines too long for 16 systems per workgroup run the partitioned
    kernel with fewer systems per workgroup, not the one-lane fallback."""
    dims = [20, n, 2] if sd == 1 else [20, 2, n]
    print(f"yz long len={n} sd={sd} {np.dtype(dtype).name}", flush=True)
    with P.Handle() as h:
        check(gen(dims, n + 3 * sd, dtype), dims, sd, h)
        cfg = h.kernel_config(sd)
        print(f"  {cfg}", flush=True)
        assert cfg.startswith("thomas-pcr"), cfg


def test_streams_and_handles():
    dims = [256, 32, 16]
    arrs = gen(dims, 3)
    ref = oracle_solve(arrs, dims, 1)
    s = torch.cuda.Stream()
    with P.Handle(stream=s) as h:
        t = [dev(a) for a in arrs]
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            P.solve(*t[:5], t[5], dims, 1, handle=h)
        s.synchronize()
        assert relerr(t[5].cpu().numpy(), ref) < FP64_TOL


@pytest.mark.parametrize("n", [1025, 1100, 3000, 4096, 8192])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_long_x_lines_on_chip(n, dtype):
    """Contiguous lines longer than 1024 rows: the LDS-staged kernel with more
    than 64 lanes per system (separators exchanged through LDS across the
    workgroup), not the one-lane kernel with its scratch passes; the same
    lines through the strided partitioned kernel at stride 1 (what a padded
    layout whose pitch breaks 16-byte alignment runs) via "thomas-pcr".
    Lengths of 128 or 256 whole chunks (4096 FP32 rows) take the wide
    partitioned kernel both ways."""
    dims = [n, 9, 3]
    print(f"long x len={n} {np.dtype(dtype).name}", flush=True)
    r = 8 if dtype == np.float64 else 16
    wide = n % r == 0 and n // r in (128, 256)
    with P.Handle() as h:
        check(gen(dims, n, dtype), dims, 0, h)
        cfg = h.kernel_config(0)
        print(f"  {cfg}", flush=True)
        if wide:  # whole chunks: the wide partitioned kernel instead
            assert cfg == f"thomas-pcr rows={r} lanes={n // r}", cfg
        else:
            assert P.kernel_that_ran(0) == "transpose"
            assert cfg.startswith("transpose rows=") and int(cfg.split("lanes=")[1]) > 64, cfg
        h.set_algo(0, "thomas-pcr")
        check(gen(dims, n + 1, dtype), dims, 0, h)
        cfg = h.kernel_config(0)
        print(f"  {cfg}", flush=True)
        if wide:
            assert cfg == f"thomas-pcr rows={r} lanes={n // r}", cfg
        else:
            assert cfg.startswith("thomas-pcr rows=") and "sys_per_wg" in cfg, cfg


@pytest.mark.parametrize("dims,dtype,forced", [
    ([1024, 9, 3], np.float64, None),      # 27 systems, one pass per workgroup
    ([1024, 467, 3], np.float64, None),    # 1401 systems, two pipelined passes
    ([2048, 1031], np.float64, None),      # one system per pass, ragged grid
    ([2048, 5, 3], np.float32, None),
    ([2048, 1027], np.float32, None),      # two pipelined passes, ragged
    ([4096, 3, 7], np.float32, None),
    ([1024, 600], np.float32, (8, 128)),   # FP32 8-row chunks by parameters
    ([2048, 40], np.float32, (8, 256)),
    ([1024, 1000], np.float64, (8, 128)),
])
def test_wide_x_lines(dims, dtype, forced):
    """x lines of 128 or 256 chunks: a system's lanes span wavefronts and
    exchange the reduced system through LDS.  Against the oracle at the
    FP64 / FP32 bars; batches that leave the last workgroup partly empty
    (clamped systems must not write), one- and two-pass launches."""
    n = dims[0]
    print(f"wide x dims={dims} {np.dtype(dtype).name} forced={forced}", flush=True)
    with P.Handle() as h:
        if forced:
            h.set_kernel_params(0, *forced)
        err = check(gen(dims, n + len(dims), dtype), dims, 0, h)
        cfg = h.kernel_config(0)
        print(f"  {cfg} err={err:.2e}", flush=True)
        r, lanes = forced or ((8 if dtype == np.float64 else 16), n // (8 if dtype == np.float64 else 16))
        assert cfg == f"thomas-pcr rows={r} lanes={lanes}", cfg


def test_wide_x_matches_transpose_bitwise():
    """FP64: the wide kernel and the LDS-staged kernel run the same
    partition (8-row chunks, 128 lanes), so their results agree bit for
    bit."""
    dims = [1024, 33, 5]
    arrs = gen(dims, 77)
    with P.Handle() as h:
        a = gpu_solve(arrs, dims, 0, h)
        assert h.kernel_config(0) == "thomas-pcr rows=8 lanes=128"
        h.set_algo(0, "transpose")
        b = gpu_solve(arrs, dims, 0, h)
        assert h.kernel_config(0) == "transpose rows=8 lanes=128"
    assert np.array_equal(a, b)


def test_long_x_lines_full_batch():
    """8192-row x lines, 512 of them, FP64, against the oracle; and one
    8192-row system alone."""
    for dims in ([8192, 64, 8], [8192]):
        print(f"long x dims={dims}", flush=True)
        with P.Handle() as h:
            check(gen(dims, 3), dims, 0, h)
            assert h.kernel_config(0) == "transpose rows=16 lanes=512", h.kernel_config(0)


@pytest.mark.parametrize("dims", [[256, 20, 6], [512, 8, 4], [64, 33, 17], [1024, 5, 3]])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_arrays_off_16_byte_boundaries(dims, dtype):
    """A caller's arrays that start one element past a 16-byte boundary
    (views into a larger allocation): x lines leave the 16-byte-vector
    kernels for ones that load element by element, and every direction
    matches the oracle with the contents around the views untouched."""
    arrs = gen(dims, 51, dtype)
    n = int(np.prod(dims))
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    for sd in range(3):
        bufs = []
        for a in arrs:
            buf = torch.full((n + 2,), 7.0, dtype=tdt, device="cuda")
            buf[1:n + 1].copy_(torch.from_numpy(a))
            bufs.append(buf)
        views = [b[1:n + 1] for b in bufs]
        assert all(v.data_ptr() % 16 != 0 for v in views)
        print(f"unaligned dims={dims} sd={sd} {np.dtype(dtype).name}", flush=True)
        with P.Handle() as h:
            P.solve(*views[:5], views[5], dims, sd, handle=h)
            torch.cuda.synchronize()
            cfg = h.kernel_config(sd)
        x = views[5].cpu().numpy()
        err = relerr(x, oracle_solve(arrs, dims, sd))
        print(f"  err={err:.2e} cfg={cfg}", flush=True)
        assert err < (FP64_TOL if dtype == np.float64 else FP32_TOL), (sd, cfg, err)
        for b in bufs:
            assert b[0].item() == 7.0 and b[n + 1].item() == 7.0


@pytest.mark.parametrize("dims", [[256, 20, 6], [512, 8, 4], [1024, 5, 3]])
def test_mixed_precision_arrays_off_16_byte_boundaries(dims):
    """The mixed-precision entry point on FP32 coefficient views and an FP64
    right-hand side view one element past 16-byte boundaries: every direction
    within 1e-12 of the FP64 oracle of the widened coefficients."""
    arrs = gen(dims, 52)
    c32 = [a.astype(np.float32) for a in arrs[:5]]
    n = int(np.prod(dims))
    for sd in range(3):
        cv = []
        for a in c32:
            buf = torch.zeros(n + 1, dtype=torch.float32, device="cuda")
            buf[1:].copy_(torch.from_numpy(a))
            cv.append(buf[1:])
        xb = torch.zeros(n + 1, dtype=torch.float64, device="cuda")
        xb[1:].copy_(torch.from_numpy(arrs[5]))
        xv = xb[1:]
        print(f"ir unaligned dims={dims} sd={sd}", flush=True)
        with P.Handle() as h:
            P.solve_ir(*cv, xv, dims, sd, handle=h)
            torch.cuda.synchronize()
            cfg = h.kernel_config(sd)
        ref = O.solve(*[a.astype(np.float64) for a in c32], arrs[5], dims, sd)
        err = relerr(xv.cpu().numpy(), ref)
        print(f"  err={err:.2e} cfg={cfg}", flush=True)
        assert err < FP64_TOL, (sd, cfg, err)
