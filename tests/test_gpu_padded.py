"""Padded-layout solves (pentadsolver_{D,S}_gpsv_batch_padded) on the GPU.

Element (i, j, k) of a field with extents dims sits at
i + pads[0] * (j + pads[1] * k) (the layout of the reference's CPU library,
src/cpu/singlenode/pentad_cpu.cpp:611-613).  Checks:
  * against the FP64 oracle on the same values packed densely: FP64
    <= 1e-12 max-norm relative forward error, FP32 the reference's rule
    (rel 1e-5 of max|x|);
  * the padding is neither read (the coefficient padding holds NaN, which
    would reach the solution) nor written (x's padding keeps its sentinel,
    bit for bit);
  * a padded solve equals the dense solve of the same values bit for bit when
    both run the same kernel (only the addressing differs);
  * where the reference's own CPU library (oracle/_ref) handles the x length,
    it agrees as well.
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SENTINEL = 7.25


def gen(dims, seed, dtype):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(2, 3, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    return [a.astype(dtype) for a in (ds, dl, d, du, dw, b)]


def run_padded(arrs, dims, pads, sd, handle):
    coef = [torch.from_numpy(O.embed_padded(a, dims, pads, np.nan)).cuda() for a in arrs[:5]]
    x = torch.from_numpy(O.embed_padded(arrs[5], dims, pads, SENTINEL)).cuda()
    nbytes = P.gpsv_batch_padded_buffer_size(*coef, x, dims, pads, sd, handle)
    buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda") if nbytes else None
    P.gpsv_batch_padded(*coef, x, dims, pads, sd, buf, handle)
    torch.cuda.synchronize()
    return x.cpu().numpy(), nbytes


def relerr(a, ref):
    return float(np.max(np.abs(a.astype(np.float64) - ref)) / np.max(np.abs(ref)))


# (dims, pads, dtype, expected kernel along x): "lanes" = the wavefront
# kernels with 16-byte vector loads, "sys_per_wg" = the workgroup-wide
# partitioned kernel with scalar loads (any pitch, any length, no scratch)
CASES = [
    ([64, 48, 40], [72, 50, 40], np.float64, "lanes"),       # pitches multiples of 16 B
    ([64, 48, 40], [67, 48, 40], np.float64, "sys_per_wg"),  # 536-byte rows
    ([128, 30, 20], [132, 33, 20], np.float32, "lanes"),
    ([128, 30, 20], [130, 30, 20], np.float32, "sys_per_wg"),
    ([256, 64], [260, 64], np.float64, "lanes"),
    ([100, 37], [104, 37], np.float64, "sys_per_wg"),        # no whole partition of 100
    ([30, 7, 3], [32, 7, 3], np.float64, "sys_per_wg"),      # the reference's own padding
    ([512, 12, 5], [512, 16, 5], np.float32, "lanes"),       # padding of the middle axis only
]


@pytest.mark.parametrize("dims,pads,dtype,xkind", CASES)
def test_padded_matches_oracle_and_keeps_padding(dims, pads, dtype, xkind):
    mask = O.padding_mask(dims, pads)
    tol = 1e-12 if dtype == np.float64 else 1e-5
    with P.Handle() as h:
        for sd in range(len(dims)):
            print(f"padded dims={dims} pads={pads} {np.dtype(dtype).name} sd={sd}", flush=True)
            arrs = gen(dims, 40 + 3 * sd + len(dims), dtype)
            xp, nbytes = run_padded(arrs, dims, pads, sd, h)
            cfg = h.kernel_config(sd if len(dims) == 3 or sd == 0 else 2)
            got = O.extract_padded(xp, dims, pads)
            ref = O.solve(*[a.astype(np.float64) for a in arrs], dims, sd)
            err = relerr(got, ref)
            print(f"  err={err:.2e} kernel={cfg} scratch={nbytes}", flush=True)
            assert err < tol
            assert np.all(xp[mask] == dtype(SENTINEL))
            if sd == 0:
                assert cfg.startswith("thomas-pcr") and xkind in cfg, cfg
                assert nbytes == 0


@pytest.mark.parametrize("dims,pads,dtype", [
    ([256, 256, 32], [264, 260, 32], np.float64),
    ([256, 128, 48], [272, 130, 48], np.float32),
])
def test_padded_bit_identical_to_dense(dims, pads, dtype):
    """The same kernel over the same values: only the addressing differs."""
    with P.Handle() as h:
        for sd in range(3):
            arrs = gen(dims, 60 + sd, dtype)
            print(f"padded vs dense dims={dims} pads={pads} sd={sd}", flush=True)
            xp, _ = run_padded(arrs, dims, pads, sd, h)
            cfg_p = h.kernel_config(sd)
            t = [torch.from_numpy(a.copy()).cuda() for a in arrs]
            P.solve(*t[:5], t[5], dims, sd, handle=h)
            torch.cuda.synchronize()
            cfg_d = h.kernel_config(sd)
            assert cfg_p == cfg_d, (cfg_p, cfg_d)
            assert np.array_equal(O.extract_padded(xp, dims, pads), t[5].cpu().numpy())


@pytest.mark.skipif(not O.ref_available("avx"), reason="oracle/_ref not built")
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_padded_matches_reference_cpu_library(dtype):
    """Reference CPU library on its own padded layout (pads[0] = dims[0]
    rounded up to its SIMD width), x lengths it handles (at least one vector,
    not one row past a whole vector)."""
    vec = O.ref_simd_vec(dtype)
    for dims in ([30, 20], [14, 9, 6], [62, 16, 5]):
        assert dims[0] >= vec and dims[0] % vec != 1
        pads = [-(-dims[0] // vec) * vec] + dims[1:]
        with P.Handle() as h:
            for sd in range(len(dims)):
                arrs = gen(dims, 80 + sd, dtype)
                xp, _ = run_padded(arrs, dims, pads, sd, h)
                padded = [O.embed_padded(a, dims, pads, 0.0) for a in arrs[:5]]
                xr = O.embed_padded(arrs[5], dims, pads, SENTINEL)
                O.ref_gpsv_batch(*padded, xr, dims, sd)
                a = O.extract_padded(xp, dims, pads).astype(np.float64)
                b = O.extract_padded(xr, dims, pads).astype(np.float64)
                err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
                print(f"vs reference CPU dims={dims} sd={sd} {err:.2e}", flush=True)
                assert err < (1e-12 if dtype == np.float64 else 1e-5)


def test_padded_rejects_bad_shapes():
    x = torch.zeros(64, dtype=torch.float64, device="cuda")
    c = [torch.ones(64, dtype=torch.float64, device="cuda") for _ in range(5)]
    with pytest.raises(P.PentadError):
        P.gpsv_batch_padded(*c, x, [8, 8], [6, 8], 0)  # pads[0] < dims[0]
    with pytest.raises(ValueError):
        P.gpsv_batch_padded(*c, x, [8, 8], [10, 8], 0)  # array smaller than the layout
    with pytest.raises(ValueError):
        P.gpsv_batch_padded(*c, x, [8, 8], [8], 0)
