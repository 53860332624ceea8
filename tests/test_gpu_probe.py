"""The bandwidth probes on the GPU (no reference counterpart; variants of the
reference's apps/roofline_probe_cuda.cu:27-49, which streams the solve's
bytes without solving).

pentadsolver_probe_strided_io stands in for the FP64 strided solve's I/O:
it must read every row of the five diagonals and of x along the lines and
write each row of x exactly once (x += ds + dl + d + du + dw, exact for the
small integers used here), whatever the split of the lines over workgroups.
pentadsolver_probe_stream computes out = in0 + ... + in5 when fmas = 0.
"""
import numpy as np
import pytest

import pentadsolver_amd as P

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,sd", [([64, 512, 4], 1), ([16, 8, 512], 2), ([8, 16, 512], 2)])
@pytest.mark.parametrize("split", [1, 2, 4])
def test_strided_io_probe_touches_every_row_once(dims, sd, split):
    print(f"probe_strided_io dims={dims} sd={sd} split={split}", flush=True)
    rng = np.random.default_rng(split * 10 + sd)
    n = int(np.prod(dims))
    coef = [torch.tensor(rng.integers(-8, 9, n).astype(np.float64), device="cuda")
            for _ in range(5)]
    x0 = rng.integers(-100, 101, n).astype(np.float64)
    x = torch.tensor(x0, device="cuda")
    with P.Handle() as h:
        P.probe_strided_io(split, coef, x, dims, sd, handle=h)
        torch.cuda.synchronize()
    want = x0 + sum(c.cpu().numpy() for c in coef)
    assert np.array_equal(x.cpu().numpy(), want)


def test_strided_io_probe_rejects_shapes_it_does_not_cover():
    n = 64 * 8 * 512
    coef = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(5)]
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    with P.Handle() as h:
        for split, dims, sd in [(3, [64, 8, 512], 2),      # split not 1, 2, 4
                                (1, [64, 8, 512], 0),      # x is not a strided direction
                                (1, [64, 512, 8], 2),      # lines of 8 rows
                                (1, [20, 8, 512], 2)]:     # 160 lines: not a multiple of 64
            with pytest.raises(P.PentadError):
                P.probe_strided_io(split, coef, x, dims, sd, handle=h)
    assert float(x.abs().max()) == 0.0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_stream_probe_sums_its_inputs(dtype):
    n = 1 << 20
    rng = np.random.default_rng(7)
    ins = [torch.tensor(rng.integers(-50, 51, n), dtype=dtype, device="cuda") for _ in range(6)]
    out = torch.empty(n, dtype=dtype, device="cuda")
    with P.Handle() as h:
        P.probe_stream(ins, out, handle=h)
        torch.cuda.synchronize()
    want = sum(t.cpu().numpy().astype(np.float64) for t in ins)
    assert np.array_equal(out.cpu().numpy().astype(np.float64), want)
