"""GPU parity of the two-workgroup strided solve against the CPU oracle.

Each line set of S systems is solved by a pair of workgroups, each holding
half of every line; the two halves trade six words per system through the
caller's scratch once, after their block PCR.  "thomas-pcr-split" asks for
the kernel by name; the automatic choice takes it for the outermost
direction of large fields when the caller passes the scratch the size query
reports.  Bars as in test_gpu_solve.py: FP64 1e-12, FP32 2e-5 relative
max-norm forward error against the FP64 oracle of the same inputs.
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from test_gpu_solve import FP32_TOL, FP64_TOL, dev, gen, oracle_solve, relerr

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TIMEOUT_WORD = 8  # int index in the scratch header: set when a wait gave up


def split_solve(arrs, dims, sd, algo="thomas-pcr-split", with_buffer=True):
    t = [dev(a) for a in arrs]
    with P.Handle() as h:
        h.set_algo(sd, algo)
        nb = P.gpsv_batch_buffer_size_ext(*t[:5], t[5], dims, sd, h)
        buf = torch.zeros(nb, dtype=torch.uint8, device="cuda") if (nb and with_buffer) else None
        P.gpsv_batch(*t[:5], t[5], dims, sd, buf, h)
        torch.cuda.synchronize()
        cfg = h.kernel_config(sd)
    timeout = int(buf[:64 * 4].view(torch.int32)[TIMEOUT_WORD].item()) if buf is not None else 0
    return t[5].cpu().numpy(), cfg, nb, timeout


# dims, solvedim, dtype, expected configuration
CASES = [
    ([33, 7, 512], 2, np.float64, "thomas-pcr split rows=16 sys_per_wg=32"),   # last set partial
    ([64, 4, 256], 2, np.float64, "thomas-pcr split rows=16 sys_per_wg=64"),
    ([16, 24, 128], 2, np.float64, "thomas-pcr split rows=16 sys_per_wg=64"),
    ([48, 5, 64], 2, np.float64, "thomas-pcr split rows=16 sys_per_wg=64"),    # two chunks per half
    ([40, 512, 3], 1, np.float64, "thomas-pcr split rows=16 sys_per_wg=32"),   # middle direction
    ([32, 9, 512], 2, np.float32, "thomas-pcr split rows=16 sys_per_wg=32"),
    ([70, 3, 256], 2, np.float32, "thomas-pcr split rows=16 sys_per_wg=64"),
    ([24, 11, 48], 2, np.float64, "thomas-pcr split rows=8 sys_per_wg=64"),    # 8-row chunks
]


@pytest.mark.parametrize("dims,sd,dtype,cfg", CASES)
def test_split_matches_oracle(dims, sd, dtype, cfg):
    print(f"split {dims} sd={sd} {np.dtype(dtype).name}", flush=True)
    arrs = gen(dims, int(np.prod(dims)) % 1000 + sd, dtype)
    x, got, nb, timeout = split_solve(arrs, dims, sd)
    assert got == cfg
    assert nb > 0 and timeout == 0
    ref = oracle_solve(arrs, dims, sd)
    assert np.all(np.isfinite(x))
    assert relerr(x, ref) < (FP64_TOL if dtype == np.float64 else FP32_TOL)
    # the same solve again: bit-identical (both halves compute the same
    # interface values, and the pairing of workgroups does not matter)
    x2, _, _, _ = split_solve(arrs, dims, sd)
    assert np.array_equal(x, x2)


def test_split_weakly_dominant():
    """Dominance 1.05: the couplings between the halves decay slowly, so the
    interface solve and the carried right-hand side matter at full weight."""
    dims = [35, 6, 512]
    arrs = gen(dims, 7, np.float64, ratio=(1.05, 1.05))
    x, cfg, _, timeout = split_solve(arrs, dims, 2)
    assert cfg.startswith("thomas-pcr split") and timeout == 0
    assert relerr(x, oracle_solve(arrs, dims, 2)) < FP64_TOL


def test_split_without_scratch_runs_one_workgroup_per_line():
    dims = [33, 7, 512]
    arrs = gen(dims, 3, np.float64)
    x, cfg, nb, _ = split_solve(arrs, dims, 2, with_buffer=False)
    assert nb > 0
    assert cfg == "thomas-pcr rows=16 sys_per_wg=16"
    assert relerr(x, oracle_solve(arrs, dims, 2)) < FP64_TOL


def test_split_automatic_choice_outermost_direction():
    """Rows 256 KB apart, lines of 512 rows and enough line sets: the automatic choice
    takes the split kernel when the caller passes the scratch."""
    dims = [256, 128, 512]
    arrs = gen(dims, 11, np.float64)
    x, cfg, nb, timeout = split_solve(arrs, dims, 2, algo="auto")
    assert cfg.startswith("thomas-pcr split") and timeout == 0
    assert relerr(x, oracle_solve(arrs, dims, 2)) < FP64_TOL


def test_split_widened_mixed_precision():
    """FP32 coefficients, FP64 right-hand side: one FP64 solve of the
    widened coefficients on the split kernel (pentadsolver_DS_gpsv_batch_ir)."""
    dims = [33, 7, 512]
    arrs = gen(dims, 13, np.float32)
    c = [dev(a) for a in arrs[:5]]
    b = arrs[5].astype(np.float64)
    x = dev(b)
    with P.Handle() as h:
        h.set_algo(2, "thomas-pcr-split")
        nb = P.gpsv_batch_ir_buffer_size_ext(*c, x, dims, 2, h)
        buf = torch.zeros(nb, dtype=torch.uint8, device="cuda")
        P.gpsv_batch_ir(*c, x, dims, 2, buffer=buf, handle=h)
        torch.cuda.synchronize()
        assert h.kernel_config(2) == "ir widened split rows=16 sys_per_wg=32"
    assert int(buf[:256].view(torch.int32)[TIMEOUT_WORD].item()) == 0
    ref = oracle_solve(arrs[:5] + [b], dims, 2)
    assert relerr(x.cpu().numpy(), ref) < FP64_TOL


def test_split_graph_replay():
    """The two-workgroup solve captured into a hipGraph (the reset is a
    kernel node, not a memset): every replay, with eager work and
    synchronisations between them, matches the eager solve bit for bit and
    leaves the time-out word clear."""
    dims = [33, 7, 512]
    arrs = gen(dims, 19, np.float64)
    t = [dev(a) for a in arrs]
    eager, _, nb, _ = split_solve(arrs, dims, 2)
    s = torch.cuda.Stream()
    with P.Handle(stream=s) as h:
        h.set_algo(2, "thomas-pcr-split")
        buf = torch.zeros(nb, dtype=torch.uint8, device="cuda")
        x = t[5].clone()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            P.gpsv_batch(*t[:5], x, dims, 2, buf, h, check=False)
        assert h.kernel_config(2) == "thomas-pcr split rows=16 sys_per_wg=32"
        for rep in range(3):
            x.copy_(t[5])
            buf[:256].fill_(7)  # stale header words: the reset node must clear them
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(x.cpu().numpy(), eager), rep
            assert int(buf[:256].view(torch.int32)[TIMEOUT_WORD].item()) == 0
