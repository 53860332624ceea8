"""The allocation-free FP64 solve, pentadsolver_D_gpsv_batch_noalloc, on the
GPU through the C ABI: it must give exactly what pentadsolver_D_gpsv_batch
gives with a caller buffer (same kernel, bit for bit) and match the oracle
(<= 1e-12), on two calls in a row with the second on a smaller grid (the
fixed scratch array keeps the first call's contents), for the automatic
choice (no scratch), the one-lane kernel (24 B of scratch per row) and the
shared factorisation (its factor table); a solve needing more scratch than
the fixed array is refused before anything runs.
"""
import ctypes

import numpy as np
import pytest

import pentadsolver_amd as P
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NOALLOC_BYTES = 16 << 20  # PENTAD_NOALLOC_SCRATCH_BYTES


def gen(dims, seed, shared_sd=None):
    """Dominant FP64 systems; with shared_sd every line along that axis has
    the same coefficients (the shared factorisation's contract)."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    if shared_sd is None:
        ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
        d = rng.uniform(2, 3, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
        coef = [ds, dl, d, du, dw]
    else:
        m = dims[shared_sd]
        ds, dl, du, dw = (rng.uniform(-1, 1, m) for _ in range(4))
        d = rng.uniform(2, 3, m) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
        shape = tuple(reversed(dims))
        view = [1] * len(dims)
        view[len(dims) - 1 - shared_sd] = m
        coef = [np.ascontiguousarray(np.broadcast_to(a.reshape(view), shape)).reshape(-1)
                for a in (ds, dl, d, du, dw)]
    return coef, rng.uniform(-1, 1, n)


def noalloc(t, x, dims, sd, h):
    lib = P.load_library()
    dd = (ctypes.c_int * len(dims))(*dims)
    return lib.pentadsolver_D_gpsv_batch_noalloc(
        h.value, *[ctypes.c_void_p(a.data_ptr()) for a in t], ctypes.c_void_p(x.data_ptr()),
        dd, len(dims), sd)


@pytest.mark.parametrize("algo", ["auto", "naive", "shared-fact"])
def test_noalloc_matches_buffered_solve(algo):
    with P.Handle() as h:
        for sd_dir in range(3):
            h.set_algo(sd_dir, algo)
        for dims in ([96, 64, 40], [32, 24, 16]):  # the second call is smaller
            for sd in range(3):
                print(f"noalloc algo={algo} dims={dims} sd={sd}", flush=True)
                coef, b = gen(dims, 7 + sd + dims[0], sd if algo == "shared-fact" else None)
                t = [torch.from_numpy(a).cuda() for a in coef]
                xb = torch.from_numpy(b.copy()).cuda()
                buf = P.solve(*t, xb, dims, sd, handle=h)
                cfg_b = h.kernel_config(sd)
                xn = torch.from_numpy(b.copy()).cuda()
                assert noalloc(t, xn, dims, sd, h) == 0
                torch.cuda.synchronize()
                cfg_n = h.kernel_config(sd)
                need = 0 if buf is None else buf.numel()
                print(f"  {cfg_n} scratch {need} B", flush=True)
                assert cfg_n == cfg_b
                if algo != "auto":
                    assert need > 0 and cfg_n.startswith(algo), (need, cfg_n)
                assert torch.equal(xn, xb)
                ref = O.solve(*coef, b.copy(), dims, sd)
                err = float(np.max(np.abs(xn.cpu().numpy() - ref)) / np.max(np.abs(ref)))
                assert err < 1e-12, err


def test_noalloc_refuses_more_scratch_than_its_array(capfd):
    dims = [128, 128, 64]  # one-lane kernel: 24 B x 2^20 rows = 24 MiB > 16 MiB
    coef, b = gen(dims, 3)
    t = [torch.from_numpy(a).cuda() for a in coef]
    x = torch.from_numpy(b.copy()).cuda()
    with P.Handle() as h:
        h.set_algo(1, "naive")
        assert P.gpsv_batch_buffer_size_ext(*t, x, dims, 1, h) > NOALLOC_BYTES
        assert noalloc(t, x, dims, 1, h) == 3  # PENTAD_STATUS_INVALID_VALUE
        assert h.last_status() == 3
        torch.cuda.synchronize()
        assert "more than the library's fixed" in capfd.readouterr().err
        assert np.array_equal(x.cpu().numpy(), b)  # nothing ran
        # the same handle still solves a shape that fits
        small = [24, 16, 8]
        coef, b = gen(small, 4)
        t = [torch.from_numpy(a).cuda() for a in coef]
        x = torch.from_numpy(b.copy()).cuda()
        assert noalloc(t, x, small, 1, h) == 0
        torch.cuda.synchronize()
        ref = O.solve(*coef, b.copy(), small, 1)
        assert float(np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref))) < 1e-12


def test_noalloc_null_handle():
    dims = [64, 32, 16]
    coef, b = gen(dims, 5)
    t = [torch.from_numpy(a).cuda() for a in coef]
    x = torch.from_numpy(b.copy()).cuda()
    lib = P.load_library()
    dd = (ctypes.c_int * 3)(*dims)
    assert lib.pentadsolver_D_gpsv_batch_noalloc(
        None, *[ctypes.c_void_p(a.data_ptr()) for a in t], ctypes.c_void_p(x.data_ptr()),
        dd, 3, 2) == 0
    torch.cuda.synchronize()
    ref = O.solve(*coef, b.copy(), dims, 2)
    assert float(np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref))) < 1e-12


def test_noalloc_two_workgroup_lines():
    """Outermost lines of 512 rows, 256 KB apart: the automatic choice shares
    each line between two workgroups, its trade words in the fixed array when
    they fit (3 MiB here); on a 512^3 field they do not (24 MiB), and the
    solve runs one workgroup per line instead of being refused."""
    with P.Handle() as h:
        for dims, split in (([256, 128, 512], True), ([512, 512, 512], False)):
            print(f"noalloc split dims={dims}", flush=True)
            coef, b = gen(dims, 17)
            t = [torch.from_numpy(a).cuda() for a in coef]
            xb = torch.from_numpy(b.copy()).cuda()
            if split:
                P.solve(*t, xb, dims, 2, handle=h)  # caller buffer: the split kernel
            else:
                P.gpsv_batch(*t, xb, dims, 2, None, h)  # no buffer: one workgroup per line
            cfg_b = h.kernel_config(2)
            xn = torch.from_numpy(b.copy()).cuda()
            assert noalloc(t, xn, dims, 2, h) == 0
            torch.cuda.synchronize()
            cfg_n = h.kernel_config(2)
            assert cfg_n == cfg_b and cfg_n.startswith("thomas-pcr split") == split, cfg_n
            assert torch.equal(xn, xb)
            if split:
                ref = O.solve(*coef, b.copy(), dims, 2)
                err = float(np.max(np.abs(xn.cpu().numpy() - ref)) / np.max(np.abs(ref)))
                assert err < 1e-12, err
            del t, xb, xn
            torch.cuda.empty_cache()
