"""The drop-in's environment contract on the GPU: the reference's
PENTA_XALGO / PENTA_YALGO / PENTA_ZALGO / PENTA_YZALGO selectors with their
precedence (a per-direction variable wins over PENTA_YZALGO,
src/cuda/singlenode/pentad_cuda.cu:1163-1170), the "legacy" alias of the
one-lane kernel, the fall-through of names a direction does not know
(:1455-1488 for x, :1316-1418 for y / z), PENTA_DEBUG_LAUNCH's report of a
kernel that declined (:17-37), and pentadsolver_freeze_env.  Every solve
goes through the C ABI and is checked against the oracle (FP64: <= 1e-12);
pentadsolver_kernel_that_ran must name the kernel that actually ran.
"""
import numpy as np
import pytest

import pentadsolver_amd as P
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ENV = ["PENTA_XALGO", "PENTA_YALGO", "PENTA_ZALGO", "PENTA_YZALGO", "PENTA_DEBUG_LAUNCH"]
DIMS = [256, 48, 40]


@pytest.fixture(autouse=True)
def clean_env(monkeypatch):
    for k in ENV:
        monkeypatch.delenv(k, raising=False)


def gen(dims, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(2, 3, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    return [ds, dl, d, du, dw], rng.uniform(-1, 1, n)


def solve_and_check(sd, dims=DIMS, handle=None, seed=3):
    """One in-place solve along sd through the C ABI, checked against the
    oracle; returns pentadsolver_kernel_that_ran for that direction."""
    coef, b = gen(dims, seed + sd)
    t = [torch.from_numpy(a).cuda() for a in coef]
    x = torch.from_numpy(b.copy()).cuda()
    P.solve(*t, x, dims, sd, handle=handle)
    torch.cuda.synchronize()
    ref = O.solve(*coef, b.copy(), dims, sd)
    err = float(np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref)))
    d = 0 if sd == 0 else (2 if sd == len(dims) - 1 else 1)
    ran = P.kernel_that_ran(d)
    print(f"  sd={sd} ran={ran} err={err:.2e}", flush=True)
    assert err < 1e-12
    return ran


def test_default_is_auto():
    assert [solve_and_check(sd) for sd in range(3)] == ["thomas-pcr"] * 3


def test_per_direction_variable_wins_over_yzalgo(monkeypatch):
    monkeypatch.setenv("PENTA_YZALGO", "naive")
    monkeypatch.setenv("PENTA_ZALGO", "thomas-pcr")
    assert solve_and_check(1) == "naive"
    assert solve_and_check(2) == "thomas-pcr"
    monkeypatch.setenv("PENTA_YALGO", "thomas-pcr")
    monkeypatch.setenv("PENTA_ZALGO", "naive")
    assert solve_and_check(1) == "thomas-pcr"
    assert solve_and_check(2) == "naive"


def test_yzalgo_applies_to_both_and_legacy_is_naive(monkeypatch):
    monkeypatch.setenv("PENTA_YZALGO", "legacy")
    assert solve_and_check(1) == "naive"
    assert solve_and_check(2) == "naive"
    # an empty per-direction variable does not override PENTA_YZALGO
    monkeypatch.setenv("PENTA_ZALGO", "")
    assert solve_and_check(2) == "naive"


@pytest.mark.parametrize("algo,want", [("naive", "naive"), ("transpose", "transpose"),
                                       ("thomas-pcr", "thomas-pcr"), ("auto", "thomas-pcr"),
                                       ("no-such-kernel", "thomas-pcr")])
def test_xalgo(monkeypatch, algo, want):
    """PENTA_XALGO selects the x kernel; an unrecognised name runs the
    automatic choice, as the reference's x dispatch does."""
    monkeypatch.setenv("PENTA_XALGO", algo)
    assert solve_and_check(0) == want


def test_unknown_strided_name_runs_the_one_lane_kernel(monkeypatch):
    """Along y / z the reference runs its one-lane kernel for any name other
    than auto, thomas-pcr and shared-fact."""
    monkeypatch.setenv("PENTA_YALGO", "transpose")
    assert solve_and_check(1) == "naive"


def test_xalgo_thomas_pcr_without_a_partition_falls_back(monkeypatch, capfd):
    """thomas-pcr asked for on a length with no partitioned kernel (37 rows):
    the LDS-staged kernel runs, and PENTA_DEBUG_LAUNCH=1 says so on stderr."""
    monkeypatch.setenv("PENTA_XALGO", "thomas-pcr")
    dims = [37, 64, 8]
    assert solve_and_check(0, dims) == "transpose"
    assert "declined" not in capfd.readouterr().err
    monkeypatch.setenv("PENTA_DEBUG_LAUNCH", "1")
    assert solve_and_check(0, dims) == "transpose"
    err = capfd.readouterr().err
    assert "[penta] thomas-pcr declined along solvedim 0 (length 37)" in err, err
    assert "falling back to transpose" in err


def test_shared_fact_on_an_uncovered_length(monkeypatch, capfd):
    """shared-fact asked for where its kernels do not fit falls back as the
    reference's dispatch does: x to the automatic choice, y / z to the
    one-lane kernel."""
    monkeypatch.setenv("PENTA_DEBUG_LAUNCH", "1")
    monkeypatch.setenv("PENTA_XALGO", "shared-fact")
    monkeypatch.setenv("PENTA_YZALGO", "shared-fact")
    dims = [37, 41, 43]
    assert solve_and_check(0, dims) == "transpose"
    assert solve_and_check(1, dims) == "naive"
    err = capfd.readouterr().err
    assert "shared-fact declined along solvedim 0" in err, err
    assert "shared-fact declined along solvedim 1" in err, err


def test_handle_setting_wins_over_environment(monkeypatch):
    monkeypatch.setenv("PENTA_YALGO", "naive")
    with P.Handle() as h:
        h.set_algo(1, "thomas-pcr")
        assert solve_and_check(1, handle=h) == "thomas-pcr"
        h.set_algo(1, None)  # cleared: the environment applies again
        assert solve_and_check(1, handle=h) == "naive"


def test_freeze_env_ignores_later_changes(monkeypatch):
    monkeypatch.setenv("PENTA_XALGO", "naive")
    monkeypatch.setenv("PENTA_ZALGO", "naive")
    with P.Handle() as h:
        h.freeze_env(True)
        monkeypatch.setenv("PENTA_XALGO", "thomas-pcr")
        monkeypatch.delenv("PENTA_ZALGO")
        assert solve_and_check(0, handle=h) == "naive"
        assert solve_and_check(2, handle=h) == "naive"
        # another handle (and the default context) reads the environment now
        assert solve_and_check(0) == "thomas-pcr"
        assert solve_and_check(2) == "thomas-pcr"
        h.freeze_env(False)
        assert solve_and_check(0, handle=h) == "thomas-pcr"
        assert solve_and_check(2, handle=h) == "thomas-pcr"
