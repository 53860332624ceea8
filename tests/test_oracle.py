"""Pins the CPU oracle before anything is checked against it.

Each check is chosen so that a plausible slip in the oracle's recurrence (a
dropped term, a wrong sign, an off-by-one index, swapped diagonals) fails at
least one of them; test_oracle_mutations shows that on planted mistakes.

* golden vectors: the reference's six shipped fixtures, at the reference's
  own tolerance rule (test/catch_utils.hpp:19-25);
* a library routine: scipy.linalg.solve_banded((2, 2)) -- LAPACK gbsv, the
  routine that wrote those fixtures (test/gen_testcases.py:27-33);
* the real reference: its CPU solvers compiled from source by oracle/Makefile
  (oracle/_ref), on seeded 1-D/2-D/3-D inputs in every direction;
* mathematics: dense LU on tiny systems (brute force), the backward residual
  |Ax - b|, the tridiagonal special case (ds = dw = 0) against
  solve_banded((1, 1)), and a manufactured solution.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import FIXTURES, FIXTURE_NAMES
from oracle import oracle as O


def _lines(dims, solvedim):
    """numpy view helper: returns (n_sys, len) gather index array."""
    idx = np.arange(int(np.prod(dims))).reshape(list(reversed(dims)))
    ax = len(dims) - 1 - solvedim
    return np.moveaxis(idx, ax, -1).reshape(-1, dims[solvedim])


def scipy_solve(ds, dl, d, du, dw, b, dims, solvedim):
    """Independent solve: LAPACK banded LU with partial pivoting per line."""
    out = np.empty_like(b, dtype=np.float64)
    for line in _lines(dims, solvedim):
        n = len(line)
        ab = np.zeros((5, n))
        # solve_banded layout: ab[2 + i - j, j] = A[i, j]
        ab[0, 2:] = dw[line][:-2]
        ab[1, 1:] = du[line][:-1]
        ab[2, :] = d[line]
        ab[3, :-1] = dl[line][1:]
        ab[4, :-2] = ds[line][2:]
        out[line] = sla.solve_banded((2, 2), ab, b[line])
    return out


def gen(dims, seed, ratio=(2.0, 3.0)):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(*ratio, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    return ds, dl, d, du, dw, b


def relerr(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", FIXTURE_NAMES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_reference_fixtures(name, dtype):
    fx = O.read_fixture(os.path.join(FIXTURES, name), dtype)
    x = fx["x"].copy()
    O.gpsv_batch(fx["ds"], fx["dl"], fx["d"], fx["du"], fx["dw"], x,
                 fx["dims"], fx["solvedim"])
    assert O.allclose_reference_rule(fx["u"], x, dtype)


SHAPES = [([37], 0), ([24, 9], 0), ([9, 24], 1), ([16, 12, 10], 0),
          ([16, 12, 10], 1), ([16, 12, 10], 2), ([8, 6, 5, 7], 2)]


@pytest.mark.parametrize("dims,sd", SHAPES)
def test_oracle_vs_lapack_banded(dims, sd):
    ds, dl, d, du, dw, b = gen(dims, 11 + sd)
    x = O.solve(ds, dl, d, du, dw, b, dims, sd)
    ref = scipy_solve(ds, dl, d, du, dw, b, dims, sd)
    assert relerr(x, ref) < 1e-13
    # the oracle's own pivoting routine agrees with LAPACK too
    piv = O.solve_pivot(ds, dl, d, du, dw, b, dims, sd, "banded")
    assert relerr(piv, ref) < 1e-13


def test_oracle_weakly_dominant_vs_lapack():
    dims = [64, 8, 8]
    ds, dl, d, du, dw, b = gen(dims, 5, ratio=(1.05, 1.05))
    for sd in range(3):
        x = O.solve(ds, dl, d, du, dw, b, dims, sd)
        ref = scipy_solve(ds, dl, d, du, dw, b, dims, sd)
        assert relerr(x, ref) < 1e-12


@pytest.mark.parametrize("n", [3, 4, 5, 7, 12])
def test_oracle_vs_dense_lu_tiny(n):
    ds, dl, d, du, dw, b = gen([n, 3], 100 + n)
    x = O.solve(ds, dl, d, du, dw, b, [n, 3], 0)
    dense = O.solve_pivot(ds, dl, d, du, dw, b, [n, 3], 0, "dense")
    # brute force: assemble with numpy and solve
    for j in range(3):
        sl = slice(j * n, (j + 1) * n)
        A = np.diag(d[sl]) + np.diag(dl[sl][1:], -1) + np.diag(du[sl][:-1], 1)
        if n > 2:
            A += np.diag(ds[sl][2:], -2) + np.diag(dw[sl][:-2], 2)
        ref = np.linalg.solve(A, b[sl])
        assert relerr(x[sl], ref) < 1e-13
        assert relerr(dense[sl], ref) < 1e-13


def test_oracle_64cube_x_vs_dense_lu():
    """The 64^3 CPU configuration: 4096 systems of length 64 against dense LU."""
    dims = [64, 64, 64]
    ds, dl, d, du, dw, b = gen(dims, 1)
    x = O.solve(ds, dl, d, du, dw, b, dims, 0)
    dense = O.solve_pivot(ds, dl, d, du, dw, b, dims, 0, "dense")
    assert relerr(x, dense) < 1e-13
    assert O.residual(ds, dl, d, du, dw, x, b, dims, 0) < 1e-13


@pytest.mark.parametrize("sd", [0, 1, 2])
def test_oracle_residual_and_outofrange_ignored(sd):
    dims = [12, 10, 9]
    ds, dl, d, du, dw, b = gen(dims, 7)
    x = O.solve(ds, dl, d, du, dw, b, dims, sd)
    assert O.residual(ds, dl, d, du, dw, x, b, dims, sd) < 1e-14
    # changing the don't-care couplings changes nothing
    lines = _lines(dims, sd)
    ds2, dl2, du2, dw2 = ds.copy(), dl.copy(), du.copy(), dw.copy()
    ds2[lines[:, :2]] = 123.0
    dl2[lines[:, 0]] = -77.0
    du2[lines[:, -1]] = 55.0
    dw2[lines[:, -2:]] = 9.0
    x2 = O.solve(ds2, dl2, d, du2, dw2, b, dims, sd)
    assert np.array_equal(x, x2)


def test_oracle_tridiagonal_special_case():
    dims = [50, 7]
    _, dl, d, du, _, b = gen(dims, 3)
    z = np.zeros_like(d)
    x = O.solve(z, dl, d, du, z, b, dims, 0)
    for j in range(7):
        sl = slice(j * 50, (j + 1) * 50)
        ab = np.zeros((3, 50))
        ab[0, 1:] = du[sl][:-1]
        ab[1] = d[sl]
        ab[2, :-1] = dl[sl][1:]
        ref = sla.solve_banded((1, 1), ab, b[sl])
        assert relerr(x[sl], ref) < 1e-14


def test_oracle_manufactured_solution():
    dims = [40, 6, 5]
    ds, dl, d, du, dw, _ = gen(dims, 9)
    xt = np.random.default_rng(4).uniform(-1, 1, int(np.prod(dims)))
    for sd in range(3):
        b = np.zeros_like(xt)
        for line in _lines(dims, sd):
            v = xt[line]
            n = len(line)
            r = d[line] * v
            r[1:] += dl[line][1:] * v[:-1]
            r[2:] += ds[line][2:] * v[:-2]
            r[:-1] += du[line][:-1] * v[1:]
            r[:-2] += dw[line][:-2] * v[2:]
            b[line] = r
        x = O.solve(ds, dl, d, du, dw, b, dims, sd)
        assert relerr(x, xt) < 1e-13


def test_oracle_long_double_agrees():
    dims = [33, 5, 4]
    ds, dl, d, du, dw, b = gen(dims, 21)
    for sd in range(3):
        x = O.solve(ds, dl, d, du, dw, b, dims, sd)
        xl = O.solve_ld(ds, dl, d, du, dw, b, dims, sd)
        assert relerr(x, xl) < 1e-14


def test_oracle_float32_matches_reference_rule_on_fp32_inputs():
    dims = [24, 9]
    ds, dl, d, du, dw, b = (a.astype(np.float32) for a in gen(dims, 2))
    x = O.solve(ds, dl, d, du, dw, b, dims, 0)
    ref = scipy_solve(*(a.astype(np.float64) for a in (ds, dl, d, du, dw, b)), dims, 0)
    assert relerr(x.astype(np.float64), ref) < 1e-5


# ---------------------------------------------------------------------------
# against the reference's own CPU solvers, compiled from its sources
# ---------------------------------------------------------------------------
REF_SHAPES = [([200], 0), ([32, 24], 0), ([32, 24], 1), ([40, 24, 36], 0),
              ([40, 24, 36], 1), ([40, 24, 36], 2), ([64, 64, 64], 0)]


@pytest.mark.skipif(not O.ref_available("avx"), reason="oracle/_ref not built")
@pytest.mark.parametrize("dims,sd", REF_SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", ["avx", "orig"])
def test_oracle_vs_reference_build(dims, sd, dtype, kind):
    ds, dl, d, du, dw, b = (a.astype(dtype) for a in gen(dims, 31 + sd))
    x = O.solve(ds, dl, d, du, dw, b, dims, sd)
    xr = b.copy()
    O.ref_gpsv_batch(ds, dl, d, du, dw, xr, dims, sd, kind=kind)
    tol = 1e-14 if dtype == np.float64 else 2e-6
    assert relerr(xr.astype(np.float64), x.astype(np.float64)) < tol


@pytest.mark.skipif(not O.ref_available("avx"), reason="oracle/_ref not built")
def test_reference_build_passes_its_own_fixtures():
    for name in FIXTURE_NAMES:
        for dtype in (np.float64, np.float32):
            fx = O.read_fixture(os.path.join(FIXTURES, name), dtype)
            if fx["dims"][0] % 8 and len(fx["dims"]) > 1:
                continue
            x = fx["x"].copy()
            O.ref_gpsv_batch(fx["ds"], fx["dl"], fx["d"], fx["du"], fx["dw"], x,
                             fx["dims"], fx["solvedim"])
            assert O.allclose_reference_rule(fx["u"], x, dtype), (name, dtype)


# ---------------------------------------------------------------------------
# the pins have power: planted mistakes in a restated recurrence fail them
# ---------------------------------------------------------------------------
def _thomas_py(ds, dl, d, du, dw, b, mut=None):
    n = len(b)
    u2, w2, x2 = np.zeros(n), np.zeros(n), np.zeros(n)
    u2[0], w2[0], x2[0] = du[0] / d[0], dw[0] / d[0], b[0] / d[0]
    ddl = dl[1]
    dd = d[1] - ddl * u2[0]
    u2[1] = (du[1] - ddl * w2[0]) / dd
    w2[1] = dw[1] / dd
    x2[1] = (b[1] - ddl * x2[0]) / dd
    for i in range(2, n):
        s = ds[i]
        ddl = dl[i] - s * u2[i - 2]
        if mut == "drop_s_dw2":
            dd = d[i] - ddl * u2[i - 1]
        elif mut == "sign":
            dd = d[i] + s * w2[i - 2] - ddl * u2[i - 1]
        else:
            dd = d[i] - s * w2[i - 2] - ddl * u2[i - 1]
        u2[i] = (du[i] - ddl * w2[i - 1]) / dd
        w2[i] = dw[i] / dd
        j = i - 1 if mut == "index" else i - 2
        x2[i] = (b[i] - s * x2[j] - ddl * x2[i - 1]) / dd
    x = np.zeros(n)
    x[-1] = x2[-1]
    x[-2] = x2[-2] - u2[-2] * x[-1]
    for i in range(n - 3, -1, -1):
        if mut == "swap_uw":
            x[i] = x2[i] - u2[i] * x[i + 2] - w2[i] * x[i + 1]
        else:
            x[i] = x2[i] - w2[i] * x[i + 2] - u2[i] * x[i + 1]
    return x


@pytest.mark.parametrize("mut", [None, "drop_s_dw2", "sign", "index", "swap_uw"])
def test_oracle_mutations(mut):
    fx = O.read_fixture(os.path.join(FIXTURES, "one_dim_large_solve0"))
    x = _thomas_py(fx["ds"], fx["dl"], fx["d"], fx["du"], fx["dw"], fx["x"], mut)
    ok = O.allclose_reference_rule(fx["u"], x, np.float64)
    assert ok == (mut is None)


def test_oracle_varlen_vs_lapack_and_reference():
    """Packed variable-length systems (lengths 3..8192): every system against
    LAPACK's banded LU; systems up to the reference CPU solver's N_MAX (1024,
    src/cpu/singlenode/pentad_cpu.cpp:11) also against the reference build,
    one 1-D solve each.  Lengths above 1024 have no reference counterpart and
    are pinned by LAPACK only."""
    rng = np.random.default_rng(31)
    lens = np.concatenate([[3, 4, 5, 1024, 1025, 8192],
                           np.rint(np.exp(rng.uniform(np.log(3), np.log(8192), 30)))]).astype(np.int64)
    off = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    n = int(off[-1])
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(1.2, 3.0, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    x = O.varlen(ds, dl, d, du, dw, b, off)
    for k in range(len(lens)):
        lo, hi = int(off[k]), int(off[k + 1])
        sl = slice(lo, hi)
        ref = scipy_solve(ds[sl], dl[sl], d[sl], du[sl], dw[sl], b[sl], [hi - lo], 0)
        assert relerr(x[sl], ref) < 1e-13, (k, hi - lo)
        if O.ref_available("avx") and hi - lo <= 1024:
            xr = b[sl].copy()
            O.ref_gpsv_batch(ds[sl].copy(), dl[sl].copy(), d[sl].copy(), du[sl].copy(),
                             dw[sl].copy(), xr, [hi - lo], 0)
            assert relerr(x[sl], xr) < 1e-13, (k, hi - lo)
