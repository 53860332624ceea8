"""Offline generator of test fixtures in the reference's text format.

Restates the published algorithm of the reference's generator
(test/gen_testcases.py:12-124; written here, the reference's script is not
imported or run):

* recipe (gen_testcases.py:83-95): with n_sys = prod(shape[:-1]) systems of
  N = shape[-1] rows drawn from numpy's legacy global generator,
  ds, dl, du, dw = -1 + 0.1 * U(0,1), d = 6 + U(0,1) (drawn in the order
  ds, dl, d, du, dw), couplings past the ends set to 0, RHS x = U(0,1);
* expected solution (:12-40): every system solved along the last numpy axis
  (the fastest dimension) with the banded LU of LAPACK gbsv
  (scipy.linalg.solve_banded, (2, 2) bands), after rounding the inputs to 15
  decimals for the FP64 line and to 6 decimals for the FP32 line;
* file per solve dimension (:43-79, :112-124): the arrays are written, then
  rotated (numpy axes (1, .., n-1, 0)) so that the solved axis becomes the
  next dimension; lines: "ndims solvedim", dims fastest first, ds, dl, d,
  du, dw, x, u (15 decimals), u for FP32 (6 decimals); values printed as
  str(round(v, 15)) / str(round(v, 6));
* order (:127-133): np.random.seed(1), then one_dim_small [5],
  one_dim_large [200], two_dim_small [8, 8], two_dim_large [32, 32],
  three_dim_large [32, 32, 32].

Run with the same seed and order, the generator rewrites the six fixtures the
reference ships, value for value (tests/test_fixture_generator.py checks
that), so the three it writes for three_dim_large are the 3-D fixtures the
reference lists as missing blobs (.MISSING_LARGE_BLOBS).

Usage:  python tests/golden/gen_fixtures.py OUTDIR [NAME ...]
(NAME among the five above; default all).  Needs numpy and scipy.
"""
from __future__ import annotations

import os
import sys
from typing import Dict, List, Sequence

import numpy as np

PRECISION = 15
PRECISION_FLOAT = 6
CASES = [("one_dim_small", [5]), ("one_dim_large", [200]), ("two_dim_small", [8, 8]),
         ("two_dim_large", [32, 32]), ("three_dim_large", [32, 32, 32])]


def solve_rows(ds, dl, d, du, dw, x, decimals: int) -> np.ndarray:
    """Solve every row-system (last axis) with LAPACK's banded LU after
    rounding the inputs to ``decimals``."""
    from scipy.linalg import solve_banded
    arrs = [np.round(a, decimals) for a in (ds, dl, d, du, dw, x)]
    ds, dl, d, du, dw, x = arrs
    n = d.shape[1]
    out = np.zeros_like(x)
    for i in range(d.shape[0]):
        # banded storage, upper bands first: ab[2 + r - c, c] = A[r, c]
        ab = np.zeros((5, n))
        ab[0, 2:] = dw[i, :n - 2]   # A[r, r+2]
        ab[1, 1:] = du[i, :n - 1]   # A[r, r+1]
        ab[2, :] = d[i]
        ab[3, :n - 1] = dl[i, 1:]   # A[r, r-1]
        ab[4, :n - 2] = ds[i, 2:]   # A[r, r-2]
        out[i] = solve_banded((2, 2), ab, x[i])
    return out


def draw(shape: Sequence[int]) -> Dict[str, np.ndarray]:
    """The recipe, drawn from numpy's legacy global generator."""
    n = int(shape[-1])
    n_sys = int(np.prod(shape[:-1])) if len(shape) > 1 else 1
    r = np.random.rand
    a = {"ds": -1 + r(n_sys, n) * 0.1, "dl": -1 + r(n_sys, n) * 0.1, "d": 6 + r(n_sys, n),
         "du": -1 + r(n_sys, n) * 0.1, "dw": -1 + r(n_sys, n) * 0.1}
    a["ds"][:, :2] = 0
    a["dl"][:, :1] = 0
    a["du"][:, -1:] = 0
    a["dw"][:, -2:] = 0
    a["x"] = r(n_sys, n)
    return a


def _line(a: np.ndarray, decimals: int) -> str:
    return " ".join(str(round(v, decimals)) for v in a.reshape(-1))


def fixture_texts(shape: Sequence[int]) -> List[str]:
    """Texts of the fixture files name_solve0 .. name_solve{n-1}."""
    a = draw(shape)
    keys = ("ds", "dl", "d", "du", "dw", "x")
    u = solve_rows(*[a[k] for k in keys], PRECISION)
    uf = solve_rows(*[a[k] for k in keys], PRECISION_FLOAT)
    arrs = [a[k].reshape(shape) for k in keys] + [u.reshape(shape), uf.reshape(shape)]
    rot = list(range(1, len(shape))) + [0]
    texts = []
    for solvedim in range(len(shape)):
        cur = arrs[0].shape
        lines = [f"{len(shape)} {solvedim}", " ".join(str(v) for v in reversed(cur))]
        lines += [_line(v, PRECISION) for v in arrs[:7]]
        lines.append(_line(arrs[7], PRECISION_FLOAT))
        texts.append("\n".join(lines) + "\n")
        arrs = [np.transpose(v, rot) for v in arrs]
    return texts


def generate(outdir: str, names: Sequence[str] = ()) -> List[str]:
    """Seed 1, then every case in order (the draws of earlier cases advance
    the generator); writes the requested ones.  Returns the paths written."""
    np.random.seed(1)
    os.makedirs(outdir, exist_ok=True)
    written = []
    for name, shape in CASES:
        texts = fixture_texts(shape)
        if names and name not in names:
            continue
        for k, t in enumerate(texts):
            path = os.path.join(outdir, f"{name}_solve{k}")
            with open(path, "w") as f:
                f.write(t)
            written.append(path)
    return written


if __name__ == "__main__":
    if len(sys.argv) < 2:
        sys.exit(__doc__)
    for p in generate(sys.argv[1], sys.argv[2:]):
        print(p)
