"""Seeded synthetic inputs for the batched pentadiagonal solve.

The reference defines no generator for these workloads; the recipes below
are this library's own, fixed so that tests and benchmarks are reproducible:

* ``dominant``: off-diagonals ds, dl, du, dw ~ U(-1, 1), main diagonal
  d = ratio * (|ds| + |dl| + |du| + |dw|) with ratio ~ U(lo, hi), RHS ~ U(-1, 1).
  Every drawn coupling enters the sum, so each row is strictly diagonally
  dominant along any of the three axes the ADI step reinterprets the arrays
  along (the same five arrays serve x, y and z, reference apps/adi_cuda.cu:167-176).
  Draw order: ds, dl, du, dw, ratio, rhs.
* ``weak``: as ``dominant`` with ratio fixed at 1.05 (weakly dominant set).
* ``hyperdiffusion``: implicit fourth-order diffusion I + s*[1,-4,6,-4,1] with
  s = nu*dt/h^4 per grid line, nu ~ U(0.5, 2).
* ``varlen_offsets``: lengths round(exp(U(ln lo, ln hi))) packed contiguously.

Host generation uses numpy (``numpy.random.default_rng(seed)``); device
generation uses torch's CUDA generator with the same recipe (different random
stream, same distribution) for sizes too large to build on the host.
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple

import numpy as np


def n_elems(dims: Sequence[int]) -> int:
    return int(np.prod([int(v) for v in dims]))


def dominant_np(dims: Sequence[int], seed: int, ratio: Tuple[float, float] = (2.0, 3.0),
                dtype=np.float64):
    """Host arrays (ds, dl, d, du, dw, rhs), flat, dims[0] fastest."""
    n = n_elems(dims)
    rng = np.random.default_rng(seed)
    ds = rng.uniform(-1.0, 1.0, n)
    dl = rng.uniform(-1.0, 1.0, n)
    du = rng.uniform(-1.0, 1.0, n)
    dw = rng.uniform(-1.0, 1.0, n)
    lo, hi = ratio
    r = rng.uniform(lo, hi, n) if hi > lo else np.full(n, lo)
    rhs = rng.uniform(-1.0, 1.0, n)
    d = r * (np.abs(ds) + np.abs(dl) + np.abs(du) + np.abs(dw))
    return tuple(a.astype(dtype) for a in (ds, dl, d, du, dw, rhs))


def dominant_torch(dims: Sequence[int], seed: int, ratio: Tuple[float, float] = (2.0, 3.0),
                   dtype=None, device="cuda"):
    """Device arrays with the ``dominant`` recipe, generated on the GPU."""
    import torch
    dtype = dtype or torch.float64
    n = n_elems(dims)
    g = torch.Generator(device=device)
    g.manual_seed(seed)

    def u(lo, hi):
        t = torch.empty(n, dtype=torch.float64, device=device)
        t.uniform_(lo, hi, generator=g)
        return t

    ds, dl, du, dw = u(-1, 1), u(-1, 1), u(-1, 1), u(-1, 1)
    lo, hi = ratio
    r = u(lo, hi) if hi > lo else torch.full((n,), lo, dtype=torch.float64, device=device)
    rhs = u(-1, 1)
    d = r * (ds.abs() + dl.abs() + du.abs() + dw.abs())
    del r
    return tuple(a.to(dtype) for a in (ds, dl, d, du, dw, rhs))


def hyperdiffusion_coeffs_np(dims: Sequence[int], seed: int, dt_over_h4: float,
                             dtype=np.float64):
    """Coefficients of I + s*D4 with s = nu*dt/h^4, nu ~ U(0.5, 2) per element.

    Rows: ds = dw = s, dl = du = -4s, d = 1 + 6s.  At the ends of a line the
    out-of-range couplings are simply not part of the system (the stencil is
    truncated: zero values beyond the boundary).  The same arrays serve all
    three directions.
    """
    n = n_elems(dims)
    rng = np.random.default_rng(seed)
    s = rng.uniform(0.5, 2.0, n) * dt_over_h4
    ds = s.copy()
    dw = s.copy()
    dl = -4.0 * s
    du = -4.0 * s
    d = 1.0 + 6.0 * s
    return tuple(a.astype(dtype) for a in (ds, dl, d, du, dw))


def varlen_offsets(nsys: int, seed: int, lo: int = 32, hi: int = 8192) -> np.ndarray:
    """int64 offsets (nsys+1) for lengths round(exp(U(ln lo, ln hi)))."""
    rng = np.random.default_rng(seed)
    lens = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), nsys))).astype(np.int64)
    lens = np.clip(lens, lo, hi)
    off = np.zeros(nsys + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    return off
