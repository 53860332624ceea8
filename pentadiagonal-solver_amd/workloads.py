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
* ``hyperdiffusion``: implicit fourth-order diffusion, one operator per
  direction: along every grid line of direction d the system is
  I + s*[1,-4,6,-4,1] with s = nu*dt/h^4 and nu ~ U(0.5, 2) drawn once per
  line (seed 7 by default; the initial field is N(0,1), seed 8).  The stencil
  is truncated at the ends of a line (values beyond the boundary are zero,
  i.e. the out-of-range couplings are zeroed), so each line's matrix is
  symmetric positive definite with eigenvalues >= 1 and every direction
  solve is a contraction in the 2-norm.  Default dt/h^4 = 0.25.
* ``varlen_offsets``: lengths round(exp(U(ln lo, ln hi))) packed contiguously.
* ``cavity``: inputs of the 2-D cavity consumer's pressure-correction step
  (reference apps/2D_cavity_cpu.cpp:183-230): face velocities u' ~ U(-1, 1)
  on (ny+2) x (nx+1), v' ~ U(-1, 1) on (ny+1) x (nx+2), momentum diagonal
  uv = A_p ~ U(1, 3) on (ny+2) x (nx+2) (positive, as the momentum
  assembly's sum of couplings makes it), all row-major; dx = 1/nx,
  dy = 1/ny, alphaU = 0.7 (the reference's velocity under-relaxation,
  2D_cavity_cpu.cpp:123).  Draw order: u', v', uv.

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


HYPERDIFFUSION_DT_OVER_H4 = 0.25


def _line_view(dims, d):
    """numpy/torch shape that broadcasts a per-line value along axis d
    (arrays are stored dims[0]-fastest, i.e. numpy order reversed)."""
    nd = len(dims)
    view = [int(v) for v in reversed(dims)]
    view[nd - 1 - d] = 1
    return view


def hyperdiffusion_np(dims: Sequence[int], seed: int = 7,
                      dt_over_h4: float = HYPERDIFFUSION_DT_OVER_H4, dtype=np.float64):
    """Per-direction coefficient sets [(ds, dl, d, du, dw) for d in axes],
    flat host arrays in the standard layout."""
    rng = np.random.default_rng(seed)
    sets = []
    nd = len(dims)
    shape = tuple(int(v) for v in reversed(dims))
    for ax in range(nd):
        nu = rng.uniform(0.5, 2.0, _line_view(dims, ax))
        s = np.broadcast_to(nu * dt_over_h4, shape)
        i = np.arange(dims[ax]).reshape([-1 if q == nd - 1 - ax else 1 for q in range(nd)])
        n = dims[ax]
        ds = np.where(i >= 2, s, 0.0)
        dl = np.where(i >= 1, -4.0 * s, 0.0)
        dd = 1.0 + 6.0 * s
        du = np.where(i + 1 < n, -4.0 * s, 0.0)
        dw = np.where(i + 2 < n, s, 0.0)
        sets.append(tuple(np.ascontiguousarray(np.broadcast_to(a, shape), dtype=dtype).reshape(-1)
                          for a in (ds, dl, dd, du, dw)))
    return sets


def hyperdiffusion_torch(dims: Sequence[int], seed: int = 7,
                         dt_over_h4: float = HYPERDIFFUSION_DT_OVER_H4, dtype=None,
                         device="cuda"):
    """hyperdiffusion_np's recipe generated on the GPU (torch generator)."""
    import torch
    dtype = dtype or torch.float64
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    nd = len(dims)
    shape = tuple(int(v) for v in reversed(dims))
    sets = []
    for ax in range(nd):
        nu = torch.empty(_line_view(dims, ax), dtype=torch.float64, device=device)
        nu.uniform_(0.5, 2.0, generator=g)
        s = (nu * dt_over_h4).expand(shape)
        idx = [-1 if q == nd - 1 - ax else 1 for q in range(nd)]
        i = torch.arange(dims[ax], device=device).reshape(idx)
        n = dims[ax]
        zero = torch.zeros((), dtype=torch.float64, device=device)
        ds = torch.where(i >= 2, s, zero)
        dl = torch.where(i >= 1, -4.0 * s, zero)
        dd = 1.0 + 6.0 * s
        du = torch.where(i + 1 < n, -4.0 * s, zero)
        dw = torch.where(i + 2 < n, s, zero)
        sets.append(tuple(a.expand(shape).contiguous().reshape(-1).to(dtype)
                          for a in (ds, dl, dd, du, dw)))
    return sets


def hyperdiffusion_scales_torch(dims: Sequence[int], seed: int = 7,
                                dt_over_h4: float = HYPERDIFFUSION_DT_OVER_H4, dtype=None,
                                device="cuda"):
    """The per-line scales s = nu*dt/h^4 of hyperdiffusion_torch (same
    generator draws, so the same operators), one flat tensor per direction
    over the remaining axes in the standard order -- the input of the
    structured-coefficient solves (weights 1, -4, 6, -4, 1, identity 1)."""
    import torch
    dtype = dtype or torch.float64
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    out = []
    for ax in range(len(dims)):
        nu = torch.empty(_line_view(dims, ax), dtype=torch.float64, device=device)
        nu.uniform_(0.5, 2.0, generator=g)
        out.append((nu * dt_over_h4).reshape(-1).to(dtype).contiguous())
    return out


HYPERDIFFUSION_WEIGHTS = (1.0, -4.0, 6.0, -4.0, 1.0)


def gaussian_field_torch(dims: Sequence[int], seed: int = 8, dtype=None, device="cuda"):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    x = torch.empty(n_elems(dims), dtype=torch.float64, device=device)
    x.normal_(0.0, 1.0, generator=g)
    return x.to(dtype or torch.float64)


def varlen_offsets(nsys: int, seed: int, lo: int = 32, hi: int = 8192) -> np.ndarray:
    """int64 offsets (nsys+1) for lengths round(exp(U(ln lo, ln hi)))."""
    rng = np.random.default_rng(seed)
    lens = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), nsys))).astype(np.int64)
    lens = np.clip(lens, lo, hi)
    off = np.zeros(nsys + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    return off


CAVITY_ALPHA_U = 0.7


def cavity_np(nx: int, ny: int, seed: int = 9):
    """(u_prime, v_prime, uv, alphaU, dx, dy) of the cavity recipe, float64."""
    rng = np.random.default_rng(seed)
    up = rng.uniform(-1, 1, (ny + 2) * (nx + 1))
    vp = rng.uniform(-1, 1, (ny + 1) * (nx + 2))
    uv = rng.uniform(1, 3, (ny + 2) * (nx + 2))
    return up, vp, uv, CAVITY_ALPHA_U, 1.0 / nx, 1.0 / ny


def cavity_torch(nx: int, ny: int, seed: int = 9, device="cuda"):
    """The same recipe drawn on the device (torch's generator)."""
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    f64 = torch.float64
    up = torch.rand((ny + 2) * (nx + 1), dtype=f64, device=device, generator=g) * 2 - 1
    vp = torch.rand((ny + 1) * (nx + 2), dtype=f64, device=device, generator=g) * 2 - 1
    uv = torch.rand((ny + 2) * (nx + 2), dtype=f64, device=device, generator=g) * 2 + 1
    return up, vp, uv, CAVITY_ALPHA_U, 1.0 / nx, 1.0 / ny
