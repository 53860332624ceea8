"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
only as the checker (or the timed CPU baseline).  The product library never
imports it.

* ``liboracle.so``: the C restatement in pentad_oracle.c (Thomas recurrence of
  src/cpu/singlenode/pentad_orig_cpu.cpp:73-122 with the routing of :158-212,
  a long-double variant, banded and dense LU with partial pivoting, residuals).
* ``_ref/libref_avx.so`` / ``_ref/libref_orig.so``: the reference's own CPU
  solvers (src/cpu/singlenode/pentad_cpu.cpp, pentad_orig_cpu.cpp) compiled
  from its sources by oracle/Makefile, when they were built.
* ``read_fixture``: reader of the reference fixture text format
  (include/util/mesh.hpp:50-80).
* ``embed_padded`` / ``extract_padded``: the padded array layout of the
  reference's CPU library (pentad_cpu.cpp:611-613).
* ``cavity_pressure_coef``: the 2-D cavity consumer's assembly of the
  pressure-correction systems (apps/preproc.hpp:248-525); with
  ``_ref/libref_cavity.so``, the reference's own code for it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_AVX = os.path.join(HERE, "_ref", "libref_avx.so")
REF_ORIG = os.path.join(HERE, "_ref", "libref_orig.so")
REF_CAVITY = os.path.join(HERE, "_ref", "libref_cavity.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_lib = None
_ref = {}


def build() -> None:
    """Compile the oracle (and _ref when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = ctypes.CDLL(LIB)
        dbl = [_P] * 6 + [_P, _I, _I]
        l.po_gpsv_batch_d.argtypes = dbl
        l.po_gpsv_batch_s.argtypes = dbl
        l.po_gpsv_batch_d.restype = _I
        l.po_gpsv_batch_s.restype = _I
        l.po_gpsv_batch_ld.argtypes = [_P] * 7 + [_P, _I, _I]
        l.po_gpsv_batch_ld.restype = _I
        l.po_gpsv_batch_pivot_d.argtypes = [_P] * 7 + [_P, _I, _I, _I]
        l.po_gpsv_batch_pivot_d.restype = _I
        l.po_residual_d.argtypes = [_P] * 7 + [_P, _I, _I]
        l.po_residual_d.restype = ctypes.c_double
        l.po_varlen_d.argtypes = [_P] * 6 + [_P, ctypes.c_int64]
        l.po_varlen_d.restype = _I
        l.po_varlen_ld.argtypes = [_P] * 7 + [_P, ctypes.c_int64]
        l.po_varlen_ld.restype = _I
        l.po_cavity_pressure_coef_d.argtypes = ([_I, _I, _I] + [ctypes.c_double] * 3
                                                + [_P] * 14)
        l.po_cavity_pressure_coef_d.restype = _I
        _lib = l
    return _lib


def _c(a: np.ndarray, dtype) -> np.ndarray:
    if a.dtype != dtype or not a.flags["C_CONTIGUOUS"]:
        raise TypeError(f"expected C-contiguous {dtype}")
    return a


def _dims(dims: Sequence[int]):
    return (ctypes.c_int * len(dims))(*[int(v) for v in dims])


def _pp(*arrs):
    return [a.ctypes.data_as(_P) for a in arrs]


def gpsv_batch(ds, dl, d, du, dw, x, dims, solvedim) -> None:
    """In-place Thomas solve along solvedim (float64 or float32 arrays)."""
    dt = x.dtype
    for a in (ds, dl, d, du, dw, x):
        _c(a, dt)
    fn = lib().po_gpsv_batch_d if dt == np.float64 else lib().po_gpsv_batch_s
    rc = fn(*_pp(ds, dl, d, du, dw, x), _dims(dims), len(dims), int(solvedim))
    if rc != 0:
        raise ValueError(f"oracle rc={rc}")


def solve(ds, dl, d, du, dw, b, dims, solvedim) -> np.ndarray:
    x = b.copy()
    gpsv_batch(ds, dl, d, du, dw, x, dims, solvedim)
    return x


def solve_ld(ds, dl, d, du, dw, b, dims, solvedim) -> np.ndarray:
    """Long-double Thomas (float64 in/out): forward-error reference."""
    arrs = [_c(np.ascontiguousarray(a, np.float64), np.float64)
            for a in (ds, dl, d, du, dw, b)]
    out = np.empty_like(arrs[5])
    rc = lib().po_gpsv_batch_ld(*_pp(*arrs, out), _dims(dims), len(dims), int(solvedim))
    if rc != 0:
        raise ValueError(f"oracle rc={rc}")
    return out


def solve_pivot(ds, dl, d, du, dw, b, dims, solvedim, method: str = "banded") -> np.ndarray:
    """Partial-pivoting solve: method 'banded' (gbsv style) or 'dense'."""
    arrs = [_c(np.ascontiguousarray(a, np.float64), np.float64)
            for a in (ds, dl, d, du, dw, b)]
    out = np.empty_like(arrs[5])
    rc = lib().po_gpsv_batch_pivot_d(*_pp(*arrs, out), _dims(dims), len(dims),
                                     int(solvedim), 0 if method == "banded" else 1)
    if rc != 0:
        raise ValueError(f"oracle pivot rc={rc}")
    return out


def residual(ds, dl, d, du, dw, x, b, dims, solvedim) -> float:
    """max |A x - b| (long double accumulation)."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (ds, dl, d, du, dw, x, b)]
    return float(lib().po_residual_d(*_pp(*arrs), _dims(dims), len(dims), int(solvedim)))


def varlen(ds, dl, d, du, dw, b, offsets, use_ld: bool = False) -> np.ndarray:
    arrs = [np.ascontiguousarray(a, np.float64) for a in (ds, dl, d, du, dw, b)]
    off = np.ascontiguousarray(offsets, np.int64)
    nsys = len(off) - 1
    if use_ld:
        out = np.empty_like(arrs[5])
        rc = lib().po_varlen_ld(*_pp(*arrs, out), off.ctypes.data_as(_P), nsys)
    else:
        out = arrs[5].copy()
        rc = lib().po_varlen_d(*_pp(*arrs[:5], out), off.ctypes.data_as(_P), nsys)
    if rc != 0:
        raise ValueError(f"oracle varlen rc={rc}")
    return out


# ---------------------------------------------------------------------------
# reference CPU solver built from its own sources (oracle/_ref)
# ---------------------------------------------------------------------------
def ref_available(kind: str = "avx") -> bool:
    return os.path.exists(REF_AVX if kind == "avx" else REF_ORIG)


def ref_lib(kind: str = "avx") -> ctypes.CDLL:
    if kind not in _ref:
        l = ctypes.CDLL(REF_AVX if kind == "avx" else REF_ORIG)
        for nm in ("ref_gpsv_batch_d", "ref_gpsv_batch_s"):
            fn = getattr(l, nm)
            fn.argtypes = [_P] * 6 + [_P, _I, _I]
            fn.restype = _I
        _ref[kind] = l
    return _ref[kind]


def ref_set_threads(n: int, kind: str = "avx") -> None:
    """OpenMP threads of the reference solver's parallel loops."""
    ref_lib(kind).ref_set_threads(int(n))


def ref_threads(kind: str = "avx") -> int:
    return int(ref_lib(kind).ref_max_threads())


def ref_proc_bind(kind: str = "avx") -> str:
    names = {0: "false", 1: "true", 2: "primary", 3: "close", 4: "spread"}
    return names.get(int(ref_lib(kind).ref_proc_bind()), "unknown")


def ref_gpsv_batch(ds, dl, d, du, dw, x, dims, solvedim, kind: str = "avx") -> None:
    """In-place solve by the reference's own CPU library.

    The AVX path needs dims[0] to be a multiple of its SIMD width for 2-D/3-D
    x-solves (reference pentad_cpu.cpp:611-613) and len <= 1024 (:11).
    """
    dt = x.dtype
    for a in (ds, dl, d, du, dw, x):
        _c(a, dt)
    l = ref_lib(kind)
    fn = l.ref_gpsv_batch_d if dt == np.float64 else l.ref_gpsv_batch_s
    fn(*_pp(ds, dl, d, du, dw, x), _dims(dims), len(dims), int(solvedim))


# ---------------------------------------------------------------------------
# 2-D cavity consumer: pressure-correction assembly (reference
# apps/preproc.hpp:248-525, pressure_coef_matrix)
# ---------------------------------------------------------------------------
CAVITY_OUT = ("x", "s_x", "l_x", "d_x", "u_x", "w_x", "s_y", "l_y", "d_y", "u_y", "w_y")


def _cavity_call(fn, nx, ny, nx_pad, alphaU, dx, dy, u_prime, v_prime, uv, out):
    ins = [np.ascontiguousarray(a, np.float64) for a in (u_prime, v_prime, uv)]
    if ins[0].size != (ny + 2) * (nx + 1) or ins[1].size != (ny + 1) * (nx + 2) \
            or ins[2].size != (ny + 2) * (nx + 2):
        raise ValueError("u_prime (ny+2)x(nx+1), v_prime (ny+1)x(nx+2), uv (ny+2)x(nx+2)")
    for k in CAVITY_OUT:
        _c(out[k], np.float64)
        if out[k].size < nx_pad * ny:
            raise ValueError(f"{k}: nx_pad * ny elements needed")
    return fn(int(nx), int(ny), int(nx_pad), float(alphaU), float(dx), float(dy),
              *_pp(*ins), *_pp(*[out[k] for k in CAVITY_OUT]))


def cavity_outputs(nx_pad, ny, fill=0.0):
    return {k: np.full(nx_pad * ny, fill, dtype=np.float64) for k in CAVITY_OUT}


def cavity_pressure_coef(nx, ny, nx_pad, alphaU, dx, dy, u_prime, v_prime, uv, out=None):
    """po_cavity_pressure_coef_d: the reference's assembly, its write order and
    index expressions included.  ``out`` (dict of CAVITY_OUT arrays of at
    least nx_pad * ny elements) is written in place; elements the assembly
    does not write keep their values."""
    out = cavity_outputs(nx_pad, ny) if out is None else out
    rc = _cavity_call(lib().po_cavity_pressure_coef_d, nx, ny, nx_pad, alphaU, dx, dy,
                      u_prime, v_prime, uv, out)
    if rc != 0:
        raise ValueError(f"oracle rc={rc}")
    return out


def ref_cavity_available() -> bool:
    return os.path.exists(REF_CAVITY)


def ref_cavity_pressure_coef(nx, ny, nx_pad, alphaU, dx, dy, u_prime, v_prime, uv, out=None):
    """The reference's own pressure_coef_matrix (oracle/_ref/libref_cavity.so)."""
    if "cavity" not in _ref:
        l = ctypes.CDLL(REF_CAVITY)
        l.ref_cavity_pressure_coef.argtypes = ([_I, _I, _I] + [ctypes.c_double] * 3
                                               + [_P] * 14)
        l.ref_cavity_pressure_coef.restype = _I
        _ref["cavity"] = l
    out = cavity_outputs(nx_pad, ny) if out is None else out
    _cavity_call(_ref["cavity"].ref_cavity_pressure_coef, nx, ny, nx_pad, alphaU, dx, dy,
                 u_prime, v_prime, uv, out)
    return out


# ---------------------------------------------------------------------------
# padded layout (reference src/cpu/singlenode/pentad_cpu.cpp:611-613: element
# (i, j, k) at i + pads[0] * (j + pads[1] * k); the last axis is not padded)
# ---------------------------------------------------------------------------
def ref_simd_vec(dtype) -> int:
    """Elements per SIMD vector of the compiled reference AVX library
    (include/pentad_simd.hpp:20-28 under -march=core-avx2: 32 bytes)."""
    return 32 // np.dtype(dtype).itemsize


def padded_extent(dims: Sequence[int], pads: Sequence[int]) -> int:
    n = int(dims[-1])
    for p in pads[:-1]:
        n *= int(p)
    return n


def _padded_shape(dims, pads):
    # numpy shape, slowest axis first; the last axis has its true extent
    shp = [int(p) for p in pads[:-1]] + [int(dims[-1])]
    return tuple(reversed(shp))


def embed_padded(a: np.ndarray, dims: Sequence[int], pads: Sequence[int],
                 fill: float) -> np.ndarray:
    """Dense flat array (dims[0] fastest) -> flat padded array, padding = fill."""
    out = np.full(_padded_shape(dims, pads), fill, dtype=a.dtype)
    view = out[tuple(slice(0, int(v)) for v in reversed(dims))]
    view[...] = a.reshape(tuple(int(v) for v in reversed(dims)))
    return out.reshape(-1)


def extract_padded(p: np.ndarray, dims: Sequence[int], pads: Sequence[int]) -> np.ndarray:
    """Flat padded array -> dense flat array."""
    v = p.reshape(_padded_shape(dims, pads))[tuple(slice(0, int(d)) for d in reversed(dims))]
    return np.ascontiguousarray(v).reshape(-1)


def padding_mask(dims: Sequence[int], pads: Sequence[int]) -> np.ndarray:
    """True at the flat positions of a padded array that hold no element."""
    m = np.ones(_padded_shape(dims, pads), dtype=bool)
    m[tuple(slice(0, int(v)) for v in reversed(dims))] = False
    return m.reshape(-1)


# ---------------------------------------------------------------------------
# fixture reader (reference include/util/mesh.hpp:50-80)
# ---------------------------------------------------------------------------
def read_fixture(path: str, dtype=np.float64) -> Dict[str, object]:
    """Returns dims, solvedim, ds, dl, d, du, dw, x (RHS) and u (solution).

    Values are parsed as float64 and cast to dtype; for float32 the expected
    solution is the 6-decimal FP32 line (the reference skips the FP64 line,
    mesh.hpp:72-79).
    """
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    ndims, solvedim = (int(v) for v in lines[0].split())
    dims = [int(v) for v in lines[1].split()]
    assert len(dims) == ndims
    n = int(np.prod(dims))

    def arr(i):
        a = np.array(lines[i].split(), dtype=np.float64)
        assert a.size == n, (path, i, a.size, n)
        return a.astype(dtype)

    keys = ["ds", "dl", "d", "du", "dw", "x"]
    out = {k: arr(2 + j) for j, k in enumerate(keys)}
    out["u"] = arr(8) if dtype == np.float64 else arr(9)
    out["dims"] = dims
    out["solvedim"] = solvedim
    return out


def allclose_reference_rule(expected, actual, dtype) -> bool:
    """Reference test rule (test/catch_utils.hpp:19-25): every element within
    rel 1e-13 or abs 1e-15 (FP64), rel 1e-5 or abs 1e-6 (FP32)."""
    e = np.asarray(expected, np.float64)
    a = np.asarray(actual, np.float64)
    if dtype == np.float64:
        rel, ab = 1e-13, 1e-15
    else:
        rel, ab = 1e-5, 1e-6
    diff = np.abs(e - a)
    ok = (diff <= ab) | (diff <= rel * np.maximum(np.abs(e), np.abs(a)))
    return bool(np.all(ok))
