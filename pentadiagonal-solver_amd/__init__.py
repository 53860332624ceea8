"""Python mirror of the pentadsolver C ABI (libpentadsolver.so, gfx950).

This module is the host-side counterpart of the reference's public interface
(include/pentadsolver.hpp:36-131 of SidhantKrAgrawal/pentadiagonal-solver):
``gpsv_batch`` solves every pentadiagonal system along ``solvedim`` in place,
``gpsv_batch_buffer_size_ext`` reports the scratch bytes, ``Handle`` wraps
``pentadsolver_create`` / ``pentadsolver_destroy``, and ``kernel_that_ran``
returns the executed-kernel record.  Device memory comes from torch tensors
(the arrays' ``data_ptr``); torch is only plumbing, every solve runs in the
HIP library.

The library must be built (``make -C pentadiagonal-solver_amd``); importing
this module does not build it and ``load_library`` raises when it is missing,
so a missing extension can never be mistaken for a result.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libpentadsolver.so")

_lib = None

# C entry points declared in include/pentadsolver.h: name -> (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_SZ = ctypes.c_size_t
_IP = ctypes.POINTER(ctypes.c_int)

_SIGNATURES = {
    "pentadsolver_create": (None, [ctypes.POINTER(_P), _P, _I, _IP]),
    "pentadsolver_destroy": (None, [ctypes.POINTER(_P)]),
    "pentadsolver_D_gpsv_batch_buffer_size_ext": (_SZ, [_P] * 7 + [_IP, _I, _I]),
    "pentadsolver_S_gpsv_batch_buffer_size_ext": (_SZ, [_P] * 7 + [_IP, _I, _I]),
    "pentadsolver_D_gpsv_batch": (None, [_P] * 7 + [_IP, _I, _I, _P]),
    "pentadsolver_S_gpsv_batch": (None, [_P] * 7 + [_IP, _I, _I, _P]),
    "pentadsolver_D_gpsv_batch_padded_buffer_size": (_SZ, [_P] * 7 + [_IP, _IP, _I, _I]),
    "pentadsolver_S_gpsv_batch_padded_buffer_size": (_SZ, [_P] * 7 + [_IP, _IP, _I, _I]),
    "pentadsolver_D_gpsv_batch_padded": (_I, [_P] * 7 + [_IP, _IP, _I, _I, _P]),
    "pentadsolver_S_gpsv_batch_padded": (_I, [_P] * 7 + [_IP, _IP, _I, _I, _P]),
    "pentadsolver_D_cavity_pressure_coef": (_I, [_P, _I, _I, _I] + [ctypes.c_double] * 3 + [_P] * 14),
    "pentadsolver_D_cavity_pressure_solve": (_I, [_P, _I, _I, _I] + [ctypes.c_double] * 3 + [_P] * 4),
    "pentadsolver_varlen_set_spread": (_I, [_P, _IP]),
    "pentadsolver_varlen_get_spread": (_I, [_P, _IP]),
    "pentadsolver_varlen_set_refine": (_I, [_P, _I]),
    "pentadsolver_varlen_get_refine": (_I, [_P]),
    "pentadsolver_D_varlen_autotune": (_I, [_P] * 7 + [_I, ctypes.c_char_p, _SZ]),
    "pentadsolver_S_varlen_autotune": (_I, [_P] * 7 + [_I, ctypes.c_char_p, _SZ]),
    "pentadsolver_DS_varlen_autotune_ir": (_I, [_P] * 7 + [_I, ctypes.c_char_p, _SZ]),
    "pentadsolver_kernel_that_ran": (ctypes.c_char_p, [_I]),
    "pentadsolver_get_last_status": (_I, [_P]),
    "pentadsolver_set_stream": (_I, [_P, _P]),
    "pentadsolver_set_algo": (_I, [_P, _I, ctypes.c_char_p]),
    "pentadsolver_freeze_env": (_I, [_P, _I]),
    "pentadsolver_kernel_config": (ctypes.c_char_p, [_P, _I]),
    "pentadsolver_set_kernel_params": (_I, [_P, _I, _I, _I]),
    "pentadsolver_DS_gpsv_batch_ir_buffer_size_ext": (_SZ, [_P] * 7 + [_IP, _I, _I]),
    "pentadsolver_DS_gpsv_batch_ir": (None, [_P] * 7 + [_IP, _I, _I, _I, ctypes.c_double, _P, _P]),
    "pentadsolver_D_autotune": (_I, [_P] * 7 + [_IP, _I, _I, _P, _SZ, _I, ctypes.c_char_p, _SZ]),
    "pentadsolver_S_autotune": (_I, [_P] * 7 + [_IP, _I, _I, _P, _SZ, _I, ctypes.c_char_p, _SZ]),
    "pentadsolver_DS_autotune_ir": (_I, [_P] * 7 + [_IP, _I, _I, _I, ctypes.c_double, _P, _SZ, _I, ctypes.c_char_p, _SZ]),
    "pentadsolver_clear_tuning": (None, [_P]),
    "pentadsolver_probe_stream": (_I, [_P, ctypes.POINTER(_P), _P, _SZ, _I, _I]),
    "pentadsolver_probe_ir_mix": (_I, [_P, _I, ctypes.POINTER(_P), _P, _P, _SZ]),
    "pentadsolver_probe_ir_layout": (_I, [_P, _I, _I, _P, _P, _P, _SZ]),
    "pentadsolver_probe_strided_io": (_I, [_P, _I, ctypes.POINTER(_P), _P, _IP, _I]),
    "pentadsolver_adi_create": (_I, [_P, ctypes.POINTER(_P), _I, _I, _IP, ctypes.POINTER(_P), _I, _I, _I, ctypes.c_double]),
    "pentadsolver_adi_create_stencil": (_I, [_P, ctypes.POINTER(_P), _I, _I, _IP, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "pentadsolver_D_gpsv_stencil_batch": (_I, [_P, _P, ctypes.POINTER(ctypes.c_double), ctypes.c_double, _P, _IP, _I, _I]),
    "pentadsolver_S_gpsv_stencil_batch": (_I, [_P, _P, ctypes.POINTER(ctypes.c_double), ctypes.c_double, _P, _IP, _I, _I]),
    "pentadsolver_adi_step": (_I, [_P, _P, _I]),
    "pentadsolver_adi_graph_capture": (_I, [_P, _P, _I]),
    "pentadsolver_adi_graph_launch": (_I, [_P, _I]),
    "pentadsolver_adi_refine_steps": (_I, [_P]),
    "pentadsolver_adi_set_refine": (_I, [_P, _I, ctypes.c_double]),
    "pentadsolver_adi_direction_ms": (_I, [_P, _P, _I, ctypes.POINTER(ctypes.c_float)]),
    "pentadsolver_adi_config": (ctypes.c_char_p, [_P, _I]),
    "pentadsolver_adi_destroy": (None, [ctypes.POINTER(_P)]),
    "pentadsolver_varlen_create": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int64), ctypes.c_int64]),
    "pentadsolver_varlen_destroy": (None, [ctypes.POINTER(_P)]),
    "pentadsolver_varlen_items": (_I, [_P]),
    "pentadsolver_D_gpsv_varlen_batch": (_I, [_P] * 7),
    "pentadsolver_S_gpsv_varlen_batch": (_I, [_P] * 7),
    "pentadsolver_DS_gpsv_varlen_batch_ir": (_I, [_P] * 7 + [_I, ctypes.c_double, _P]),
    "pentadsolver_D_gpsv_batch_noalloc": (_I, [_P] * 7 + [_IP, _I, _I]),
    "pentadsolver_D_dist_forward": (_I, [_P] * 7 + [_IP, _IP, _I, _I, _I, _I, _P]),
    "pentadsolver_S_dist_forward": (_I, [_P] * 7 + [_IP, _IP, _I, _I, _I, _I, _P]),
    "pentadsolver_D_dist_reduced": (_I, [_P, _P, ctypes.c_int64, _I, _P]),
    "pentadsolver_S_dist_reduced": (_I, [_P, _P, ctypes.c_int64, _I, _P]),
    "pentadsolver_D_dist_backward_buffer_size": (_SZ, [_P] * 7 + [_IP, _IP, _I, _I]),
    "pentadsolver_S_dist_backward_buffer_size": (_SZ, [_P] * 7 + [_IP, _IP, _I, _I]),
    "pentadsolver_D_dist_backward": (_I, [_P] * 7 + [_IP, _IP, _I, _I, _I, _I, _P, _P]),
    "pentadsolver_S_dist_backward": (_I, [_P] * 7 + [_IP, _IP, _I, _I, _I, _I, _P, _P]),
}

STATUS_NAMES = {
    0: "SUCCESS", 1: "NOT_INITIALIZED", 2: "ALLOC_FAILED", 3: "INVALID_VALUE",
    4: "ARCH_MISMATCH", 5: "MAPPING_ERROR", 6: "EXECUTION_FAILED",
    7: "INTERNAL_ERROR", 8: "MATRIX_TYPE_NOT_SUPPORTED", 9: "ZERO_PIVOT",
}


class PentadError(RuntimeError):
    """A library call reported a non-success pentadStatus_t."""


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libpentadsolver.so and bind the declared signatures.

    Raises FileNotFoundError when the library has not been built.
    """
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `make -C {PKG_DIR}` "
            "(no CPU fallback exists)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def exported_symbols() -> Sequence[str]:
    return tuple(_SIGNATURES)


def _dims_arg(dims: Sequence[int]):
    arr = (ctypes.c_int * len(dims))(*[int(v) for v in dims])
    return arr


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def _prefix(dtype) -> str:
    import torch
    if dtype == torch.float64:
        return "D"
    if dtype == torch.float32:
        return "S"
    raise TypeError(f"unsupported dtype {dtype}: float32 or float64")


class Handle:
    """pentadsolver_create / pentadsolver_destroy as a context manager."""

    def __init__(self, stream=None):
        lib = load_library()
        self._h = _P()
        lib.pentadsolver_create(ctypes.byref(self._h), None, 0, None)
        if stream is not None:
            self.set_stream(stream)

    @property
    def value(self):
        return self._h

    def set_stream(self, stream) -> None:
        """stream: a torch.cuda.Stream (or raw hipStream_t integer)."""
        raw = getattr(stream, "cuda_stream", stream)
        _check(load_library().pentadsolver_set_stream(self._h, _P(int(raw))))

    def set_algo(self, direction: int, algo: Optional[str]) -> None:
        enc = None if algo is None else algo.encode()
        _check(load_library().pentadsolver_set_algo(self._h, direction, enc))

    def freeze_env(self, frozen: bool = True) -> None:
        _check(load_library().pentadsolver_freeze_env(self._h, int(frozen)))

    def set_kernel_params(self, direction: int, rows: int, width: int) -> None:
        _check(load_library().pentadsolver_set_kernel_params(
            self._h, direction, rows, width))

    def kernel_config(self, direction: int) -> str:
        return load_library().pentadsolver_kernel_config(self._h, direction).decode()

    def last_status(self) -> int:
        return load_library().pentadsolver_get_last_status(self._h)

    def close(self) -> None:
        if self._h:
            load_library().pentadsolver_destroy(ctypes.byref(self._h))
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _hval(handle) -> Optional[int]:
    if handle is None:
        return None
    return handle.value if isinstance(handle, Handle) else handle


def _check(status: int) -> None:
    if status != 0:
        raise PentadError(f"pentadsolver status {STATUS_NAMES.get(status, status)}")


def gpsv_batch_buffer_size_ext(ds, dl, d, du, dw, x, dims: Sequence[int],
                               solvedim: int, handle=None) -> int:
    """Bytes of scratch the solve needs (reference include/pentadsolver.hpp:64-73)."""
    lib = load_library()
    fn = getattr(lib, f"pentadsolver_{_prefix(x.dtype)}_gpsv_batch_buffer_size_ext")
    return int(fn(_hval(handle), _ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw),
                  _ptr(x), _dims_arg(dims), len(dims), int(solvedim)))


def gpsv_batch(ds, dl, d, du, dw, x, dims: Sequence[int], solvedim: int,
               buffer=None, handle=None, check: bool = True) -> None:
    """Solve in place along ``solvedim`` (reference include/pentadsolver.hpp:107-118).

    All arrays are device tensors of one dtype with prod(dims) elements;
    dims[0] is the fastest dimension.  ``buffer`` is a device tensor of at
    least ``gpsv_batch_buffer_size_ext`` bytes, or None.  Asynchronous on the
    handle's stream.  Raises PentadError when the library reports an error.
    """
    lib = load_library()
    for t in (ds, dl, d, du, dw):
        if t.dtype != x.dtype:
            raise TypeError("all arrays must share one dtype")
    fn = getattr(lib, f"pentadsolver_{_prefix(x.dtype)}_gpsv_batch")
    fn(_hval(handle), _ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw), _ptr(x),
       _dims_arg(dims), len(dims), int(solvedim), _ptr(buffer))
    if check:
        _check(lib.pentadsolver_get_last_status(_hval(handle)))


def solve(ds, dl, d, du, dw, x, dims: Sequence[int], solvedim: int, handle=None):
    """gpsv_batch with the scratch buffer allocated as needed (convenience)."""
    import torch
    nbytes = gpsv_batch_buffer_size_ext(ds, dl, d, du, dw, x, dims, solvedim, handle)
    buf = None
    if nbytes:
        buf = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
    gpsv_batch(ds, dl, d, du, dw, x, dims, solvedim, buf, handle)
    return buf


def padded_extent(dims: Sequence[int], pads: Sequence[int]) -> int:
    """Elements a padded array of extents ``pads`` (>= dims) occupies: the
    last axis needs only dims[-1] (its padding moves no element)."""
    n = int(dims[-1])
    for p in pads[:-1]:
        n *= int(p)
    return n


def gpsv_batch_padded_buffer_size(ds, dl, d, du, dw, x, dims: Sequence[int],
                                  pads: Sequence[int], solvedim: int, handle=None) -> int:
    """Scratch bytes of the padded-layout solve (0 for the on-chip kernels)."""
    if len(pads) != len(dims):
        raise ValueError("pads must have one entry per dimension")
    fn = getattr(load_library(), f"pentadsolver_{_prefix(x.dtype)}_gpsv_batch_padded_buffer_size")
    return int(fn(_hval(handle), _ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw), _ptr(x),
                  _dims_arg(dims), _dims_arg(pads), len(dims), int(solvedim)))


def gpsv_batch_padded(ds, dl, d, du, dw, x, dims: Sequence[int], pads: Sequence[int],
                      solvedim: int, buffer=None, handle=None) -> None:
    """Solve in place along ``solvedim`` over padded arrays: element
    (i, j, k) at i + pads[0] * (j + pads[1] * k), pads[k] >= dims[k] -- the
    layout of the reference's CPU library (src/cpu/singlenode/
    pentad_cpu.cpp:611-613) and its 2-D consumer
    (apps/2D_cavity_cpu.cpp:70-72).  Every array holds at least
    ``padded_extent(dims, pads)`` elements; the padding is neither read nor
    written.  Raises PentadError on an error status."""
    if len(pads) != len(dims):
        raise ValueError("pads must have one entry per dimension")
    need = padded_extent(dims, pads)
    for t in (ds, dl, d, du, dw, x):
        if t.dtype != x.dtype:
            raise TypeError("all arrays must share one dtype")
        if t.numel() < need:
            raise ValueError(f"array of {t.numel()} elements; the padded layout needs {need}")
    fn = getattr(load_library(), f"pentadsolver_{_prefix(x.dtype)}_gpsv_batch_padded")
    _check(fn(_hval(handle), _ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw), _ptr(x),
              _dims_arg(dims), _dims_arg(pads), len(dims), int(solvedim), _ptr(buffer)))


def _cavity_inputs(nx, ny, nx_pad, u_prime, v_prime, uv, x):
    import torch
    for t, n, name in ((u_prime, (ny + 2) * (nx + 1), "u_prime"),
                       (v_prime, (ny + 1) * (nx + 2), "v_prime"),
                       (uv, (ny + 2) * (nx + 2), "uv"), (x, nx_pad * ny, "x")):
        if t.dtype != torch.float64:
            raise TypeError(f"{name} must be float64")
        if t.numel() < n:
            raise ValueError(f"{name}: {n} elements needed, got {t.numel()}")


def cavity_pressure_coef(nx, ny, nx_pad, alpha_u, dx, dy, u_prime, v_prime, uv, x,
                         xset, yset=None, handle=None) -> None:
    """The 2-D cavity consumer's pressure-correction assembly on the GPU (the
    drop-in for apps/preproc.hpp:248-525, pressure_coef_matrix).

    u_prime (ny+2)x(nx+1), v_prime (ny+1)x(nx+2), uv (ny+2)x(nx+2) row-major
    float64 device tensors; x and the arrays of ``xset`` (and ``yset``, or
    None) = (ds, dl, d, du, dw) hold ny rows of nx_pad elements.  Point
    (i, j) goes to (i-1)*nx_pad + (j-1)."""
    _cavity_inputs(nx, ny, nx_pad, u_prime, v_prime, uv, x)
    arrs = list(xset) + (list(yset) if yset is not None else [None] * 5)
    for t in arrs:
        if t is not None:
            _cavity_inputs(nx, ny, nx_pad, u_prime, v_prime, uv, t)
    _check(load_library().pentadsolver_D_cavity_pressure_coef(
        _hval(handle), int(nx), int(ny), int(nx_pad), float(alpha_u), float(dx), float(dy),
        _ptr(u_prime), _ptr(v_prime), _ptr(uv), _ptr(x), *[_ptr(t) for t in arrs]))


def cavity_pressure_solve(nx, ny, nx_pad, alpha_u, dx, dy, u_prime, v_prime, uv, x,
                          handle=None) -> None:
    """Assembly + x solve + y solve of the cavity consumer's pressure step
    (apps/2D_cavity_cpu.cpp:195-219) with the coefficients generated inside
    the solve kernels; the solution overwrites x (ny rows of nx_pad)."""
    _cavity_inputs(nx, ny, nx_pad, u_prime, v_prime, uv, x)
    _check(load_library().pentadsolver_D_cavity_pressure_solve(
        _hval(handle), int(nx), int(ny), int(nx_pad), float(alpha_u), float(dx), float(dy),
        _ptr(u_prime), _ptr(v_prime), _ptr(uv), _ptr(x)))


def gpsv_batch_ir_buffer_size_ext(ds, dl, d, du, dw, x, dims: Sequence[int],
                                  solvedim: int, handle=None) -> int:
    """Scratch bytes of the mixed-precision solve (0 for the on-chip kernels)."""
    return int(load_library().pentadsolver_DS_gpsv_batch_ir_buffer_size_ext(
        _hval(handle), _ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw), _ptr(x),
        _dims_arg(dims), len(dims), int(solvedim)))


def gpsv_batch_ir(ds, dl, d, du, dw, x, dims: Sequence[int], solvedim: int,
                  max_refine: int = 5, tol: float = 1e-13, steps=None,
                  buffer=None, handle=None, check: bool = True) -> None:
    """FP32 solve + FP64 iterative refinement, in place on the float64 x.

    ds..dw: float32 device tensors; x: float64 device tensor (RHS in,
    solution out); steps: optional int32 device tensor (1 element) that
    receives the largest number of refinement steps taken.
    """
    import torch
    for t in (ds, dl, d, du, dw):
        if t.dtype != torch.float32:
            raise TypeError("coefficients must be float32")
    if x.dtype != torch.float64:
        raise TypeError("x must be float64")
    lib = load_library()
    lib.pentadsolver_DS_gpsv_batch_ir(
        _hval(handle), _ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw), _ptr(x),
        _dims_arg(dims), len(dims), int(solvedim), int(max_refine), float(tol),
        _ptr(steps), _ptr(buffer))
    if check:
        _check(lib.pentadsolver_get_last_status(_hval(handle)))


def solve_ir(ds, dl, d, du, dw, x, dims, solvedim, max_refine=5, tol=1e-13,
             steps=None, handle=None):
    """gpsv_batch_ir with the scratch buffer allocated as needed."""
    import torch
    nbytes = gpsv_batch_ir_buffer_size_ext(ds, dl, d, du, dw, x, dims, solvedim, handle)
    buf = torch.empty(nbytes, dtype=torch.uint8, device=x.device) if nbytes else None
    gpsv_batch_ir(ds, dl, d, du, dw, x, dims, solvedim, max_refine, tol, steps, buf, handle)
    return buf


def autotune(ds, dl, d, du, dw, x, dims, solvedim, handle, reps=3, buffer=None,
             ir=False, max_refine=5, tol=1e-13):
    """Time every configuration for this shape; the handle keeps the fastest.

    Returns the list of (config, ms) the library measured.  x is unchanged.
    """
    lib = load_library()
    rep = ctypes.create_string_buffer(1 << 14)
    nb = 0 if buffer is None else buffer.numel() * buffer.element_size()
    dd = _dims_arg(dims)
    args = [_ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw), _ptr(x), dd, len(dims), int(solvedim)]
    if ir:
        n = lib.pentadsolver_DS_autotune_ir(_hval(handle), *args, int(max_refine), float(tol),
                                            _ptr(buffer), nb, int(reps), rep, len(rep))
    else:
        fn = getattr(lib, f"pentadsolver_{_prefix(x.dtype)}_autotune")
        n = fn(_hval(handle), *args, _ptr(buffer), nb, int(reps), rep, len(rep))
    if n < 0:
        raise PentadError("autotune failed")
    out = []
    for line in rep.value.decode().splitlines():
        cfg, ms, _ = line.rsplit(" ", 2)
        out.append((cfg, float(ms)))
    return out


def probe_stream(ins, out, fmas=0, handle=None):
    """Launch the 6-read/1-write bandwidth probe on device tensors."""
    arr = (_P * 6)(*[_ptr(t) for t in ins])
    _check(load_library().pentadsolver_probe_stream(
        _hval(handle), arr, _ptr(out), out.numel(), out.element_size(), int(fmas)))


def probe_ir_mix(streams, coef, b, x, handle=None):
    """The refinement solve's byte mix (20 B FP32 in + 8 B FP64 in + 8 B FP64
    out per row) as 7, 3 or 2 streams; coef: five float32 tensors (7) or one
    interleaved float32 tensor (3: 5n floats, 2: 7n floats)."""
    ptrs = [_ptr(t) for t in coef] + [0] * (5 - len(coef))
    arr = (_P * 5)(*ptrs)
    _check(load_library().pentadsolver_probe_ir_mix(_hval(handle), int(streams), arr, _ptr(b),
                                                    _ptr(x), x.numel()))


def probe_ir_layout(layout, rounds, coef, b, x, handle=None):
    """The refinement kernel's I/O without its solve: layout 0 (the kernel's
    per-lane loads, coalesced writes in line order) or 1 (the headline
    kernel's workgroup structure and rotated tile store, 512^3 rows only);
    rounds x 8 FP64 FMAs per thread stand in for the solve.  coef: one
    float32 tensor of 5n floats (interleaved per 4 rows)."""
    _check(load_library().pentadsolver_probe_ir_layout(
        _hval(handle), int(layout), int(rounds), _ptr(coef), _ptr(b), _ptr(x), x.numel()))


def probe_strided_io(split, coef, x, dims, solvedim, handle=None):
    """The FP64 strided solve's I/O without the solve, lines of 512 rows along
    y (solvedim 1) or z (2) of a 3-D grid: x += sum of the five float64
    diagonals in place, by workgroups of 16 * split lines x 512 / split
    rows (split 1, 2 or 4)."""
    arr = (_P * 5)(*[_ptr(t) for t in coef])
    d = (ctypes.c_int * 3)(*dims)
    _check(load_library().pentadsolver_probe_strided_io(
        _hval(handle), int(split), arr, _ptr(x), d, int(solvedim)))


ADI_FP64, ADI_FP32, ADI_IR, ADI_SHARED_FP64, ADI_SHARED_FP32 = 0, 1, 2, 3, 4


class Adi:
    """ADI sweep driver (pentadsolver_adi_*): x, then y, then z in place.

    mode: ADI_FP64 (float64 field and coefficients), ADI_FP32 (float32),
    ADI_IR (float32 coefficients solved in FP32 with FP64 refinement on a
    float64 field), ADI_SHARED_FP64 / ADI_SHARED_FP32 (shared factorisation:
    each direction's system 0 is factored once and applied to every line;
    coefficient dtype must match).  coeff_sets: one (ds, dl, d, du, dw)
    tuple of device tensors for every direction, or one tuple per direction;
    float32 or float64 (converted as the mode needs, except shared modes).
    """

    def __init__(self, mode, dims, coeff_sets, handle=None, max_refine=5, tol=1e-13):
        lib = load_library()
        self._a = _P()
        self.ndim = len(dims)
        flat = [t for cs in coeff_sets for t in cs]
        eb = flat[0].element_size()
        if any(t.element_size() != eb for t in flat):
            raise TypeError("coefficient tensors must share one dtype")
        ptrs = (_P * len(flat))(*[_ptr(t) for t in flat])
        self._keep = flat
        _check(lib.pentadsolver_adi_create(_hval(handle), ctypes.byref(self._a), int(mode),
                                           len(dims), _dims_arg(dims), ptrs, len(coeff_sets),
                                           eb, int(max_refine), float(tol)))
        self._keep = None  # the driver holds its own copies

    def step(self, x, nsteps=1):
        _check(load_library().pentadsolver_adi_step(self._a, _ptr(x), int(nsteps)))

    def capture(self, x, nsteps=1):
        """Capture nsteps steps on x as one hipGraph."""
        _check(load_library().pentadsolver_adi_graph_capture(self._a, _ptr(x), int(nsteps)))

    def replay(self, reps=1):
        """Replay the captured steps reps times."""
        _check(load_library().pentadsolver_adi_graph_launch(self._a, int(reps)))

    def direction_ms(self, x, reps=5):
        ms = (ctypes.c_float * 3)()
        _check(load_library().pentadsolver_adi_direction_ms(self._a, _ptr(x), int(reps), ms))
        return [float(v) for v in ms[: self.ndim]]

    def set_refine(self, max_refine: int, tol: float) -> None:
        _check(load_library().pentadsolver_adi_set_refine(self._a, int(max_refine), float(tol)))

    def refine_steps(self) -> int:
        return load_library().pentadsolver_adi_refine_steps(self._a)

    def config(self, direction: int) -> str:
        return load_library().pentadsolver_adi_config(self._a, direction).decode()

    def close(self):
        if self._a:
            load_library().pentadsolver_adi_destroy(ctypes.byref(self._a))
            self._a = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AdiStencil(Adi):
    """ADI driver over structured coefficients (pentadsolver_adi_create_stencil).

    Along every line of direction d the system is
    id_weights[d] * I + s * weights[d] (ds, dl, d - id, du, dw), couplings
    past the ends of the line dropped, with s the line's entry of scales[d]:
    a device tensor with one value per line over the remaining axes in the
    standard order.  float64 scales give the FP64 driver, float32 the FP32
    one (field of the same dtype); refine=True with float64 scales solves
    every line in FP32 with FP64 refinement against the FP64 operator (mode
    2).  weights: ndim sequences of 5 floats (or one for every direction),
    id_weights: ndim floats (or one).
    """

    def __init__(self, dims, scales, weights, id_weights=1.0, handle=None, refine=False):
        lib = load_library()
        self._a = _P()
        self.ndim = len(dims)
        if len(scales) != self.ndim:
            raise ValueError("one scale tensor per direction")
        eb = scales[0].element_size()
        if any(t.element_size() != eb for t in scales) or eb not in (4, 8):
            raise TypeError("scales must all be float32 or all float64")
        w = [list(weights)] * self.ndim if not hasattr(weights[0], "__len__") else \
            [list(v) for v in weights]
        ids = [float(id_weights)] * self.ndim if not hasattr(id_weights, "__len__") else \
            [float(v) for v in id_weights]
        if len(w) != self.ndim or any(len(v) != 5 for v in w) or len(ids) != self.ndim:
            raise ValueError("weights: 5 per direction; id_weights: one per direction")
        wa = (ctypes.c_double * (5 * self.ndim))(*[float(v) for r in w for v in r])
        ia = (ctypes.c_double * self.ndim)(*ids)
        ptrs = (_P * self.ndim)(*[_ptr(t) for t in scales])
        if refine and eb != 8:
            raise TypeError("the refinement mode takes float64 scales")
        mode = 2 if refine else (0 if eb == 8 else 1)
        _check(lib.pentadsolver_adi_create_stencil(_hval(handle), ctypes.byref(self._a),
                                                   mode, self.ndim,
                                                   _dims_arg(dims), ptrs, wa, ia))


def gpsv_stencil_batch(scale, weights, id_weight, x, dims: Sequence[int], solvedim: int,
                       handle=None) -> None:
    """Structured-coefficient solve in place along solvedim
    (pentadsolver_{D,S}_gpsv_stencil_batch): every line's system is
    id_weight * I + scale[line] * weights, couplings past the ends dropped.
    scale: device tensor (x's dtype), one value per line over the remaining
    axes in the standard order; weights: 5 floats."""
    lib = load_library()
    if scale.dtype != x.dtype:
        raise TypeError("scale and x must share one dtype")
    if len(weights) != 5:
        raise ValueError("5 weights (ds, dl, d, du, dw)")
    wa = (ctypes.c_double * 5)(*[float(v) for v in weights])
    fn = getattr(lib, f"pentadsolver_{_prefix(x.dtype)}_gpsv_stencil_batch")
    _check(fn(_hval(handle), _ptr(scale), wa, float(id_weight), _ptr(x), _dims_arg(dims),
              len(dims), int(solvedim)))


DIST_RECORD_WORDS = 23  # PENTAD_DIST_RECORD_WORDS


def dist_forward(ds, dl, d, du, dw, x, dims, solvedim, rank, nranks, record, pads=None,
                 handle=None) -> None:
    """Phase 1 of the distributed line solve on one rank's arrays (shape
    ``dims``, optional ``pads``): writes the rank's record
    (DIST_RECORD_WORDS * lines elements of x's dtype) into ``record``."""
    fn = getattr(load_library(), f"pentadsolver_{_prefix(x.dtype)}_dist_forward")
    _check(fn(_hval(handle), _ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw), _ptr(x),
              _dims_arg(dims), None if pads is None else _dims_arg(pads), len(dims),
              int(solvedim), int(rank), int(nranks), _ptr(record)))


def dist_reduced(records, n_sys, nranks, seps, handle=None) -> None:
    """Phase 2 for every rank on one device: ``records`` holds the nranks
    records back to back; ``seps`` receives 4 * n_sys values per rank."""
    fn = getattr(load_library(), f"pentadsolver_{_prefix(records.dtype)}_dist_reduced")
    _check(fn(_hval(handle), _ptr(records), int(n_sys), int(nranks), _ptr(seps)))


def dist_backward(ds, dl, d, du, dw, x, dims, solvedim, rank, nranks, sep, pads=None,
                  buffer=None, handle=None) -> None:
    """Phase 3 on one rank's arrays: x (right-hand side) becomes the rank's
    part of the solution, from its 4 * lines separator values ``sep``."""
    fn = getattr(load_library(), f"pentadsolver_{_prefix(x.dtype)}_dist_backward")
    _check(fn(_hval(handle), _ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw), _ptr(x),
              _dims_arg(dims), None if pads is None else _dims_arg(pads), len(dims),
              int(solvedim), int(rank), int(nranks), _ptr(sep), _ptr(buffer)))


def dist_backward_buffer_size(ds, dl, d, du, dw, x, dims, solvedim, pads=None,
                              handle=None) -> int:
    fn = getattr(load_library(), f"pentadsolver_{_prefix(x.dtype)}_dist_backward_buffer_size")
    return int(fn(_hval(handle), _ptr(ds), _ptr(dl), _ptr(d), _ptr(du), _ptr(dw), _ptr(x),
                  _dims_arg(dims), None if pads is None else _dims_arg(pads), len(dims),
                  int(solvedim)))


class DistSlabs:
    """Every line of a field along ``solvedim`` split into ``nranks`` slabs of
    m = dims[solvedim] / nranks rows on one GPU, slab k standing for rank
    k's arrays (a view of the field: the rank's local shape, the field's
    extents as pads).  ``solve`` runs the three phases of the distributed
    line solve in place on x; the buffers (records, separators, scratch) are
    allocated once here."""

    def __init__(self, dims, solvedim, nranks, dtype, handle=None, device="cuda"):
        import torch
        dims = [int(v) for v in dims]
        if dims[solvedim] % nranks:
            raise ValueError("the line length must split into nranks equal slabs")
        self.dims, self.sd, self.P, self.handle = dims, int(solvedim), int(nranks), handle
        self.m = dims[solvedim] // nranks
        self.local = list(dims)
        self.local[solvedim] = self.m
        self.stride = int(np.prod(dims[:solvedim])) if solvedim else 1
        self.n_sys = int(np.prod(dims)) // dims[solvedim]
        self.pads = None if solvedim == len(dims) - 1 else dims
        self.records = torch.empty(nranks * DIST_RECORD_WORDS * self.n_sys, dtype=dtype,
                                   device=device)
        self.seps = torch.empty(nranks * 4 * self.n_sys, dtype=dtype, device=device)
        self._buf = None

    def _view(self, t, k):
        return t.reshape(-1)[k * self.m * self.stride:]

    def forward(self, ds, dl, d, du, dw, x, k):
        rec = self.records[k * DIST_RECORD_WORDS * self.n_sys:]
        v = [self._view(t, k) for t in (ds, dl, d, du, dw, x)]
        dist_forward(*v, self.local, self.sd, k, self.P, rec, self.pads, self.handle)

    def reduced(self):
        dist_reduced(self.records, self.n_sys, self.P, self.seps, self.handle)

    def backward(self, ds, dl, d, du, dw, x, k):
        import torch
        v = [self._view(t, k) for t in (ds, dl, d, du, dw, x)]
        if self._buf is None:  # the largest need over the ranks (alignment differs)
            nb = max(dist_backward_buffer_size(*[self._view(t, j) for t in (ds, dl, d, du, dw, x)],
                                               self.local, self.sd, self.pads, self.handle)
                     for j in range(self.P))
            self._buf = torch.empty(max(nb, 1), dtype=torch.uint8, device=x.device)
        dist_backward(*v, self.local, self.sd, k, self.P, self.seps[k * 4 * self.n_sys:],
                      self.pads, self._buf, self.handle)

    def solve(self, ds, dl, d, du, dw, x):
        for k in range(self.P):
            self.forward(ds, dl, d, du, dw, x, k)
        self.reduced()
        for k in range(self.P):
            self.backward(ds, dl, d, du, dw, x, k)


def kernel_that_ran(direction: int) -> str:
    """Reference pentadsolver_kernel_that_ran (pentad_cuda.cu:1159-1161)."""
    return load_library().pentadsolver_kernel_that_ran(direction).decode()


class Varlen:
    """Variable-length batch (pentadsolver_varlen_*): systems packed back to
    back, system k owning elements offsets[k]..offsets[k+1]-1.

    offsets: host sequence of n_sys + 1 int64 values.  solve(ds, dl, d, du,
    dw, x) dispatches on the dtypes: float64 -> FP64, float32 -> FP32,
    float32 coefficients with a float64 x -> FP32 solve + FP64 refinement
    (max_refine, tol, steps as in gpsv_batch_ir).
    """

    def __init__(self, offsets, handle=None):
        import numpy as _np
        lib = load_library()
        off = _np.ascontiguousarray(offsets, dtype=_np.int64)
        self._plan = _P()
        self._handle = handle  # the plan uses the handle: keep it alive
        _check(lib.pentadsolver_varlen_create(
            _hval(handle), ctypes.byref(self._plan),
            off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(off) - 1))

    def items(self) -> int:
        return int(load_library().pentadsolver_varlen_items(self._plan))

    def spread(self):
        """Extra lanes per system (log2) of the 10 length buckets."""
        out = (ctypes.c_int * 10)()
        _check(load_library().pentadsolver_varlen_get_spread(self._plan, out))
        return list(out)

    def set_spread(self, spread) -> None:
        arr = (ctypes.c_int * 10)(*[int(v) for v in spread])
        _check(load_library().pentadsolver_varlen_set_spread(self._plan, arr))

    @property
    def refine(self) -> bool:
        """float32 coefficients with a float64 x: True runs the FP32 solve +
        FP64 refinement kernel, False (the default) one FP64 solve of the
        coefficients widened on load (no refinement sweep, steps untouched)."""
        return load_library().pentadsolver_varlen_get_refine(self._plan) == 1

    def set_refine(self, on: bool) -> None:
        _check(load_library().pentadsolver_varlen_set_refine(self._plan, 1 if on else 0))

    def autotune(self, ds, dl, d, du, dw, x, reps=3):
        """Time spreads 0..2 per length bucket on these arrays (x unchanged)
        and keep the fastest; returns [(bucket, spread, ms)]."""
        import torch
        lib = load_library()
        rep = ctypes.create_string_buffer(1 << 14)
        args = [self._plan] + [_ptr(t) for t in (ds, dl, d, du, dw, x)] + [int(reps), rep,
                                                                          len(rep)]
        if ds.dtype == torch.float32 and x.dtype == torch.float64:
            n = lib.pentadsolver_DS_varlen_autotune_ir(*args)
        elif ds.dtype == x.dtype == torch.float64:
            n = lib.pentadsolver_D_varlen_autotune(*args)
        elif ds.dtype == x.dtype == torch.float32:
            n = lib.pentadsolver_S_varlen_autotune(*args)
        else:
            raise TypeError("varlen: float64, float32, or float32 coefficients with float64 x")
        if n < 0:
            raise PentadError(f"varlen autotune failed ({n})")
        out = []
        for line in rep.value.decode().splitlines():
            w = line.split()
            if w[0] == "bucket":
                out.append((int(w[1]), int(w[6]), float(w[7])))
        return out

    def solve(self, ds, dl, d, du, dw, x, max_refine=5, tol=1e-13, steps=None):
        import torch
        lib = load_library()
        cdt = ds.dtype
        args = [self._plan] + [_ptr(t) for t in (ds, dl, d, du, dw, x)]
        if cdt == torch.float32 and x.dtype == torch.float64:
            _check(lib.pentadsolver_DS_gpsv_varlen_batch_ir(*args, int(max_refine), float(tol),
                                                            _ptr(steps)))
        elif cdt == x.dtype == torch.float64:
            _check(lib.pentadsolver_D_gpsv_varlen_batch(*args))
        elif cdt == x.dtype == torch.float32:
            _check(lib.pentadsolver_S_gpsv_varlen_batch(*args))
        else:
            raise TypeError("varlen: float64, float32, or float32 coefficients with float64 x")

    def close(self):
        if self._plan:
            load_library().pentadsolver_varlen_destroy(ctypes.byref(self._plan))
            self._plan = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
