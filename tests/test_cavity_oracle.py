"""The oracle's restatement of the 2-D cavity consumer's pressure-correction
assembly (reference apps/preproc.hpp:248-525, pressure_coef_matrix) against
the reference's own code compiled from its header (oracle/_ref).

Bit-for-bit on all eleven output arrays (right-hand side, the x set and the y
set of five diagonals), prefilled with a sentinel so that elements the
assembly leaves alone are compared too.  Shapes with nx == nx_pad (the
layout in which the reference's index expressions address row (i-1)*nx_pad
+ (j-1)) and with nx < nx_pad (where some of its writes land elsewhere; the
restatement reproduces them).  A planted mistake (dx and dy exchanged) must
be seen.
"""
import importlib.util
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location(
    "_workloads", os.path.join(HERE, "..", "pentadiagonal-solver_amd", "workloads.py"))
W = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(W)

SHAPES = [(16, 16, 16), (12, 9, 12), (3, 3, 3), (3, 40, 4), (14, 9, 16), (30, 7, 32),
          (64, 5, 64), (9, 33, 12)]


def fresh(nx_pad, ny):
    return O.cavity_outputs(nx_pad, ny, fill=-3.5)


@pytest.mark.skipif(not O.ref_cavity_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("nx,ny,nx_pad", SHAPES)
def test_restatement_bit_identical_to_reference(nx, ny, nx_pad):
    up, vp, uv, aU, dx, dy = W.cavity_np(nx, ny, seed=nx * 100 + ny)
    ours = O.cavity_pressure_coef(nx, ny, nx_pad, aU, dx, dy, up, vp, uv, fresh(nx_pad, ny))
    ref = O.ref_cavity_pressure_coef(nx, ny, nx_pad, aU, dx, dy, up, vp, uv,
                                     fresh(nx_pad, ny))
    for k in O.CAVITY_OUT:
        assert np.array_equal(ours[k], ref[k]), k


@pytest.mark.skipif(not O.ref_cavity_available(), reason="oracle/_ref not built")
def test_planted_mistake_is_seen():
    nx, ny = 12, 9
    up, vp, uv, aU, dx, dy = W.cavity_np(nx, ny, seed=5)
    ours = O.cavity_pressure_coef(nx, ny, nx, aU, dy, dx, up, vp, uv, fresh(nx, ny))
    ref = O.ref_cavity_pressure_coef(nx, ny, nx, aU, dx, dy, up, vp, uv, fresh(nx, ny))
    assert not np.array_equal(ours["d_x"], ref["d_x"])


def test_structure_without_padding():
    """nx == nx_pad: every point written once, the y set equals the x set,
    d is the sum of the four other entries, all entries positive."""
    nx, ny = 10, 7
    up, vp, uv, aU, dx, dy = W.cavity_np(nx, ny, seed=3)
    out = O.cavity_pressure_coef(nx, ny, nx, aU, dx, dy, up, vp, uv, fresh(nx, ny))
    for a, b in zip(("s_x", "l_x", "d_x", "u_x", "w_x"), ("s_y", "l_y", "d_y", "u_y", "w_y")):
        assert np.array_equal(out[a], out[b])
    for k in O.CAVITY_OUT:
        assert not np.any(out[k] == -3.5), k
    s, l, d, u, w = (out[k] for k in ("s_x", "l_x", "d_x", "u_x", "w_x"))
    assert np.array_equal(d, u + l + s + w)
    assert np.all(s > 0) and np.all(l > 0) and np.all(u > 0) and np.all(w > 0)
    # an interior point: the four couplings of (i, j) = (3, 4)
    i, j, ldv = 3, 4, nx + 2
    c = lambda a, b: 0.5 * aU * (1.0 / uv[a] + 1.0 / uv[b])
    k = (i - 1) * nx + (j - 1)
    assert u[k] == c(i * ldv + j, i * ldv + j + 1) * (dy * dy)
    assert l[k] == c(i * ldv + j, i * ldv + j - 1) * (dy * dy)
    assert s[k] == c(i * ldv + j, (i - 1) * ldv + j) * (dx * dx)
    assert w[k] == c(i * ldv + j, (i + 1) * ldv + j) * (dx * dx)
