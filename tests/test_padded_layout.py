"""Padded array layout (CPU): the reference's own CPU library against the oracle.

The reference's AVX library addresses element (i, j, k) at
i + pads[0] * (j + pads[1] * k) with pads[0] = dims[0] rounded up to its
SIMD width (src/cpu/singlenode/pentad_cpu.cpp:611-613, :457, :569); its 2-D
consumer allocates rows of nx_pad elements (apps/2D_cavity_cpu.cpp:70-72).
Running that library (compiled from its own sources into oracle/_ref) on
padded arrays and the oracle on the same values packed densely must agree:
this pins the layout that pentadsolver_{D,S}_gpsv_batch_padded implements
and the embed / extract helpers the GPU tests use.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available("avx"),
                                reason="oracle/_ref not built (reference sources absent)")

# x lengths that the reference's padded x path handles: at least one SIMD
# vector, and not one row past a whole vector.  Its pentad_x_transpose
# writes row len-2 to x_reg[-1] when len % SIMD_VEC == 1
# (src/cpu/singlenode/pentad_cpu.cpp:201-203, an out-of-bounds stack write,
# after which the backward sweep starts from the wrong rows), and reads the
# padding as rows of the system when len < SIMD_VEC (:77-129); those cases
# are undefined behaviour and are not run.  (14, 10, 30, 11 are 2, 2, 2, 3
# past a multiple of 4 and 6, 2, 6, 3 past a multiple of 8.)
SHAPES = [[14, 9], [10, 20], [30, 7], [11, 9, 5], [30, 7, 3], [14, 11, 4]]


def gen(dims, seed, dtype):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    ds, dl, du, dw = (rng.uniform(-1, 1, n) for _ in range(4))
    d = rng.uniform(2, 3, n) * (abs(ds) + abs(dl) + abs(du) + abs(dw))
    b = rng.uniform(-1, 1, n)
    return [a.astype(dtype) for a in (ds, dl, d, du, dw, b)]


def padded_pads(dims, dtype):
    v = O.ref_simd_vec(dtype)
    return [-(-dims[0] // v) * v] + list(dims[1:])


def test_embed_extract_round_trip():
    dims, pads = [5, 3, 2], [8, 4, 2]
    a = np.arange(30, dtype=np.float64)
    p = O.embed_padded(a, dims, pads, -1.0)
    assert p.size == O.padded_extent(dims, pads) == 64
    # element (i, j, k) sits at i + 8 * (j + 4 * k)
    assert p[4 + 8 * (2 + 4 * 1)] == a[4 + 5 * (2 + 3 * 1)]
    assert np.array_equal(O.extract_padded(p, dims, pads), a)
    assert np.all(p[O.padding_mask(dims, pads)] == -1.0)
    assert int(O.padding_mask(dims, pads).sum()) == 64 - 30


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims", SHAPES)
def test_reference_cpu_solver_on_padded_arrays_matches_oracle(dims, dtype):
    pads = padded_pads(dims, dtype)
    vec = O.ref_simd_vec(dtype)
    assert pads[0] != dims[0]  # the shapes exercise real padding
    assert dims[0] >= vec and dims[0] % vec != 1
    for sd in range(len(dims)):
        arrs = gen(dims, 100 + 7 * sd + len(dims), dtype)
        ref = O.solve(*[a.astype(np.float64) for a in arrs], dims, sd)
        padded = [O.embed_padded(a, dims, pads, 0.0) for a in arrs[:5]]
        x = O.embed_padded(arrs[5], dims, pads, 7.25)
        O.ref_gpsv_batch(*padded, x, dims, sd)
        got = O.extract_padded(x, dims, pads).astype(np.float64)
        err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        assert err < (1e-13 if dtype == np.float64 else 1e-5), (sd, err)
