"""Seeded workload generators (CPU): the hyperdiffusion operator's structure
and the contraction the ADI step relies on, checked with the oracle."""
import numpy as np

from pentadsolver_amd import workloads as W
from oracle import oracle as O


def test_hyperdiffusion_lines_are_symmetric_and_truncated():
    dims = [7, 6, 9]
    sets = W.hyperdiffusion_np(dims, 7)
    shape = tuple(reversed(dims))
    for ax, (ds, dl, d, du, dw) in enumerate(sets):
        a = [v.reshape(shape) for v in (ds, dl, d, du, dw)]
        mv = [np.moveaxis(v, 2 - ax, -1) for v in a]   # lines along the last numpy axis
        s = mv[0][..., 2]                               # the line's s
        n = dims[ax]
        # per line: ds = dw = s, dl = du = -4 s, d = 1 + 6 s, constant along the line
        assert np.allclose(mv[2], 1 + 6 * s[..., None])
        assert np.all(mv[0][..., :2] == 0) and np.all(mv[1][..., :1] == 0)
        assert np.all(mv[3][..., n - 1:] == 0) and np.all(mv[4][..., n - 2:] == 0)
        # symmetry: coupling (i, i+1) == (i+1, i), (i, i+2) == (i+2, i)
        assert np.array_equal(mv[3][..., :-1], mv[1][..., 1:])
        assert np.array_equal(mv[4][..., :-2], mv[0][..., 2:])
        nu = s / W.HYPERDIFFUSION_DT_OVER_H4
        assert nu.min() >= 0.5 and nu.max() <= 2.0


def test_hyperdiffusion_direction_solves_contract():
    """(I + s T)^-1 with T = D2^2 + e1 e1^T + eN eN^T positive definite:
    every direction solve shrinks the 2-norm (the energy decay the run shows)."""
    dims = [24, 16, 20]
    sets = W.hyperdiffusion_np(dims, 3)
    x = np.random.default_rng(0).standard_normal(int(np.prod(dims)))
    for _ in range(3):
        for sd in range(3):
            y = O.solve(*sets[sd], x, dims, sd)
            assert np.linalg.norm(y) < np.linalg.norm(x)
            x = y
