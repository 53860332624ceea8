"""The micro-benchmark covers the reference's benchmark matrix
(benchmarks/cuda/my_bm.cu:55-71): N = 32..256 by doubling, ndims 1 (solvedim
0), 2 (0, 1) and 3 (0, 1, 2).  CPU only: no solve is run."""
import importlib


def test_matrix_is_the_reference_matrix():
    mb = importlib.import_module("microbench")
    cases = mb.matrix(256)
    assert len(cases) == 4 * 6
    assert sorted({c[0] for c in cases}) == [32, 64, 128, 256]
    for n in (32, 64, 128, 256):
        assert sorted((nd, sd) for (m, nd, sd) in cases if m == n) == \
            [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2)]
    assert mb.matrix(64) == [c for c in cases if c[0] <= 64]
