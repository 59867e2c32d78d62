"""The tile visiting orders of block_compute_task::run
(matrix/sparse_matrix.cpp:120-314): seq_exec_order and hilbert_exec_order.
The Hilbert position is pinned by the curve's defining properties -- a
bijection onto 0..n^2-1 that starts at the origin, whose consecutive cells
are grid neighbours, and whose every aligned run of 4^m positions fills an
aligned 2^m x 2^m square (the property matrix/unit-test/test-hilbert.cpp
checks) -- not by retyping it.  The order only moves FP32 rounding: integer
products agree exactly under both orders, and FP32 ones stay within the
usual bound of the exact sum."""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32])
def test_hilbert_curve_properties(orc, n):
    pos = {}
    for x in range(n):
        for y in range(n):
            pos[orc.hilbert_xy2d(n, x, y)] = (x, y)
    assert sorted(pos) == list(range(n * n))
    assert pos[0] == (0, 0)
    for d in range(1, n * n):
        (a, b), (c, e) = pos[d - 1], pos[d]
        assert abs(a - c) + abs(b - e) == 1, (n, d)
    size = 1
    while size * size <= n * n:
        for start in range(0, n * n, size * size):
            cells = [pos[d] for d in range(start, start + size * size)]
            xs = {c[0] // size for c in cells}
            ys = {c[1] // size for c in cells}
            assert len(xs) == 1 and len(ys) == 1, (n, size, start)
        size *= 2


@pytest.fixture(scope="module")
def image(orc):
    s, d = orc.rmat_edges(12, 16 << 12, seed=13)
    return orc.build_2d(s, d, n=1 << 12, log2h=6, log2w=6, flags=orc.DEDUP)


def test_seq_order_is_the_reference_order(orc, image):
    X = np.random.default_rng(1).random((1 << 12, 8)).astype(np.float32)
    ref = orc.spmm(image, X)
    for t in (1, 3, 4, 16):
        assert np.array_equal(orc.spmm_exec(image, X, t, False), ref), t


def test_hilbert_order_moves_only_rounding(orc, image):
    rng = np.random.default_rng(2)
    Xi = rng.integers(-50, 50, (1 << 12, 5)).astype(np.int64)
    ref_i = orc.spmm(image, Xi)
    X = rng.random((1 << 12, 8)).astype(np.float32)
    seq = orc.spmm_exec(image, X, 8, False)
    exact = orc.spmm(image, X.astype(np.float64))
    for t in (2, 8, 16):
        assert np.array_equal(orc.spmm_exec(image, Xi, t, True), ref_i), t
        hil = orc.spmm_exec(image, X, t, True)
        nz = exact != 0
        assert (np.abs(hil - exact)[nz] / exact[nz]).max() <= 1e-5
        assert np.all(hil[~nz] == 0)
    # the visiting order really changed some FP32 sums
    assert not np.array_equal(orc.spmm_exec(image, X, 8, True), seq)


def test_hilbert_needs_full_square_tasks(orc, image):
    X = np.zeros((1 << 12, 2), np.float32)
    with pytest.raises(ValueError):
        orc.spmm_exec(image, X, 3, True)  # not a power of two
    with pytest.raises(ValueError):
        orc.spmm_exec(image, X, 128, True)  # more than 64 block rows per task
