"""The oracle's NMF (Lee-Seung multiplicative updates,
matrix/libmatrix-algs/NMF.cpp:110-126,164-230,259-316) pinned by mathematics
rather than by retyping it: the distance it reports equals ||A - W H||_F^2
formed densely, the Frobenius distance never increases under the updates
(Lee and Seung, "Algorithms for non-negative matrix factorization", NIPS
2000, Theorem 1), factors stay non-negative, and a matrix that is an exact
rank-k non-negative product (disjoint all-ones blocks) is factorised to a
distance near zero."""
import numpy as np


def dense(orc, img, n):
    return orc.spmm(img, np.eye(n))


def graph(orc, n, m_edges, seed):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m_edges).astype(np.uint32)
    d = rng.integers(0, n, m_edges).astype(np.uint32)
    # symmetric: the matrix is its own transpose
    return orc.build_2d(s, d, n=n, log2h=5, log2w=6, flags=orc.DEDUP | orc.SYMMETRIZE)


def test_distance_is_the_dense_frobenius_norm(orc):
    n, k = 200, 5
    img = graph(orc, n, 1500, 1)
    A = dense(orc, img, n)
    rng = np.random.default_rng(2)
    W0, tH0 = rng.random((n, k)).astype(np.float32), rng.random((n, k)).astype(np.float32)
    for iters in (1, 6):
        W, tH, dist = orc.nmf(img, img, W0, tH0, iters, int(A.sum()))
        want = float(np.sum((A - W.astype(np.float64) @ tH.astype(np.float64).T) ** 2))
        assert abs(dist[-1] - want) <= 1e-4 * want, (iters, dist[-1], want)


def test_distance_never_increases_and_factors_stay_nonnegative(orc):
    n, k = 150, 4
    img = graph(orc, n, 900, 3)
    nnz = int(dense(orc, img, n).sum())
    rng = np.random.default_rng(4)
    W, tH = rng.random((n, k)).astype(np.float32), rng.random((n, k)).astype(np.float32)
    prev = None
    for _ in range(30):
        W, tH, dist = orc.nmf(img, img, W, tH, 1, nnz)
        assert np.all(W > 0) and np.all(tH > 0)
        if prev is not None:
            assert dist[0] <= prev * (1 + 1e-6) + 1e-6, (dist[0], prev)
        prev = dist[0]


def test_exact_block_factorisation(orc):
    """A = disjoint all-ones blocks (self loops kept) has the exact rank-k
    non-negative factorisation W = H^T = block indicators."""
    k, b = 4, 20
    n = k * b
    s, d = [], []
    for blk in range(k):
        for i in range(b):
            for j in range(b):
                s.append(blk * b + i)
                d.append(blk * b + j)
    img = orc.build_2d(np.array(s, np.uint32), np.array(d, np.uint32), n=n, log2h=4, log2w=5,
                       flags=orc.DEDUP | orc.KEEP_SELF)
    nnz = k * b * b
    rng = np.random.default_rng(7)
    W, tH = rng.random((n, k)).astype(np.float32), rng.random((n, k)).astype(np.float32)
    W, tH, dist = orc.nmf(img, img, W, tH, 301, nnz)
    assert dist[0] > 1.0
    assert dist[-1] < 1e-3 * nnz, dist[-1]


def test_asymmetric_uses_the_transpose(orc):
    """For a directed graph the H update needs A^T: the oracle with the true
    transpose matches the dense formulas; with A in its place it does not."""
    n, k = 120, 3
    rng = np.random.default_rng(9)
    s = rng.integers(0, n, 700).astype(np.uint32)
    d = rng.integers(0, n, 700).astype(np.uint32)
    A_img = orc.build_2d(s, d, n=n, log2h=5, log2w=6, flags=orc.DEDUP)
    At_img = orc.build_2d(d, s, n=n, log2h=5, log2w=6, flags=orc.DEDUP)
    A = dense(orc, A_img, n)
    W0, tH0 = rng.random((n, k)).astype(np.float32), rng.random((n, k)).astype(np.float32)
    W, tH, dist = orc.nmf(A_img, At_img, W0, tH0, 1, int(A.sum()))
    # one update in dense FP64 arithmetic
    eps = 1e-8
    W64, H64 = W0.astype(np.float64), tH0.astype(np.float64).T
    H1 = np.maximum(H64 * (W64.T @ A), eps) / (W64.T @ W64 @ H64 + eps)
    W1 = np.maximum(W64 * (A @ H1.T), eps) / (W64 @ (H1 @ H1.T) + eps)
    assert np.allclose(tH, H1.T, rtol=1e-5, atol=1e-7)
    assert np.allclose(W, W1, rtol=1e-5, atol=1e-7)
    Wb, tHb, _ = orc.nmf(A_img, A_img, W0, tH0, 1, int(A.sum()))
    assert not np.allclose(tHb, H1.T, rtol=1e-3)
