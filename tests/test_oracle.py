"""CPU tests that pin the oracle (oracle/liboracle.so) before it is trusted.

Pins used (each chosen so that a plausible slip fails one of them):
  * a hand-written byte image (tests/golden/tiny_kat.py) -> writer layout;
  * an independent naive Python writer (tests/golden/py_writer.py) on random
    graphs at many tile shapes -> writer rules (COO vs SCSR, ordering, empty
    block rows, headers, offsets);
  * decoding every non-zero back and comparing with the normalised edge set;
  * brute-force CSR products (scipy) on integer operands -> SpMM exactly;
  * the reference's known-answer test A*[0..n-1] = sum of column ids
    (matrix/test/rand_mat_gen.cpp:178-199);
  * the exactly representable pattern X[i][j] = (i%100)*(j+1)
    (matrix/test/test-2d_multiply.cpp:35-41);
  * a Python loop over sorted neighbours for the FP accumulation order;
  * closed forms for the dense ops (numpy) and invariants for R-MAT.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import py_writer
import tiny_kat


def csr_of(orc, img):
    r, c = orc.nonzeros(img)
    n = img.n
    return sp.csr_matrix((np.ones(len(r), np.int64), (r.astype(np.int64), c.astype(np.int64))),
                         shape=(n, n))


def test_tiny_kat_bytes(orc):
    img = orc.build_2d(np.array(tiny_kat.SRC), np.array(tiny_kat.DST), n=tiny_kat.N,
                       log2h=tiny_kat.LOG2H, log2w=tiny_kat.LOG2W, flags=orc.DEDUP)
    assert bytes(img.data) == tiny_kat.IMAGE
    assert list(img.offs) == tiny_kat.OFFS
    assert orc.verify(img) == 0


def test_tiny_kat_spmm(orc):
    img = orc.Image(np.frombuffer(tiny_kat.IMAGE, np.uint8), tiny_kat.OFFS, tiny_kat.N)
    for dt in (np.int32, np.int64, np.float32, np.float64):
        Y = orc.spmm(img, np.array(tiny_kat.X, dt))
        assert np.array_equal(Y, np.array(tiny_kat.Y, dt))
        Y2 = orc.spmm(img, np.array(tiny_kat.X, dt), threads=3)
        assert np.array_equal(Y2, np.array(tiny_kat.Y, dt))


def test_tiny_kat_flags(orc):
    # keep self edges: row 2 gets (2,2) as a COO entry; no dedup: (0,1) twice
    src, dst = np.array(tiny_kat.SRC), np.array(tiny_kat.DST)
    for flags, kw in ((orc.KEEP_SELF | orc.DEDUP, dict(keep_self=True)),
                      (0, dict(dedup=False)),
                      (orc.SYMMETRIZE | orc.DEDUP, dict(symmetrize=True))):
        img = orc.build_2d(src, dst, n=6, log2h=1, log2w=1, flags=flags)
        ref, offs, n = py_writer.write(src, dst, 1, 1, n=6, **kw)
        assert bytes(img.data) == ref and list(img.offs) == offs
        assert orc.verify(img) == 0


@pytest.mark.parametrize("scale,lh,lw", [(6, 1, 1), (7, 2, 5), (8, 3, 3), (9, 8, 4), (10, 4, 10),
                                         (10, 15, 15), (9, 5, 7)])
def test_writer_matches_independent_writer(orc, scale, lh, lw):
    rng = np.random.default_rng(scale * 100 + lh)
    n = 1 << scale
    m = 6 * n
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = (rng.zipf(1.6, m) - 1).clip(0, n - 1).astype(np.uint32)  # skewed columns
    src[:50] = dst[:50]  # self edges
    for flags, kw in ((orc.DEDUP, {}), (0, dict(dedup=False)),
                      (orc.DEDUP | orc.SYMMETRIZE, dict(symmetrize=True))):
        try:
            ref, offs, nn = py_writer.write(src, dst, lh, lw, n=n + 3, **kw)
        except ValueError:  # a run longer than W without dedup: both refuse
            with pytest.raises(ValueError):
                orc.build_2d(src, dst, n=n + 3, log2h=lh, log2w=lw, flags=flags)
            continue
        img = orc.build_2d(src, dst, n=n + 3, log2h=lh, log2w=lw, flags=flags)
        assert img.n == nn
        assert bytes(img.data) == ref
        assert list(img.offs) == offs
        assert orc.verify(img) == 0


@pytest.mark.parametrize("lh,lw", [(15, 15), (10, 10), (6, 12), (12, 6)])
def test_roundtrip_nonzeros(orc, lh, lw):
    src, dst = orc.rmat_edges(12, 16 << 12, seed=5)
    img = orc.build_2d(src, dst, log2h=lh, log2w=lw)
    assert orc.verify(img) == 0
    r, c = orc.nonzeros(img)
    keys, n = orc.normalize(src, dst)
    assert n == img.n and len(keys) == img.nnz == len(r)
    assert np.array_equal(np.sort((r << 32) | c), keys)
    # normalised set == python set of non-self edges
    want = {(int(s), int(d)) for s, d in zip(src, dst) if s != d}
    assert len(want) == len(keys)


def test_spmm_integer_bruteforce(orc):
    src, dst = orc.rmat_edges(12, 16 << 12, seed=9)
    for lh, lw in ((15, 15), (8, 9), (4, 4)):
        img = orc.build_2d(src, dst, log2h=lh, log2w=lw)
        A = csr_of(orc, img)
        rng = np.random.default_rng(1)
        X = rng.integers(0, 8, (img.n, 8)).astype(np.int32)
        assert np.array_equal(orc.spmm(img, X), (A @ X.astype(np.int64)).astype(np.int32))
        X64 = rng.integers(-1000, 1000, (img.n, 5)).astype(np.int64)
        assert np.array_equal(orc.spmm(img, X64), A @ X64)


def test_spmm_seq_kat(orc):
    # rand_mat_gen.cpp:178-199: A * [0..n-1] = sum of the column ids of each row
    src, dst = orc.rmat_edges(11, 16 << 11, seed=3)
    img = orc.build_2d(src, dst, log2h=9, log2w=10)
    Y = orc.spmm(img, np.arange(img.n, dtype=np.int64))
    r, c = orc.nonzeros(img)
    want = np.bincount(r.astype(np.int64), weights=c.astype(np.float64), minlength=img.n)
    assert np.array_equal(Y[:, 0], want.astype(np.int64))


def test_spmm_exact_fp_pattern(orc):
    # test-2d_multiply.cpp:35-41: X[i][j] = (i%100)*(j+1) is exact in FP
    src, dst = orc.rmat_edges(11, 16 << 11, seed=4)
    img = orc.build_2d(src, dst, log2h=15, log2w=15)
    i = np.arange(img.n)[:, None]
    j = np.arange(4)[None, :]
    Xi = (i % 100) * (j + 1)
    A = csr_of(orc, img)
    for dt in (np.float32, np.float64):
        assert np.array_equal(orc.spmm(img, Xi.astype(dt)), (A @ Xi).astype(dt))


def test_spmm_fp_order_is_ascending_columns(orc):
    # the reference accumulates each row in ascending column order
    src, dst = orc.rmat_edges(8, 16 << 8, seed=11)
    img = orc.build_2d(src, dst, log2h=2, log2w=3)
    r, c = orc.nonzeros(img)
    order = np.lexsort((c, r))
    r, c = r[order], c[order]
    rng = np.random.default_rng(2)
    for dt in (np.float32, np.float64):
        X = rng.random((img.n, 3)).astype(dt)
        Y = orc.spmm(img, X)
        want = np.zeros((img.n, 3), dt)
        for rr, cc in zip(r, c):
            want[rr] = want[rr] + X[cc]
        assert np.array_equal(Y, want)


@pytest.mark.parametrize("threads,copy", [(1, 1), (4, 1), (7, 0)])
def test_mt_baseline_matches_sequential(orc, threads, copy):
    src, dst = orc.rmat_edges(13, 16 << 13, seed=8)
    rng = np.random.default_rng(3)
    for lh in (15, 8):
        img = orc.build_2d(src, dst, log2h=lh, log2w=15)
        for dt, k in ((np.float32, 16), (np.float64, 8), (np.int32, 3)):
            X = (rng.random((img.n, k)) * 8).astype(dt)
            assert np.array_equal(orc.spmm(img, X, threads=threads, copy=copy), orc.spmm(img, X))


def test_verify_rejects_corruption(orc):
    img = orc.Image(np.frombuffer(tiny_kat.IMAGE, np.uint8).copy(), tiny_kat.OFFS, tiny_kat.N)
    assert orc.verify(img) == 0
    bad = orc.Image(img.data.copy(), img.offs, img.n)
    bad.data[0] ^= 1  # magic
    assert orc.verify(bad) == 4
    bad = orc.Image(img.data.copy(), img.offs, img.n)
    bad.data[4096 + 16 + 2] = 0x05  # COO row header without its high bit
    assert orc.verify(bad) == 4
    bad = orc.Image(img.data.copy(), np.array([4096, 4192, 4146, 4208]), img.n)
    assert orc.verify(bad) == 4
    bad = orc.Image(np.frombuffer(tiny_kat.BAD_EMPTY_IMAGE, np.uint8).copy(), tiny_kat.BAD_EMPTY_OFFS, tiny_kat.N)
    assert orc.verify(bad) == 4  # empty tile in a non-empty block row


def test_normalize_errors(orc):
    with pytest.raises(ValueError):
        orc.build_2d(np.array([0, 5]), np.array([1, 2]), n=4)  # id >= n


def test_rmat_invariants(orc):
    scale, m, seed = 12, 1 << 16, 21
    s1, d1 = orc.rmat_edges(scale, m, seed, permute=False)
    s2, d2 = orc.rmat_edges(scale, m, seed, permute=False)
    assert np.array_equal(s1, s2) and np.array_equal(d1, d2)  # reproducible
    assert len(s1) == m and s1.max() < (1 << scale) and d1.max() < (1 << scale)
    # every level draws its quadrant with (a, b, c, d) = (.57, .19, .19, .05)
    for lvl in (0, 5, 11):
        sb = (s1 >> (scale - 1 - lvl)) & 1
        db = (d1 >> (scale - 1 - lvl)) & 1
        q = 2 * sb + db
        freq = np.bincount(q, minlength=4) / m
        exp = np.array([0.57, 0.19, 0.19, 0.05])
        chi2 = (((freq - exp) ** 2) / exp).sum() * m
        assert chi2 < 25, (lvl, freq)  # 3 dof: p ~ 1.6e-5
    # the relabelling is a bijection and is what permute=True applies
    perm = np.array([orc.rmat_perm(v, scale, seed) for v in range(1 << scale)])
    assert np.array_equal(np.sort(perm), np.arange(1 << scale))
    s3, d3 = orc.rmat_edges(scale, m, seed, permute=True)
    assert np.array_equal(s3, perm[s1]) and np.array_equal(d3, perm[d1])
    # splitmix64 known values (Vigna's reference splitmix64.c, seed 0 -> state += gamma)
    assert orc.splitmix64(0, 0) == 0xE220A8397B1DCDAF
    assert orc.splitmix64(0, 1) == 0x6E789E6AA1B965F4


def test_dense_ops_oracle(orc):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((1000, 16))
    B = rng.standard_normal((1000, 7))
    assert np.allclose(orc.gram_f64(A, B), A.T @ B, rtol=1e-12, atol=1e-12)
    C = rng.standard_normal((16, 5))
    Y0 = rng.standard_normal((1000, 5))
    assert np.allclose(orc.gemm_tall_f64(A, C, 2.0, 0.5, Y0.copy()), 2.0 * A @ C + 0.5 * Y0,
                       rtol=1e-12, atol=1e-12)
    assert np.allclose(orc.colnorm_f64(A), np.linalg.norm(A, axis=0), rtol=1e-13)
    # integer-valued inputs: exact (test-dense_matrix.cpp:653-713 naive multiply)
    Ai = rng.integers(-5, 5, (300, 4)).astype(np.float64)
    Bi = rng.integers(-5, 5, (300, 3)).astype(np.float64)
    assert np.array_equal(orc.gram_f64(Ai, Bi), Ai.T @ Bi)


def last_row_edges():
    # row 32767 of a 32768-row tile: its row-part header is 0xFFFF, the same
    # value as the end-of-row-parts marker (sparse_matrix_format.cpp:119-123)
    src = np.array([32767, 32767, 32766, 32767, 0, 32767], np.uint32)
    dst = np.array([5, 9, 3, 40000, 32767, 32767], np.uint32)
    return src, dst


def test_row_32767_header_collides_with_end_marker(orc):
    src, dst = last_row_edges()
    img = orc.build_2d(src, dst, n=40001, log2h=15, log2w=15)
    ref, offs, _ = py_writer.write(src, dst, 15, 15, n=40001)
    assert bytes(img.data) == ref and list(img.offs) == offs
    words = np.frombuffer(bytes(img.data[4096 + 16:4096 + 16 + 8]), np.uint16)
    assert words[0] == 0xFFFF and list(words[1:3]) == [5, 9] and words[3] == 0xFFFF
    assert orc.verify(img) == 0
    X = np.arange(40001 * 2, dtype=np.int64).reshape(-1, 2)
    Y = orc.spmm(img, X)
    assert list(Y[32767]) == list(X[5] + X[9] + X[40000])
    assert list(Y[32766]) == list(X[3]) and list(Y[0]) == list(X[32767])


def test_spmm_range_matches_full(orc):
    s, d = orc.rmat_edges(12, 16 << 12, 3)
    img = orc.build_2d(s, d, n=1 << 12, log2h=7, log2w=10)
    X = np.random.default_rng(0).random((img.n, 8))
    Y = orc.spmm(img, X)
    for br0, br1 in ((0, 1), (5, 9), (30, 32)):
        assert np.array_equal(orc.spmm_range(img, X, br0, br1), Y[br0 * 128:br1 * 128])


def test_weighted_tiny_kat(orc):
    """The weighted writer reproduces the hand-written weighted image, and
    the reader multiplies with its values."""
    img = orc.build_2d_w(tiny_kat.SRC, tiny_kat.DST, np.array(tiny_kat.W_VALS, np.float32), n=tiny_kat.N,
                         log2h=tiny_kat.LOG2H, log2w=tiny_kat.LOG2W, flags=orc.DEDUP)
    assert bytes(img.data) == tiny_kat.W_IMAGE
    assert list(img.offs) == tiny_kat.W_OFFS
    assert orc.verify(img) == 0
    Y = orc.spmm(img, np.array(tiny_kat.X, np.float32))
    assert Y.tolist() == tiny_kat.W_Y


@pytest.mark.parametrize("dt,lh,lw,flags", [(np.float32, 5, 6, 1), (np.float64, 7, 4, 1 | 4),
                                            (np.int32, 3, 9, 0), (np.int64, 9, 9, 1 | 2)])
def test_weighted_against_brute_force(orc, dt, lh, lw, flags):
    """Weighted images of R-MAT graphs: SpMM equals a dense brute force built
    straight from the edge list (first value of a duplicate kept, self edges
    per KEEP_SELF, (dst, src) added with the same value for SYMMETRIZE)."""
    rng = np.random.default_rng(lh)
    s, d = orc.rmat_edges(10, 8 << 10, lh)
    v = (rng.random(len(s)) * 100).astype(dt)
    img = orc.build_2d_w(s, d, v, n=1 << 10, log2h=lh, log2w=lw, flags=flags)
    assert orc.verify(img) == 0
    A = np.zeros((1 << 10, 1 << 10))
    seen = set()
    for a, b, w in zip(s.tolist(), d.tolist(), v.astype(np.float64).tolist()):
        pairs = [(a, b)] + ([(b, a)] if flags & 4 and a != b else [])
        for p in pairs:
            if p[0] == p[1] and not flags & 2:
                continue
            if flags & 1:
                if p in seen:
                    continue
                seen.add(p)
            A[p] += w
    X = (rng.random((1 << 10, 3)) * 8).astype(np.float64)
    assert np.allclose(orc.spmm(img, X), A @ X, rtol=1e-12, atol=1e-9)


def test_pagerank_closed_forms(orc):
    """The oracle's PageRank (PageRank.cpp restated) on graphs whose ranks
    are known: a directed cycle and an undirected regular graph (every vertex
    has the same in and out degree) keep the uniform start 1/n and stop after
    one iteration; on a star with a hub the hub ranks above the leaves by the
    ratio the fixed point fixes."""
    n = 1000
    s = np.arange(n, dtype=np.uint32)
    d = (s + 1) % n
    A = orc.build_2d(s, d, n=n, log2h=6, log2w=8)
    At = orc.build_2d(d, s, n=n, log2h=6, log2w=8)
    pr, it = orc.pagerank(A, At, 50, 0.85)
    assert it == 1 and np.all(pr == np.float32(1.0 / n))
    # 4-regular undirected ring (i <-> i+1, i <-> i+2)
    s2 = np.concatenate([s, s]).astype(np.uint32)
    d2 = np.concatenate([(s + 1) % n, (s + 2) % n]).astype(np.uint32)
    R = orc.build_2d(s2, d2, n=n, log2h=7, log2w=7, flags=orc.DEDUP | orc.SYMMETRIZE)
    pr, it = orc.pagerank(R, R, 50, 0.85)
    assert it == 1 and np.allclose(pr, 1.0 / n, rtol=1e-6)
    # star: leaves 1..m point at hub 0 and the hub at every leaf (undirected).
    # Fixed point: h = d*m*l/1 + c, l = d*h/m + c with c = (1-d)/(m+1).
    m = 99
    leaves = np.arange(1, m + 1, dtype=np.uint32)
    S = orc.build_2d(leaves, np.zeros(m, np.uint32), n=m + 1, log2h=4, log2w=4,
                     flags=orc.DEDUP | orc.SYMMETRIZE)
    pr, it = orc.pagerank(S, S, 200, 0.85)
    dmp, c = 0.85, 0.15 / (m + 1)
    h = (c + dmp * m * c) / (1 - dmp * dmp)
    l = dmp * h / m + c
    assert abs(pr[0] - h) < 1e-5 and np.allclose(pr[1:], l, rtol=1e-4)
