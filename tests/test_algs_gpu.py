"""PageRank on the SpMV path (matrix/libmatrix-algs/PageRank.cpp:49-137)
against the oracle's FP32 restatement of the same iteration on the same
images.  The element-wise steps round identically; the SpMV sums differ
only in rounding (FP64 accumulation on the GPU, FP32 in reference order on
the CPU), so a fixed number of iterations agrees to 1e-5 relative, and a
converged run stops within one iteration of the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def directed(fmx, orc):
    s, d = fmx.rmat_edges(14, 8 << 14, 77)
    n = 1 << 14
    A = fmx.SparseMatrix.build(s, d, n=n, log2_h=9, log2_w=12)
    At = fmx.SparseMatrix.build(d, s, n=n, log2_h=9, log2_w=12)
    sh, dh = s.cpu().numpy().view(np.uint32), d.cpu().numpy().view(np.uint32)
    oA = orc.build_2d(sh, dh, n=n, log2h=9, log2w=12)
    oAt = orc.build_2d(dh, sh, n=n, log2h=9, log2w=12)
    return A, At, oA, oAt, n


def test_pagerank_fixed_iterations(fmx, orc, directed):
    A, At, oA, oAt, n = directed
    pr, it = fmx.pagerank(A, At, max_iters=10, damping=0.85)
    ref, rit = orc.pagerank(oA, oAt, 10, 0.85)
    assert it == rit == 10
    assert np.abs(pr.cpu().numpy() - ref).max() <= 1e-5 * np.abs(ref).max()


def test_pagerank_converged(fmx, orc, directed):
    A, At, oA, oAt, n = directed
    pr, it = fmx.pagerank(A, At, max_iters=200, damping=0.85)
    ref, rit = orc.pagerank(oA, oAt, 200, 0.85)
    assert abs(it - rit) <= 1 and rit < 200
    assert np.abs(pr.cpu().numpy() - ref).max() <= 2e-3 / n


def test_pagerank_symmetric_and_errors(fmx, orc):
    n = 1000
    s = np.arange(n, dtype=np.uint32)
    import torch
    src = torch.from_numpy(np.concatenate([s, s]).view(np.int32)).cuda()
    dst = torch.from_numpy(np.concatenate([(s + 1) % n, (s + 2) % n]).astype(np.uint32).view(np.int32)).cuda()
    R = fmx.SparseMatrix.build(src, dst, n=n, log2_h=7, log2_w=7, flags=fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE)
    pr, it = fmx.pagerank(R)
    assert it == 1 and np.allclose(pr.cpu().numpy(), 1.0 / n, rtol=1e-6)
    with pytest.raises(fmx.FmxError) as e:
        fmx.pagerank(R, max_iters=-1)
    assert e.value.status == fmx.EINVAL


# ---------------------------------------------------------------------------
# NMF (matrix/libmatrix-algs/NMF.cpp): fmx_nmf against the oracle's
# restatement from the same start.  One iteration agrees to FP32 rounding;
# after 11 iterations the factors agree to 1e-3 relative and the reported
# distances to 1e-6 relative (the updates are contractive, and the remaining
# differences are FP64 summation orders inside the Gram and SpMM products).

def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def undirected(fmx, orc):
    s, d = orc.rmat_edges(13, 8 << 13, 5)
    n = 1 << 13
    img = orc.build_2d(s, d, n=n, log2h=8, log2w=11, flags=orc.DEDUP | orc.SYMMETRIZE)
    return fmx.SparseMatrix.upload(img.data, img.offs), img, n


@pytest.mark.parametrize("k,iters", [(8, 1), (8, 11), (3, 6), (16, 11)])
def test_nmf_matches_oracle(fmx, orc, undirected, k, iters):
    import torch
    A, img, n = undirected
    rng = np.random.default_rng(k * 10 + iters)
    W0 = rng.random((n, k)).astype(np.float32)
    tH0 = rng.random((n, k)).astype(np.float32)
    W, tH, dist = fmx.nmf(A, k, iters, W=torch.from_numpy(W0).cuda(), tH=torch.from_numpy(tH0).cuda())
    rW, rtH, rdist = orc.nmf(img, img, W0, tH0, iters, A.get_nnz())
    tol = 1e-5 if iters == 1 else 1e-3
    assert _rel(W.cpu().numpy(), rW) <= tol and _rel(tH.cpu().numpy(), rtH) <= tol
    assert len(dist) == len(rdist) == (iters + 4) // 5
    for a, b in zip(dist, rdist):
        assert abs(a - b) <= 1e-6 * abs(b), (a, b)
    assert all(x >= y * (1 - 1e-9) for x, y in zip(dist, dist[1:]))  # never increases


def test_nmf_directed_streamed_and_errors(fmx, orc, directed):
    """A directed graph needs its transpose; a streamed matrix gives the same
    factors; weighted matrices and bad ranks are refused."""
    import torch
    A, At, oA, oAt, n = directed
    k = 6
    rng = np.random.default_rng(1)
    W0 = rng.random((n, k)).astype(np.float32)
    tH0 = rng.random((n, k)).astype(np.float32)
    W, tH, dist = fmx.nmf(A, k, 6, At=At, W=torch.from_numpy(W0).cuda(), tH=torch.from_numpy(tH0).cuda())
    rW, rtH, rdist = orc.nmf(oA, oAt, W0, tH0, 6, A.get_nnz())
    assert _rel(W.cpu().numpy(), rW) <= 1e-3 and _rel(tH.cpu().numpy(), rtH) <= 1e-3
    assert all(abs(a - b) <= 1e-6 * abs(b) for a, b in zip(dist, rdist))
    data, offs = A.download()
    tdata, toffs = At.download()
    As = fmx.SparseMatrix.upload_streamed(data, offs, chunk_bytes=1 << 16)
    Ats = fmx.SparseMatrix.upload_streamed(tdata, toffs, chunk_bytes=1 << 16)
    W2, tH2, dist2 = fmx.nmf(As, k, 6, At=Ats, W=torch.from_numpy(W0).cuda(), tH=torch.from_numpy(tH0).cuda())
    assert torch.equal(W2, W) and torch.equal(tH2, tH) and dist2 == dist
    with pytest.raises(fmx.FmxError) as e:
        fmx.nmf(A, 65, 1, At=At)
    assert e.value.status == fmx.EDIM
    vals = np.ones(8, np.float32)
    wimg = orc.build_2d_w(np.arange(8, dtype=np.uint32), (np.arange(8, dtype=np.uint32) + 1) % 8, vals, n=8,
                          log2h=3, log2w=3)
    Wm = fmx.SparseMatrix.upload(wimg.data, wimg.offs)
    with pytest.raises(fmx.FmxError) as e:
        fmx.nmf(Wm, 2, 1, At=Wm)
    assert e.value.status == fmx.EUNSUPPORTED
