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
