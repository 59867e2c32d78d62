"""What the learned warm start is asked to predict (CPU, scipy): the top-16
eigenvectors of D^-1/2 A D^-1/2 for the DC-SBM workload (32 blocks,
p_in/p_out 20, power-law degree weights with exponent 2.5, mean degree 16),
here at 2^16 vertices.  Only a few of them lie in the span of the 32
community directions D^1/2 1_block; most are localised on a handful of
vertices (n * sum v^4 in the thousands, against about 3 for a vector spread
over the graph), and the eigenvalues below the first form a continuum with
no gap at 16.  A network that smooths random features over 3 hops can
approximate the community directions, but not which of thousands of similar
local structures carry the 16 largest eigenvalues: this bounds what a
learned start can save the Krylov solver on this workload."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def test_dcsbm_top16_mostly_localised(orc):
    import flashx_amd.fmx as F
    n, seed = 1 << 16, 900016
    bstart, wtab = F.dcsbm_params(n, 32, 2.5, 1.06)
    s, d = orc.dcsbm_edges(n, bstart, wtab, 20, n * 8, seed)
    keep = s != d
    A = sp.coo_matrix((np.ones(int(keep.sum())), (s[keep], d[keep])), shape=(n, n)).tocsr()
    A = ((A + A.T) > 0).astype(np.float64)
    deg = np.asarray(A.sum(1)).ravel()
    dinv = np.where(deg > 0, 1.0 / np.sqrt(np.maximum(deg, 1.0)), 0.0)
    N = sp.diags(dinv) @ A @ sp.diags(dinv)
    ev, V = sla.eigsh(N, k=20, which="LA", tol=1e-9)
    order = np.argsort(-ev)
    ev, V = ev[order], V[:, order]
    # block of each vertex: blocks are contiguous in the generator's ids,
    # which the R-MAT permutation relabels
    perm = np.array([orc.rmat_perm(v, 16, seed) for v in range(n)], np.int64)
    blk = np.empty(n, np.int64)
    blk[perm] = np.searchsorted(bstart[1:], np.arange(n), side="right")
    C = np.zeros((n, 32))
    C[np.arange(n), blk] = np.sqrt(deg)
    Q, _ = np.linalg.qr(C)
    in_span = (np.linalg.norm(Q.T @ V[:, :16], axis=0) ** 2)
    n_ipr = (V[:, :16] ** 4).sum(0) * n
    assert abs(ev[0] - 1.0) < 1e-8 and in_span[0] > 0.99
    localised = n_ipr > 1000.0
    assert localised.sum() >= 8, n_ipr
    assert np.median(in_span[localised]) < 0.1, in_span
    assert (in_span > 0.5).sum() <= 6, in_span  # few community-like directions
    # no gap at 16: lambda_16 - lambda_17 is small against lambda_2 - lambda_16
    assert ev[15] - ev[16] < 0.2 * (ev[1] - ev[15])
