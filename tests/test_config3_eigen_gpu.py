"""The mixed-precision spectral embedding workload itself: the top 16
eigenpairs of N = D^-1/2 A D^-1/2 for the symmetrised R-MAT scale-22 graph
(edge factor 16, dedup, 128,312,530 non-zeros), block 16, tolerance 1e-6,
in the default mode (FP32 SpMM, FP64-MFMA Grams) and the mixed mode
(BF16-MFMA first-pass Grams with FP32 accumulation).

Checked against the oracle, not the solver's own report:
  * every pair's residual ||N x - theta x|| / |theta| recomputed in FP64 from
    the returned vectors, with N applied by the oracle's reference-order CPU
    SpMM on the downloaded tile image (the reference's operator,
    matrix/test/test-eigen.cpp:112-127);
  * the vectors are orthonormal;
  * the eigenvalues: this graph has many connected components, and for each
    component with an edge, D^1/2 times its indicator is an eigenvector of N
    with eigenvalue 1, so eigenvalue 1 has multiplicity = that component
    count (counted here with scipy).  That count exceeds 16, so all 16
    wanted eigenvalues are 1 and the solve finds 16 orthonormal vectors of
    that eigenspace.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem(fmx, orc):
    import scipy.sparse as sp
    import scipy.sparse.csgraph as csg
    n = 1 << 22
    s, d = fmx.rmat_edges(22, 16 << 22, 22)
    spm = fmx.SparseMatrix.build(s, d, n=n, log2_h=6, log2_w=15, flags=fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE)
    del s, d
    data, offs = spm.download()
    img = orc.Image(data, offs, n)
    deg = orc.spmm(img, np.ones((n, 1)), threads=8).ravel()
    dinv = np.where(deg > 0, 1.0 / np.sqrt(np.maximum(deg, 1.0)), 0.0)
    r, c = orc.nonzeros(img)
    A = sp.csr_matrix((np.ones(r.size, np.int8), (r.astype(np.int64), c.astype(np.int64))), shape=(n, n))
    del r, c
    ncomp, labels = csg.connected_components(A, directed=False)
    del A
    # components with at least one edge (isolated vertices have D^-1/2 := 0)
    nontrivial = int(np.unique(labels[deg > 0]).size)
    return spm, img, dinv, nontrivial


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_config3_top16(fmx, orc, problem, precision):
    spm, img, dinv, nontrivial = problem
    prec = fmx.EIG_FP32 if precision == "fp32" else fmx.EIG_MIXED
    evals, vecs, resid, st = fmx.compute_eigen(spm, nev=16, block_size=16, tol=1e-6, precision=prec, seed=1,
                                               max_restarts=200)
    assert st["converged"] == 1
    V = vecs.cpu().numpy()
    NV = dinv[:, None] * orc.spmm(img, np.ascontiguousarray(dinv[:, None] * V), threads=8)
    rel = np.linalg.norm(NV - V * evals[None, :], axis=0) / np.abs(evals)
    print("%s: eigenvalues %s; FP64 residuals (oracle) %s; %d block SpMMs; %d components with an edge"
          % (precision, np.array2string(evals, precision=9), np.array2string(rel, precision=2),
             st["spmm_blocks"], nontrivial))
    assert np.all(rel <= 1e-6 * 1.01), rel.max()
    assert np.abs(V.T @ V - np.eye(16)).max() < 1e-6
    assert nontrivial > 16
    assert np.all(np.abs(evals - 1.0) <= 1e-6)
