"""The eigensolver's other operators (matrix/test/test-eigen.cpp): A^T A
(SVD_Operator, :138-177) over a directed R-MAT graph and its transpose, and A
itself (eigen_Operator, :32-58) over a symmetrised one; fmx_eigen with
FMX_EIG_OP_SVD / FMX_EIG_OP_PLAIN.

Pinned by dense linear algebra on the oracle-decoded matrices, not by the
solver's own report:
  * the top nev eigenvalues equal numpy's: the squared singular values of A
    for the SVD operator (np.linalg.svd), the largest algebraic eigenvalues
    of A for the plain one (np.linalg.eigvalsh), and for the SVD operator of
    a symmetric A (no transpose given) the largest squares of A's
    eigenvalues; at R-MAT scale 16, scipy's ARPACK singular values;
  * each returned pair's residual ||N x - theta x|| / |theta| is recomputed in
    FP64 from the returned vectors, with N applied by the oracle's CPU SpMM of
    the downloaded images (A, then A's transpose for the SVD operator);
  * the vectors are orthonormal."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCALE = 11
N = 1 << SCALE
NEV = 8


def decode(fmx, orc, spm):
    data, offs = spm.download()
    img = orc.Image(data, offs, spm.get_num_rows())
    r, c = orc.nonzeros(img)
    A = np.zeros((spm.get_num_rows(), spm.get_num_cols()))
    A[r.astype(np.int64), c.astype(np.int64)] = 1.0  # binary, deduplicated
    return img, A


@pytest.fixture(scope="module")
def directed(fmx, orc):
    s, d = fmx.rmat_edges(SCALE, 8 << SCALE, 5)
    A = fmx.SparseMatrix.build(s, d, n=N, log2_h=6, log2_w=10, flags=fmx.BUILD_DEDUP)
    At = fmx.SparseMatrix.build(d, s, n=N, log2_h=6, log2_w=10, flags=fmx.BUILD_DEDUP)
    img, Ad = decode(fmx, orc, A)
    img_t, Atd = decode(fmx, orc, At)
    assert np.array_equal(Atd, Ad.T)  # the transpose image holds A^T
    return A, At, img, img_t, Ad


@pytest.fixture(scope="module")
def symmetric(fmx, orc):
    s, d = fmx.rmat_edges(SCALE, 8 << SCALE, 6)
    S = fmx.SparseMatrix.build(s, d, n=N, log2_h=6, log2_w=10, flags=fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE)
    img, Sd = decode(fmx, orc, S)
    assert np.array_equal(Sd, Sd.T)
    return S, img, Sd


def check_pairs(evals, vecs, apply, tol):
    V = vecs.cpu().numpy()
    NV = apply(V)
    rel = np.linalg.norm(NV - V * evals[None, :], axis=0) / np.abs(evals)
    assert np.all(rel <= tol * 1.01), rel.max()
    assert np.abs(V.T @ V - np.eye(V.shape[1])).max() < 1e-6
    return rel


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-6)])
def test_svd_operator(fmx, orc, directed, precision, tol):
    A, At, img, img_t, Ad = directed
    prec = fmx.EIG_FP64 if precision == "fp64" else fmx.EIG_FP32
    evals, vecs, resid, st = fmx.compute_eigen(A, nev=NEV, block_size=8, num_blocks=6, tol=tol, precision=prec,
                                               op="svd", At=At, max_restarts=300)
    assert st["converged"] == 1
    assert vecs.shape == (N, NEV)
    rel = check_pairs(evals, vecs, lambda V: orc.spmm(img_t, np.ascontiguousarray(orc.spmm(img, V))), tol)
    want = np.linalg.svd(Ad, compute_uv=False)[:NEV] ** 2
    print("%s: squared singular values %s (numpy %s), residuals %s, %d block SpMMs"
          % (precision, np.array2string(evals, precision=10), np.array2string(want, precision=10),
             np.array2string(rel, precision=2), st["spmm_blocks"]))
    assert np.all(np.abs(evals - want) <= 2 * tol * want + 1e-12 * want[0])
    assert st["spmm_blocks"] % 2 == 0  # two SpMMs per application of A^T A


def test_svd_operator_of_symmetric_matrix(fmx, orc, symmetric):
    S, img, Sd = symmetric
    tol = 1e-10
    evals, vecs, resid, st = fmx.compute_eigen(S, nev=NEV, block_size=8, num_blocks=6, tol=tol,
                                               precision=fmx.EIG_FP64, op="svd", max_restarts=300)
    assert st["converged"] == 1
    check_pairs(evals, vecs, lambda V: orc.spmm(img, np.ascontiguousarray(orc.spmm(img, V))), tol)
    want = np.sort(np.linalg.eigvalsh(Sd) ** 2)[::-1][:NEV]
    assert np.all(np.abs(evals - want) <= 2 * tol * want + 1e-12 * want[0])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-6)])
def test_plain_operator(fmx, orc, symmetric, precision, tol):
    S, img, Sd = symmetric
    prec = fmx.EIG_FP64 if precision == "fp64" else fmx.EIG_FP32
    evals, vecs, resid, st = fmx.compute_eigen(S, nev=NEV, block_size=8, num_blocks=6, tol=tol, precision=prec,
                                               op="plain", max_restarts=300)
    assert st["converged"] == 1
    check_pairs(evals, vecs, lambda V: orc.spmm(img, V), tol)
    want = np.linalg.eigvalsh(Sd)[::-1][:NEV]
    print("%s: eigenvalues %s (numpy %s)" % (precision, np.array2string(evals, precision=10),
                                             np.array2string(want, precision=10)))
    assert np.all(np.abs(evals - want) <= 2 * tol * np.abs(want) + 1e-12 * abs(want[0]))


def test_operator_errors(fmx, directed, symmetric):
    A, At, img, img_t, Ad = directed
    S = symmetric[0]
    with pytest.raises(fmx.FmxError):
        fmx.compute_eigen(S, nev=4, block_size=8, op="plain", At=S)  # only the SVD takes a transpose
    with pytest.raises(fmx.FmxError):
        fmx.compute_eigen(S, nev=4, block_size=8, op="bogus")


def test_svd_operator_rmat16_vs_arpack(fmx, orc):
    """A larger directed case: R-MAT scale 16 (65,536 vertices, about 1 M
    non-zeros) against scipy's ARPACK singular values of the oracle-decoded
    matrix, with every pair's residual recomputed through the oracle's SpMM of
    both images."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    n = 1 << 16
    s, d = fmx.rmat_edges(16, 16 << 16, 9)
    A = fmx.SparseMatrix.build(s, d, n=n, log2_h=6, log2_w=15, flags=fmx.BUILD_DEDUP)
    At = fmx.SparseMatrix.build(d, s, n=n, log2_h=6, log2_w=15, flags=fmx.BUILD_DEDUP)
    del s, d
    data, offs = A.download()
    img = orc.Image(data, offs, n)
    tdata, toffs = At.download()
    timg = orc.Image(tdata, toffs, n)
    r, c = orc.nonzeros(img)
    As = sp.csr_matrix((np.ones(r.size), (r.astype(np.int64), c.astype(np.int64))), shape=(n, n))
    tol = 1e-8
    evals, vecs, resid, st = fmx.compute_eigen(A, nev=NEV, block_size=8, num_blocks=6, tol=tol,
                                               precision=fmx.EIG_FP64, op="svd", At=At, max_restarts=400)
    assert st["converged"] == 1
    check_pairs(evals, vecs, lambda V: orc.spmm(timg, np.ascontiguousarray(orc.spmm(img, V))), tol)
    sv = np.sort(sla.svds(As, k=NEV, tol=1e-12, return_singular_vectors=False))[::-1]
    want = sv ** 2
    print("scale 16: squared singular values %s (ARPACK %s), %d block SpMMs"
          % (np.array2string(evals, precision=8), np.array2string(want, precision=8), st["spmm_blocks"]))
    assert np.all(np.abs(evals - want) <= 1e-7 * want)
