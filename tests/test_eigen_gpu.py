"""GPU block Krylov eigensolver (fmx_eigen_normalized) for N = D^-1/2 A D^-1/2.

The reference delegates the numerics to Trilinos/Anasazi, which is not
available, so eigen parity is pinned by mathematics instead:
  * closed form: a p x q torus has normalised-adjacency eigenvalues
    (cos(2 pi i/p) + cos(2 pi j/q)) / 2;
  * a library routine: scipy.sparse.linalg.eigsh on the FP64 operator built
    from the oracle's decoding of the same tile image;
  * invariants: every returned pair has ||N x - theta x|| / |theta| <= tol
    recomputed here in FP64 with scipy, and the vectors are orthonormal.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

pytestmark = pytest.mark.gpu


def dev_u32(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).cuda()


def normalized_operator(orc, img):
    r, c = orc.nonzeros(img)
    n = img.n
    A = sp.csr_matrix((np.ones(len(r)), (r.astype(np.int64), c.astype(np.int64))), shape=(n, n))
    deg = np.asarray(A.sum(1)).ravel()
    dinv = np.where(deg > 0, 1.0 / np.sqrt(np.maximum(deg, 1)), 0.0)
    return sp.diags(dinv) @ A @ sp.diags(dinv)


def check_pairs(N, evals, vecs, tol):
    V = vecs.cpu().numpy()
    R = N @ V - V * evals[None, :]
    rel = np.linalg.norm(R, axis=0) / np.abs(evals)
    assert np.all(rel <= tol * 1.01), rel.max()
    G = V.T @ V
    assert np.abs(G - np.eye(len(evals))).max() < 1e-6


def torus_edges(p, q):
    idx = np.arange(p * q).reshape(p, q)
    right = np.stack([idx.ravel(), np.roll(idx, -1, axis=1).ravel()], 1)
    down = np.stack([idx.ravel(), np.roll(idx, -1, axis=0).ravel()], 1)
    e = np.concatenate([right, down])
    return e[:, 0].astype(np.uint32), e[:, 1].astype(np.uint32)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_torus_closed_form(fmx, orc, precision):
    p, q = 40, 37
    src, dst = torus_edges(p, q)
    flags = fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE
    spm = fmx.SparseMatrix.build(dev_u32(src), dev_u32(dst), n=p * q, log2_h=8, log2_w=10, flags=flags)
    nev = 8
    tol = 1e-10 if precision == 2 else 1e-6  # an FP32 operator limits the FP64 residual
    evals, vecs, resid, st = fmx.compute_eigen(spm, nev=nev, block_size=16, num_blocks=6, tol=tol,
                                               precision=precision, max_restarts=300)
    assert st["converged"] == 1
    lam = np.sort(((np.cos(2 * np.pi * np.arange(p) / p)[:, None] +
                    np.cos(2 * np.pi * np.arange(q) / q)[None, :]) / 2).ravel())[::-1]
    assert np.allclose(evals, lam[:nev], atol=10 * tol)
    img = orc.build_2d(src, dst, n=p * q, log2h=8, log2w=10, flags=orc.DEDUP | orc.SYMMETRIZE)
    check_pairs(normalized_operator(orc, img), evals, vecs, tol)


@pytest.mark.parametrize("precision", [0, 2])
def test_rmat_vs_eigsh(fmx, orc, precision):
    s, d = fmx.rmat_edges(13, 16 << 13, 77)
    flags = fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE
    spm = fmx.SparseMatrix.build(s, d, n=1 << 13, log2_h=8, log2_w=13, flags=flags)
    evals, vecs, resid, st = fmx.compute_eigen(spm, nev=16, tol=1e-6, precision=precision)
    assert st["converged"] == 1 and np.all(resid <= 1e-6)
    img = orc.build_2d(s.cpu().numpy().view(np.uint32), d.cpu().numpy().view(np.uint32),
                       n=1 << 13, log2h=8, log2w=13, flags=orc.DEDUP | orc.SYMMETRIZE)
    N = normalized_operator(orc, img)
    ref = np.sort(sla.eigsh(N, k=16, which="LA", tol=1e-12, return_eigenvectors=False))[::-1]
    # residual <= 1e-6 bounds the eigenvalue error by ~1e-6 (and ~1e-12 when separated)
    assert np.allclose(evals, ref, atol=2e-6)
    check_pairs(N, evals, vecs, 1e-6)


@pytest.mark.parametrize("streamed", [False, True])
def test_rmat_split_bands_and_streamed(fmx, orc, monkeypatch, streamed):
    """Hub bands split into partial slots (a small work cap), and the same
    matrix streamed from host memory: the operator's outer D^-1/2, applied
    by the SpMM to each row's sum, also reaches rows written by the split
    reduction and by every streamed chunk."""
    monkeypatch.setenv("FMX_ITEM_CAP", "300")
    s, d = fmx.rmat_edges(13, 16 << 13, 77)
    flags = fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE
    spm = fmx.SparseMatrix.build(s, d, n=1 << 13, log2_h=8, log2_w=13, flags=flags)
    if streamed:
        data, offs = spm.download()
        spm = fmx.SparseMatrix.upload_streamed(data, offs, chunk_bytes=1 << 14)
        assert spm.streamed
    evals, vecs, resid, st = fmx.compute_eigen(spm, nev=16, tol=1e-6, precision=0)
    assert st["converged"] == 1 and np.all(resid <= 1e-6)
    img = orc.build_2d(s.cpu().numpy().view(np.uint32), d.cpu().numpy().view(np.uint32),
                       n=1 << 13, log2h=8, log2w=13, flags=orc.DEDUP | orc.SYMMETRIZE)
    N = normalized_operator(orc, img)
    ref = np.sort(sla.eigsh(N, k=16, which="LA", tol=1e-12, return_eigenvectors=False))[::-1]
    assert np.allclose(evals, ref, atol=2e-6)
    check_pairs(N, evals, vecs, 1e-6)


def test_warm_start_from_solution(fmx, orc):
    import torch
    p, q = 30, 33
    src, dst = torus_edges(p, q)
    flags = fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE
    spm = fmx.SparseMatrix.build(dev_u32(src), dev_u32(dst), n=p * q, log2_h=8, log2_w=10, flags=flags)
    ev, vecs, _, st0 = fmx.compute_eigen(spm, nev=16, block_size=16, tol=1e-8, precision=2,
                                         max_restarts=300)
    ev2, _, _, st1 = fmx.compute_eigen(spm, nev=16, block_size=16, tol=1e-8, precision=2,
                                       X0=vecs.contiguous())
    assert st1["converged"] == 1 and st1["spmm_blocks"] <= st0["spmm_blocks"]
    assert st1["restarts"] == 0
    assert np.allclose(ev, ev2, atol=1e-9)


def test_ill_conditioned_start_block(fmx, orc):
    """A start block whose 16 columns are one random vector plus 1e-4
    relative perturbations (condition number about 1e4-1e5, above the bound
    under which an FP32 block's Cholesky-QR stops after one pass) takes the
    second pass and must still give a converged, orthonormal solve that
    agrees with eigsh."""
    import torch
    s, d = fmx.rmat_edges(13, 16 << 13, 77)
    flags = fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE
    n = 1 << 13
    spm = fmx.SparseMatrix.build(s, d, n=n, log2_h=8, log2_w=13, flags=flags)
    rng = np.random.default_rng(5)
    base = rng.random(n)
    X0 = base[:, None] * (1.0 + 1e-4 * rng.standard_normal((n, 16)))
    G = X0.T @ X0
    assert np.linalg.cond(G) > 1e7  # condition of the block above 1e3.5
    X0 = torch.from_numpy(X0.astype(np.float32)).cuda()
    evals, vecs, resid, st = fmx.compute_eigen(spm, nev=16, tol=1e-6, precision=fmx.EIG_FP32, X0=X0,
                                               max_restarts=300)
    assert st["converged"] == 1 and np.all(resid <= 1e-6)
    img = orc.build_2d(s.cpu().numpy().view(np.uint32), d.cpu().numpy().view(np.uint32),
                       n=n, log2h=8, log2w=13, flags=orc.DEDUP | orc.SYMMETRIZE)
    N = normalized_operator(orc, img)
    ref = np.sort(sla.eigsh(N, k=16, which="LA", tol=1e-12, return_eigenvectors=False))[::-1]
    assert np.allclose(evals, ref, atol=2e-6)
    check_pairs(N, evals, vecs, 1e-6)


@pytest.mark.parametrize("num_blocks", [7, 8])
def test_largest_bases(fmx, orc, num_blocks):
    """Bases of 112 and 128 vectors (the largest Rayleigh-Ritz problem the
    one-workgroup Jacobi takes) against the torus's closed form in FP64."""
    p, q = 40, 37
    src, dst = torus_edges(p, q)
    flags = fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE
    spm = fmx.SparseMatrix.build(dev_u32(src), dev_u32(dst), n=p * q, log2_h=8, log2_w=10, flags=flags)
    nev, tol = 8, 1e-10
    evals, vecs, resid, st = fmx.compute_eigen(spm, nev=nev, block_size=16, num_blocks=num_blocks, tol=tol,
                                               precision=fmx.EIG_FP64, max_restarts=300)
    assert st["converged"] == 1
    lam = np.sort(((np.cos(2 * np.pi * np.arange(p) / p)[:, None] +
                    np.cos(2 * np.pi * np.arange(q) / q)[None, :]) / 2).ravel())[::-1]
    assert np.allclose(evals, lam[:nev], atol=10 * tol)
    img = orc.build_2d(src, dst, n=p * q, log2h=8, log2w=10, flags=orc.DEDUP | orc.SYMMETRIZE)
    check_pairs(normalized_operator(orc, img), evals, vecs, tol)
