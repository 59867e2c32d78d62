"""CPU pins of the learned-warm-start checker (oracle): the network's
gradient against central finite differences, the subspace loss against its
closed form, and the DC-SBM generator's invariants."""
import numpy as np
import pytest
import scipy.sparse as sp


def tiny_graph(n=60, seed=0):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=0.08, random_state=seed, format="csr")
    A = ((A + A.T) > 0).astype(np.float64)
    A.setdiag(0)
    A.eliminate_zeros()
    A = A.tocsr()
    deg = np.asarray(A.sum(1)).ravel()
    dinv = np.where(deg > 0, 1 / np.sqrt(np.maximum(deg, 1)), 0.0)
    return A, dinv


def test_gnn_gradient_finite_differences(orc):
    A, dinv = tiny_graph()
    n = A.shape[0]
    rng = np.random.default_rng(1)
    X0 = rng.standard_normal((n, 16))
    U, _ = np.linalg.qr(rng.standard_normal((n, 16)))
    npar = orc.gnn_param_count()
    assert npar == 2 * 16 * 64 + 64 + 2 * (2 * 64 * 64 + 64) + 64 * 16
    P = rng.standard_normal(npar) * 0.2
    loss, g, _ = orc.gnn_loss_grad(A.indptr, A.indices, dinv, X0, U, P)
    assert 0 <= loss <= 16
    h = 1e-6
    idx = rng.choice(npar, 40, replace=False)
    for i in idx:
        Pp = P.copy(); Pp[i] += h
        Pm = P.copy(); Pm[i] -= h
        fd = (orc.gnn_loss_grad(A.indptr, A.indices, dinv, X0, U, Pp, False)[0] -
              orc.gnn_loss_grad(A.indptr, A.indices, dinv, X0, U, Pm, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_subspace_loss_closed_form(orc):
    rng = np.random.default_rng(2)
    n, k = 200, 16
    U, _ = np.linalg.qr(rng.standard_normal((n, k)))
    # Y spanning U exactly (any invertible mixing) -> loss 0
    Y = U @ rng.standard_normal((k, k))
    assert abs(orc.subspace_loss(Y, U)[0]) < 1e-9
    # Y orthogonal to U -> loss k
    Q, _ = np.linalg.qr(np.hstack([U, rng.standard_normal((n, k))]))
    assert abs(orc.subspace_loss(Q[:, k:], U)[0] - k) < 1e-9
    # general: k - ||Q^T U||_F^2 with Q an orthonormal basis of Y
    Y = rng.standard_normal((n, k))
    Qy, _ = np.linalg.qr(Y)
    loss, dY = orc.subspace_loss(Y, U)
    assert abs(loss - (k - np.linalg.norm(Qy.T @ U) ** 2)) < 1e-9
    # gradient by central differences
    for _ in range(10):
        i, j = rng.integers(n), rng.integers(k)
        Yp = Y.copy(); Yp[i, j] += 1e-6
        Ym = Y.copy(); Ym[i, j] -= 1e-6
        fd = (orc.subspace_loss(Yp, U)[0] - orc.subspace_loss(Ym, U)[0]) / 2e-6
        assert abs(fd - dY[i, j]) < 1e-6


def test_dcsbm_invariants(orc):
    import flashx_amd.fmx as F
    n, B = 1 << 14, 32
    bstart, wtab = F.dcsbm_params(n, B, 2.5, 1.06)
    assert bstart[0] == 0 and bstart[-1] == n and np.all(np.diff(bstart.astype(np.int64)) > 0)
    assert np.all(np.diff(wtab.astype(np.int64)) >= 0) and wtab[0] >= 1024
    m = 8 * n
    s1, d1 = orc.dcsbm_edges(n, bstart, wtab, 20, m, 9)
    s2, d2 = orc.dcsbm_edges(n, bstart, wtab, 20, m, 9)
    assert np.array_equal(s1, s2) and np.array_equal(d1, d2)
    assert s1.max() < n and d1.max() < n
    # undo the relabelling to recover blocks: intra-block fraction ~ ratio-weighted
    inv = np.empty(n, np.int64)
    perm = np.array([orc.rmat_perm(v, 14, 9) for v in range(n)])
    inv[perm] = np.arange(n)
    blk = np.searchsorted(bstart, inv, side="right") - 1
    intra = np.mean(blk[s1] == blk[d1])
    assert 0.3 < intra < 0.5, intra  # 20x odds over 31 other blocks -> ~0.39


def test_bf16_rounding_matches_torch(orc):
    """The checker's BF16 rounding against torch's float32 -> bfloat16
    conversion (round to nearest even), including exact ties."""
    import torch
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.standard_normal(5000) * 10.0 ** rng.integers(-6, 6, 5000),
                        [1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -(1 + 2.0 ** -8), 0.0, -0.0, 65504.0]])
    x = x.astype(np.float32).astype(np.float64)
    want = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    got = orc.bf16_round(x)
    assert np.array_equal(got, want)
    assert got[-6] == 1.0 and got[-5] == 1 + 2.0 ** -6  # ties to even


def test_gnn_bf16_operand_mode(orc):
    """Rounding the product operands to BF16 moves the output by a BF16-sized
    amount; the gradient moves more (the subspace loss inverts Y^T Y), which
    is why the GPU's BF16 mode is checked against this mode, not FP64."""
    A, dinv = tiny_graph()
    n = A.shape[0]
    rng = np.random.default_rng(5)
    X0 = rng.standard_normal((n, 16))
    U, _ = np.linalg.qr(rng.standard_normal((n, 16)))
    P = rng.standard_normal(orc.gnn_param_count()) * 0.2
    l64, g64, Y64 = orc.gnn_loss_grad(A.indptr, A.indices, dinv, X0, U, P)
    l16, g16, Y16 = orc.gnn_loss_grad(A.indptr, A.indices, dinv, X0, U, P, bf16=True)
    rel = np.abs(Y16 - Y64).max() / np.abs(Y64).max()
    assert 1e-5 < rel < 5e-2
    assert abs(l16 - l64) < 5e-2 * max(1.0, abs(l64))
    assert 1e-4 * np.linalg.norm(g64) < np.linalg.norm(g16 - g64) < 0.3 * np.linalg.norm(g64)
