"""GPU parity of the tall-skinny dense operations against the oracle
(oracle/fmx_oracle.c: orc_gram_f64, orc_gemm_tall_f64, orc_colnorm_f64)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def bf16_round(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("n,ka,kb", [(100003, 128, 16), (5000, 16, 16), (777, 20, 3), (65536, 64, 64),
                                     (1, 5, 7), (1000003, 96, 16), (70001, 33, 50), (4099, 250, 16),
                                     (3001, 70, 20), (513, 250, 40), (20011, 40, 16), (65537, 80, 9),
                                     (3001, 48, 1)])
def test_gram_f64(fmx, orc, n, ka, kb):
    rng = np.random.default_rng(n + ka)
    for dt in (np.float32, np.float64):
        A = rng.standard_normal((n, ka)).astype(dt)
        B = rng.standard_normal((n, kb)).astype(dt)
        G = fmx.gram(dev(A), dev(B)).cpu().numpy()
        ref = orc.gram_f64(A.astype(np.float64), B.astype(np.float64))
        scale = np.abs(A.astype(np.float64)).T @ np.abs(B.astype(np.float64))
        assert np.all(np.abs(G - ref) <= 1e-13 * scale + 1e-300)


@pytest.mark.parametrize("n,ka,kb", [(100003, 128, 16), (4096, 16, 16), (999, 20, 3), (65536, 64, 48)])
def test_gram_bf16_mfma(fmx, orc, n, ka, kb):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, ka)).astype(np.float32)
    B = rng.standard_normal((n, kb)).astype(np.float32)
    G = fmx.gram(dev(A), dev(B), mode=fmx.GRAM_BF16).cpu().numpy()
    # products of BF16 values are exact in FP32; FP32 accumulation bounds the rest
    Ab = bf16_round(A).astype(np.float64)
    Bb = bf16_round(B).astype(np.float64)
    ref = orc.gram_f64(Ab, Bb)
    scale = np.abs(Ab).T @ np.abs(Bb)
    assert np.all(np.abs(G - ref) <= 2e-6 * scale)
    # and close to the FP64 Gram of the unrounded data (BF16 rounding ~ 2^-9)
    exact = orc.gram_f64(A.astype(np.float64), B.astype(np.float64))
    assert np.all(np.abs(G - exact) <= 1e-2 * scale)


@pytest.mark.parametrize("cols", [72, 96, 16])
def test_gram_bf16_strided_window(fmx, orc, cols):
    """BF16 Gram of a column window of a wider row-major basis (the
    eigensolver's first projection pass: lda = 96, kb = 16)."""
    rng = np.random.default_rng(cols)
    Vfull = rng.standard_normal((70001, 96)).astype(np.float32)
    W = rng.standard_normal((70001, 16)).astype(np.float32)
    G = fmx.gram(dev(Vfull)[:, :cols], dev(W), mode=fmx.GRAM_BF16).cpu().numpy()
    Ab = bf16_round(Vfull[:, :cols]).astype(np.float64)
    Bb = bf16_round(W).astype(np.float64)
    ref = orc.gram_f64(Ab, Bb)
    scale = np.abs(Ab).T @ np.abs(Bb)
    assert np.all(np.abs(G - ref) <= 2e-6 * scale)


@pytest.mark.parametrize("n,k,ld", [(100003, 16, 16), (4099, 32, 32), (70001, 16, 96), (777, 32, 40)])
def test_gram_f64_symmetric(fmx, orc, n, k, ld):
    """A^T A with B the same operand (one staged copy, the Cholesky-QR Grams),
    also for column windows of wider arrays; equal to the general kernel's
    result on a copy of A."""
    rng = np.random.default_rng(n + k)
    for dt in (np.float32, np.float64):
        Afull = rng.standard_normal((n, ld)).astype(dt)
        A = dev(Afull)[:, :k]
        G = fmx.gram(A, A).cpu().numpy()
        Gc = fmx.gram(A, A.clone()).cpu().numpy()  # distinct pointer: the general kernel
        ref = orc.gram_f64(Afull[:, :k].astype(np.float64), Afull[:, :k].astype(np.float64))
        scale = np.abs(Afull[:, :k].astype(np.float64)).T @ np.abs(Afull[:, :k].astype(np.float64))
        assert np.all(np.abs(G - ref) <= 1e-13 * scale)
        assert np.array_equal(G, Gc)


def test_gram_f64_strided(fmx, orc):
    """Operands that are column windows of wider row-major arrays (lda > ka)."""
    import torch
    rng = np.random.default_rng(4)
    Vfull = rng.standard_normal((30011, 96)).astype(np.float32)
    W = rng.standard_normal((30011, 16)).astype(np.float32)
    V = dev(Vfull)[:, :72]
    G = fmx.gram(V, dev(W)).cpu().numpy()
    ref = orc.gram_f64(Vfull[:, :72].astype(np.float64), W.astype(np.float64))
    scale = np.abs(Vfull[:, :72].astype(np.float64)).T @ np.abs(W.astype(np.float64))
    assert np.all(np.abs(G - ref) <= 1e-13 * scale)


@pytest.mark.parametrize("n,ka,kc", [(100003, 128, 16), (3000, 16, 48), (10, 3, 5), (1000003, 96, 16),
                                     (20011, 7, 40), (5003, 96, 80),
                                     # FP32 whole-row vector loads at each staged width, a width
                                     # short of its 16-column tile, and fewer rows than a stage
                                     (65537, 16, 16), (4099, 32, 12), (30001, 64, 16), (777, 60, 16),
                                     (13, 16, 16)])
def test_gemm_tall(fmx, orc, n, ka, kc):
    rng = np.random.default_rng(ka * kc)
    C = rng.standard_normal((ka, kc))
    for dt in (np.float32, np.float64):
        A = rng.standard_normal((n, ka)).astype(dt)
        Y0 = rng.standard_normal((n, kc)).astype(dt)
        Y = fmx.gemm_tall(dev(A), dev(C), alpha=0.5, beta=-2.0, Y=dev(Y0)).cpu().numpy()
        ref = orc.gemm_tall_f64(A.astype(np.float64), C, 0.5, -2.0, Y0.astype(np.float64).copy())
        tol = 1e-6 if dt == np.float32 else 1e-12
        scale = 0.5 * np.abs(A.astype(np.float64)) @ np.abs(C) + 2 * np.abs(Y0)
        assert np.all(np.abs(Y - ref) <= tol * scale)


@pytest.mark.parametrize("n,ka,kc", [(70001, 96, 32), (4099, 64, 20), (33, 16, 24), (20011, 128, 32)])
def test_gemm_tall_new_output(fmx, orc, n, ka, kc):
    """Y = A C into a fresh output (beta = 0, no old values read), two
    16-column output tiles, against the oracle's FP64 product."""
    rng = np.random.default_rng(n + ka * kc)
    C = rng.standard_normal((ka, kc))
    for dt in (np.float32, np.float64):
        A = rng.standard_normal((n, ka)).astype(dt)
        Y = fmx.gemm_tall(dev(A), dev(C)).cpu().numpy()
        ref = orc.gemm_tall_f64(A.astype(np.float64), C, 1.0, 0.0, np.zeros((n, kc)))
        tol = 1e-6 if dt == np.float32 else 1e-12
        scale = np.abs(A.astype(np.float64)) @ np.abs(C)
        assert np.all(np.abs(Y - ref) <= tol * scale + 1e-300)


@pytest.mark.parametrize("m,ka,kc", [(96, 96, 16), (96, 80, 32), (96, 96, 48), (33, 33, 8), (16, 16, 16)])
def test_gemm_tall_in_place(fmx, orc, m, ka, kc):
    """Y = A C written over A's own first kc columns (the thick restart's
    V = V S): every row is read in full before it is overwritten."""
    import torch
    rng = np.random.default_rng(m + ka + kc)
    C = rng.standard_normal((ka, kc))
    for dt in (np.float32, np.float64):
        T0 = rng.standard_normal((40009, m)).astype(dt)
        T = dev(T0)
        fmx.gemm_tall(T[:, :ka], dev(C), Y=T[:, :kc])
        got = T.cpu().numpy()
        ref = T0[:, :ka].astype(np.float64) @ C
        tol = 1e-6 if dt == np.float32 else 1e-12
        scale = np.abs(T0[:, :ka].astype(np.float64)) @ np.abs(C)
        assert np.all(np.abs(got[:, :kc] - ref) <= tol * scale + 1e-300)
        assert np.array_equal(got[:, kc:], T0[:, kc:])


def test_colnorm_and_row_scale(fmx, orc):
    rng = np.random.default_rng(0)
    for dt in (np.float32, np.float64):
        A = rng.standard_normal((123457, 16)).astype(dt)
        nrm = fmx.colnorm(dev(A)).cpu().numpy()
        assert np.allclose(nrm, orc.colnorm_f64(A.astype(np.float64)), rtol=1e-13)
        d = rng.standard_normal(123457)
        Y = fmx.row_scale(dev(d), dev(A)).cpu().numpy()
        assert np.array_equal(Y, (A.astype(np.float64) * d[:, None]).astype(dt))
