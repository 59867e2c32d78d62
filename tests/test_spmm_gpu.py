"""GPU parity of fmx_spmm (HIP, through the C ABI) against the oracle.

Bars: integer operands bit-exact; FP64 within 1e-12 relative, FP32 within
1e-5 relative of the oracle's reference-order product (the tolerances the
path is specified with); rows without non-zeros exactly 0.
"""
import os

import numpy as np
import pytest

import tiny_kat

pytestmark = pytest.mark.gpu

TOL = {np.float32: 1e-5, np.float64: 1e-12}


def to_dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def rel_err(y, ref):
    y = np.asarray(y, np.float64)
    ref = np.asarray(ref, np.float64)
    zero = ref == 0
    assert np.all(y[zero] == 0), "rows without non-zeros must be exactly 0"
    d = np.abs(y - ref)[~zero] / np.abs(ref[~zero])
    return float(d.max()) if d.size else 0.0


def check(fmx, orc, img, X, spm=None):
    spm = spm or fmx.SparseMatrix.upload(img.data, img.offs)
    Y = spm.multiply(to_dev(X)).cpu().numpy()
    ref = orc.spmm(img, X)
    if np.issubdtype(X.dtype, np.integer):
        assert np.array_equal(Y, ref)
    else:
        e = rel_err(Y, ref)
        assert e <= TOL[X.dtype.type], e
    return Y, ref


@pytest.fixture(scope="module")
def rmat16(orc):
    # config 1: R-MAT scale 16, edge factor 16, dedup, self edges removed
    src, dst = orc.rmat_edges(16, 16 << 16, seed=2024)
    return src, dst


def test_tiny_kat(fmx, orc):
    img = orc.Image(np.frombuffer(tiny_kat.IMAGE, np.uint8), tiny_kat.OFFS, tiny_kat.N)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    assert spm.get_nnz() == 7 and spm.get_num_rows() == 6
    for dt in (np.int32, np.int64, np.float32, np.float64):
        Y = spm.multiply(to_dev(np.array(tiny_kat.X, dt))).cpu().numpy()
        assert np.array_equal(Y, np.array(tiny_kat.Y, dt)), dt
    # download round trip
    data, offs = spm.download()
    assert bytes(data) == tiny_kat.IMAGE and list(offs) == tiny_kat.OFFS


@pytest.mark.parametrize("lh,lw", [(15, 15), (10, 10), (8, 15), (4, 12)])
def test_config1_int32_exact_fp64(fmx, orc, rmat16, lh, lw):
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=lh, log2w=lw)
    rng = np.random.default_rng(7)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    Xi = rng.integers(0, 8, (img.n, 8)).astype(np.int32)
    check(fmx, orc, img, Xi, spm)
    Xd = rng.random((img.n, 8))
    check(fmx, orc, img, Xd, spm)


def test_seq_kat(fmx, orc, rmat16):
    # rand_mat_gen.cpp:178-199: A*[0..n-1] = sum of column ids per row
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=12, log2w=15)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    import torch
    Y = spm.multiply(torch.arange(img.n, dtype=torch.int64, device="cuda")).cpu().numpy()
    r, c = orc.nonzeros(img)
    want = np.bincount(r.astype(np.int64), weights=c.astype(np.float64), minlength=img.n)
    assert np.array_equal(Y, want.astype(np.int64))


def test_exact_fp_pattern(fmx, orc, rmat16):
    # test-2d_multiply.cpp:35-41: X[i][j] = (i%100)*(j+1) -> exact in FP
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=8, log2w=15)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    i = np.arange(img.n)[:, None]
    for k in (4, 16):
        Xi = (i % 100) * (np.arange(k)[None, :] + 1)
        for dt in (np.float32, np.float64):
            Y = spm.multiply(to_dev(Xi.astype(dt))).cpu().numpy()
            assert np.array_equal(Y, orc.spmm(img, Xi.astype(dt)))


@pytest.mark.parametrize("k", [1, 2, 3, 4, 7, 8, 16, 32, 64])
def test_widths(fmx, orc, rmat16, k):
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=8, log2w=15)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    rng = np.random.default_rng(k)
    check(fmx, orc, img, rng.random((img.n, k)).astype(np.float32), spm)
    check(fmx, orc, img, rng.random((img.n, k)), spm)
    check(fmx, orc, img, rng.integers(-9, 9, (img.n, k)).astype(np.int64), spm)
    check(fmx, orc, img, rng.integers(-9, 9, (img.n, k)).astype(np.int32), spm)


def test_strided_operands(fmx, orc, rmat16):
    import torch
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=8, log2w=15)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    rng = np.random.default_rng(5)
    Xfull = rng.random((img.n, 24)).astype(np.float32)
    Xd = torch.from_numpy(Xfull).cuda()[:, 3:19]  # ldx = 24, misaligned start
    Yfull = torch.full((img.n, 40), -1.0, device="cuda")
    Y = Yfull[:, 8:24]
    spm.multiply_async(Xd, Y)
    torch.cuda.synchronize()
    ref = orc.spmm(img, np.ascontiguousarray(Xfull[:, 3:19]))
    assert rel_err(Y.cpu().numpy(), ref) <= 1e-5
    assert torch.all(Yfull[:, :8] == -1) and torch.all(Yfull[:, 24:] == -1)


def test_split_bands_and_determinism(fmx, orc, rmat16):
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=10, log2w=15)
    rng = np.random.default_rng(9)
    X = rng.random((img.n, 16)).astype(np.float32)
    Xi = rng.integers(0, 8, (img.n, 16)).astype(np.int32)
    os.environ["FMX_ITEM_CAP"] = "300"  # force every band to be split
    try:
        spm = fmx.SparseMatrix.upload(img.data, img.offs)
        Y1, _ = check(fmx, orc, img, X, spm)
        check(fmx, orc, img, Xi, spm)
        # widths drained in 256-word chunks
        check(fmx, orc, img, rng.random((img.n, 32)).astype(np.float32), spm)
        check(fmx, orc, img, rng.random((img.n, 16)), spm)
        check(fmx, orc, img, rng.integers(-9, 9, (img.n, 16)).astype(np.int64), spm)
    finally:
        del os.environ["FMX_ITEM_CAP"]
    spm2 = fmx.SparseMatrix.upload(img.data, img.offs)
    Ya = spm2.multiply(to_dev(X)).cpu().numpy()
    Yb = spm2.multiply(to_dev(X)).cpu().numpy()
    assert np.array_equal(Ya, Yb)  # deterministic call to call
    assert rel_err(Y1, Ya) <= 1e-5


def test_empty_rows_and_ragged(fmx, orc):
    # n not a multiple of H, isolated vertices, empty block rows
    rng = np.random.default_rng(3)
    n = 5000
    src = rng.integers(0, 1200, 20000).astype(np.uint32)
    dst = rng.integers(0, n - 7, 20000).astype(np.uint32)
    img = orc.build_2d(src, dst, n=n, log2h=7, log2w=9)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    X = rng.random((n, 5))
    Y, ref = check(fmx, orc, img, X, spm)
    assert np.all(Y[1200:] == 0)


def test_errors(fmx, orc):
    import torch
    img = orc.Image(np.frombuffer(tiny_kat.IMAGE, np.uint8), tiny_kat.OFFS, tiny_kat.N)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    with pytest.raises(fmx.FmxError) as e:
        spm.multiply(torch.zeros((5, 2), device="cuda"))
    assert e.value.status == fmx.EDIM
    assert not spm.multiply_into(torch.zeros((6, 2), device="cuda"),
                                 torch.zeros((6, 2), device="cuda", dtype=torch.float64))
    with pytest.raises(fmx.FmxError) as e:
        spm.multiply(torch.zeros((6, 2), device="cuda", dtype=torch.int16))
    assert e.value.status == fmx.ETYPE
    bad = bytearray(tiny_kat.IMAGE)
    bad[4096 + 16 + 2] = 0x05
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.upload(bytes(bad), tiny_kat.OFFS)
    assert e.value.status == fmx.EFORMAT
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.upload(tiny_kat.IMAGE, tiny_kat.OFFS[:-1])
    assert e.value.status == fmx.EFORMAT
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.upload(tiny_kat.BAD_EMPTY_IMAGE, tiny_kat.BAD_EMPTY_OFFS)
    assert e.value.status == fmx.EFORMAT


def test_rmat_generator_matches_oracle(fmx, orc):
    for scale, m, seed, perm in ((16, 1 << 20, 2024, True), (9, 5000, 1, False), (21, 333, 5, True)):
        s, d = fmx.rmat_edges(scale, m, seed, permute=perm)
        so, do = orc.rmat_edges(scale, m, seed, permute=perm)
        assert np.array_equal(s.cpu().numpy().view(np.uint32), so)
        assert np.array_equal(d.cpu().numpy().view(np.uint32), do)


def test_row_32767_header(fmx, orc):
    from test_oracle import last_row_edges
    src, dst = last_row_edges()
    img = orc.build_2d(src, dst, n=40001, log2h=15, log2w=15)
    X = np.arange(40001 * 2, dtype=np.int64).reshape(-1, 2)
    check(fmx, orc, img, X)
    spm = fmx.SparseMatrix.build(to_dev(src.view(np.int32)), to_dev(dst.view(np.int32)), n=40001,
                                 log2_h=15, log2_w=15, flags=fmx.BUILD_DEDUP)
    data, offs = spm.download()
    assert np.array_equal(data, img.data) and np.array_equal(offs, img.offs)


def test_fp32_within_bound_of_either_tile_order(fmx, orc, rmat16):
    """With hilbert_order the reference visits each task's tiles along a
    Hilbert curve (sparse_matrix.cpp:120-186) and its FP32 row sums change
    rounding; the FP64-accumulated GPU result is within 1e-5 of the
    reference's sums in both orders."""
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=8, log2w=8)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    X = np.random.default_rng(21).random((img.n, 16)).astype(np.float32)
    Y = spm.multiply(to_dev(X)).cpu().numpy()
    for hil in (False, True):
        ref = orc.spmm_exec(img, X, 16, hil)
        assert rel_err(Y, ref) <= 1e-5, hil


@pytest.mark.parametrize("lh", [6, 7])
def test_config2_rmat22_spmm_parity(fmx, orc, lh):
    """The scale-22 workload itself: R-MAT scale 22, edge factor 16, dedup,
    in the autotuned 64 x 32768 tiles and in 128 x 32768 tiles, times
    4,194,304 x 16 operands; the FP32 operand is the benchmark's own
    (torch.rand on the device, seed 22).
    - int32: bit-exact.
    - FP32, sequential mode (FMX_SPMM_SEQ): bit-identical to the oracle's
      reference-order FP32 sums.
    - FP32, default mode (FP64 sums rounded once): within 2^-24 relative of
      the exact sums (the FP64 sums of FP32 terms, exact here: multiples of
      2^-24 below 2^29), and elementwise within |ref - exact| + 2^-24 |exact|
      of the reference's FP32 sums, whose own error reaches about 1.3e-5 on
      this operand's hub rows.
    - FP64: within 1e-12 (default) and bit-identical (sequential).
    Rows without non-zeros are exactly 0 in every mode."""
    import torch
    n = 1 << 22
    s, d = fmx.rmat_edges(22, 16 << 22, 22)
    spm = fmx.SparseMatrix.build(s, d, n=n, log2_h=lh, log2_w=15, flags=fmx.BUILD_DEDUP)
    del s, d
    data, offs = spm.download()
    img = orc.Image(data, offs, n)
    rng = np.random.default_rng(22 + lh)
    Xi = rng.integers(0, 8, (n, 16)).astype(np.int32)
    assert np.array_equal(spm.multiply(to_dev(Xi)).cpu().numpy(), orc.spmm(img, Xi, threads=8))
    del Xi
    g = torch.Generator(device="cuda")
    g.manual_seed(22)
    Xt = torch.rand((n, 16), dtype=torch.float32, device="cuda", generator=g)
    Xf = Xt.cpu().numpy()
    ref = orc.spmm(img, Xf, threads=8)
    exact = orc.spmm(img, Xf.astype(np.float64), threads=8)
    Y = spm.multiply(Xt).cpu().numpy()
    assert rel_err(Y, exact) <= 2.0 ** -24
    nz = exact != 0
    bound = np.abs(ref.astype(np.float64) - exact) + 2.0 ** -24 * np.abs(exact)
    assert np.all(np.abs(Y.astype(np.float64) - ref)[nz] <= bound[nz])
    Ys = torch.empty_like(Xt)
    spm.multiply_async(Xt, Ys, seq=True)
    assert np.array_equal(Ys.cpu().numpy(), ref)
    del Xt, Xf, Y, Ys, ref, exact, bound, nz
    Xd = rng.random((n, 16))
    refd = orc.spmm(img, Xd, threads=8)
    Xdt = to_dev(Xd)
    assert rel_err(spm.multiply(Xdt).cpu().numpy(), refd) <= 1e-12
    Yd = torch.empty_like(Xdt)
    spm.multiply_async(Xdt, Yd, seq=True)
    assert np.array_equal(Yd.cpu().numpy(), refd)
