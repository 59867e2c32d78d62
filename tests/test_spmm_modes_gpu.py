"""GPU parity of the SpMM's options, through the C ABI:

- the sequential floating-point order (FMX_SPMM_SEQ): bit-identical to the
  oracle's reference-order sums (row_part_func / coo_func,
  matrix/sparse_matrix.h:219-331) for FP32 and FP64, binary and weighted
  images, every tile shape, runtime widths, split-band plans and streamed
  images;
- column-major X and Y (matrix/sparse_matrix.cpp:720-736; the L_COL output
  of matrix/test/test-eigen.cpp:115-118): the same bits as row-major;
- size_t operands (spmm_creator<size_t, size_t>, sparse_matrix.cpp:568-569):
  sums modulo 2^64, checked against Python integers;
- the scaled SpMM of a custom task creator (fmx_spmm_scaled, the
  normalised-adjacency operator of test-eigen.cpp:112-127, including
  spmm_creator<double, float> over FP32 values);
- operand views that overlap (an expanded X) are refused before any launch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(15, 15), (6, 15), (10, 10), (8, 12), (4, 12)]


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def col_major(t):
    """The same matrix stored column-major (element (i, j) at j * rows + i)."""
    return t.t().contiguous().t()


@pytest.fixture(scope="module")
def rmat16(orc):
    return orc.rmat_edges(16, 16 << 16, seed=2024)


def run(spm, X, seq=False, ycol=False):
    import torch
    Xt = X if isinstance(X, torch.Tensor) else dev(X)
    Y = torch.empty((spm.get_num_rows(), Xt.shape[1]), dtype=Xt.dtype, device="cuda")
    if ycol:
        Y = col_major(Y)
    spm.multiply_async(Xt, Y, seq=seq)
    torch.cuda.synchronize()
    return Y.cpu().numpy()


@pytest.mark.parametrize("lh,lw", SHAPES)
def test_seq_bit_exact_config1(fmx, orc, rmat16, lh, lw):
    """R-MAT scale 16 (the CPU-runnable configuration) in five tile shapes,
    including the reference's default 32768 x 32768: FP32 and FP64 at
    compile-time widths 1, 8, 16 and the runtime width 3."""
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=lh, log2w=lw)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    rng = np.random.default_rng(lh * 100 + lw)
    for dt in (np.float32, np.float64):
        for k in (1, 3, 8, 16):
            X = rng.random((img.n, k)).astype(dt)
            assert np.array_equal(run(spm, X, seq=True), orc.spmm(img, X)), (dt, k)


def test_seq_ignores_split_plans_and_streams(fmx, orc, rmat16, monkeypatch):
    """With every band split (a work cap of 300 words) the default mode sums
    pieces in a fixed order; the sequential mode keeps one item per band and
    stays bit-identical, also on a streamed image (4 KiB chunks)."""
    monkeypatch.setenv("FMX_ITEM_CAP", "300")
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=10, log2w=15)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    st = fmx.SparseMatrix.upload_streamed(img.data, img.offs, chunk_bytes=4096)
    X = np.random.default_rng(3).random((img.n, 16)).astype(np.float32)
    ref = orc.spmm(img, X)
    assert np.array_equal(run(spm, X, seq=True), ref)
    assert np.array_equal(run(st, X, seq=True), ref)
    exact = orc.spmm(img, X.astype(np.float64))
    nz = exact != 0
    d = run(spm, X).astype(np.float64)
    assert np.max(np.abs(d - exact)[nz] / exact[nz]) <= 2.0 ** -24


@pytest.mark.parametrize("vdt", [np.float32, np.float64])
def test_seq_weighted(fmx, orc, rmat16, vdt):
    """Weighted images: each product is rounded in the data type before it
    is added (src_row[j] * data), compile-time and runtime widths."""
    src, dst = rmat16
    vals = (np.random.default_rng(5).random(src.size) + 0.5).astype(vdt)
    img = orc.build_2d_w(src, dst, vals, n=1 << 16, log2h=7, log2w=15)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    rng = np.random.default_rng(6)
    for k in (4, 5, 16):
        X = rng.random((img.n, k)).astype(vdt)
        ref = orc.spmm(img, X)
        assert np.array_equal(run(spm, X, seq=True), ref), k
        # the default mode: FP64 products and sums, one rounding
        exact = orc.spmm(img, X.astype(np.float64))
        y = run(spm, X).astype(np.float64)
        nz = exact != 0
        tol = 2.0 ** -24 + 1e-15 if vdt == np.float32 else 1e-12
        assert np.max(np.abs(y - exact)[nz] / exact[nz]) <= tol, k


def test_mixed_creator_double_float(fmx, orc, rmat16):
    """spmm_creator<double, float>: FP64 data over an image with FP32 values
    (test-eigen.cpp:47-49,122-124); each product is the FP64 product of the
    element and the widened value.  Plain multiply refuses the pairing, as
    the reference's type check does (sparse_matrix.cpp:754-766)."""
    import torch
    src, dst = rmat16
    vals = np.random.default_rng(8).random(src.size).astype(np.float32)
    img = orc.build_2d_w(src, dst, vals, n=1 << 16, log2h=8, log2w=15)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    X = np.random.default_rng(9).random((img.n, 16))
    ref = orc.spmm(img, X)  # the oracle widens stored FP32 values to the FP64 data type
    Xt = dev(X)
    Y = torch.empty_like(Xt)
    with pytest.raises(fmx.FmxError):
        spm.multiply_async(Xt, Y)
    spm.multiply_scaled_async(Xt, Y, seq=True)
    assert np.array_equal(Y.cpu().numpy(), ref)
    spm.multiply_scaled_async(Xt, Y)
    nz = ref != 0
    assert np.max(np.abs(Y.cpu().numpy() - ref)[nz] / np.abs(ref[nz])) <= 1e-12


@pytest.mark.parametrize("cap", [None, "300"])
def test_column_major_operands(fmx, orc, rmat16, cap, monkeypatch):
    """Column-major X is restaged row-major and column-major Y is written in
    place; both give the row-major result bit for bit (int32, FP32, FP64,
    compile-time and runtime widths), also when bands are split and reduced
    (cap 300) and in the sequential mode."""
    if cap:
        monkeypatch.setenv("FMX_ITEM_CAP", cap)
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=8, log2w=15)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    rng = np.random.default_rng(11)
    for X in (rng.integers(0, 8, (img.n, 8)).astype(np.int32), rng.random((img.n, 16)).astype(np.float32),
              rng.random((img.n, 5))):
        Xt = dev(X)
        base = run(spm, Xt)
        for xc, yc in ((True, False), (False, True), (True, True)):
            got = run(spm, col_major(Xt) if xc else Xt, ycol=yc)
            assert np.array_equal(got, base), (X.dtype, xc, yc)
        if X.dtype != np.int32:
            assert np.array_equal(run(spm, col_major(Xt), seq=True, ycol=True), orc.spmm(img, X))


def test_size_t_operands(fmx, orc):
    """size_t data sum modulo 2^64 (spmm_creator<size_t, size_t>): values
    near 2^63 wrap, checked against exact Python integers; a P_ULONG
    weighted image multiplies the same way, and its bytes match the
    oracle's weighted writer."""
    import torch
    rng = np.random.default_rng(12)
    n = 3000
    src = rng.integers(0, n, 20000).astype(np.uint32)
    dst = rng.integers(0, n, 20000).astype(np.uint32)
    img = orc.build_2d(src, dst, n=n, log2h=6, log2w=10)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    X = rng.integers(0, 2 ** 63, (n, 4), dtype=np.uint64) * np.uint64(3)  # wraps modulo 2^64
    rows, cols = orc.nonzeros(img)[:2]
    want = [[0] * 4 for _ in range(n)]
    for r, c in zip(rows.tolist(), cols.tolist()):
        for j in range(4):
            want[r][j] = (want[r][j] + int(X[c, j])) % (1 << 64)
    got = run(spm, dev(X))
    assert got.dtype == np.uint64
    assert [[int(v) for v in row] for row in got] == want
    # weighted, P_ULONG values
    vals = rng.integers(0, 2 ** 30, src.size, dtype=np.uint64)
    wimg = orc.build_2d_w(src, dst, vals, n=n, log2h=6, log2w=10)
    wspm = fmx.SparseMatrix.build_weighted(dev(src.view(np.int32)), dev(dst.view(np.int32)),
                                          torch.from_numpy(vals).cuda(), n=n, log2_h=6, log2_w=10,
                                          flags=fmx.BUILD_DEDUP)
    data, offs = wspm.download()
    assert np.array_equal(data, wimg.data) and np.array_equal(offs, wimg.offs)
    Xs = rng.integers(0, 2 ** 20, (n, 4), dtype=np.uint64)
    assert np.array_equal(run(wspm, dev(Xs)), orc.spmm(wimg, Xs))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_scaled_creator(fmx, orc, rmat16, dt):
    """Y = diag(post) A diag(pre) X: pre is applied to X with one rounding
    (staged), post to each row's sum before its rounding; in the sequential
    mode the reference-order sum is rounded, then scaled."""
    import torch
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=6, log2w=15, flags=orc.DEDUP)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    rng = np.random.default_rng(13)
    pre = rng.random(img.n) + 0.5
    post = rng.random(img.n) + 0.5
    for k in (1, 5, 16):
        X = rng.random((img.n, k)).astype(dt)
        Xs = (X.astype(np.float64) * pre[:, None]).astype(dt)
        Xt = dev(X)
        Y = torch.empty_like(Xt)
        spm.multiply_scaled_async(Xt, Y, pre=dev(pre), post=dev(post), seq=True)
        seq_ref = (orc.spmm(img, Xs).astype(np.float64) * post[:, None]).astype(dt)
        assert np.array_equal(Y.cpu().numpy(), seq_ref), k
        spm.multiply_scaled_async(col_major(Xt), Y, pre=dev(pre), post=dev(post))
        exact = orc.spmm(img, Xs.astype(np.float64)) * post[:, None]
        nz = exact != 0
        tol = 2.0 ** -24 + 1e-15 if dt == np.float32 else 1e-12
        assert np.max(np.abs(Y.cpu().numpy() - exact)[nz] / exact[nz]) <= tol, k
        assert np.all(Y.cpu().numpy()[~nz] == 0)


def test_overlapping_views_refused(fmx, orc, rmat16):
    """An expanded X (row stride 0) or Y, or an expanded Gram operand, is
    refused with EINVAL before anything runs."""
    import torch
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=8, log2w=15)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    Xe = torch.zeros(1, 16, device="cuda").expand(img.n, 16)
    Y = torch.empty(img.n, 16, device="cuda")
    for a, b in ((Xe, Y), (torch.zeros(img.n, 16, device="cuda"), torch.zeros(1, 16, device="cuda").expand(img.n, 16))):
        with pytest.raises(fmx.FmxError) as e:
            spm.multiply_async(a, b)
        assert e.value.status == fmx.EINVAL
    with pytest.raises(fmx.FmxError) as e:
        fmx.gram(Xe, Y)
    assert e.value.status == fmx.EINVAL
