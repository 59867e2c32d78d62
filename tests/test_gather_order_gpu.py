"""GPU parity of the gather-order image (flashx_amd/csrc/hotcols.cpp,
image_dev.hip): the
default-mode SpMM of a matrix whose X is large runs on a copy of the matrix
with its most used columns relabelled into a dense front block of X that is
restaged every call.  FMX_HOT_MIN_BYTES=0 makes small matrices take that
path too.  Bars as for the plain path: integers bit-exact; FP32 within one
rounding of the exact (FP64) sums; FP64 within 1e-12; rows without
non-zeros exactly 0; streamed equal to resident; the sequential mode
unchanged (it never uses the relabelled image)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def rel(y, ref):
    y = np.asarray(y, np.float64)
    ref = np.asarray(ref, np.float64)
    z = ref == 0
    assert np.all(y[z] == 0)
    return float((np.abs(y - ref)[~z] / np.abs(ref[~z])).max())


@pytest.fixture
def forced(monkeypatch):
    monkeypatch.setenv("FMX_HOT_MIN_BYTES", "0")


@pytest.fixture(scope="module")
def rmat16(orc):
    return orc.rmat_edges(16, 16 << 16, seed=2024)


@pytest.mark.parametrize("lh,lw,cap", [(6, 15, None), (15, 15, None), (8, 12, "300"), (4, 10, None)])
def test_gather_order_matches_oracle(fmx, orc, rmat16, forced, monkeypatch, lh, lw, cap):
    if cap:
        monkeypatch.setenv("FMX_ITEM_CAP", cap)
    src, dst = rmat16
    img = orc.build_2d(src, dst, n=1 << 16, log2h=lh, log2w=lw)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    rng = np.random.default_rng(lh + 10 * lw)
    for k in (1, 4, 8, 16):
        Xi = rng.integers(-9, 10, (img.n, k)).astype(np.int32)
        assert np.array_equal(spm.multiply(dev(Xi)).cpu().numpy(), orc.spmm(img, Xi)), k
        Xf = rng.random((img.n, k)).astype(np.float32)
        assert rel(spm.multiply(dev(Xf)).cpu().numpy(), orc.spmm(img, Xf.astype(np.float64))) <= 2.0 ** -24
        Xd = rng.random((img.n, 3))  # 24-byte rows: the runtime-width path
        assert rel(spm.multiply(dev(Xd)).cpu().numpy(), orc.spmm(img, Xd)) <= 1e-12
        # the sequential mode keeps the reference's column order
        import torch
        Y = torch.empty((img.n, k), dtype=torch.float32, device="cuda")
        spm.multiply_async(dev(Xf), Y, seq=True)
        assert np.array_equal(Y.cpu().numpy(), orc.spmm(img, Xf))


def test_gather_order_weighted_streamed_layouts(fmx, orc, rmat16, forced):
    """Weighted values follow their entries; a streamed matrix streams its own
    gather-order image and gives the resident bits; column-major operands and
    row scales go through the same path."""
    import torch
    src, dst = rmat16
    vals = (np.random.default_rng(4).random(src.size) + 0.5).astype(np.float32)
    img = orc.build_2d_w(src, dst, vals, n=1 << 16, log2h=7, log2w=15)
    res = fmx.SparseMatrix.upload(img.data, img.offs)
    st = fmx.SparseMatrix.upload_streamed(img.data, img.offs, chunk_bytes=1 << 16)
    # the resident image is relabelled on the device, the streamed one from
    # its host copy: the same hot columns
    assert res.info.hot_columns == st.info.hot_columns > 0
    X = np.random.default_rng(5).random((img.n, 8)).astype(np.float32)
    Yr = res.multiply(dev(X)).cpu().numpy()
    assert rel(Yr, orc.spmm(img, X.astype(np.float64))) <= 2.0 ** -24 + 1e-15
    assert np.array_equal(st.multiply(dev(X)).cpu().numpy(), Yr)
    Xt = dev(X)
    Yc = torch.empty((8, img.n), dtype=torch.float32, device="cuda").t()
    res.multiply_async(Xt.t().contiguous().t(), Yc)
    assert np.array_equal(Yc.cpu().numpy(), Yr)
    post = torch.rand(img.n, dtype=torch.float64, device="cuda") + 0.5
    Ys = torch.empty_like(Xt)
    res.multiply_scaled_async(Xt, Ys, post=post)
    want = orc.spmm(img, X.astype(np.float64)) * post.cpu().numpy()[:, None]
    assert rel(Ys.cpu().numpy(), want) <= 2.0 ** -24 + 1e-15
