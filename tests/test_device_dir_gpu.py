"""A matrix built on the device keeps no host copy of its image: its tile
directory comes from the construction (build_tiles.hip), the carries of split
work items from the image in HBM (image_dev.hip), and an image in tiles taller
than an SpMM band is re-cut on the device.  Each is checked against the host
route of the same image: the image downloaded and uploaded again, where the
host analyses it, re-cuts it and reads the carries from its copy.

  * the same image bytes, non-zeros, tiles and block rows;
  * SpMM results bit-identical: int32 and FP64 integer-valued data, FP32 in
    the sequential mode (the reference's order), FP32 in the default mode,
    for k = 1 and 16, with hub rows split into several work items
    (FMX_ITEM_CAP) so that carries are exercised;
  * empty block rows (n above the largest id: the 16-byte empty tiles),
    weighted images (value words in the directory), the reference's default
    32768 x 32768 tiles (the device re-cut)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def same_products(fmx, a, b, n, seed):
    import torch
    rng = np.random.default_rng(seed)
    for k in (1, 16):
        Xi = rng.integers(-9, 10, (n, k)).astype(np.int32)
        assert np.array_equal(a.multiply(dev(Xi)).cpu().numpy(), b.multiply(dev(Xi)).cpu().numpy()), k
        Xd = rng.integers(-9, 10, (n, k)).astype(np.float64)
        assert np.array_equal(a.multiply(dev(Xd)).cpu().numpy(), b.multiply(dev(Xd)).cpu().numpy()), k
        Xf = dev(rng.random((n, k)).astype(np.float32))
        for seq in (True, False):
            ya = torch.empty_like(Xf)
            yb = torch.empty_like(Xf)
            a.multiply_async(Xf, ya, seq=seq)
            b.multiply_async(Xf, yb, seq=seq)
            assert torch.equal(ya, yb), (k, seq)


def host_route(fmx, spm):
    data, offs = spm.download()
    return fmx.SparseMatrix.upload(data, offs), data


@pytest.mark.parametrize("lh,lw,cap,hot", [(6, 15, "64", False), (5, 12, None, True), (15, 15, "512", True)])
def test_device_directory_equals_host_analysis(fmx, orc, monkeypatch, lh, lw, cap, hot):
    if cap:
        monkeypatch.setenv("FMX_ITEM_CAP", cap)
    if hot:  # the gather-order image for this small X too
        monkeypatch.setenv("FMX_HOT_MIN_BYTES", "0")
    scale = 14
    n = (1 << scale) + 1000  # block rows past the largest id stay empty
    src, dst = orc.rmat_edges(scale, 16 << scale, seed=77)
    a = fmx.SparseMatrix.build(dev(src.astype(np.int32)), dev(dst.astype(np.int32)), n=n, log2_h=lh, log2_w=lw,
                               flags=fmx.BUILD_DEDUP)
    b, data = host_route(fmx, a)
    ia, ib = a.info, b.info
    for f in ("nrows", "ncols", "nnz", "image_bytes", "ntiles", "nbr", "block_h", "block_w", "hot_columns"):
        assert getattr(ia, f) == getattr(ib, f), f
    data2, _ = b.download()
    assert np.array_equal(data, data2)
    same_products(fmx, a, b, n, lh + lw)


@pytest.mark.parametrize("lh", [6, 15])
def test_device_directory_weighted(fmx, orc, monkeypatch, lh):
    monkeypatch.setenv("FMX_ITEM_CAP", "128")
    scale = 13
    n = (1 << scale) + 300
    src, dst = orc.rmat_edges(scale, 16 << scale, seed=78)
    vals = (np.random.default_rng(5).integers(1, 9, src.size)).astype(np.float32)
    import torch
    a = fmx.SparseMatrix.build_weighted(dev(src.astype(np.int32)), dev(dst.astype(np.int32)), dev(vals), n=n,
                                        log2_h=lh, log2_w=15, flags=fmx.BUILD_DEDUP)
    b, _ = host_route(fmx, a)
    assert a.info.ntiles == b.info.ntiles and a.info.nnz == b.info.nnz
    rng = np.random.default_rng(lh)
    for k in (1, 8):
        X = dev(rng.integers(-9, 10, (n, k)).astype(np.float32))
        for seq in (True, False):
            ya = torch.empty_like(X)
            yb = torch.empty_like(X)
            a.multiply_async(X, ya, seq=seq)
            b.multiply_async(X, yb, seq=seq)
            assert torch.equal(ya, yb), (k, seq)
