"""Out-of-core matrices (the SAFS-backed sparse_matrix::create,
matrix/sparse_matrix.h:534-535, and the block-row streaming of
matrix/sparse_matrix.cpp:796-830, matrix_io.cpp:188-236): the image stays in
pinned host memory and each SpMM streams it through two HBM staging buffers
in block-row chunks.  The streamed product must equal the resident one bit
for bit (same kernels, same work items), and the oracle's within the usual
bounds, for every chunking, including chunks of a single block row, split
hub bands, weighted images and back-to-back calls that reuse the staging
buffers."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def to_dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def write_pair(tmp, image, offs, name="m"):
    mat, idx = os.path.join(tmp, name + ".mat"), os.path.join(tmp, name + ".mat_idx")
    image = np.asarray(image, np.uint8)
    with open(mat, "wb") as f:
        f.write(image.tobytes())
    with open(idx, "wb") as f:
        f.write(image[:4096].tobytes())
        f.write(np.asarray(offs, "<i8").tobytes())
    return mat, idx


def close(Y, ref, dt):
    if np.issubdtype(dt, np.integer):
        return np.array_equal(Y, ref)
    tol = 1e-5 if dt == np.float32 else 1e-12
    nz = ref != 0
    return np.all(Y[~nz] == 0) and (np.abs(Y.astype(np.float64) - ref)[nz] / np.abs(ref[nz])).max() <= tol


@pytest.fixture(scope="module")
def edges(orc):
    return orc.rmat_edges(15, 16 << 15, seed=77)


@pytest.mark.parametrize("lh,lw,chunk", [(7, 15, 4096), (8, 12, 65536), (10, 10, 1 << 20), (6, 15, 1 << 30)])
def test_streamed_equals_resident(fmx, orc, edges, lh, lw, chunk):
    s, d = edges
    n = 1 << 15
    img = orc.build_2d(s, d, n=n, log2h=lh, log2w=lw, flags=orc.DEDUP)
    res = fmx.SparseMatrix.upload(img.data, img.offs)
    st = fmx.SparseMatrix.upload_streamed(img.data, img.offs, chunk_bytes=chunk)
    assert st.streamed and not res.streamed
    if chunk == 4096:
        assert st.info.stream_chunks > 100
    rng = np.random.default_rng(lh * 100 + lw)
    for dt, k in ((np.int32, 8), (np.float32, 16), (np.float64, 16), (np.float32, 3), (np.int64, 1)):
        X = (rng.integers(-9, 10, (n, k)) if np.issubdtype(dt, np.integer) else rng.random((n, k))).astype(dt)
        Yr = res.multiply(to_dev(X)).cpu().numpy()
        Ys = st.multiply(to_dev(X)).cpu().numpy()
        assert np.array_equal(Ys, Yr), (dt, k)
        assert close(Ys, orc.spmm(img, X), dt), (dt, k)


def test_open_file_and_download(fmx, orc, edges, tmp_path):
    s, d = edges
    img = orc.build_2d(s, d, n=1 << 15, log2h=7, log2w=15, flags=orc.DEDUP)
    mat, idx = write_pair(str(tmp_path), img.data, img.offs)
    st = fmx.SparseMatrix.open(mat, idx, chunk_bytes=1 << 16)
    assert st.streamed and st.get_nnz() == img.nnz
    data, offs = st.download()
    assert np.array_equal(data, img.data) and np.array_equal(offs, img.offs)
    X = np.random.default_rng(3).integers(0, 8, (1 << 15, 16)).astype(np.int32)
    assert np.array_equal(st.multiply(to_dev(X)).cpu().numpy(), orc.spmm(img, X))


def test_streamed_split_bands(fmx, orc, edges, monkeypatch):
    """Hub bands split into word ranges: their partial slots are reduced
    once, after the last chunk."""
    s, d = edges
    img = orc.build_2d(s, d, n=1 << 15, log2h=10, log2w=15, flags=orc.DEDUP)
    monkeypatch.setenv("FMX_ITEM_CAP", "300")
    res = fmx.SparseMatrix.upload(img.data, img.offs)
    st = fmx.SparseMatrix.upload_streamed(img.data, img.offs, chunk_bytes=8192)
    rng = np.random.default_rng(11)
    for X in (rng.random((1 << 15, 16)).astype(np.float32), rng.integers(0, 8, (1 << 15, 32)).astype(np.int32)):
        Yr = res.multiply(to_dev(X)).cpu().numpy()
        Ys = st.multiply(to_dev(X)).cpu().numpy()
        assert np.array_equal(Ys, Yr)
        assert close(Ys, orc.spmm(img, X), X.dtype.type)


def test_back_to_back_calls_reuse_buffers(fmx, orc, edges):
    """Several multiplies queued without a host wait: each chunk copy waits
    for the kernel that last read its staging buffer."""
    import torch
    s, d = edges
    n = 1 << 15
    img = orc.build_2d(s, d, n=n, log2h=7, log2w=13, flags=orc.DEDUP)
    st = fmx.SparseMatrix.upload_streamed(img.data, img.offs, chunk_bytes=16384)
    rng = np.random.default_rng(5)
    Xs = [rng.integers(-5, 6, (n, 8)).astype(np.int32) for _ in range(4)]
    Xd = [to_dev(X) for X in Xs]
    Yd = [torch.empty_like(x) for x in Xd]
    for x, y in zip(Xd, Yd):
        st.multiply_async(x, y)
    torch.cuda.synchronize()
    for X, y in zip(Xs, Yd):
        assert np.array_equal(y.cpu().numpy(), orc.spmm(img, X))


def test_streamed_weighted(fmx, orc, edges):
    s, d = edges
    n = 1 << 15
    v = (np.random.default_rng(8).random(len(s)) * 4 + 0.25).astype(np.float64)
    img = orc.build_2d_w(s, d, v, n=n, log2h=8, log2w=14, flags=1)
    res = fmx.SparseMatrix.upload(img.data, img.offs)
    st = fmx.SparseMatrix.upload_streamed(img.data, img.offs, chunk_bytes=32768)
    X = np.random.default_rng(9).random((n, 4))
    Ys = st.multiply(to_dev(X)).cpu().numpy()
    assert np.array_equal(Ys, res.multiply(to_dev(X)).cpu().numpy())
    assert close(Ys, orc.spmm(img, X), np.float64)


def test_streamed_pagerank(fmx, orc, edges):
    """Callers of SpMM run unchanged on a streamed matrix."""
    s, d = edges
    img = orc.build_2d(s, d, n=1 << 15, log2h=8, log2w=15, flags=orc.DEDUP)
    res = fmx.SparseMatrix.upload(img.data, img.offs)
    st = fmx.SparseMatrix.upload_streamed(img.data, img.offs, chunk_bytes=1 << 16)
    pr_r, it_r = fmx.pagerank(res, max_iters=8)
    pr_s, it_s = fmx.pagerank(st, max_iters=8)
    assert it_r == it_s and np.array_equal(pr_r.cpu().numpy(), pr_s.cpu().numpy())


def test_streamed_errors(fmx, orc, tmp_path):
    s, d = orc.rmat_edges(10, 16 << 10, seed=1)
    img = orc.build_2d(s, d, n=1 << 10, log2h=6, log2w=8, flags=orc.DEDUP)
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.upload_streamed(img.data, img.offs, chunk_bytes=100)
    assert e.value.status == fmx.EINVAL
    mat, idx = write_pair(str(tmp_path), np.concatenate([img.data, np.zeros(16, np.uint8)]), img.offs)
    with pytest.raises(fmx.FmxError) as e:  # trailing bytes after the last block row
        fmx.SparseMatrix.open(mat, idx)
    assert e.value.status == fmx.EFORMAT
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.open(str(tmp_path / "missing.mat"), idx)
    assert e.value.status == fmx.EINVAL


@pytest.mark.parametrize("k,dt", [(37, np.int32), (16, np.float32), (5, np.float64)])
def test_host_operands_in_column_panels(fmx, orc, edges, k, dt):
    """X and Y in host memory, columns through HBM in panels (the column-panel
    branch of multiply(dense_matrix, mem_size), sparse_matrix.cpp:796-830):
    every column equals the device SpMM's (FP64: within 1e-12, since panels
    of other widths may order the FP64 adds of one row differently), also for
    a streamed matrix."""
    s, d = edges
    n = 1 << 15
    img = orc.build_2d(s, d, n=n, log2h=8, log2w=13, flags=orc.DEDUP)
    res = fmx.SparseMatrix.upload(img.data, img.offs)
    st = fmx.SparseMatrix.upload_streamed(img.data, img.offs, chunk_bytes=1 << 16)
    rng = np.random.default_rng(k)
    X = (rng.integers(-9, 10, (n, k)) if np.issubdtype(dt, np.integer) else rng.random((n, k))).astype(dt)
    Yd = res.multiply(to_dev(X)).cpu().numpy()
    es = X.dtype.itemsize
    for mem in (2 * n * es * 3, 2 * n * es * k, 1 << 40):  # panels of 3 columns, exactly k, all
        for A in (res, st):
            Yh = A.multiply_host(X, mem)
            if dt == np.float64:
                assert np.allclose(Yh, Yd, rtol=1e-12, atol=0), (mem, A.streamed)
            else:
                assert np.array_equal(Yh, Yd), (mem, A.streamed)
    assert close(Yd, orc.spmm(img, X), dt)
    with pytest.raises(fmx.FmxError) as e:  # not one column fits
        res.multiply_host(X, 2 * n * es - 1)
    assert e.value.status == fmx.ENOMEM
