"""Weighted tile images (entry_size 4 or 8): GPU SpMM through the C ABI
against the oracle's reference-order product over the same image
(matrix/sparse_matrix.h:219-331 with values, sparse_matrix_format.h:233-241).
Integer values and operands are exact; FP64 within 1e-12 and FP32 within
1e-5 relative, rows without non-zeros exactly 0."""
import os

import numpy as np
import pytest

import tiny_kat

pytestmark = pytest.mark.gpu


def to_dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def compare(Y, ref):
    if np.issubdtype(ref.dtype, np.integer):
        assert np.array_equal(Y, ref)
        return
    tol = 1e-5 if ref.dtype == np.float32 else 1e-12
    y, r = Y.astype(np.float64), ref.astype(np.float64)
    zero = r == 0
    assert np.all(y[zero] == 0)
    assert (np.abs(y - r)[~zero] / np.abs(r[~zero])).max() <= tol


def test_weighted_tiny_kat(fmx):
    spm = fmx.SparseMatrix.upload(tiny_kat.W_IMAGE, tiny_kat.W_OFFS)
    assert spm.get_entry_size() == 4 and spm.get_nnz() == 7
    Y = spm.multiply(to_dev(np.array(tiny_kat.X, np.float32))).cpu().numpy()
    assert Y.tolist() == tiny_kat.W_Y
    with pytest.raises(fmx.FmxError) as e:  # float values need float operands
        spm.multiply(to_dev(np.array(tiny_kat.X, np.float64)))
    assert e.value.status == fmx.ETYPE


@pytest.fixture(scope="module")
def edges(orc):
    return orc.rmat_edges(14, 16 << 14, seed=41)


@pytest.mark.parametrize("dt,lh,lw,flags", [(np.int32, 7, 15, 1), (np.int64, 10, 8, 1 | 4),
                                            (np.float32, 6, 12, 1), (np.float64, 9, 9, 0),
                                            (np.float32, 14, 14, 1 | 2)])
def test_weighted_rmat(fmx, orc, edges, dt, lh, lw, flags):
    s, d = edges
    n = 1 << 14
    rng = np.random.default_rng(lh * 31 + lw)
    if np.issubdtype(dt, np.integer):
        v = rng.integers(-9, 10, len(s)).astype(dt)
    else:
        # positive weights: the relative bounds are for sums without
        # cancellation, as with the U[0,1) operands
        v = (rng.random(len(s)) * 4 + 0.25).astype(dt)
    img = orc.build_2d_w(s, d, v, n=n, log2h=lh, log2w=lw, flags=flags)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    for k in (1, 3, 8, 16):
        if np.issubdtype(dt, np.integer):
            X = rng.integers(-7, 8, (n, k)).astype(dt)
        else:
            X = rng.random((n, k)).astype(dt)
        Y = spm.multiply(to_dev(X)).cpu().numpy()
        compare(Y, orc.spmm(img, X))


def test_weighted_split_bands_and_rotation(fmx, orc, edges, monkeypatch):
    """Small work cap: hub bands are split at tile starts and summed in
    order; two calls agree bit for bit."""
    s, d = edges
    n = 1 << 14
    v = np.random.default_rng(5).random(len(s)).astype(np.float64)
    img = orc.build_2d_w(s, d, v, n=n, log2h=12, log2w=10)
    monkeypatch.setenv("FMX_ITEM_CAP", "300")
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    X = np.random.default_rng(6).random((n, 16))
    Y1 = spm.multiply(to_dev(X)).cpu().numpy()
    Y2 = spm.multiply(to_dev(X)).cpu().numpy()
    assert np.array_equal(Y1, Y2)
    compare(Y1, orc.spmm(img, X))


def test_weighted_file_round_trip(fmx, orc, edges, tmp_path):
    s, d = edges
    v = np.arange(len(s), dtype=np.int32) % 97 - 48
    img = orc.build_2d_w(s, d, v, n=1 << 14, log2h=8, log2w=12)
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    mat, idx = str(tmp_path / "w.mat"), str(tmp_path / "w.mat_idx")
    spm.dump(mat, idx)
    assert np.array_equal(np.fromfile(mat, np.uint8), img.data)
    back = fmx.SparseMatrix.load(mat, idx)
    X = np.random.default_rng(2).integers(-5, 6, (1 << 14, 4)).astype(np.int32)
    assert np.array_equal(back.multiply(to_dev(X)).cpu().numpy(), orc.spmm(img, X))


@pytest.mark.parametrize("dt,lh,lw,flags", [(np.float32, 7, 12, 1), (np.float64, 5, 9, 1 | 4),
                                            (np.int32, 10, 10, 0), (np.int64, 6, 15, 1 | 2 | 4)])
def test_weighted_device_build_bit_exact(fmx, orc, dt, lh, lw, flags):
    """fmx_spm_build_w writes the same bytes as the oracle's weighted writer
    (duplicates keep their first value, symmetrised edges share it)."""
    import torch
    s, d = orc.rmat_edges(13, 16 << 13, lh)
    rng = np.random.default_rng(lw)
    v = (rng.random(len(s)) * 100).astype(dt) if not np.issubdtype(dt, np.integer) else \
        rng.integers(-1000, 1000, len(s)).astype(dt)
    img = orc.build_2d_w(s, d, v, n=1 << 13, log2h=lh, log2w=lw, flags=flags)
    spm = fmx.SparseMatrix.build_weighted(to_dev(s.view(np.int32)), to_dev(d.view(np.int32)), to_dev(v),
                                          n=1 << 13, log2_h=lh, log2_w=lw, flags=flags)
    data, offs = spm.download()
    assert np.array_equal(offs, img.offs)
    assert np.array_equal(data, img.data)


def tile_starts_at_chunk_edges(img, esz, chunk_words):
    """Tiles (after a chunk holding entries) that start exactly where an
    unsplit block row's chunk of `chunk_words` words begins: the SpMM decode
    must not carry value positions across such an edge."""
    data, offs = img.data, img.offs
    hits = 0
    for br in range(len(offs) - 1):
        p, end = int(offs[br]), int(offs[br + 1])
        bw = p // 2
        starts = []
        while p < end:
            tnnz = int(data[p + 8:p + 12].view("<u4")[0])
            tnrow = int(data[p + 12:p + 14].view("<u2")[0])
            starts.append(p // 2)
            p += 16 + (2 * tnrow + 2 * tnnz + 2 + esz * tnnz if tnnz else 0)
        edges = set(range((bw & ~7) + chunk_words, end // 2, chunk_words))
        hits += sum(1 for s in starts[1:] if s in edges)
    return hits


@pytest.mark.parametrize("dt,lh,lw", [(np.float32, 6, 8), (np.float64, 7, 7), (np.int32, 6, 10),
                                      (np.int64, 8, 6)])
def test_weighted_tiles_starting_at_chunk_edges(fmx, orc, dt, lh, lw):
    """Narrow tiles put many tile starts exactly on the decoder's chunk
    edges (512- and 1024-word chunks); every value must still be the one
    stored for its non-zero."""
    s, d = orc.rmat_edges(15, 16 << 15, 77)
    n = 1 << 15
    rng = np.random.default_rng(lh * 16 + lw)
    if np.issubdtype(dt, np.integer):
        v = rng.integers(-9, 10, len(s)).astype(dt)
    else:
        v = (rng.random(len(s)) * 4 + 0.25).astype(dt)
    img = orc.build_2d_w(s, d, v, n=n, log2h=lh, log2w=lw, flags=1)
    esz = np.dtype(dt).itemsize
    assert tile_starts_at_chunk_edges(img, esz, 512) > 10
    assert tile_starts_at_chunk_edges(img, esz, 1024) > 5
    spm = fmx.SparseMatrix.upload(img.data, img.offs)
    for k in (1, 4, 16):
        X = (rng.integers(-7, 8, (n, k)) if np.issubdtype(dt, np.integer) else rng.random((n, k))).astype(dt)
        compare(spm.multiply(to_dev(X)).cpu().numpy(), orc.spmm(img, X))
