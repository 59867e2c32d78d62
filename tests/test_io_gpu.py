"""The .mat / .mat_idx file pair (SpM_2d_storage::load/dump,
SpM_2d_index::load/dump, matrix/sparse_matrix_format.cpp:164-181,238-306,
417-420): files written in the reference layout by this test load and
multiply like the in-memory image, a dump is byte-identical to the oracle's
image and index, and malformed pairs are refused."""
import os

import numpy as np
import pytest

import tiny_kat

pytestmark = pytest.mark.gpu


def write_pair(tmp, image, offs, name="m"):
    """The reference layout: the matrix file is the image; the index file is
    the image's 4096-byte header page followed by the int64 offsets."""
    mat, idx = os.path.join(tmp, name + ".mat"), os.path.join(tmp, name + ".mat_idx")
    image = np.asarray(image, np.uint8)
    with open(mat, "wb") as f:
        f.write(image.tobytes())
    with open(idx, "wb") as f:
        f.write(image[:4096].tobytes())
        f.write(np.asarray(offs, "<i8").tobytes())
    return mat, idx


def test_load_tiny_kat(fmx, tmp_path):
    import torch
    mat, idx = write_pair(str(tmp_path), np.frombuffer(tiny_kat.IMAGE, np.uint8), tiny_kat.OFFS)
    spm = fmx.SparseMatrix.load(mat, idx)
    assert spm.get_num_rows() == tiny_kat.N and spm.get_nnz() == 7
    Y = spm.multiply(torch.tensor(tiny_kat.X, dtype=torch.int32, device="cuda"))
    assert Y.cpu().tolist() == tiny_kat.Y


def test_dump_matches_oracle_and_round_trips(fmx, orc, tmp_path):
    import torch
    s, d = fmx.rmat_edges(14, 16 << 14, 5)
    spm = fmx.SparseMatrix.build(s, d, n=1 << 14, log2_h=9, log2_w=12)
    img = orc.build_2d(s.cpu().numpy().view(np.uint32), d.cpu().numpy().view(np.uint32), n=1 << 14,
                       log2h=9, log2w=12, flags=orc.DEDUP)
    mat, idx = str(tmp_path / "a.mat"), str(tmp_path / "a.mat_idx")
    spm.dump(mat, idx)
    raw = np.fromfile(mat, np.uint8)
    assert np.array_equal(raw, img.data)
    ix = np.fromfile(idx, np.uint8)
    assert np.array_equal(ix[:4096], img.data[:4096])
    assert np.array_equal(ix[4096:].view("<i8"), img.offs)
    back = fmx.SparseMatrix.load(mat, idx)
    X = torch.randint(0, 8, (1 << 14, 4), dtype=torch.int32, device="cuda")
    assert torch.equal(back.multiply(X), spm.multiply(X))
    ref = orc.spmm(img, X.cpu().numpy())
    assert np.array_equal(back.multiply(X).cpu().numpy(), ref)


def test_load_errors(fmx, tmp_path):
    img = np.frombuffer(tiny_kat.IMAGE, np.uint8)
    t = str(tmp_path)
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.load(os.path.join(t, "missing.mat"), os.path.join(t, "missing.mat_idx"))
    assert e.value.status == fmx.EINVAL
    # index header disagreeing with the matrix header (other row count)
    mat, idx = write_pair(t, img, tiny_kat.OFFS, "h")
    raw = bytearray(open(idx, "rb").read())
    raw[24] ^= 1
    open(idx, "wb").write(bytes(raw))
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.load(mat, idx)
    assert e.value.status == fmx.EFORMAT
    # truncated index (offsets not a multiple of 8 bytes)
    mat, idx = write_pair(t, img, tiny_kat.OFFS, "t")
    raw = open(idx, "rb").read()
    open(idx, "wb").write(raw[:-3])
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.load(mat, idx)
    assert e.value.status == fmx.EFORMAT
    # matrix file with trailing bytes past the last block row
    mat, idx = write_pair(t, np.concatenate([img, np.zeros(8, np.uint8)]), tiny_kat.OFFS, "x")
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.load(mat, idx)
    assert e.value.status == fmx.EFORMAT
