"""GPU tile construction (fmx_spm_build) is bit-exact with the oracle writer.

Every case compares the downloaded image and offsets byte for byte with the
oracle's restatement of fm::create_2d_matrix, itself pinned by the
hand-written image and the independent Python writer (test_oracle.py).
"""
import numpy as np
import pytest

import tiny_kat

pytestmark = pytest.mark.gpu


def dev_u32(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).cuda()


def build_both(fmx, orc, src, dst, n, lh, lw, flags):
    spm = fmx.SparseMatrix.build(dev_u32(src), dev_u32(dst), n=n, log2_h=lh, log2_w=lw, flags=flags)
    img = orc.build_2d(src, dst, n=n, log2h=lh, log2w=lw, flags=flags)
    data, offs = spm.download()
    return spm, img, data, offs


def test_tiny_kat(fmx, orc):
    spm = fmx.SparseMatrix.build(dev_u32(tiny_kat.SRC), dev_u32(tiny_kat.DST), n=tiny_kat.N,
                                 log2_h=1, log2_w=1, flags=fmx.BUILD_DEDUP)
    data, offs = spm.download()
    assert bytes(data) == tiny_kat.IMAGE
    assert list(offs) == tiny_kat.OFFS


@pytest.mark.parametrize("scale,lh,lw", [(8, 1, 1), (10, 3, 5), (12, 15, 15), (12, 8, 12),
                                         (13, 4, 15), (11, 10, 2)])
def test_random_graphs_all_flags(fmx, orc, scale, lh, lw):
    rng = np.random.default_rng(scale + lh)
    n = 1 << scale
    m = 5 * n
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = (rng.zipf(1.5, m) - 1).clip(0, n - 1).astype(np.uint32)
    src[:40] = dst[:40]
    for flags in (fmx.BUILD_DEDUP, fmx.BUILD_DEDUP | fmx.BUILD_KEEP_SELF,
                  fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE, 0):
        try:
            img = orc.build_2d(src, dst, n=n + 5, log2h=lh, log2w=lw, flags=flags)
        except ValueError:
            continue  # duplicate runs longer than W without dedup (the reference refuses too)
        spm = fmx.SparseMatrix.build(dev_u32(src), dev_u32(dst), n=n + 5, log2_h=lh, log2_w=lw,
                                     flags=flags)
        data, offs = spm.download()
        assert np.array_equal(offs, img.offs), flags
        assert np.array_equal(data, img.data), flags
        assert spm.get_nnz() == img.nnz


def test_n_from_max_id(fmx, orc):
    rng = np.random.default_rng(1)
    src = rng.integers(0, 700, 3000).astype(np.uint32)
    dst = rng.integers(0, 900, 3000).astype(np.uint32)
    spm, img, data, offs = build_both(fmx, orc, src, dst, 0, 6, 7, fmx.BUILD_DEDUP)
    assert spm.get_num_rows() == img.n == max(src.max(), dst.max()) + 1
    assert np.array_equal(data, img.data)


@pytest.mark.parametrize("lh,lw", [(15, 15), (8, 15)])
def test_config1_rmat16(fmx, orc, lh, lw):
    s, d = fmx.rmat_edges(16, 16 << 16, 2024)
    src = s.cpu().numpy().view(np.uint32)
    dst = d.cpu().numpy().view(np.uint32)
    spm, img, data, offs = build_both(fmx, orc, src, dst, 1 << 16, lh, lw, fmx.BUILD_DEDUP)
    assert np.array_equal(offs, img.offs)
    assert np.array_equal(data, img.data)


def test_config2_rmat22_construction_bit_exact(fmx, orc):
    # R-MAT scale 22, edge factor 16 (67,108,864 edges), dedup, default tiles
    s, d = fmx.rmat_edges(22, 16 << 22, 22)
    spm = fmx.SparseMatrix.build(s, d, n=1 << 22, log2_h=15, log2_w=15, flags=fmx.BUILD_DEDUP)
    data, offs = spm.download()
    src = s.cpu().numpy().view(np.uint32)
    dst = d.cpu().numpy().view(np.uint32)
    del s, d
    img = orc.build_2d(src, dst, n=1 << 22, log2h=15, log2w=15, flags=orc.DEDUP)
    assert np.array_equal(offs, img.offs)
    assert np.array_equal(data, img.data)


def test_build_errors(fmx):
    import torch
    src = dev_u32([0, 1, 9])
    dst = dev_u32([1, 2, 3])
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.build(src, dst, n=5)
    assert e.value.status == fmx.EDIM
    empty = torch.empty(0, dtype=torch.int32, device="cuda")
    with pytest.raises(fmx.FmxError) as e:
        fmx.SparseMatrix.build(empty, empty, n=0)
    assert e.value.status == fmx.EINVAL
    spm = fmx.SparseMatrix.build(empty, empty, n=10, log2_h=2, log2_w=2)
    data, offs = spm.download()
    assert len(offs) == 4 and spm.get_nnz() == 0  # three empty block rows
    Y = spm.multiply(torch.ones((10, 3), device="cuda"))
    assert torch.all(Y == 0)
