"""CPU check of the work image (flashx_amd/csrc/retile.cpp): an image with
tiles taller than an SpMM band is re-cut, at creation, into 64-row tiles of
the same width.  Every row part and single-entry row keeps its columns,
values and kind, so the re-cut image must equal, byte for byte, the image
the oracle's writer (the restatement of fm_utils.cpp:228-433) produces from
the same edges at the shorter tile height -- for binary and weighted
images, dedup on and off, and empty block rows.  Runs the host code of
libfmx without a GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "retile_check.cpp")
LIBDIR = os.path.join(ROOT, "flashx_amd", "lib")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libfmx.so")):
        pytest.skip("libfmx.so not built")
    exe = str(tmp_path_factory.mktemp("retile") / "retile_check")
    subprocess.check_call(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17",
                           "-I" + os.path.join(ROOT, "include"), SRC,
                           "-L" + LIBDIR, "-lfmx", "-Wl,-rpath," + LIBDIR, "-o", exe],
                          stdout=subprocess.DEVNULL, stderr=subprocess.STDOUT)
    return exe


def recut(checker, tmp_path, img, h2):
    a, b, c, d = (str(tmp_path / x) for x in ("i", "o", "ri", "ro"))
    img.data.tofile(a)
    np.asarray(img.offs, np.int64).tofile(b)
    out = subprocess.run([checker, a, b, str(h2), c, d], capture_output=True, text=True)
    assert out.returncode == 0, (out.returncode, out.stdout, out.stderr)
    return np.fromfile(c, np.uint8), np.fromfile(d, np.int64)


@pytest.mark.parametrize("lh,lw,h2", [(15, 15, 64), (12, 10, 64), (10, 15, 16), (8, 8, 64)])
def test_recut_equals_direct_build(orc, checker, tmp_path, lh, lw, h2):
    src, dst = orc.rmat_edges(14, 16 << 14, seed=77)
    n = (1 << 14) + 300  # trailing rows without non-zeros: empty block rows
    for flags in (orc.DEDUP, 0):
        tall = orc.build_2d(src, dst, n=n, log2h=lh, log2w=lw, flags=flags)
        data, offs = recut(checker, tmp_path, tall, h2)
        direct = orc.build_2d(src, dst, n=n, log2h=int(np.log2(h2)), log2w=lw, flags=flags)
        assert np.array_equal(offs, direct.offs)
        assert np.array_equal(data, direct.data)


@pytest.mark.parametrize("vdt", [np.float32, np.int64])
def test_recut_weighted(orc, checker, tmp_path, vdt):
    src, dst = orc.rmat_edges(13, 16 << 13, seed=78)
    vals = (np.arange(src.size) % 1000).astype(vdt)
    tall = orc.build_2d_w(src, dst, vals, n=1 << 13, log2h=15, log2w=12)
    data, offs = recut(checker, tmp_path, tall, 64)
    direct = orc.build_2d_w(src, dst, vals, n=1 << 13, log2h=6, log2w=12)
    assert np.array_equal(offs, direct.offs) and np.array_equal(data, direct.data)
