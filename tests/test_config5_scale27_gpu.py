"""The large single-GPU SpMM itself: R-MAT scale 27 (2^27 vertices, 2^31
edges, dedup: 2.13 G non-zeros), built on the device in 16 x 32768 tiles
(the autotuned shape), times a 2^27 x 32 operand, everything resident in
HBM.  A seeded sample of 2^20 rows (whole block rows) is checked against the
oracle's reference-order CPU SpMM of the same block rows of the downloaded
image:
  * int32 {0..7}: bit-exact;
  * FP32 U[0,1), sequential mode (FMX_SPMM_SEQ): bit-identical to the
    reference's FP32 sums;
  * FP32, default mode: within one rounding (2^-24 relative) of the FP64
    sums of the same terms;
rows without non-zeros exactly 0."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_config5_rmat27_k32(fmx, orc):
    import torch
    scale, k, lh = 27, 32, 4
    n = 1 << scale
    s, d = fmx.rmat_edges(scale, 16 << scale, scale)
    spm = fmx.SparseMatrix.build(s, d, n=n, log2_h=lh, log2_w=15, flags=fmx.BUILD_DEDUP)
    del s, d
    torch.cuda.empty_cache()
    H = 1 << lh
    nb = (1 << 20) // H
    br0 = int(np.random.default_rng(scale).integers(0, spm.info.nbr - nb))
    r0, r1 = br0 * H, (br0 + nb) * H
    data, offs = spm.download()
    img = orc.Image(data, offs, n)
    del data
    g = torch.Generator(device="cuda")
    g.manual_seed(scale)
    X = torch.randint(0, 8, (n, k), dtype=torch.int32, device="cuda", generator=g)
    Y = torch.empty_like(X)
    spm.multiply_async(X, Y)
    got = Y[r0:r1].cpu().numpy()
    Xh = X.cpu().numpy()
    del X, Y
    torch.cuda.empty_cache()
    assert np.array_equal(got, orc.spmm_range(img, Xh, br0, br0 + nb))
    del Xh, got
    X = torch.rand((n, k), dtype=torch.float32, device="cuda", generator=g)
    Y = torch.empty_like(X)
    spm.multiply_async(X, Y, seq=True)
    got_seq = Y[r0:r1].cpu().numpy()
    spm.multiply_async(X, Y)
    got = Y[r0:r1].cpu().numpy().astype(np.float64)
    Xh = X.cpu().numpy()
    del X, Y
    torch.cuda.empty_cache()
    assert np.array_equal(got_seq, orc.spmm_range(img, Xh, br0, br0 + nb))
    exact = orc.spmm_range(img, Xh.astype(np.float64), br0, br0 + nb)
    z = exact == 0
    assert np.all(got[z] == 0)
    rel = float((np.abs(got - exact)[~z] / exact[~z]).max())
    print("scale 27 sample rows [%d, %d): int32 and sequential FP32 bit-exact, default FP32 within %.3g of "
          "the FP64 sums" % (r0, r1, rel))
    assert rel <= 2.0 ** -24
