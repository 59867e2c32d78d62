"""fmx_gram_noalloc, the Gram for callers that cannot allocate: through the C
ABI, two calls in a row (the second with a smaller input, so stale partial
sums of the first would show) against fmx_gram on the same operands.  On
integer-valued data every partial sum is exact, so the results must be
bit-identical whatever the block count; on U[0,1) data they agree within
the Gram's FP64 bound (1e-13 of sum |a b|).  BF16 mode is checked the same
way against fmx_gram's BF16 mode (bit-identical where the block counts
agree)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def gram_abi(fmx, A, B, mode, noalloc):
    import torch
    ctx = fmx.default_context()
    G = torch.empty((A.shape[1], B.shape[1]), dtype=torch.float64, device="cuda")
    fn = fmx.lib().fmx_gram_noalloc if noalloc else fmx.lib().fmx_gram
    st = fn(ctx.handle, fmx.dtype_code(A.dtype), mode, A.shape[0], A.shape[1], B.shape[1], A.data_ptr(),
            A.stride(0), B.data_ptr(), B.stride(0), G.data_ptr(), ctx.stream_ptr())
    ctx.check(st, "gram")
    torch.cuda.synchronize()
    return G.cpu().numpy()


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_gram_noalloc_matches_gram(fmx, dt):
    import torch
    tdt = getattr(torch, dt)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    # eigensolver shapes first (96 x 16 over 2^22 rows: the partials of the
    # default block count would not fit, so fewer blocks run), then smaller
    for n, ka, kb in ((1 << 22, 96, 16), (100003, 16, 16), (5000, 40, 24)):
        A = torch.randint(-8, 9, (n, ka), device="cuda", generator=g).to(tdt)
        B = A if ka == kb else torch.randint(-8, 9, (n, kb), device="cuda", generator=g).to(tdt)
        want = gram_abi(fmx, A, B, fmx.GRAM_F64, False)
        got = gram_abi(fmx, A, B, fmx.GRAM_F64, True)
        assert np.array_equal(got, want), (n, ka, kb)
        R = torch.rand((n, ka), device="cuda", generator=g, dtype=tdt)
        S = torch.rand((n, kb), device="cuda", generator=g, dtype=tdt)
        want = gram_abi(fmx, R, S, fmx.GRAM_F64, False)
        got = gram_abi(fmx, R, S, fmx.GRAM_F64, True)
        bound = 1e-13 * (R.double().abs().T.contiguous() @ S.double().abs()).cpu().numpy()
        assert np.all(np.abs(got - want) <= bound), (n, ka, kb)
        if dt == "float32" and ka <= 128 and kb <= 128:
            # BF16 products with FP32 partial sums: the same blocks give the
            # same bits; fewer blocks (the 2^22-row case) sum longer runs in
            # FP32, within 1e-4 of the total
            want = gram_abi(fmx, R, S, fmx.GRAM_BF16, False)
            got = gram_abi(fmx, R, S, fmx.GRAM_BF16, True)
            if n < (1 << 20):
                assert np.array_equal(got, want), (n, ka, kb)
            else:
                assert np.all(np.abs(got - want) <= 1e-4 * np.abs(want)), (n, ka, kb)
