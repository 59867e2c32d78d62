"""GPU kernels of the learned warm start against the FP64 oracle
(orc_gnn_loss_grad, itself pinned by finite differences in
test_gnn_oracle.py): forward output, loss and every parameter gradient, in
the FP32 mode against FP64 and the BF16-MFMA mode against the checker with
the same operands rounded to BF16."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev(x, dt=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x))
    if dt is not None:
        t = t.to(dt)
    return t.cuda()


def setup_graph(fmx, orc, scale=10, seed=3):
    import torch
    s, d = fmx.rmat_edges(scale, 8 << scale, seed)
    flags = fmx.BUILD_DEDUP | fmx.BUILD_SYMMETRIZE
    spm = fmx.SparseMatrix.build(s, d, n=1 << scale, log2_h=7, log2_w=12, flags=flags)
    img = orc.build_2d(s.cpu().numpy().view(np.uint32), d.cpu().numpy().view(np.uint32), n=1 << scale,
                       log2h=7, log2w=12, flags=orc.DEDUP | orc.SYMMETRIZE)
    r, c = orc.nonzeros(img)
    order = np.lexsort((c, r))
    r, c = r[order].astype(np.int64), c[order].astype(np.int64)
    n = img.n
    rowptr = np.zeros(n + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr)
    return spm, img, rowptr, c


# mode 1: FP32 products against the FP64 checker.  mode 0: BF16 matrix-core
# products against the checker with the same operands rounded to BF16 (what
# remains is FP32 accumulation and the odd operand that rounds the other way).
@pytest.mark.parametrize("mode,tol", [(1, 2e-4), (0, 1e-2)])
def test_forward_backward_vs_oracle(fmx, orc, mode, tol):
    import torch
    spm, img, rowptr, col = setup_graph(fmx, orc)
    n = img.n
    X0 = torch.empty((n, 16), dtype=torch.float32, device="cuda")
    dinv = torch.empty(n, dtype=torch.float64, device="cuda")
    L = fmx.lib()
    L.fmx_gnn_features(fmx.default_context().handle, spm.handle, 7, X0.data_ptr(), dinv.data_ptr(), None)
    torch.cuda.synchronize()
    deg = np.diff(rowptr)
    assert np.allclose(X0[:, 0].cpu().numpy(), np.log1p(deg), rtol=1e-6)
    assert np.allclose(dinv.cpu().numpy(), np.where(deg > 0, 1 / np.sqrt(np.maximum(deg, 1)), 0))
    rng = np.random.default_rng(0)
    npar = L.fmx_gnn_param_count()
    assert npar == orc.gnn_param_count()
    P = (rng.standard_normal(npar) * 0.15).astype(np.float32)
    U, _ = np.linalg.qr(rng.standard_normal((n, 16)))
    acts = torch.empty(L.fmx_gnn_acts_floats(n), dtype=torch.float32, device="cuda")
    work = torch.empty(L.fmx_gnn_work_floats(n), dtype=torch.float32, device="cuda")
    Y = torch.empty((n, 16), dtype=torch.float32, device="cuda")
    Pd = dev(P)
    ctx = fmx.default_context()
    ctx.check(L.fmx_gnn_forward(ctx.handle, spm.handle, dinv.data_ptr(), Pd.data_ptr(), X0.data_ptr(),
                                acts.data_ptr(), Y.data_ptr(), work.data_ptr(), mode, None), "forward")
    Ud = dev(U.astype(np.float32))
    dY = torch.empty_like(Y)
    import ctypes
    loss = ctypes.c_double()
    ctx.check(L.fmx_subspace_loss(ctx.handle, n, 16, Y.data_ptr(), Ud.data_ptr(), ctypes.byref(loss),
                                  dY.data_ptr(), None), "loss")
    G = torch.zeros(npar, dtype=torch.float32, device="cuda")
    ctx.check(L.fmx_gnn_backward(ctx.handle, spm.handle, dinv.data_ptr(), Pd.data_ptr(), X0.data_ptr(),
                                 acts.data_ptr(), dY.data_ptr(), G.data_ptr(), work.data_ptr(), mode,
                                 None), "backward")
    torch.cuda.synchronize()
    ref_loss, ref_g, ref_Y = orc.gnn_loss_grad(rowptr, col, dinv.cpu().numpy(),
                                               X0.cpu().numpy().astype(np.float64),
                                               U.astype(np.float32).astype(np.float64),
                                               P.astype(np.float64), bf16=(mode == 0))
    Yg = Y.cpu().numpy()
    assert np.abs(Yg - ref_Y).max() <= tol * np.abs(ref_Y).max()
    assert abs(loss.value - ref_loss) <= tol * max(1.0, abs(ref_loss))
    g = G.cpu().numpy()
    # per parameter block, relative to the block's largest gradient
    offs = [0]
    for l in range(3):
        din = 16 if l == 0 else 64
        offs += [offs[-1] + din * 64, offs[-1] + 2 * din * 64, offs[-1] + 2 * din * 64 + 64]
    offs.append(npar)
    for a, b in zip(offs[:-1], offs[1:]):
        scale = np.abs(ref_g[a:b]).max()
        assert np.abs(g[a:b] - ref_g[a:b]).max() <= tol * scale + 1e-12, (a, b)


def test_adam_step(fmx):
    import torch
    import ctypes
    L = fmx.lib()
    ctx = fmx.default_context()
    rng = np.random.default_rng(1)
    p = rng.standard_normal(1000).astype(np.float32)
    g = rng.standard_normal(1000).astype(np.float32)
    m = np.zeros(1000, np.float32)
    v = np.zeros(1000, np.float32)
    pd, gd, md, vd = dev(p), dev(g), dev(m), dev(v)
    for step in (1, 2):
        ctx.check(L.fmx_adam(ctx.handle, 1000, pd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(),
                             1e-2, 0.9, 0.999, 1e-8, step, None), "adam")
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    torch.cuda.synchronize()
    assert np.allclose(pd.cpu().numpy(), p, rtol=1e-5, atol=1e-6)


def test_dcsbm_generator_matches_oracle(fmx, orc):
    for n, seed in ((1 << 12, 5), (1 << 16, 11)):
        s, d = fmx.dcsbm_edges(n, seed)
        bstart, wtab = fmx.dcsbm_params(n)
        so, do = orc.dcsbm_edges(n, bstart, wtab, 20, n * 8, seed)
        assert np.array_equal(s.cpu().numpy().view(np.uint32), so)
        assert np.array_equal(d.cpu().numpy().view(np.uint32), do)
