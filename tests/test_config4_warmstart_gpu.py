"""The learned warm start on its evaluation workload: held-out degree-corrected
SBM graphs of 2^20 vertices (32 blocks, p_in/p_out 20, power-law degree
weights with exponent 2.5, mean degree 16; the first two of the 20 held-out
seeds, 900000 and 900001), top-16 eigenpairs of D^-1/2 A D^-1/2 to a
per-pair FP64 residual of 1e-6, from a random start and from the committed
network's start.

Both starts must converge, every pair's residual is recomputed in FP64 from
the returned vectors with the oracle's CPU SpMM on the downloaded image, the
two starts must agree on the eigenvalues, and the learned start must need
fewer block SpMMs than the random one -- the gain the trained network brings
(the totals over all 20 held-out graphs are reported by bench.py)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def check(orc, g, evals, vecs):
    data, offs = g.spm.download()
    img = orc.Image(data, offs, g.n)
    deg = orc.spmm(img, np.ones((g.n, 1)), threads=8).ravel()
    dinv = np.where(deg > 0, 1.0 / np.sqrt(np.maximum(deg, 1.0)), 0.0)
    V = vecs.cpu().numpy()
    NV = dinv[:, None] * orc.spmm(img, np.ascontiguousarray(dinv[:, None] * V), threads=8)
    rel = np.linalg.norm(NV - V * evals[None, :], axis=0) / np.abs(evals)
    assert np.all(rel <= 1e-6 * 1.01), rel.max()
    assert np.abs(V.T @ V - np.eye(V.shape[1])).max() < 1e-6
    return rel


@pytest.mark.parametrize("seed", [900000, 900001])
def test_config4_learned_start_gains(fmx, orc, seed):
    from flashx_amd import learned as Lm
    if not os.path.exists(Lm.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights")
    net = Lm.WarmStartNet(params=np.load(Lm.DEFAULT_WEIGHTS))
    g = Lm.Graph(1 << 20, seed)
    ev_r, v_r, res_r, st_r = g.solve(tol=1e-6, seed=seed)
    Y = net.forward(g)
    ev_l, v_l, res_l, st_l = g.solve(X0=Y, tol=1e-6, seed=seed)
    assert st_r["converged"] == 1 and st_l["converged"] == 1
    rr = check(orc, g, ev_r, v_r)
    rl = check(orc, g, ev_l, v_l)
    print("seed %d: random start %d block SpMMs (max residual %.2e), learned start %d (max residual %.2e)"
          % (seed, st_r["spmm_blocks"], rr.max(), st_l["spmm_blocks"], rl.max()))
    assert np.allclose(ev_r, ev_l, atol=2e-6)
    assert st_l["spmm_blocks"] < st_r["spmm_blocks"]
