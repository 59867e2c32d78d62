"""Learned warm start for the block Krylov eigensolver.

A 3-layer message-passing network (hidden 64; kernels in csrc/gnn.hip)
predicts the n x 16 starting block of fmx_eigen_normalized from the graph
alone.  It is trained on degree-corrected SBM graphs (32 blocks, p_in/p_out
20, power-law degree weights with exponent 2.5, mean degree 16) against the
top-16 eigenvectors the plain GPU solver finds, with the rotation-invariant
subspace loss k - tr((Y^T Y)^-1 (U^T Y)^T (U^T Y)).  Every step - features,
aggregation (fmx_spmm), BF16-MFMA dense layers, backward pass, loss and Adam
- runs in libfmx on the GPU; this module only orchestrates.

There is no reference counterpart: the reference's solver starts from a
random block (matrix/eigensolver/FM_MultiVector.h:164-168).

    python -m flashx_amd.learned train --graphs 200 --out flashx_amd/weights/warmstart.npy
    python -m flashx_amd.learned eval --weights flashx_amd/weights/warmstart.npy --graphs 20
    python -m flashx_amd.learned analyse --graphs 3 --report warmstart_analysis.json
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import fmx as F

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "warmstart.npy")
LOG2_H, LOG2_W = 7, 15
NEV = 16


def _torch():
    import torch
    return torch


def init_params(seed=0):
    """He-style initialisation in the flat layout of fmx_gnn_param_count()."""
    rng = np.random.default_rng(seed)
    parts = []
    for l in range(3):
        din = 16 if l == 0 else 64
        s = np.sqrt(2.0 / (2 * din))
        parts += [rng.standard_normal(din * 64) * s, rng.standard_normal(din * 64) * s, np.zeros(64)]
    parts.append(rng.standard_normal(64 * 16) * np.sqrt(1.0 / 64))
    p = np.concatenate(parts).astype(np.float32)
    assert len(p) == F.lib().fmx_gnn_param_count()
    return p


class Graph:
    """A DC-SBM graph resident in HBM with its features and (optionally) target."""

    def __init__(self, n, seed, feat_seed=None):
        torch = _torch()
        self.n, self.seed = n, seed
        src, dst = F.dcsbm_edges(n, seed)
        self.spm = F.SparseMatrix.build(src, dst, n=n, log2_h=LOG2_H, log2_w=LOG2_W,
                                        flags=F.BUILD_DEDUP | F.BUILD_SYMMETRIZE)
        del src, dst
        ctx = self.spm.ctx
        self.X0 = torch.empty((n, 16), dtype=torch.float32, device="cuda")
        self.dinv = torch.empty(n, dtype=torch.float64, device="cuda")
        fs = seed * 7919 + 1 if feat_seed is None else feat_seed
        torch.cuda.synchronize()
        ctx.check(F.lib().fmx_gnn_features(ctx.handle, self.spm.handle, fs, self.X0.data_ptr(),
                                           self.dinv.data_ptr(), None), "fmx_gnn_features")
        self.U = None

    def warmup(self):
        """A short loose solve, so that one-time set-up (SpMM work plans for
        the FP32 and FP64 widths) is not charged to the first timed solve."""
        self.solve(tol=1e-1, seed=12345)
        torch = _torch()
        torch.cuda.synchronize()

    def solve(self, X0=None, tol=1e-6, seed=1, verbose=False):
        return F.compute_eigen(self.spm, nev=NEV, block_size=16, num_blocks=6, tol=tol, X0=X0, seed=seed,
                               precision=F.EIG_FP32, verbose=verbose, max_restarts=400)


class WarmStartNet:
    """Parameters on the GPU plus the forward/backward/Adam entry points."""

    def __init__(self, params=None, seed=0, mode=0):
        torch = _torch()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.P = torch.from_numpy(p).cuda()
        self.m = torch.zeros_like(self.P)
        self.v = torch.zeros_like(self.P)
        self.G = torch.zeros_like(self.P)
        self.step = 0
        self.mode = mode  # 0: BF16 MFMA, FP32 accumulation
        self._ws = {}

    def _buffers(self, n):
        torch = _torch()
        if n not in self._ws:
            L = F.lib()
            self._ws = {n: (torch.empty(L.fmx_gnn_acts_floats(n), dtype=torch.float32, device="cuda"),
                            torch.empty(L.fmx_gnn_work_floats(n), dtype=torch.float32, device="cuda"),
                            torch.empty((n, 16), dtype=torch.float32, device="cuda"),
                            torch.empty((n, 16), dtype=torch.float32, device="cuda"))}
        return self._ws[n]

    def forward(self, g):
        acts, work, Y, _ = self._buffers(g.n)
        ctx = g.spm.ctx
        ctx.check(F.lib().fmx_gnn_forward(ctx.handle, g.spm.handle, g.dinv.data_ptr(), self.P.data_ptr(),
                                          g.X0.data_ptr(), acts.data_ptr(), Y.data_ptr(), work.data_ptr(),
                                          self.mode, None), "fmx_gnn_forward")
        return Y

    def train_step(self, g, lr):
        acts, work, Y, dY = self._buffers(g.n)
        ctx = g.spm.ctx
        L = F.lib()
        self.forward(g)
        loss = ctypes.c_double()
        ctx.check(L.fmx_subspace_loss(ctx.handle, g.n, NEV, Y.data_ptr(), g.U.data_ptr(), ctypes.byref(loss),
                                      dY.data_ptr(), None), "fmx_subspace_loss")
        ctx.check(L.fmx_gnn_backward(ctx.handle, g.spm.handle, g.dinv.data_ptr(), self.P.data_ptr(),
                                     g.X0.data_ptr(), acts.data_ptr(), dY.data_ptr(), self.G.data_ptr(),
                                     work.data_ptr(), self.mode, None), "fmx_gnn_backward")
        self.step += 1
        ctx.check(L.fmx_adam(ctx.handle, self.P.numel(), self.P.data_ptr(), self.G.data_ptr(),
                             self.m.data_ptr(), self.v.data_ptr(), lr, 0.9, 0.999, 1e-8, self.step, None),
                  "fmx_adam")
        return loss.value

    def loss(self, g):
        _, _, Y, _ = self._buffers(g.n)
        self.forward(g)
        out = ctypes.c_double()
        g.spm.ctx.check(F.lib().fmx_subspace_loss(g.spm.ctx.handle, g.n, NEV, Y.data_ptr(), g.U.data_ptr(),
                                                  ctypes.byref(out), None, None), "fmx_subspace_loss")
        return out.value

    def params(self):
        return self.P.cpu().numpy()


def training_sizes(count, lo=16, hi=20):
    """Vertex counts 2^lo .. 2^hi, most of them small (graphs are seeded)."""
    rng = np.random.default_rng(12345)
    e = rng.choice(np.arange(lo, hi + 1), size=count, p=None)
    return [1 << int(x) for x in np.sort(e)]


def build_training_set(count, seed0=1000, log=print, lo=16, hi=20):
    torch = _torch()
    graphs = []
    t0 = time.time()
    for i, n in enumerate(training_sizes(count, lo, hi)):
        g = Graph(n, seed0 + i)
        ev, vecs, res, st = g.solve(tol=1e-5, seed=seed0 + i)
        g.U = vecs.to(torch.float32).contiguous()
        del vecs
        graphs.append(g)
        if (i + 1) % 20 == 0:
            log("  target %d/%d (n=%d): %d SpMM blocks, %.2fs elapsed" % (i + 1, count, n, st["spmm_blocks"],
                                                                       time.time() - t0))
    return graphs


def train(graphs, epochs=40, lr=2e-3, seed=0, log=print, halve_every=10):
    net = WarmStartNet(seed=seed)
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(graphs))
        tot = 0.0
        t0 = time.time()
        cur_lr = lr * (0.5 ** (ep // halve_every))
        for i in order:
            tot += net.train_step(graphs[i], cur_lr)
        hist.append(tot / len(graphs))
        log("epoch %d: mean subspace loss %.4f (lr %.1e, %.1fs)" % (ep, hist[-1], cur_lr, time.time() - t0))
    return net, hist


def evaluate(net, count=20, n=1 << 20, seed0=900000, tol=1e-6, log=print):
    """Seconds and operator applications to residual `tol`, random vs learned start."""
    torch = _torch()
    rows = []
    for i in range(count):
        g = Graph(n, seed0 + i)
        g.warmup()
        net.forward(g)
        torch.cuda.synchronize()
        t0 = time.time()
        _, _, res_r, st_r = g.solve(tol=tol, seed=seed0 + i)
        t_rand = time.time() - t0
        torch.cuda.synchronize()
        t0 = time.time()
        Y = net.forward(g)
        torch.cuda.synchronize()
        t_net = time.time() - t0
        _, _, res_l, st_l = g.solve(X0=Y, tol=tol, seed=seed0 + i)
        t_learn = time.time() - t0
        rows.append({"seed": seed0 + i, "n": n, "random": {"sec": t_rand, "spmm_blocks": st_r["spmm_blocks"],
                                                         "restarts": st_r["restarts"],
                                                         "converged": st_r["converged"],
                                                         "max_resid": float(res_r.max())},
                     "learned": {"sec": t_learn, "net_sec": t_net, "spmm_blocks": st_l["spmm_blocks"],
                                 "restarts": st_l["restarts"], "converged": st_l["converged"],
                                 "max_resid": float(res_l.max())}})
        log("  held-out %d: random %.3fs (%d blocks), learned %.3fs (net %.3fs, %d blocks)" % (
            i, t_rand, st_r["spmm_blocks"], t_learn, t_net, st_l["spmm_blocks"]))
        del g
    sr = sum(r["random"]["sec"] for r in rows)
    sl = sum(r["learned"]["sec"] for r in rows)
    summary = {"graphs": count, "n": n, "tol": tol, "random_sec_total": sr, "learned_sec_total": sl,
               "speedup": sr / sl if sl > 0 else None,
               "random_spmm_blocks": int(sum(r["random"]["spmm_blocks"] for r in rows)),
               "learned_spmm_blocks": int(sum(r["learned"]["spmm_blocks"] for r in rows)),
               "all_converged": all(r["random"]["converged"] and r["learned"]["converged"] for r in rows)}
    return summary, rows


def analyse(net, count=3, n=1 << 20, seed0=900000, tol=1e-6, log=print):
    """Why the learned start gains what it gains, per held-out graph: the
    residual history of both solves (largest relative residual after each
    restart), and for each of the 16 converged eigenvectors v_i how
    localised it is (n * sum v_i^4: about 3 for a vector spread over the
    graph, n / (sites) for one concentrated on a few vertices) and how much
    of it the learned start spans (||Q^T v_i||^2 with Q an orthonormal basis
    of the start; a random start spans about 16/n of any vector)."""
    torch = _torch()
    rows = []
    for i in range(count):
        g = Graph(n, seed0 + i)
        g.warmup()
        ev, V, res_r, st_r = F.compute_eigen(g.spm, nev=NEV, block_size=16, num_blocks=6, tol=tol, seed=seed0 + i,
                                             precision=F.EIG_FP32, max_restarts=400, history=True)
        Y = net.forward(g).double()
        _, _, res_l, st_l = F.compute_eigen(g.spm, nev=NEV, block_size=16, num_blocks=6, tol=tol, seed=seed0 + i,
                                            precision=F.EIG_FP32, max_restarts=400, history=True, X0=Y.float())
        Q, _ = torch.linalg.qr(Y)
        cap = (Q.T.contiguous() @ V).pow(2).sum(0).cpu().numpy()
        ipr = (V.pow(4).sum(0) * n).cpu().numpy()
        loc = ipr > 1000.0
        row = {"seed": seed0 + i, "n": n, "eigenvalues": [round(float(x), 6) for x in ev],
               "n_ipr": [round(float(x), 1) for x in ipr], "learned_capture": [round(float(x), 3) for x in cap],
               "localised": int(loc.sum()),
               "capture_localised": round(float(cap[loc].mean()), 3) if loc.any() else None,
               "capture_spread": round(float(cap[~loc].mean()), 3) if (~loc).any() else None,
               "subspace_loss_of_start": round(float(NEV - cap.sum()), 3),
               "random": {"spmm_blocks": st_r["spmm_blocks"], "history": [float("%.3g" % x) for x in st_r["history"]]},
               "learned": {"spmm_blocks": st_l["spmm_blocks"],
                           "history": [float("%.3g" % x) for x in st_l["history"]]}}
        log("  held-out %d: %d of 16 eigenvectors localised (n*IPR > 1000); the learned start spans %.2f of the "
            "localised ones and %.2f of the spread ones on average; random %d blocks, learned %d" % (
                i, row["localised"], row["capture_localised"] or 0, row["capture_spread"] or 0,
                st_r["spmm_blocks"], st_l["spmm_blocks"]))
        rows.append(row)
        del g
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval", "analyse"])
    ap.add_argument("--graphs", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--halve-every", type=int, default=10, help="epochs between learning-rate halvings")
    ap.add_argument("--min-log2", type=int, default=16, help="smallest training graph, log2 of its vertices")
    ap.add_argument("--max-log2", type=int, default=20, help="largest training graph, log2 of its vertices")
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--eval-n", type=int, default=1 << 20)
    ap.add_argument("--report", default=None)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, flush=True)
    if args.cmd == "train":
        t0 = time.time()
        graphs = build_training_set(args.graphs, log=log, lo=args.min_log2, hi=args.max_log2)
        log("targets ready in %.1fs" % (time.time() - t0))
        net, hist = train(graphs, args.epochs, args.lr, log=log, halve_every=args.halve_every)
        os.makedirs(os.path.dirname(args.out), exist_ok=True)
        np.save(args.out, net.params())
        rep = {"graphs": args.graphs, "epochs": args.epochs, "sizes_log2": [args.min_log2, args.max_log2],
               "halve_every": args.halve_every, "lr": args.lr, "loss_history": hist,
               "seconds": time.time() - t0}
        if args.report:
            json.dump(rep, open(args.report, "w"), indent=1)
        log(json.dumps(rep))
    elif args.cmd == "analyse":
        net = WarmStartNet(params=np.load(args.weights))
        rows = analyse(net, args.graphs, args.eval_n, log=log)
        if args.report:
            json.dump(rows, open(args.report, "w"), indent=1)
        log(json.dumps(rows))
    else:
        net = WarmStartNet(params=np.load(args.weights))
        summary, rows = evaluate(net, args.graphs, args.eval_n, log=log)
        if args.report:
            json.dump({"summary": summary, "rows": rows}, open(args.report, "w"), indent=1)
        log(json.dumps(summary))


if __name__ == "__main__":
    main()
