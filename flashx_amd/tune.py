"""Tile-shape autotuner for the SpMM path.

The tile height H trades image size (taller tiles: fewer tile headers, more
rows per row part) against the LDS band each wave holds (H rows x k
accumulators) and hence the number of resident waves.  The tuner builds the
matrix on the device for each candidate (log2 H, log2 W), times a few SpMMs
with the caller's width and type, and returns the fastest.  Everything it
runs is the product path (fmx_spm_build, fmx_spmm)."""
import time

from . import fmx as F

DEFAULT_CANDIDATES = ((5, 15), (6, 15), (7, 15))


def _time_spmm(spm, X, Y, reps):
    torch = F._torch()
    spm.multiply_async(X, Y)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        spm.multiply_async(X, Y)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def autotune_tile(src, dst, n, k, dtype, flags=F.BUILD_DEDUP, candidates=DEFAULT_CANDIDATES, reps=10, X=None,
                  keep=False):
    """Returns ((log2_h, log2_w), {candidate: ms per SpMM}) for the edge list
    src, dst (device tensors) multiplied by an n x k operand of `dtype`.  With
    keep=True it also returns the fastest candidate's matrix and its build
    seconds, so the caller need not build it again."""
    torch = F._torch()
    Y = None
    times, best, best_spm, best_build = {}, None, None, None
    for lh, lw in candidates:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        spm = F.SparseMatrix.build(src, dst, n=n, log2_h=lh, log2_w=lw, flags=flags)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - t0
        if X is None:  # the operands after the first construction, which needs the most memory
            X = torch.rand((n, k), device=src.device, dtype=torch.float32).to(dtype)
        if Y is None:
            Y = torch.empty((n, k), device=src.device, dtype=dtype)
        times[(lh, lw)] = _time_spmm(spm, X, Y, reps)
        if best is None or times[(lh, lw)] < times[best]:
            best, best_build = (lh, lw), build_s
            best_spm = spm if keep else None
        del spm
    del Y
    if keep:
        return best, times, best_spm, best_build
    return best, times
