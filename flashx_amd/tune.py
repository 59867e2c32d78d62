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
                  keep=False, lean=False):
    """Returns ((log2_h, log2_w), {candidate: ms per SpMM}) for the edge list
    src, dst (device tensors) multiplied by an n x k operand of `dtype`.  With
    keep=True it also returns the fastest candidate's matrix, its build
    seconds and its construction kernels' milliseconds (fmx_spm_info.build_ns),
    so the caller need not build it again.

    lean=True is for matrices whose construction needs most of HBM: every
    construction runs with nothing else of the tuner's allocated (its own
    operands and the best matrix so far are freed first), and with keep=True
    the fastest candidate is built once more at the end (the seconds
    returned are still those of its construction in the sweep).  A caller's
    X stays allocated throughout, so a lean caller passes none."""
    torch = F._torch()
    Y = None
    own_x = X is None
    times, best, best_spm, best_build, best_dev_ms = {}, None, None, None, None
    for lh, lw in candidates:
        if lean:
            Y = best_spm = None
            if own_x:
                X = None
            torch.cuda.empty_cache()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        spm = F.SparseMatrix.build(src, dst, n=n, log2_h=lh, log2_w=lw, flags=flags)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - t0
        if X is None:  # the operands after the construction, which needs the most memory
            X = torch.rand((n, k), device=src.device, dtype=torch.float32).to(dtype)
        if Y is None:
            Y = torch.empty((n, k), device=src.device, dtype=dtype)
        times[(lh, lw)] = _time_spmm(spm, X, Y, reps)
        if best is None or times[(lh, lw)] < times[best]:
            best, best_build, best_dev_ms = (lh, lw), build_s, spm.info.build_ns / 1e6
            best_spm = spm if keep and not lean else None
        del spm
    del Y
    if own_x:
        del X
    if keep and lean:
        # (the build figures stay those of the winner's construction in the
        # sweep: this one follows other constructions' frees and is not a
        # fresh build)
        torch.cuda.empty_cache()
        best_spm = F.SparseMatrix.build(src, dst, n=n, log2_h=best[0], log2_w=best[1], flags=flags)
    if keep:
        return best, times, best_spm, best_build, best_dev_ms
    return best, times
