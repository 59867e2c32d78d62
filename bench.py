#!/usr/bin/env python3
"""SpMM throughput on one MI355X: Y = A*X with A the R-MAT scale-22,
edge-factor-16 binary adjacency (duplicates and self edges removed) in the
reference's 2D tile format, X a 4,194,304 x 16 FP32 operand U[0,1).

One step = one fmx_spmm over the whole matrix with A, X and Y resident in
HBM.  value = algorithmic bytes / step time, where algorithmic bytes =
tile bytes + 8*(block rows + 1) + X read once + Y written once.  The CPU
baseline is the oracle's multi-threaded restatement of the reference engine
on the same image and X, on this host's cores.

python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--no-cpu]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("SpMM HBM GB/s (% peak); eigensolve sec to 1e-6 residual, learned vs plain")
PEAK_HBM_GBS = 8000.0
SCALE, EDGE_FACTOR, K, SEED = 22, 16, 16, 22


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_traffic(workload, tile):
    """HBM bytes per launch measured with rocprofv3 --pmc (committed summary)."""
    p = os.path.join(ROOT, "profiles", "spmm_traffic.json")
    try:
        d = json.load(open(p))
        e = d.get("%s/%s" % (workload, tile))
        return e["hbm_bytes_per_launch"] if e else None
    except Exception:
        return None


def cpu_baseline(data, offs, n, X_host, threads, reps):
    from oracle import oracle as O
    img = O.Image(data, offs, n)
    O.spmm(img, X_host, threads=threads)  # warm-up
    times = []
    Y = None
    for _ in range(reps):
        t0 = time.perf_counter()
        Y = O.spmm(img, X_host, threads=threads)
        times.append(time.perf_counter() - t0)
    return statistics.median(times), Y


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log2-h", type=int, default=None,
                    help="tile height (log2); default: autotuned over 32/64/128 rows")
    ap.add_argument("--log2-w", type=int, default=15)
    ap.add_argument("--large", action="store_true",
                    help="also run the R-MAT scale-27, k=32 SpMM with a sampled CPU check")
    ap.add_argument("--large-log2-h", type=int, default=None,
                    help="tile height (log2) of the scale-27 leg; default: autotuned over 16/32/64 rows")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stream", action="store_true", help="skip the out-of-core (streamed image) leg")
    ap.add_argument("--no-floor", action="store_true", help="skip the gather-floor leg")
    ap.add_argument("--stream-chunk", type=int, default=64 << 20,
                    help="block-row chunk bytes of the streamed leg")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-eigen", action="store_true", help="skip the eigensolver legs")
    ap.add_argument("--eigen-only", action="store_true",
                    help="only the scale-22 eigensolve leg (one JSON line), e.g. under a profiler")
    ap.add_argument("--only-spmm", action="store_true",
                    help="only the timed SpMM steps (no other legs), e.g. under a profiler; "
                         "takes the tile shape from --log2-h/--log2-w (default 32 x 32768)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 1 GiB device copies (the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE into bytes)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if args.calibrate_copy:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 3 x %d bytes" % (a.numel() * 4))
        return
    import flashx_amd.fmx as F

    if args.eigen_only:
        print(json.dumps({"eigen": eigen_leg(F, torch, 6, 15)}), flush=True)
        return
    ctx = F.default_context(0)
    n = 1 << SCALE
    m = EDGE_FACTOR * n
    from flashx_amd.tune import autotune_tile
    src, dst = F.rmat_edges(SCALE, m, SEED)
    torch.cuda.synchronize()
    tune = None
    if args.only_spmm and args.log2_h is None:
        args.log2_h = 5
    if args.log2_h is None:
        (args.log2_h, args.log2_w), tt = autotune_tile(src, dst, n, K, torch.float32)
        tune = {"%dx%d" % (1 << a, 1 << b): round(v, 4) for (a, b), v in tt.items()}
        log("autotune (ms per SpMM): %s" % tune)
    t0 = time.perf_counter()
    spm = F.SparseMatrix.build(src, dst, n=n, log2_h=args.log2_h, log2_w=args.log2_w,
                               flags=F.BUILD_DEDUP)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    cols_seq = None if (args.no_floor or args.only_spmm) else stream_columns(torch, src, dst, n, args.log2_h,
                                                                             args.log2_w)
    big_image = None
    if not (args.no_cpu or args.only_spmm):
        # the same matrix in the reference's default 32768 x 32768 tiles, for the CPU baseline
        big = F.SparseMatrix.build(src, dst, n=n, log2_h=15, log2_w=15, flags=F.BUILD_DEDUP)
        big_image = big.download()
        big.close()
        del big
    del src, dst
    info = spm.info
    log("built: n=%d nnz=%d tiles=%d image=%d B (%.2f s for the whole build call)"
        % (info.nrows, info.nnz, info.ntiles, info.image_bytes, t_build))
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED)
    X = torch.rand((n, K), dtype=torch.float32, device="cuda", generator=g)
    Y = torch.empty((n, K), dtype=torch.float32, device="cuda")
    alg_bytes = spm.alg_bytes(torch.float32, K)

    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        spm.multiply_async(X, Y)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for i in range(args.steps):
        ev[i][0].record(stream)
        spm.multiply_async(X, Y)
        ev[i][1].record(stream)
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    per = [a.elapsed_time(b) for a, b in ev]
    ms_per_step = total_ms / args.steps
    avg_launch_ms = sum(per) / len(per)
    value = alg_bytes / (ms_per_step * 1e-3) / 1e9
    achieved = alg_bytes / (avg_launch_ms * 1e-3) / 1e9
    log("spmm: %.1f us/step (events per launch avg %.1f us, min %.1f)" %
        (ms_per_step * 1e3, avg_launch_ms * 1e3, min(per) * 1e3))

    workload = "rmat%d_ef%d_f32_k%d" % (SCALE, EDGE_FACTOR, K)
    tile = "%dx%d" % (1 << args.log2_h, 1 << args.log2_w)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded R-MAT (a,b,c,d)=(.57,.19,.19,.05), X U[0,1) seeded",
        "config": {"workload": "spmm " + workload, "scale": SCALE, "edge_factor": EDGE_FACTOR,
                   "k": K, "tile": tile, "tile_autotune_ms": tune, "nnz": int(info.nnz),
                   "build_s": round(t_build, 3),
                   "image_bytes": int(info.image_bytes), "alg_bytes_per_step": int(alg_bytes),
                   "pct_hbm_peak": round(100.0 * value / PEAK_HBM_GBS, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": load_traffic(workload, tile)},
        "cpu_baseline": None,
    }

    if args.only_spmm:
        print(json.dumps(out), flush=True)
        return
    out["operand_variants"] = variants_leg(F, torch, spm, n, check=not args.no_cpu)
    out["scale16"] = scale16_leg(F, torch, check=not args.no_cpu)
    if cols_seq is not None:
        out["gather_floor"] = gather_floor_leg(F, torch, X, cols_seq, ms_per_step)
        del cols_seq
    if not args.no_stream:
        out["streamed"] = streamed_leg(F, torch, spm, X, Y, args.stream_chunk)

    out["construction"] = construction_leg(F, torch, spm, t_build, check=not args.no_cpu)
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        data, offs = spm.download()
        Xh = X.cpu().numpy()
        t_cpu, Ycpu = cpu_baseline(data, offs, n, Xh, threads, args.cpu_reps)
        Yg = Y.cpu().numpy().astype(np.float64)
        ref = Ycpu.astype(np.float64)
        from oracle import oracle as O
        exact = O.spmm(O.Image(data, offs, n), Xh.astype(np.float64), threads=threads)
        nz = exact != 0
        rel = float(np.max(np.abs(Yg - ref)[nz] / np.abs(ref[nz])))
        rel64 = float(np.max(np.abs(Yg - exact)[nz] / np.abs(exact[nz])))
        cpu64 = float(np.max(np.abs(ref - exact)[nz] / np.abs(exact[nz])))
        zeros_ok = bool(np.all(Yg[~nz] == 0) and np.all(ref[~nz] == 0))
        extra = cpu_baseline_extra(F, torch, data, offs, big_image, n, Xh, threads, alg_bytes)
        del big_image
        out["cpu_baseline"] = {
            "value": round(alg_bytes / t_cpu / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port",
            "sample": "full image and X, 1 warm-up + median of %d multiply() calls, "
                      "seconds per SpMM %.4f" % (args.cpu_reps, t_cpu)}
        out["cpu_baseline"].update(extra)
        out["check"] = {"max_rel_err_vs_cpu": rel, "max_rel_err_vs_fp64_exact": rel64,
                        "cpu_fp32_max_rel_err_vs_fp64_exact": cpu64, "zero_rows_exact": zeros_ok}
    if not args.no_eigen or args.large:
        del spm, X, Y
        torch.cuda.empty_cache()
    if args.large:
        out["large"] = large_leg(F, torch, candidates=((args.large_log2_h, 15),) if args.large_log2_h is not None
                                 else ((4, 15), (5, 15), (6, 15)))
        torch.cuda.empty_cache()
    if not args.no_eigen:
        out["eigen"] = eigen_leg(F, torch, args.log2_h, args.log2_w)
        out["learned_warm_start"] = learned_leg(F, torch)
    print(json.dumps(out), flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_extra(F, torch, data, offs, big, n, Xh, threads, alg_bytes, reps=3):
    """The rest of the CPU baseline: one thread, the per-task block-row copy
    off (the reference's in_mem_io memcpy, in_mem_io.cpp:498-508), and the
    reference's default 32768 x 32768 tiles (fg2fm.cpp:155-156; `big`, the
    same matrix built in those tiles) on all threads, each the median of
    `reps` multiply() calls after a warm-up."""
    from oracle import oracle as O
    img = O.Image(data, offs, n)
    big = O.Image(big[0], big[1], n)

    def med(image, th, copy):
        O.spmm(image, Xh, threads=th, copy=copy)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            O.spmm(image, Xh, threads=th, copy=copy)
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    t1 = med(img, 1, True)
    toff = med(img, threads, False)
    tbig = med(big, threads, True)
    big_bytes = (len(big.data) - 4096) + 8 * (big.nbr + 1) + 2 * n * Xh.shape[1] * Xh.itemsize
    return {"cpu_model": cpu_model(), "threads_available": os.cpu_count(),
            "t1": {"value": round(alg_bytes / t1 / 1e9, 3), "sec": round(t1, 4)},
            "memcpy_off": {"value": round(alg_bytes / toff / 1e9, 3), "sec": round(toff, 4), "cores": threads},
            "default_32768_tiles": {"value": round(big_bytes / tbig / 1e9, 3), "sec": round(tbig, 4),
                                    "cores": threads, "alg_bytes": int(big_bytes),
                                    "tasks": int(big.nbr),
                                    "note": "32768 x 32768 tiles: one task per block row, as the reference "
                                            "runs it (%d tasks for %d threads)" % (big.nbr, threads)}}


def construction_leg(F, torch, spm, t_build, check=True):
    """Tile construction, counted as algorithmic bytes: 8*m (the edge list
    read) + the tile bytes + 8*(block rows + 1) (written), over the device
    time of the construction kernels (fmx_spm_info.build_ns); beside it the
    whole fmx_spm_build call (host validation and the work images
    included), and the oracle's single-threaded restatement of the
    reference writer (fm_utils.cpp:228-433,617-715) on a bounded sample of
    the same edge list (one 2^24-edge prefix)."""
    info = spm.info
    m = int(info.edges_in)
    alg = 8 * m + int(info.tile_bytes) + 8 * (int(info.nbr) + 1)
    dev_s = info.build_ns / 1e9
    out = {"edges": m, "alg_bytes": alg, "device_ms": round(dev_s * 1e3, 3),
           "device_GBps": round(alg / dev_s / 1e9, 1) if dev_s else None,
           "pct_hbm_peak": round(100 * alg / dev_s / 1e9 / PEAK_HBM_GBS, 2) if dev_s else None,
           "call_s": round(t_build, 3),
           "call_note": "the whole call: construction kernels (with the tile directory and its check), the "
                        "block-row offsets' download, and the gather-order image the SpMM runs on (decoded, "
                        "ranked and built on the device)"}
    if check:
        from oracle import oracle as O
        import numpy as np
        sample = 1 << 24
        s, d = F.rmat_edges(SCALE, sample, SEED)  # the first 2^24 edges of the same seeded list
        sh = s.cpu().numpy().view(np.uint32)
        dh = d.cpu().numpy().view(np.uint32)
        del s, d
        lh = int(np.log2(info.block_h))
        lw = int(np.log2(info.block_w))
        t0 = time.perf_counter()
        img = O.build_2d(sh, dh, n=1 << SCALE, log2h=lh, log2w=lw, flags=O.DEDUP)
        t_cpu = time.perf_counter() - t0
        calg = 8 * sample + (len(img.data) - 4096) + 8 * (img.nbr + 1)
        out["cpu_baseline"] = {"value": round(calg / t_cpu / 1e9, 3), "unit": "GB/s", "cores": 1,
                               "kind": "port", "sec": round(t_cpu, 3),
                               "sample": "the first 2^24 edges of the same seeded list, same tiles"}
    log("construction: %.2f ms of kernels (%.1f GB/s, %s%% of peak), %.2f s per call"
        % (out["device_ms"], out["device_GBps"] or 0, out["pct_hbm_peak"], t_build))
    return out


def variants_leg(F, torch, spm, n, steps=10, check=True):
    """The scale-22 matrix times its other named operands: 4,194,304 x 16 FP64
    U[0,1) and int32 {0..7} (seeded), timed like the FP32 product; with
    check, int32 is compared bit for bit and FP64 within 1e-12 relative with
    the oracle's reference-order CPU product."""
    import numpy as np
    from oracle import oracle as O
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED + 1)
    out = {}
    img = None
    if check:
        data, offs = spm.download()
        img = O.Image(data, offs, n)
    threads = len(os.sched_getaffinity(0))
    for name, X in (("f64_k16", torch.rand((n, K), dtype=torch.float64, device="cuda", generator=g)),
                    ("i32_k16", torch.randint(0, 8, (n, K), dtype=torch.int32, device="cuda", generator=g))):
        Y = torch.empty_like(X)
        spm.multiply_async(X, Y)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            spm.multiply_async(X, Y)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        r = {"ms": round(ms, 4), "GBps": round(spm.alg_bytes(X.dtype, K) / ms / 1e6, 1)}
        if img is not None:
            ref = O.spmm(img, X.cpu().numpy(), threads=threads)
            got = Y.cpu().numpy()
            if X.dtype == torch.int32:
                r["bit_exact_vs_cpu"] = bool(np.array_equal(got, ref))
            else:
                nz = ref != 0
                r["max_rel_err_vs_cpu"] = float((np.abs(got - ref)[nz] / np.abs(ref[nz])).max())
                r["zero_rows_exact"] = bool(np.all(got[~nz] == 0))
        log("spmm %s: %.3f ms, %.1f GB/s %s" % (name, ms, r["GBps"],
                                               {k: v for k, v in r.items() if k not in ("ms", "GBps")}))
        out[name] = r
        del X, Y
    return out


def scale16_leg(F, torch, check=True, steps=20):
    """The CPU-runnable case: R-MAT scale 16, edge factor 16, dedup, times
    65,536 x 8 FP64 U[0,1) and int32 {0..7} operands, in the reference's
    default 32768 x 32768 tiles and in 64 x 32768 tiles; int32 bit-exact and
    FP64 within 1e-12 relative of the oracle's reference-order CPU product.
    The default tiles leave 2 block rows: each of the 256 LDS bands (hs = 256
    rows at k = 8 FP64) decodes its whole block row, so that shape shows the
    cost of tiles much taller than a band, not the kernel's throughput."""
    import numpy as np
    from oracle import oracle as O
    scale, k = 16, 8
    n = 1 << scale
    src, dst = F.rmat_edges(scale, EDGE_FACTOR * n, scale)
    g = torch.Generator(device="cuda")
    g.manual_seed(scale)
    Xs = (("f64_k8", torch.rand((n, k), dtype=torch.float64, device="cuda", generator=g)),
          ("i32_k8", torch.randint(0, 8, (n, k), dtype=torch.int32, device="cuda", generator=g)))
    out = {"workload": "spmm rmat16_ef16, k=8"}
    for lh, lw in ((15, 15), (6, 15)):
        spm = F.SparseMatrix.build(src, dst, n=n, log2_h=lh, log2_w=lw, flags=F.BUILD_DEDUP)
        tile = "%dx%d" % (1 << lh, 1 << lw)
        res = {"nnz": int(spm.info.nnz)}
        img = None
        if check:
            data, offs = spm.download()
            img = O.Image(data, offs, n)
        for name, X in Xs:
            Y = torch.empty_like(X)
            spm.multiply_async(X, Y)
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(steps):
                spm.multiply_async(X, Y)
            b.record()
            torch.cuda.synchronize()
            r = {"us": round(a.elapsed_time(b) / steps * 1e3, 1)}
            if img is not None:
                ref = O.spmm(img, X.cpu().numpy())
                got = Y.cpu().numpy()
                if X.dtype == torch.int32:
                    r["bit_exact_vs_cpu"] = bool(np.array_equal(got, ref))
                else:
                    nz = ref != 0
                    r["max_rel_err_vs_cpu"] = float((np.abs(got - ref)[nz] / np.abs(ref[nz])).max())
            log("scale 16, %s tiles, %s: %.1f us %s" % (tile, name, r["us"],
                                                       {kk: v for kk, v in r.items() if kk != "us"}))
            res[name] = r
        out[tile] = res
        del spm
    return out


def stream_columns(torch, src, dst, n, lh, lw, chunk_edges=1 << 28):
    """Columns of the non-zeros in the tiles' order -- block row, block
    column, then row and column inside the tile (the image's order up to the
    split of each tile into its multi-entry and single-entry rows) -- from
    the edge list after self-edge removal and dedup, as the build does.
    Edge lists beyond chunk_edges are taken in ranges of whole block rows,
    which the order puts one after the other."""
    m = src.numel()
    nbr = (n + (1 << lh) - 1) >> lh
    nchunks = max(1, (m + chunk_edges - 1) // chunk_edges)
    step = ((nbr + nchunks - 1) // nchunks) << lh  # rows per range, whole block rows
    bits = max(1, (n - 1).bit_length())
    out = []
    piece = 1 << 29  # torch indexing of 2^31 elements or more overflows: work on slices
    for lo in range(0, n, step):
        ss, ds = [], []
        for p0 in range(0, m, piece):
            sp = src[p0:p0 + piece].long() & 0xFFFFFFFF
            dp = dst[p0:p0 + piece].long() & 0xFFFFFFFF
            keep = (sp >= lo) & (sp < lo + step)
            ss.append(sp[keep])
            ds.append(dp[keep])
            del sp, dp, keep
        s = torch.cat(ss)
        d = torch.cat(ds)
        del ss, ds
        keep = s != d
        u = torch.unique(s[keep] * n + d[keep])  # sorted by (row, column)
        del s, d, keep
        # a stable sort by (block row, block column) keeps (row, column) order inside each tile
        tile = (((u // n) >> lh) << max(0, bits - lw)) | ((u % n) >> lw)
        order = torch.sort(tile, stable=True).indices
        del tile
        out.append((u[order] % n).to(torch.int32))
        del u, order
    return torch.cat(out).contiguous() if len(out) > 1 else out[0].contiguous()


def gather_floor_leg(F, torch, X, cols, spmm_ms, steps=10):
    """The gather-only pass over the SpMM's rows of X in the SpMM's order:
    the memory floor of any kernel fetching the same rows in that order."""
    for _ in range(2):
        F.gather_floor(X, cols)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        F.gather_floor(X, cols)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    log("gather floor: %.3f ms for %d rows in tile order (SpMM %.2fx of it)" % (ms, cols.numel(), spmm_ms / ms))
    return {"order": "tile order of the non-zeros", "rows": int(cols.numel()), "ms": round(ms, 4),
            "rows_per_s": round(cols.numel() / ms * 1e3), "spmm_over_floor": round(spmm_ms / ms, 3)}


def streamed_leg(F, torch, spm, X, Y_ref, chunk_bytes, steps=10):
    """The same SpMM with the image kept in pinned host memory and streamed
    through HBM in block-row chunks (out of core): time per SpMM with the
    host-to-device copies inside it, and bit-equality with the resident
    result."""
    data, offs = spm.download()
    st = F.SparseMatrix.upload_streamed(data, offs, chunk_bytes=chunk_bytes)
    del data
    Y = torch.empty_like(Y_ref)
    for _ in range(2):
        st.multiply_async(X, Y)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        st.multiply_async(X, Y)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    # the resident result of the same X (Y_ref holds the last timed product)
    same = bool(torch.equal(Y, Y_ref))
    tile_bytes = st.info.tile_bytes
    log("streamed: %.3f ms per SpMM in %d chunks, image %.1f GB/s over the host link, same as resident: %s"
        % (ms, st.info.stream_chunks, tile_bytes / ms / 1e6, same))
    out = {"chunk_bytes": chunk_bytes, "chunks": int(st.info.stream_chunks), "ms_per_step": round(ms, 4),
           "image_GBps_host_to_hbm": round(tile_bytes / ms / 1e6, 2), "equal_to_resident": same}
    st.close()
    return out


def large_leg(F, torch, scale=27, k=32, steps=5, candidates=((4, 15), (5, 15), (6, 15))):
    """R-MAT scale 27 (2^27 vertices, 2^31 edges), X 2^27 x 32 FP32, all in
    HBM; the tile height is autotuned over 16/32/64 rows (the fastest
    candidate's matrix is kept); FP32 check of a seeded sample of >= 1M rows
    (whole block rows) against the oracle's reference-order CPU SpMM."""
    import numpy as np
    from oracle import oracle as O
    from flashx_amd.tune import autotune_tile
    n = 1 << scale
    g = torch.Generator(device="cuda")
    g.manual_seed(scale)
    log("large: generating %d edges" % (EDGE_FACTOR * n))
    src, dst = F.rmat_edges(scale, EDGE_FACTOR * n, scale)
    if len(candidates) == 1:
        # one shape: build before X exists, so the construction has the memory
        (lh, lw), = candidates
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        spm = F.SparseMatrix.build(src, dst, n=n, log2_h=lh, log2_w=lw, flags=F.BUILD_DEDUP)
        torch.cuda.synchronize()
        t_build = time.perf_counter() - t0
        dev_ms = spm.info.build_ns / 1e6
        tune = None
        X = torch.rand((n, k), dtype=torch.float32, device="cuda", generator=g)
    else:
        # every candidate built with nothing else resident but the edge list,
        # the fastest built again at the end; the seeded X after that
        log("large: autotuning the tile shape over %s" % (candidates,))
        (lh, lw), tt, spm, t_build, dev_ms = autotune_tile(src, dst, n, k, torch.float32, candidates=candidates,
                                                           keep=True, lean=True)
        tune = {"%dx%d" % (1 << a, 1 << b): round(v, 2) for (a, b), v in tt.items()}
        log("large: autotune (ms per SpMM): %s" % tune)
        X = torch.rand((n, k), dtype=torch.float32, device="cuda", generator=g)
    log("large: %dx%d tiles built in %.1f s, nnz %d, image %.2f GB" % (1 << lh, 1 << lw, t_build, spm.info.nnz,
                                                                      spm.info.image_bytes / 1e9))
    cols = stream_columns(torch, src, dst, n, lh, lw)
    del src, dst
    torch.cuda.empty_cache()
    Y = torch.empty((n, k), dtype=torch.float32, device="cuda")
    for _ in range(2):
        spm.multiply_async(X, Y)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        spm.multiply_async(X, Y)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    alg = spm.alg_bytes(torch.float32, k)
    floor = gather_floor_leg(F, torch, X, cols, ms, steps=3)
    del cols
    log("large: %.3f ms per SpMM; checking a sample on the CPU" % ms)
    H = 1 << lh
    rng = np.random.default_rng(scale)
    nb = (1 << 20) // H
    br0 = int(rng.integers(0, spm.info.nbr - nb))
    data, offs = spm.download()
    img = O.Image(data, offs, n)
    Xh = X.cpu().numpy()
    t0 = time.perf_counter()
    ref = O.spmm_range(img, Xh, br0, br0 + nb)
    t_cpu = time.perf_counter() - t0
    ref64 = O.spmm_range(img, Xh.astype(np.float64), br0, br0 + nb)
    del Xh
    got = Y[br0 * H:(br0 + nb) * H].cpu().numpy().astype(np.float64)
    nzm = ref64 != 0
    rel = float((np.abs(got - ref)[nzm] / np.abs(ref[nzm])).max())
    rel64 = float((np.abs(got - ref64)[nzm] / np.abs(ref64[nzm])).max())
    log("large: %.3f ms per SpMM, %.1f GB/s, sample rel err %.2e" % (ms, alg / ms / 1e6, rel))
    return {"workload": "spmm rmat%d_ef%d_f32_k%d" % (scale, EDGE_FACTOR, k), "nnz": int(spm.info.nnz),
            "image_bytes": int(spm.info.image_bytes), "tile": "%dx%d" % (H, 1 << lw), "tile_autotune_ms": tune,
            "build_s": round(t_build, 2), "build_device_ms": round(dev_ms, 1),
            "ms_per_step": round(ms, 4), "GBps": round(alg / ms / 1e6, 1),
            "pct_hbm_peak": round(100 * alg / ms / 1e6 / PEAK_HBM_GBS, 2), "gather_floor": floor,
            "sample": {"rows": nb * H, "first_row": br0 * H, "max_rel_err_vs_cpu_fp32": rel,
                       "max_rel_err_vs_fp64": rel64,
                       "cpu_sec_single_thread": round(t_cpu, 2)}}


def eigen_leg(F, torch, log2_h, log2_w):
    """Top-16 eigenpairs of D^-1/2 A D^-1/2 for the symmetrised scale-22
    R-MAT graph: block 16, tolerance 1e-6 on every pair's FP64 residual.
    The default mode (FP32 SpMM, FP64-MFMA Gram and tall products, FP64
    Rayleigh-Ritz and residuals) is timed, then the mixed mode whose first
    projection pass uses BF16-MFMA Grams with FP32 accumulation; each
    result lists every pair's FP64 residual ||Nx - theta x|| / |theta|.

    What the solve finds: for every connected component with an edge,
    D^1/2 times its indicator is an eigenvector with eigenvalue 1, so
    eigenvalue 1 has one eigenvector per such component (counted here with
    scipy on the edge list).  This graph has hundreds, so all 16 wanted
    eigenvalues are 1 and the solve finds 16 orthonormal vectors of that
    eigenspace.  The same solve on the largest connected component alone
    (vertices relabelled 0..n_c-1) has a simple eigenvalue 1 and then
    distinct ones; it is timed the same way and reported with lambda_17
    (from an untimed solve for 20 pairs) and the gap lambda_16 - lambda_17."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.csgraph as csg
    n = 1 << SCALE
    src, dst = F.rmat_edges(SCALE, EDGE_FACTOR * n, SEED)
    spm = F.SparseMatrix.build(src, dst, n=n, log2_h=log2_h, log2_w=log2_w,
                               flags=F.BUILD_DEDUP | F.BUILD_SYMMETRIZE)
    sh = src.cpu().numpy().view(np.uint32)
    dh = dst.cpu().numpy().view(np.uint32)
    del src, dst
    F.compute_eigen(spm, nev=16, block_size=16, tol=1e-2, precision=F.EIG_FP32, seed=3)  # warm-up

    def solve(A, precision, nev=16):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev, vecs, res, st = F.compute_eigen(A, nev=nev, block_size=16, tol=1e-6, precision=precision,
                                            seed=1, max_restarts=400)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, ev, res, st

    def row(sec, ev, res, st):
        return {"sec": round(sec, 4), "spmm_blocks": int(st["spmm_blocks"]), "restarts": int(st["restarts"]),
                "converged": int(st["converged"]), "max_resid_fp64": float(res.max()),
                "resid_fp64_per_pair": [float("%.3g" % x) for x in res],
                "eigenvalues": [float("%.10f" % x) for x in ev]}

    sec, ev, res, st = solve(spm, F.EIG_FP32)
    log("eigen: %.3f s, %d block SpMMs, max residual %.2e" % (sec, st["spmm_blocks"], res.max()))
    msec, mev, mres, mst = solve(spm, F.EIG_MIXED)
    log("eigen (bf16 first-pass grams): %.3f s, %d block SpMMs, max residual %.2e"
        % (msec, mst["spmm_blocks"], mres.max()))
    out = {"workload": "rmat22_ef16 symmetrised, top-16 of D^-1/2 A D^-1/2, block 16, tol 1e-6",
           "precision": "fp32 spmm, fp64-mfma gram and tall products, fp64 rr/residual",
           "nnz": int(spm.info.nnz)}
    out.update(row(sec, ev, res, st))
    out["mixed"] = {"precision": "fp32 spmm, bf16-mfma first-pass gram (fp32 accumulation), fp64 second pass, "
                                 "rr and residual"}
    out["mixed"].update(row(msec, mev, mres, mst))
    out["mixed"]["max_abs_eigenvalue_diff_vs_default"] = float(abs(mev - ev).max())
    del spm
    torch.cuda.empty_cache()
    # the eigenvalue-1 eigenspace: one vector per component with an edge
    keep = sh != dh
    A = sp.coo_matrix((np.ones(int(keep.sum()), np.int8), (sh[keep], dh[keep])), shape=(n, n)).tocsr()
    ncomp, labels = csg.connected_components(A, directed=False)
    del A
    touched = np.zeros(n, bool)
    touched[sh[keep]] = True
    touched[dh[keep]] = True
    comp_edge = np.unique(labels[touched])
    sizes = np.bincount(labels[touched], minlength=ncomp)
    lcc = int(np.argmax(sizes))
    out["eigenvalue_1_multiplicity"] = int(comp_edge.size)
    out["what_is_measured"] = ("eigenvalue 1 has %d eigenvectors (one per component with an edge), so the 16 "
                               "wanted pairs are 16 orthonormal vectors of that eigenspace" % comp_edge.size)
    # the largest connected component alone
    inl = labels == lcc
    newid = np.cumsum(inl) - 1
    sel = keep & inl[sh]
    nl = int(inl.sum())
    s2 = torch.from_numpy(newid[sh[sel]].astype(np.int32)).cuda()
    d2 = torch.from_numpy(newid[dh[sel]].astype(np.int32)).cuda()
    del sh, dh, keep, labels, newid, sel
    L = F.SparseMatrix.build(s2, d2, n=nl, log2_h=log2_h, log2_w=log2_w, flags=F.BUILD_DEDUP | F.BUILD_SYMMETRIZE)
    del s2, d2
    F.compute_eigen(L, nev=16, block_size=16, tol=1e-2, precision=F.EIG_FP32, seed=3)  # warm-up
    lsec, lev, lres, lst = solve(L, F.EIG_FP32)
    lmsec, lmev, lmres, lmst = solve(L, F.EIG_MIXED)
    _, ev20, _, st20 = solve(L, F.EIG_FP32, nev=20)
    lc = {"workload": "largest connected component of the same graph: %d vertices, %d non-zeros" % (nl, L.info.nnz),
          "vertices": nl, "nnz": int(L.info.nnz)}
    lc.update(row(lsec, lev, lres, lst))
    lc["mixed"] = row(lmsec, lmev, lmres, lmst)
    lc["lambda_17"] = float("%.10f" % ev20[16]) if st20["converged"] else None
    lc["gap_16_17"] = float("%.3g" % (ev20[15] - ev20[16])) if st20["converged"] else None
    log("eigen, largest component (%d vertices): %.3f s (%d block SpMMs), mixed %.3f s; lambda_2 %.6f, "
        "lambda_16 %.6f, lambda_17 %s" % (nl, lsec, lst["spmm_blocks"], lmsec, lev[1], lev[15], lc["lambda_17"]))
    out["largest_component"] = lc
    del L
    torch.cuda.empty_cache()
    return out


def learned_leg(F, torch, graphs=20):
    """The 20 held-out DC-SBM graphs (2^20 vertices, seeds 900000..): seconds
    and block SpMMs to residual 1e-6 from a random start and from the trained
    network's start (its forward pass included)."""
    from flashx_amd import learned as Lm
    if not os.path.exists(Lm.DEFAULT_WEIGHTS):
        return {"status": "no trained weights at flashx_amd/weights/warmstart.npy"}
    import numpy as np
    net = Lm.WarmStartNet(params=np.load(Lm.DEFAULT_WEIGHTS))
    summary, rows = Lm.evaluate(net, count=graphs, n=1 << 20, seed0=900000, tol=1e-6, log=log)
    rs = sum(r["random"]["sec"] for r in rows)
    ls = sum(r["learned"]["sec"] for r in rows)
    return {"workload": "DC-SBM n=2^20, 32 blocks, ratio 20, gamma 2.5, mean degree 16, %d held-out graphs "
                        "(seeds 900000..%d)" % (graphs, 900000 + graphs - 1),
            "random_sec": round(rs, 4), "learned_sec": round(ls, 4),
            "random_spmm_blocks": sum(r["random"]["spmm_blocks"] for r in rows),
            "learned_spmm_blocks": sum(r["learned"]["spmm_blocks"] for r in rows),
            "speedup": round(rs / ls, 3),
            "all_converged": all(r["random"]["converged"] and r["learned"]["converged"] for r in rows),
            "max_resid": max(max(r["random"]["max_resid"], r["learned"]["max_resid"]) for r in rows)}


if __name__ == "__main__":
    main()
