"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py as the checker / CPU baseline.  The
product package (flashx_amd) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

I32, I64, F32, F64 = 0, 1, 2, 3
DEDUP, KEEP_SELF, SYMMETRIZE = 1, 2, 4
# size_t operands (P_ULONG) reuse the int64 path: two's-complement sums and
# products modulo 2^64 are the same bits (the C restatement keeps the test
# values below the int64 range, where no signed overflow occurs)
_NP2DT = {np.dtype(np.int32): I32, np.dtype(np.int64): I64, np.dtype(np.uint64): I64,
          np.dtype(np.float32): F32, np.dtype(np.float64): F64}

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_rmat_perm.restype = ctypes.c_uint32
        L.orc_rmat_perm.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64]
        L.orc_rmat_edges.restype = None
        L.orc_rmat_edges.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_int, _u32p, _u32p]
        L.orc_build_2d.restype = ctypes.c_int
        L.orc_build_2d.argtypes = [_u32p, _u32p, ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                   ctypes.POINTER(_u8p), _u64p, ctypes.POINTER(_i64p),
                                   _u64p, _u64p, _u64p]
        L.orc_normalize.restype = ctypes.c_int
        L.orc_normalize.argtypes = [_u32p, _u32p, ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.c_uint, ctypes.POINTER(_u64p), _u64p, _u64p]
        L.orc_free.restype = None
        L.orc_free.argtypes = [_vp]
        L.orc_verify_image.restype = ctypes.c_int
        L.orc_verify_image.argtypes = [_u8p, ctypes.c_uint64, _i64p, ctypes.c_uint64]
        L.orc_image_nonzeros.restype = ctypes.c_int
        L.orc_image_nonzeros.argtypes = [_u8p, ctypes.c_uint64, _i64p, ctypes.c_uint64,
                                         _u64p, _u64p, ctypes.c_uint64, _u64p]
        L.orc_spmm.restype = ctypes.c_int
        L.orc_spmm.argtypes = [_u8p, ctypes.c_uint64, _i64p, ctypes.c_uint64,
                               ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, _vp,
                               ctypes.c_int64]
        L.orc_spmm_mt.restype = ctypes.c_int
        L.orc_spmm_mt.argtypes = [_u8p, ctypes.c_uint64, _i64p, ctypes.c_uint64,
                                  ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, _vp,
                                  ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        for nm in ("orc_gram_f64",):
            getattr(L, nm).restype = None
        L.orc_gram_f64.argtypes = [_vp, ctypes.c_int64, _vp, ctypes.c_int64,
                                   ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _vp]
        L.orc_gemm_tall_f64.restype = None
        L.orc_gemm_tall_f64.argtypes = [_vp, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int,
                                        _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        _vp, ctypes.c_int64]
        L.orc_colnorm_f64.restype = None
        L.orc_colnorm_f64.argtypes = [_vp, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, _vp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def rmat_thresholds(a=0.57, b=0.19, c=0.19):
    """floor(cumulative probability * 2^32), clamped to u32."""
    f = lambda x: min(int(np.floor(x * 4294967296.0)), 0xFFFFFFFF)
    return f(a), f(a + b), f(a + b + c)


def splitmix64(key, ctr):
    return lib().orc_splitmix64(key, ctr)


def rmat_perm(v, scale, seed):
    return lib().orc_rmat_perm(v, scale, seed)


def rmat_edges(scale, m, seed, a=0.57, b=0.19, c=0.19, permute=True):
    ta, tab, tabc = rmat_thresholds(a, b, c)
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    lib().orc_rmat_edges(scale, m, seed, ta, tab, tabc, int(permute),
                         _p(src, _u32p), _p(dst, _u32p))
    return src, dst


class Image:
    """A reference-format tile image plus its block-row offsets."""

    def __init__(self, data, offs, n, nnz=None):
        self.data = np.ascontiguousarray(data, dtype=np.uint8)
        self.offs = np.ascontiguousarray(offs, dtype=np.int64)
        self.nbr = len(self.offs) - 1
        self.n = n
        self.nnz = nnz

    @property
    def H(self):
        return int(self.data[48:56].view(np.uint64)[0])

    @property
    def W(self):
        return int(self.data[56:64].view(np.uint64)[0])


def build_2d(src, dst, n=0, log2h=15, log2w=15, flags=DEDUP):
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    img = _u8p()
    nbytes = ctypes.c_uint64()
    offs = _i64p()
    nbr = ctypes.c_uint64()
    n_out = ctypes.c_uint64()
    nnz = ctypes.c_uint64()
    rc = lib().orc_build_2d(_p(src, _u32p), _p(dst, _u32p), len(src), n, log2h, log2w,
                            flags, ctypes.byref(img), ctypes.byref(nbytes),
                            ctypes.byref(offs), ctypes.byref(nbr), ctypes.byref(n_out),
                            ctypes.byref(nnz))
    if rc:
        raise ValueError("orc_build_2d failed with status %d" % rc)
    data = np.ctypeslib.as_array(img, shape=(nbytes.value,)).copy()
    o = np.ctypeslib.as_array(offs, shape=(nbr.value + 1,)).copy()
    lib().orc_free(ctypes.cast(img, _vp))
    lib().orc_free(ctypes.cast(offs, _vp))
    return Image(data, o, n_out.value, nnz.value)


_PRIM = {np.dtype(np.int32): 4, np.dtype(np.int64): 6, np.dtype(np.uint64): 7, np.dtype(np.float32): 8,
         np.dtype(np.float64): 9}


def build_2d_w(src, dst, vals, n=0, log2h=15, log2w=15, flags=DEDUP):
    """Weighted image: vals (int32/int64/float32/float64, one per edge)
    become the tiles' values regions; duplicates keep the first value."""
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    vals = np.ascontiguousarray(vals)
    prim = _PRIM[vals.dtype]
    L = lib()
    L.orc_build_2d_w.restype = ctypes.c_int
    L.orc_build_2d_w.argtypes = [_u32p, _u32p, _vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                 ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                 ctypes.POINTER(_u8p), _u64p, ctypes.POINTER(_i64p), _u64p, _u64p, _u64p]
    img = _u8p()
    nbytes = ctypes.c_uint64()
    offs = _i64p()
    nbr = ctypes.c_uint64()
    n_out = ctypes.c_uint64()
    nnz = ctypes.c_uint64()
    rc = L.orc_build_2d_w(_p(src, _u32p), _p(dst, _u32p), vals.ctypes.data, vals.itemsize, prim, len(src), n,
                          log2h, log2w, flags, ctypes.byref(img), ctypes.byref(nbytes), ctypes.byref(offs),
                          ctypes.byref(nbr), ctypes.byref(n_out), ctypes.byref(nnz))
    if rc:
        raise ValueError("orc_build_2d_w failed with status %d" % rc)
    data = np.ctypeslib.as_array(img, shape=(nbytes.value,)).copy()
    o = np.ctypeslib.as_array(offs, shape=(nbr.value + 1,)).copy()
    L.orc_free(ctypes.cast(img, _vp))
    L.orc_free(ctypes.cast(offs, _vp))
    return Image(data, o, n_out.value, nnz.value)


def normalize(src, dst, n=0, flags=DEDUP):
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    keys = _u64p()
    nk = ctypes.c_uint64()
    n_out = ctypes.c_uint64()
    rc = lib().orc_normalize(_p(src, _u32p), _p(dst, _u32p), len(src), n, flags,
                             ctypes.byref(keys), ctypes.byref(nk), ctypes.byref(n_out))
    if rc:
        raise ValueError("orc_normalize failed with status %d" % rc)
    out = np.ctypeslib.as_array(keys, shape=(max(nk.value, 1),))[:nk.value].copy()
    lib().orc_free(ctypes.cast(keys, _vp))
    return out, n_out.value


def verify(image):
    return lib().orc_verify_image(_p(image.data, _u8p), len(image.data),
                                  _p(image.offs, _i64p), image.nbr)


def nonzeros(image, cap=None):
    if cap is None:
        cap = len(image.data)  # at most one non-zero per two image bytes
    rows = np.empty(cap, np.uint64)
    cols = np.empty(cap, np.uint64)
    cnt = ctypes.c_uint64()
    rc = lib().orc_image_nonzeros(_p(image.data, _u8p), len(image.data),
                                  _p(image.offs, _i64p), image.nbr,
                                  _p(rows, _u64p), _p(cols, _u64p), cap,
                                  ctypes.byref(cnt))
    if rc:
        raise ValueError("orc_image_nonzeros failed with status %d" % rc)
    return rows[:cnt.value], cols[:cnt.value]


def spmm(image, X, threads=0, copy=True):
    """Y = A*X in the reference order (threads=0) or with the multi-threaded
    restatement of the reference engine (threads>=1)."""
    X = np.ascontiguousarray(X)
    if X.ndim == 1:
        X = X.reshape(-1, 1)
    dt = _NP2DT[X.dtype]
    n_rows = int(image.data[24:32].view(np.uint64)[0])
    k = X.shape[1]
    Y = np.empty((n_rows, k), X.dtype)
    if threads <= 0:
        rc = lib().orc_spmm(_p(image.data, _u8p), len(image.data), _p(image.offs, _i64p),
                            image.nbr, dt, k, X.ctypes.data, k, Y.ctypes.data, k)
    else:
        rc = lib().orc_spmm_mt(_p(image.data, _u8p), len(image.data),
                               _p(image.offs, _i64p), image.nbr, dt, k, X.ctypes.data,
                               k, Y.ctypes.data, k, threads, int(copy))
    if rc:
        raise ValueError("orc_spmm failed with status %d" % rc)
    return Y


def spmm_range(image, X, br0, br1):
    """Rows of block rows [br0, br1) of A*X, in the reference order."""
    X = np.ascontiguousarray(X)
    dt = _NP2DT[X.dtype]
    H = int(image.data[48:56].view(np.uint64)[0])
    n_rows = int(image.data[24:32].view(np.uint64)[0])
    rows = min(br1 * H, n_rows) - br0 * H
    k = X.shape[1]
    Y = np.empty((rows, k), X.dtype)
    L = lib()
    L.orc_spmm_range.restype = ctypes.c_int
    L.orc_spmm_range.argtypes = [_u8p, ctypes.c_uint64, _i64p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                 ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, _vp, ctypes.c_int64]
    rc = L.orc_spmm_range(_p(image.data, _u8p), len(image.data), _p(image.offs, _i64p), image.nbr, br0, br1,
                          dt, k, X.ctypes.data, k, Y.ctypes.data, k)
    if rc:
        raise ValueError("orc_spmm_range failed with status %d" % rc)
    return Y


def hilbert_xy2d(n, x, y):
    L = lib()
    L.orc_hilbert_xy2d.restype = ctypes.c_uint64
    L.orc_hilbert_xy2d.argtypes = [ctypes.c_uint64] * 3
    return int(L.orc_hilbert_xy2d(n, x, y))


def spmm_exec(image, X, task_rows, hilbert):
    """Sequential Y = A*X with tasks of task_rows block rows whose super
    blocks are visited in seq or Hilbert order (block_compute_task::run)."""
    X = np.ascontiguousarray(X)
    if X.ndim == 1:
        X = X.reshape(-1, 1)
    dt = _NP2DT[X.dtype]
    n_rows = int(image.data[24:32].view(np.uint64)[0])
    k = X.shape[1]
    Y = np.empty((n_rows, k), X.dtype)
    L = lib()
    L.orc_spmm_exec.restype = ctypes.c_int
    L.orc_spmm_exec.argtypes = [_u8p, ctypes.c_uint64, _i64p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _vp,
                                ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int]
    rc = L.orc_spmm_exec(_p(image.data, _u8p), len(image.data), _p(image.offs, _i64p), image.nbr, dt, k,
                         X.ctypes.data, k, Y.ctypes.data, k, task_rows, int(bool(hilbert)))
    if rc:
        raise ValueError("orc_spmm_exec failed with status %d" % rc)
    return Y


def pagerank(A, At, max_iters=30, damping=0.85):
    """(pr float32 array, iterations) as matrix/libmatrix-algs/PageRank.cpp computes them."""
    L = lib()
    L.orc_pagerank.restype = ctypes.c_int
    L.orc_pagerank.argtypes = [_u8p, ctypes.c_uint64, _i64p, ctypes.c_uint64, _u8p, ctypes.c_uint64, _i64p,
                               ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, _vp,
                               ctypes.POINTER(ctypes.c_int)]
    n = int(A.data[24:32].view(np.uint64)[0])
    pr = np.empty(n, np.float32)
    it = ctypes.c_int()
    rc = L.orc_pagerank(_p(A.data, _u8p), len(A.data), _p(A.offs, _i64p), A.nbr, _p(At.data, _u8p),
                        len(At.data), _p(At.offs, _i64p), At.nbr, n, max_iters, damping, pr.ctypes.data,
                        ctypes.byref(it))
    if rc:
        raise ValueError("orc_pagerank failed with status %d" % rc)
    return pr, it.value


def nmf(A, At, W, tH, iters, nnz):
    """NMF by Lee-Seung multiplicative updates, restating
    matrix/libmatrix-algs/NMF.cpp:164-230 (update_lee), 259-316 (the loop) and
    110-126 (Fnorm), with H kept transposed (tH = H^T, m x k):
        tD  = tH (W^T W)                              (:168-173)
        tH  = max(tH * (A^T W), eps) / (tD + eps)     (:176-193)
        D2  = W (tH^T tH)                             (:201-208)
        W   = max(W * (A tH), eps) / (D2 + eps)       (:211-224)
    eps = 10e-9 (:166).  Distances after iterations 0, 5, ...:
    nnz - 2 sum((A^T W) * tH) + sum((tH W^T W) * tH).  Numerics follow
    fmx_nmf: sparse products are FP64 sums of the FP32 operands rounded once
    to FP32, Gram and k x k products FP64 (rounded to FP32 where a matrix is
    stored), the element-wise update FP32.  A, At: Images (At = A for a
    symmetric matrix)."""
    f32, f64 = np.float32, np.float64
    eps = f32(10e-9)
    W = np.array(W, f32)
    tH = np.array(tH, f32)
    dist = []
    WtW = W.astype(f64).T @ W.astype(f64)
    for it in range(iters):
        tD = (tH.astype(f64) @ WtW).astype(f32)
        AtW = spmm(At, W.astype(f64)).astype(f32)
        tH = np.maximum(tH * AtW, eps) / (tD + eps)
        HHt = tH.astype(f64).T @ tH.astype(f64)
        D2 = (W.astype(f64) @ HHt).astype(f32)
        AH = spmm(A, tH.astype(f64)).astype(f32)
        W = np.maximum(W * AH, eps) / (D2 + eps)
        WtW = W.astype(f64).T @ W.astype(f64)
        if it % 5 == 0:
            AtW = spmm(At, W.astype(f64)).astype(f32)
            Q = (tH.astype(f64) @ WtW).astype(f32)
            t1 = float(np.sum(AtW.astype(f64) * tH.astype(f64)))
            t2 = float(np.sum(Q.astype(f64) * tH.astype(f64)))
            dist.append(float(nnz) - 2.0 * t1 + t2)
    return W, tH, dist


def gram_f64(A, B):
    A = np.ascontiguousarray(A, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    G = np.empty((A.shape[1], B.shape[1]), np.float64)
    lib().orc_gram_f64(A.ctypes.data, A.shape[1], B.ctypes.data, B.shape[1], A.shape[0],
                       A.shape[1], B.shape[1], G.ctypes.data)
    return G


def gemm_tall_f64(A, C, alpha=1.0, beta=0.0, Y=None):
    A = np.ascontiguousarray(A, np.float64)
    C = np.ascontiguousarray(C, np.float64)
    if Y is None:
        Y = np.zeros((A.shape[0], C.shape[1]), np.float64)
    lib().orc_gemm_tall_f64(A.ctypes.data, A.shape[1], A.shape[0], A.shape[1],
                            C.ctypes.data, C.shape[1], alpha, beta, Y.ctypes.data, Y.shape[1])
    return Y


def colnorm_f64(A):
    A = np.ascontiguousarray(A, np.float64)
    out = np.empty(A.shape[1], np.float64)
    lib().orc_colnorm_f64(A.ctypes.data, A.shape[1], A.shape[0], A.shape[1], out.ctypes.data)
    return out


def _bind_gnn():
    L = lib()
    if getattr(L, "_gnn_bound", False):
        return L
    L.orc_dcsbm_edges.restype = ctypes.c_int
    L.orc_dcsbm_edges.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _u32p, _u32p,
                                  ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p]
    L.orc_gnn_param_count.restype = ctypes.c_uint64
    L.orc_gnn_param_count.argtypes = []
    L.orc_subspace_loss.restype = ctypes.c_int
    L.orc_subspace_loss.argtypes = [ctypes.c_uint64, ctypes.c_int, _vp, _vp, ctypes.POINTER(ctypes.c_double), _vp]
    L.orc_gnn_loss_grad.restype = ctypes.c_int
    L.orc_gnn_loss_grad.argtypes = [ctypes.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp,
                                    ctypes.POINTER(ctypes.c_double), _vp, _vp]
    L.orc_gnn_loss_grad_ex.restype = ctypes.c_int
    L.orc_gnn_loss_grad_ex.argtypes = [ctypes.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp,
                                       ctypes.POINTER(ctypes.c_double), _vp, _vp, ctypes.c_int]
    L.orc_bf16_round.restype = None
    L.orc_bf16_round.argtypes = [_vp, _vp, ctypes.c_uint64]
    L._gnn_bound = True
    return L


def dcsbm_edges(n, bstart, wtab, ratio, m, seed):
    L = _bind_gnn()
    scale = int(n).bit_length() - 1
    bstart = np.ascontiguousarray(bstart, np.uint32)
    wtab = np.ascontiguousarray(wtab, np.uint32)
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    rc = L.orc_dcsbm_edges(n, scale, len(bstart) - 1, _p(bstart, _u32p), _p(wtab, _u32p), ratio, m, seed,
                           _p(src, _u32p), _p(dst, _u32p))
    if rc:
        raise ValueError("orc_dcsbm_edges failed %d" % rc)
    return src, dst


def gnn_param_count():
    return _bind_gnn().orc_gnn_param_count()


def subspace_loss(Y, U):
    L = _bind_gnn()
    Y = np.ascontiguousarray(Y, np.float64)
    U = np.ascontiguousarray(U, np.float64)
    dY = np.empty_like(Y)
    loss = ctypes.c_double()
    rc = L.orc_subspace_loss(Y.shape[0], Y.shape[1], Y.ctypes.data, U.ctypes.data, ctypes.byref(loss),
                             dY.ctypes.data)
    if rc:
        raise ValueError("orc_subspace_loss failed")
    return loss.value, dY


def bf16_round(x):
    """Round-to-nearest-even to bfloat16 (through FP32), returned as float64."""
    L = _bind_gnn()
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    L.orc_bf16_round(x.ctypes.data, out.ctypes.data, x.size)
    return out


def gnn_loss_grad(rowptr, col, dinv, X0, U, params, want_grad=True, bf16=False):
    """FP64 forward, subspace loss and gradient of the warm-start network;
    bf16=True rounds every dense-product operand to BF16 as the GPU's BF16
    matrix-core mode does."""
    L = _bind_gnn()
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int64)
    dinv = np.ascontiguousarray(dinv, np.float64)
    X0 = np.ascontiguousarray(X0, np.float64)
    U = np.ascontiguousarray(U, np.float64)
    params = np.ascontiguousarray(params, np.float64)
    n = len(rowptr) - 1
    grad = np.zeros_like(params) if want_grad else None
    Y = np.empty((n, 16), np.float64)
    loss = ctypes.c_double()
    rc = L.orc_gnn_loss_grad_ex(n, rowptr.ctypes.data, col.ctypes.data, dinv.ctypes.data, X0.ctypes.data,
                                U.ctypes.data, params.ctypes.data, ctypes.byref(loss),
                                grad.ctypes.data if want_grad else None, Y.ctypes.data, int(bool(bf16)))
    if rc:
        raise ValueError("orc_gnn_loss_grad failed")
    return loss.value, grad, Y
