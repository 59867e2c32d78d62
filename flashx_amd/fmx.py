"""ctypes binding of libfmx (include/fmx.h) and the Python mirror of the
reference's sparse-matrix operator interface.

The classes mirror fm::sparse_matrix (matrix/sparse_matrix.h:475-651):
``SparseMatrix.multiply(X)`` returns a new Y = A*X like
``sparse_matrix::multiply(dense_matrix::ptr)`` (sparse_matrix.cpp:773-831) and
``SparseMatrix.multiply_into(X, Y)`` fills a caller-owned Y and returns a bool
like ``sparse_matrix::multiply(in, out)`` (sparse_matrix.cpp:751-771).  Dense
operands are torch tensors on the GPU (row-major, any leading dimension);
torch only provides device memory and streams here.  There is no CPU path:
every call goes to the HIP library and raises if it is not built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfmx.so")

OK, EINVAL, ETYPE, EDIM, EFORMAT, ENOMEM, EHIP, EUNSUPPORTED = range(8)
I32, I64, F32, F64, BF16, U64 = range(6)
BUILD_DEDUP, BUILD_KEEP_SELF, BUILD_SYMMETRIZE = 1, 2, 4
SPMM_SEQ, SPMM_X_COLMAJOR, SPMM_Y_COLMAJOR = 1, 2, 4

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64


class FmxError(RuntimeError):
    """A libfmx call failed; ``status`` is the fmx_status code."""

    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


class fmx_spm_info(ctypes.Structure):
    _fields_ = [("nrows", _u64), ("ncols", _u64), ("nnz", _u64), ("block_h", _u64),
                ("block_w", _u64), ("nbr", _u64), ("image_bytes", _u64),
                ("tile_bytes", _u64), ("entry_size", _u64), ("ntiles", _u64),
                ("stream_chunks", _u64), ("build_ns", _u64), ("edges_in", _u64), ("hot_columns", _u64)]


class fmx_eigen_opts(ctypes.Structure):
    _fields_ = [("nev", ctypes.c_int), ("block_size", ctypes.c_int), ("num_blocks", ctypes.c_int),
                ("max_restarts", ctypes.c_int), ("tol", ctypes.c_double),
                ("precision", ctypes.c_int), ("verbose", ctypes.c_int), ("seed", _u64),
                ("history", ctypes.POINTER(ctypes.c_double)), ("history_cap", ctypes.c_int)]


class fmx_eigen_stats(ctypes.Structure):
    _fields_ = [("spmm_blocks", ctypes.c_int), ("fp64_spmm_blocks", ctypes.c_int),
                ("restarts", ctypes.c_int), ("converged", ctypes.c_int),
                ("max_resid", ctypes.c_double), ("seconds", ctypes.c_double),
                ("history_len", ctypes.c_int)]


# name -> (restype, argtypes); every symbol include/fmx.h declares.
SIGNATURES = {
    "fmx_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "fmx_ctx_destroy": (ctypes.c_int, [_vp]),
    "fmx_ctx_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "fmx_ctx_sync": (ctypes.c_int, [_vp]),
    "fmx_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "fmx_ctx_last_error": (ctypes.c_char_p, [_vp]),
    "fmx_spm_upload": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t,
                                      ctypes.POINTER(_vp)]),
    "fmx_spm_build": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t, _u64, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "fmx_spm_info_get": (ctypes.c_int, [_vp, ctypes.POINTER(fmx_spm_info)]),
    "fmx_spm_download": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t), _vp]),
    "fmx_spm_destroy": (ctypes.c_int, [_vp]),
    "fmx_pagerank": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_float, _vp,
                                    ctypes.POINTER(ctypes.c_int), _vp]),
    "fmx_spm_build_w": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_size_t, ctypes.c_uint64,
                                       ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "fmx_spm_load": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "fmx_spm_dump": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "fmx_spm_open": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p, _u64, ctypes.POINTER(_vp)]),
    "fmx_spm_upload_streamed": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, _u64,
                                               ctypes.POINTER(_vp)]),
    "fmx_spmm_host": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _i64, _vp, _i64, _u64, _vp]),
    "fmx_gather_floor": (ctypes.c_int, [_vp, _u64, _vp, _i64, _u64, _vp, _u64, _vp]),
    "fmx_nmf": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "fmx_spm_open_index": (ctypes.c_int, [_vp, ctypes.c_char_p, _vp, _vp, ctypes.c_size_t, _u64,
                                          ctypes.POINTER(_vp)]),
    "fmx_spmm": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _i64, _vp, _i64, _vp]),
    "fmx_spmm_ex": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _i64, _vp, _i64, ctypes.c_uint32,
                                   _vp]),
    "fmx_spmm_scaled": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _i64, _vp, _i64,
                                       ctypes.c_uint32, _vp]),
    "fmx_spmm_alg_bytes": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(_u64)]),
    "fmx_gram": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _u64, ctypes.c_int, ctypes.c_int,
                                _vp, _i64, _vp, _i64, _vp, _vp]),
    "fmx_gram_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _u64, ctypes.c_int, ctypes.c_int,
                                        _vp, _i64, _vp, _i64, _vp, _vp]),
    "fmx_gemm_tall": (ctypes.c_int, [_vp, ctypes.c_int, _u64, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, _vp, _i64, _vp, ctypes.c_double, _vp, _i64,
                                     _vp]),
    "fmx_colnorm": (ctypes.c_int, [_vp, ctypes.c_int, _u64, ctypes.c_int, _vp, _i64, _vp, _vp]),
    "fmx_row_scale": (ctypes.c_int, [_vp, ctypes.c_int, _u64, ctypes.c_int, _vp, _vp, _i64, _vp,
                                     _i64, _vp]),
    "fmx_eigen_opts_init": (ctypes.c_int, [ctypes.POINTER(fmx_eigen_opts), ctypes.c_int]),
    "fmx_eigen_normalized": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(fmx_eigen_opts), _vp, _i64,
                                            ctypes.c_int, _vp, _vp, _i64, _vp,
                                            ctypes.POINTER(fmx_eigen_stats)]),
    "fmx_eigen": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.POINTER(fmx_eigen_opts), _vp, _i64,
                                 ctypes.c_int, _vp, _vp, _i64, _vp, ctypes.POINTER(fmx_eigen_stats)]),
    "fmx_gnn_param_count": (_u64, []),
    "fmx_gnn_acts_floats": (_u64, [_u64]),
    "fmx_gnn_work_floats": (_u64, [_u64]),
    "fmx_gnn_features": (ctypes.c_int, [_vp, _vp, _u64, _vp, _vp, _vp]),
    "fmx_gnn_forward": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "fmx_gnn_backward": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int,
                                        _vp]),
    "fmx_subspace_loss": (ctypes.c_int, [_vp, _u64, ctypes.c_int, _vp, _vp,
                                         ctypes.POINTER(ctypes.c_double), _vp, _vp]),
    "fmx_adam": (ctypes.c_int, [_vp, _u64, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, ctypes.c_float, ctypes.c_int, _vp]),
    "fmx_dcsbm_params": (ctypes.c_int, [_u64, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp,
                                        _vp]),
    "fmx_dcsbm_edges": (ctypes.c_int, [_vp, _u64, ctypes.c_int, _vp, _vp, ctypes.c_uint32, _u64, _u64,
                                       _vp, _vp, _vp]),
    "fmx_rmat_edges": (ctypes.c_int, [_vp, ctypes.c_int, _u64, _u64, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _vp,
                                      _vp, _vp]),
}

_lib = None


def lib():
    """Loads libfmx.so; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libfmx.so is not built (%s); run python -m flashx_amd.build"
                              % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _torch():
    import torch
    return torch


def dtype_code(t):
    torch = _torch()
    m = {torch.int32: I32, torch.int64: I64, torch.float32: F32, torch.float64: F64,
         torch.bfloat16: BF16, torch.uint64: U64}
    if t not in m:
        raise FmxError(ETYPE, "unsupported dtype %s" % t)
    return m[t]


class Context:
    """A device and its stream (replaces fm::init_flash_matrix)."""

    def __init__(self, device=0):
        self.handle = _vp()
        st = lib().fmx_ctx_create(device, ctypes.byref(self.handle))
        if st:
            raise FmxError(st, "fmx_ctx_create(%d): %s" % (device, lib().fmx_strerror(st).decode()))
        self.device = device

    def check(self, st, what):
        if st:
            detail = lib().fmx_ctx_last_error(self.handle).decode()
            raise FmxError(st, "%s: %s (%s)" % (what, lib().fmx_strerror(st).decode(), detail))

    def stream_ptr(self):
        torch = _torch()
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def sync(self):
        self.check(lib().fmx_ctx_sync(self.handle), "fmx_ctx_sync")

    def close(self):
        if self.handle:
            lib().fmx_ctx_destroy(self.handle)
            self.handle = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def default_context(device=0):
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


class SparseMatrix:
    """fm::sparse_matrix over a 2D tile image, resident in HBM or (open,
    upload_streamed) kept in host memory and streamed through HBM."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle
        info = fmx_spm_info()
        ctx.check(lib().fmx_spm_info_get(handle, ctypes.byref(info)), "fmx_spm_info_get")
        self.info = info

    # -- factories (sparse_matrix::create, fm::create_2d_matrix) -----------
    @classmethod
    def upload(cls, image, offs, ctx=None):
        """From a reference-format image (bytes/uint8 array) and offsets."""
        ctx = ctx or default_context()
        img = np.ascontiguousarray(np.frombuffer(image, np.uint8) if isinstance(image, (bytes, bytearray))
                                   else image, dtype=np.uint8)
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        h = _vp()
        ctx.check(lib().fmx_spm_upload(ctx.handle, img.ctypes.data, img.nbytes, offs.ctypes.data,
                                       len(offs) - 1, ctypes.byref(h)), "fmx_spm_upload")
        return cls(ctx, h)

    @classmethod
    def load(cls, mat_file, idx_file, ctx=None):
        """From a .mat image file and its .mat_idx index file
        (SpM_2d_index::load + SpM_2d_storage::load + sparse_matrix::create)."""
        ctx = ctx or default_context()
        h = _vp()
        ctx.check(lib().fmx_spm_load(ctx.handle, os.fsencode(mat_file), os.fsencode(idx_file), ctypes.byref(h)),
                  "fmx_spm_load")
        return cls(ctx, h)

    @classmethod
    def open(cls, mat_file, idx_file, chunk_bytes=1 << 30, ctx=None):
        """Out-of-core: the image of a .mat / .mat_idx pair stays in pinned
        host memory and each multiply streams it through HBM in block-row
        chunks of at most chunk_bytes (the SAFS-backed sparse_matrix::create)."""
        ctx = ctx or default_context()
        h = _vp()
        ctx.check(lib().fmx_spm_open(ctx.handle, os.fsencode(mat_file), os.fsencode(idx_file), int(chunk_bytes),
                                     ctypes.byref(h)), "fmx_spm_open")
        return cls(ctx, h)

    @classmethod
    def upload_streamed(cls, image, offs, chunk_bytes=1 << 30, ctx=None):
        """Out-of-core form of upload(): the image is kept in pinned host memory."""
        ctx = ctx or default_context()
        img = np.ascontiguousarray(np.frombuffer(image, np.uint8) if isinstance(image, (bytes, bytearray))
                                   else image, dtype=np.uint8)
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        h = _vp()
        ctx.check(lib().fmx_spm_upload_streamed(ctx.handle, img.ctypes.data, img.nbytes, offs.ctypes.data,
                                                len(offs) - 1, int(chunk_bytes), ctypes.byref(h)),
                  "fmx_spm_upload_streamed")
        return cls(ctx, h)

    @property
    def streamed(self):
        return self.info.stream_chunks > 0

    def dump(self, mat_file, idx_file):
        """Writes the .mat / .mat_idx pair (SpM_2d_storage::dump, SpM_2d_index::dump)."""
        self.ctx.check(lib().fmx_spm_dump(self.handle, os.fsencode(mat_file), os.fsencode(idx_file)),
                       "fmx_spm_dump")

    @classmethod
    def build(cls, src, dst, n=0, log2_h=15, log2_w=15, flags=BUILD_DEDUP, ctx=None):
        """On-device tile construction from device tensors src, dst (int32/uint32)."""
        torch = _torch()
        ctx = ctx or default_context()
        if src.device.type != "cuda" or dst.device.type != "cuda" or src.numel() != dst.numel():
            raise FmxError(EINVAL, "src and dst must be device tensors of equal length")
        if src.dtype not in (torch.int32, torch.uint32) or dst.dtype != src.dtype:
            raise FmxError(ETYPE, "src and dst must be 32-bit vertex ids")
        src = src.contiguous()
        dst = dst.contiguous()
        h = _vp()
        ctx.check(lib().fmx_spm_build(ctx.handle, src.data_ptr(), dst.data_ptr(), src.numel(), n,
                                      log2_h, log2_w, flags, ctypes.byref(h)), "fmx_spm_build")
        return cls(ctx, h)

    @classmethod
    def build_weighted(cls, src, dst, vals, n=0, log2_h=15, log2_w=15, flags=BUILD_DEDUP, ctx=None):
        """Weighted image from device tensors src, dst (32-bit ids) and vals
        (int32 / int64 / float32 / float64, one per edge)."""
        ctx = ctx or default_context()
        if vals.device.type != "cuda" or vals.numel() != src.numel() or src.numel() != dst.numel():
            raise FmxError(EINVAL, "src, dst and vals must be device tensors of equal length")
        src, dst, vals = src.contiguous(), dst.contiguous(), vals.contiguous()
        h = _vp()
        ctx.check(lib().fmx_spm_build_w(ctx.handle, src.data_ptr(), dst.data_ptr(), vals.data_ptr(),
                                        dtype_code(vals.dtype), src.numel(), n, log2_h, log2_w, flags,
                                        ctypes.byref(h)), "fmx_spm_build_w")
        return cls(ctx, h)

    # -- reference accessors -------------------------------------------------
    def get_num_rows(self):
        return self.info.nrows

    def get_num_cols(self):
        return self.info.ncols

    def get_nnz(self):
        return self.info.nnz

    def get_block_size(self):
        return (self.info.block_h, self.info.block_w)

    def get_entry_size(self):
        return self.info.entry_size

    def download(self):
        """(image uint8 array, offsets int64 array) copied back to the host."""
        img = np.empty(self.info.image_bytes, np.uint8)
        offs = np.empty(self.info.nbr + 1, np.int64)
        nb = ctypes.c_size_t()
        self.ctx.check(lib().fmx_spm_download(self.handle, img.ctypes.data, img.nbytes,
                                              ctypes.byref(nb), offs.ctypes.data),
                       "fmx_spm_download")
        return img, offs

    def alg_bytes(self, dtype, k):
        out = _u64()
        self.ctx.check(lib().fmx_spmm_alg_bytes(self.handle, dtype_code(dtype), k,
                                                ctypes.byref(out)), "fmx_spmm_alg_bytes")
        return out.value

    # -- SpMM -------------------------------------------------------------------
    def _operands(self, X, Y):
        if X.dim() == 1:
            X = X.view(-1, 1)
        if Y.dim() == 1:
            Y = Y.view(-1, 1)
        if X.device.type != "cuda" or Y.device.type != "cuda":
            raise FmxError(EINVAL, "dense operands must be device tensors")
        if X.dtype != Y.dtype:
            raise FmxError(ETYPE, "wrong matrix type in spmm creator")
        if X.shape[0] != self.info.ncols or Y.shape[0] != self.info.nrows or X.shape[1] != Y.shape[1]:
            raise FmxError(EDIM, "dimension mismatch: A is %dx%d, X %s, Y %s"
                           % (self.info.nrows, self.info.ncols, tuple(X.shape), tuple(Y.shape)))
        ldx, xcol = dense_layout(X, "X")
        ldy, ycol = dense_layout(Y, "Y")
        flags = (SPMM_X_COLMAJOR if xcol else 0) | (SPMM_Y_COLMAJOR if ycol else 0)
        return X, Y, ldx, ldy, flags

    def multiply_async(self, X, Y, stream=None, seq=False):
        """Enqueues Y = A*X on `stream` (default: torch's current stream).
        X and Y may each be row-major or column-major (any leading
        dimension); seq=True sums floating-point rows in the reference's
        sequential order (FMX_SPMM_SEQ: bit-identical to the reference's
        CPU SpMM)."""
        X, Y, ldx, ldy, flags = self._operands(X, Y)
        k = X.shape[1]
        sp = ctypes.c_void_p(stream) if stream is not None else self.ctx.stream_ptr()
        if flags == 0 and not seq:
            self.ctx.check(lib().fmx_spmm(self.ctx.handle, self.handle, dtype_code(X.dtype), k,
                                          X.data_ptr(), ldx, Y.data_ptr(), ldy, sp), "fmx_spmm")
        else:
            self.ctx.check(lib().fmx_spmm_ex(self.ctx.handle, self.handle, dtype_code(X.dtype), k, X.data_ptr(),
                                             ldx, Y.data_ptr(), ldy, flags | (SPMM_SEQ if seq else 0), sp),
                           "fmx_spmm_ex")
        return Y

    def multiply_scaled_async(self, X, Y, pre=None, post=None, seq=False, stream=None):
        """Y = diag(post) A diag(pre) X (fmx_spmm_scaled): the SpMM of a
        custom task creator with row scalings, as the reference's
        normalised-adjacency operator runs it.  pre (ncols) and post (nrows)
        are FP64 device vectors or None; an FP32-weighted matrix also
        multiplies FP64 data (spmm_creator<double, float>)."""
        torch = _torch()
        X, Y, ldx, ldy, flags = self._operands(X, Y)
        vec = []
        for v, n, what in ((pre, self.info.ncols, "pre"), (post, self.info.nrows, "post")):
            if v is None:
                vec.append(None)
                continue
            if v.dtype != torch.float64 or v.device.type != "cuda" or v.numel() != n:
                raise FmxError(EINVAL, "%s must be an FP64 device vector of %d values" % (what, n))
            vec.append(v.contiguous())
        sp = ctypes.c_void_p(stream) if stream is not None else self.ctx.stream_ptr()
        self.ctx.check(lib().fmx_spmm_scaled(self.ctx.handle, self.handle, dtype_code(X.dtype), X.shape[1],
                                             vec[0].data_ptr() if vec[0] is not None else None,
                                             vec[1].data_ptr() if vec[1] is not None else None,
                                             X.data_ptr(), ldx, Y.data_ptr(), ldy,
                                             flags | (SPMM_SEQ if seq else 0), sp), "fmx_spmm_scaled")
        return Y

    def multiply_into(self, X, Y):
        """sparse_matrix::multiply(in, out): True on success, False on a
        type or dimension error (logged as an exception message)."""
        try:
            self.multiply_async(X, Y)
        except FmxError:
            return False
        _torch().cuda.current_stream(self.ctx.device).synchronize()
        return True

    def multiply(self, X):
        """sparse_matrix::multiply(dense_matrix): a new Y = A*X (synchronous)."""
        torch = _torch()
        squeeze = X.dim() == 1
        X2 = X.view(-1, 1) if squeeze else X
        Y = torch.empty((self.info.nrows, X2.shape[1]), dtype=X.dtype, device=X.device)
        self.multiply_async(X2, Y)
        torch.cuda.current_stream(self.ctx.device).synchronize()
        return Y.view(-1) if squeeze else Y

    def multiply_host(self, X, mem_bytes):
        """Y = A*X for a host numpy X, through HBM in column panels that fit
        mem_bytes (sparse_matrix::multiply(dense_matrix, mem_size) with an
        external-memory X); returns a host numpy Y."""
        X = np.ascontiguousarray(X)
        if X.ndim == 1:
            X = X.reshape(-1, 1)
        dt = {np.dtype(np.int32): I32, np.dtype(np.int64): I64, np.dtype(np.float32): F32,
              np.dtype(np.float64): F64, np.dtype(np.uint64): U64}.get(X.dtype)
        if dt is None:
            raise FmxError(ETYPE, "wrong matrix type in spmm creator")
        if X.shape[0] != self.info.ncols:
            raise FmxError(EDIM, "dimension mismatch: A is %dx%d, X %s"
                           % (self.info.nrows, self.info.ncols, X.shape))
        k = X.shape[1]
        Y = np.empty((self.info.nrows, k), X.dtype)
        self.ctx.check(lib().fmx_spmm_host(self.ctx.handle, self.handle, dt, k, X.ctypes.data, k, Y.ctypes.data, k,
                                           int(mem_bytes), None), "fmx_spmm_host")
        return Y

    def close(self):
        if self.handle:
            lib().fmx_spm_destroy(self.handle)
            self.handle = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


GRAM_F64, GRAM_BF16 = 0, 1


def dense_layout(t, what="operand"):
    """(leading dimension, column-major?) of a 2-D tensor whose rows (or
    columns) do not overlap: row-major with stride(0) >= columns, or
    column-major with stride(1) >= rows.  Expanded or overlapping views are
    refused before anything is launched."""
    if t.dim() != 2:
        raise FmxError(EINVAL, "%s must be a 2-D tensor" % what)
    r, c = t.shape
    s0, s1 = t.stride()
    if (c == 1 or s1 == 1) and (r <= 1 or s0 >= c):
        return (s0 if r > 1 else c), False
    if (r == 1 or s0 == 1) and (c <= 1 or s1 >= r):
        return (s1 if c > 1 else r), True
    raise FmxError(EINVAL, "%s (shape %s, strides %s) is neither row-major nor column-major with "
                   "non-overlapping rows" % (what, tuple(t.shape), t.stride()))


def _ld(t):
    ld, col = dense_layout(t)
    if col and t.shape[1] > 1:
        raise FmxError(EINVAL, "dense operands must be 2-D row-major device tensors")
    return ld


def gram(A, B, mode=GRAM_F64, ctx=None):
    """G = A^T B (FP64 result), MvTransMv."""
    torch = _torch()
    ctx = ctx or default_context()
    if A.shape[0] != B.shape[0] or A.dtype != B.dtype:
        raise FmxError(EDIM, "Gram operands must share rows and dtype")
    G = torch.empty((A.shape[1], B.shape[1]), dtype=torch.float64, device=A.device)
    ctx.check(lib().fmx_gram(ctx.handle, dtype_code(A.dtype), mode, A.shape[0], A.shape[1],
                             B.shape[1], A.data_ptr(), _ld(A), B.data_ptr(), _ld(B), G.data_ptr(),
                             ctx.stream_ptr()), "fmx_gram")
    return G


def gemm_tall(A, C, alpha=1.0, beta=0.0, Y=None, ctx=None):
    """Y = alpha*A*C + beta*Y with C a small FP64 matrix (MvTimesMatAddMv)."""
    torch = _torch()
    ctx = ctx or default_context()
    C = C.to(device=A.device, dtype=torch.float64).contiguous()
    if Y is None:
        Y = torch.empty((A.shape[0], C.shape[1]), dtype=A.dtype, device=A.device)
        beta = 0.0
    if C.shape[0] != A.shape[1] or Y.shape != (A.shape[0], C.shape[1]) or Y.dtype != A.dtype:
        raise FmxError(EDIM, "gemm_tall shape mismatch")
    ctx.check(lib().fmx_gemm_tall(ctx.handle, dtype_code(A.dtype), A.shape[0], A.shape[1],
                                  C.shape[1], alpha, A.data_ptr(), _ld(A), C.data_ptr(), beta,
                                  Y.data_ptr(), _ld(Y), ctx.stream_ptr()), "fmx_gemm_tall")
    return Y


def colnorm(A, ctx=None):
    """Column 2-norms (FP64), MvNorm."""
    torch = _torch()
    ctx = ctx or default_context()
    out = torch.empty(A.shape[1], dtype=torch.float64, device=A.device)
    ctx.check(lib().fmx_colnorm(ctx.handle, dtype_code(A.dtype), A.shape[0], A.shape[1],
                                A.data_ptr(), _ld(A), out.data_ptr(), ctx.stream_ptr()),
              "fmx_colnorm")
    return out


def row_scale(d, X, Y=None, ctx=None):
    """Y = diag(d) X with d FP64."""
    torch = _torch()
    ctx = ctx or default_context()
    d = d.to(dtype=torch.float64).contiguous()
    if Y is None:
        Y = torch.empty_like(X)
    ctx.check(lib().fmx_row_scale(ctx.handle, dtype_code(X.dtype), X.shape[0], X.shape[1],
                                  d.data_ptr(), X.data_ptr(), _ld(X), Y.data_ptr(), _ld(Y),
                                  ctx.stream_ptr()), "fmx_row_scale")
    return Y


EIG_MIXED, EIG_FP32, EIG_FP64 = 0, 1, 2
EIG_OP_NORMALIZED, EIG_OP_PLAIN, EIG_OP_SVD = 0, 1, 2
_EIG_OPS = {"normalized": EIG_OP_NORMALIZED, "plain": EIG_OP_PLAIN, "svd": EIG_OP_SVD}


def compute_eigen(spm, nev=16, block_size=16, num_blocks=6, tol=1e-6, max_restarts=100,
                  precision=EIG_FP32, X0=None, seed=1, verbose=False, history=False, op="normalized", At=None):
    """Top-nev eigenpairs of an operator of spm (fm::eigen::compute_eigen):
    op "normalized" is D^-1/2 A D^-1/2 (NA_eigen_Operator), "plain" A itself
    (eigen_Operator), "svd" A^T A (SVD_Operator; At is A's transpose, None
    for a symmetric A; the eigenvalues are the squared singular values and
    the vectors, ncols x nev, the right singular vectors).  Returns (evals,
    evecs n x nev FP64 device tensor, relative residuals, stats dict); with
    history=True the stats carry the largest relative residual after each
    restart ("history")."""
    torch = _torch()
    ctx = spm.ctx
    o = fmx_eigen_opts()
    lib().fmx_eigen_opts_init(ctypes.byref(o), nev)
    o.block_size, o.num_blocks, o.tol = block_size, num_blocks, tol
    o.max_restarts, o.precision, o.seed, o.verbose = max_restarts, precision, seed, int(verbose)
    hist = (ctypes.c_double * (max_restarts + 2))() if history else None
    if hist is not None:
        o.history = ctypes.cast(hist, ctypes.POINTER(ctypes.c_double))
        o.history_cap = max_restarts + 2
    if op not in _EIG_OPS:
        raise FmxError(EINVAL, "unknown eigen operator %r" % (op,))
    n = spm.get_num_cols() if op == "svd" else spm.get_num_rows()
    evals = np.zeros(nev, np.float64)
    resid = np.zeros(nev, np.float64)
    vecs = torch.empty((n, nev), dtype=torch.float64, device="cuda:%d" % ctx.device)
    st = fmx_eigen_stats()
    x0p, ldx0, x0dt = None, block_size, F32
    if X0 is not None:
        if X0.shape != (n, block_size) or X0.stride(1) != 1:
            raise FmxError(EDIM, "warm start must be n x block_size row-major")
        x0p, ldx0, x0dt = X0.data_ptr(), X0.stride(0), dtype_code(X0.dtype)
    torch.cuda.synchronize()
    ctx.check(lib().fmx_eigen(ctx.handle, _EIG_OPS[op], spm.handle, At.handle if At is not None else None,
                              ctypes.byref(o), x0p, ldx0, x0dt, evals.ctypes.data, vecs.data_ptr(), nev,
                              resid.ctypes.data, ctypes.byref(st)),
              "fmx_eigen")
    stats = {f: getattr(st, f) for f, _ in fmx_eigen_stats._fields_}
    if hist is not None:
        stats["history"] = [hist[i] for i in range(st.history_len)]
    return evals, vecs, resid, stats


def gather_floor(X, cols, stream=None, ctx=None):
    """Enqueues the gather-only pass over X's rows cols (int32/uint32 device
    tensor): the memory floor of an SpMM visiting the same rows in the same
    order (fmx_gather_floor)."""
    torch = _torch()
    ctx = ctx or default_context()
    if X.dim() != 2 or X.stride(1) != 1 or cols.dtype not in (torch.int32, torch.uint32) or not cols.is_contiguous():
        raise FmxError(EINVAL, "X must be a row-major matrix and cols a contiguous 32-bit tensor")
    es = X.element_size()
    st = stream if stream is not None else torch.cuda.current_stream(ctx.device).cuda_stream
    ctx.check(lib().fmx_gather_floor(ctx.handle, X.shape[1] * es, X.data_ptr(), X.stride(0) * es, X.shape[0],
                                     cols.data_ptr(), cols.numel(), st), "fmx_gather_floor")


def nmf(A, k, max_iters=10, At=None, W=None, tH=None, seed=0):
    """fm::alg::NMF (matrix/libmatrix-algs/NMF.cpp): returns (W n x k, tH =
    H^T m x k, distances after iterations 0, 5, 10, ...).  W and tH start
    U[0,1) from `seed` unless given (device float32 tensors, updated in
    place)."""
    torch = _torch()
    ctx = A.ctx
    dev = "cuda:%d" % ctx.device
    n, m = A.get_num_rows(), A.get_num_cols()
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    if W is None:
        W = torch.rand((n, k), dtype=torch.float32, device=dev, generator=g)
    if tH is None:
        tH = torch.rand((m, k), dtype=torch.float32, device=dev, generator=g)
    if W.shape != (n, k) or tH.shape != (m, k) or W.dtype != torch.float32 or tH.dtype != torch.float32 \
            or not W.is_contiguous() or not tH.is_contiguous():
        raise FmxError(EDIM, "W must be n x k and tH m x k contiguous float32")
    nd = (max_iters + 4) // 5
    dist = (ctypes.c_double * max(nd, 1))()
    torch.cuda.synchronize()
    ctx.check(lib().fmx_nmf(ctx.handle, A.handle, At.handle if At is not None else None, k, max_iters,
                            W.data_ptr(), tH.data_ptr(), dist, ctx.stream_ptr()), "fmx_nmf")
    torch.cuda.synchronize()
    return W, tH, [dist[i] for i in range(nd)]


def pagerank(A, At=None, max_iters=30, damping=0.85):
    """PageRank (matrix/libmatrix-algs/PageRank.cpp): returns (pr float32
    device tensor, iterations).  At is A's transpose; None for symmetric A."""
    torch = _torch()
    ctx = A.ctx
    pr = torch.empty(A.get_num_rows(), dtype=torch.float32, device="cuda:%d" % ctx.device)
    it = ctypes.c_int()
    ctx.check(lib().fmx_pagerank(ctx.handle, A.handle, At.handle if At is not None else None, max_iters,
                                 damping, pr.data_ptr(), ctypes.byref(it), ctx.stream_ptr()), "fmx_pagerank")
    torch.cuda.synchronize()
    return pr, it.value


def dcsbm_params(n, nblocks=32, gamma=2.5, size_decay=1.06):
    """Block boundaries and degree-weight table of the DC-SBM generator."""
    bstart = np.zeros(nblocks + 1, np.uint32)
    wtab = np.zeros(1024, np.uint32)
    st = lib().fmx_dcsbm_params(n, nblocks, gamma, size_decay, bstart.ctypes.data, wtab.ctypes.data)
    if st:
        raise FmxError(st, "fmx_dcsbm_params: %s" % lib().fmx_strerror(st).decode())
    return bstart, wtab


def dcsbm_edges(n, seed, nblocks=32, ratio=20, gamma=2.5, mean_degree=16, size_decay=1.06, ctx=None):
    """Seeded DC-SBM edge list (m = n * mean_degree / 2) on the device."""
    torch = _torch()
    ctx = ctx or default_context()
    bstart, wtab = dcsbm_params(n, nblocks, gamma, size_decay)
    m = int(n * mean_degree // 2)
    src = torch.empty(m, dtype=torch.int32, device="cuda:%d" % ctx.device)
    dst = torch.empty(m, dtype=torch.int32, device="cuda:%d" % ctx.device)
    ctx.check(lib().fmx_dcsbm_edges(ctx.handle, n, nblocks, bstart.ctypes.data, wtab.ctypes.data, ratio, m,
                                    seed, src.data_ptr(), dst.data_ptr(), ctx.stream_ptr()),
              "fmx_dcsbm_edges")
    return src, dst


def rmat_edges(scale, m, seed, a=0.57, b=0.19, c=0.19, permute=True, ctx=None):
    """Seeded R-MAT edge list generated on the device (two int32 tensors)."""
    torch = _torch()
    ctx = ctx or default_context()
    f = lambda x: min(int(np.floor(x * 4294967296.0)), 0xFFFFFFFF)
    src = torch.empty(m, dtype=torch.int32, device="cuda:%d" % ctx.device)
    dst = torch.empty(m, dtype=torch.int32, device="cuda:%d" % ctx.device)
    ctx.check(lib().fmx_rmat_edges(ctx.handle, scale, m, seed, f(a), f(a + b), f(a + b + c),
                                   int(permute), src.data_ptr(), dst.data_ptr(),
                                   ctx.stream_ptr()), "fmx_rmat_edges")
    return src, dst
