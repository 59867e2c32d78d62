/*
 * fmx.h -- C ABI of the MI355X-native FlashMatrix SpMM path.
 *
 * Plain C types only: device buffers are `void *` / typed pointers into HBM
 * (allocated by the caller with hipMalloc or any allocator on the same
 * device), host buffers are ordinary pointers, sizes are counts.  Streams are
 * hipStream_t passed as `void *` and used as given (NULL = HIP's null
 * stream); fmx_ctx_stream returns a stream the context owns.
 *
 * Every call returns an fmx_status.  Calls taking a stream are stream-ordered
 * (asynchronous); the C++ layer (include/fm/sparse_matrix.h) synchronises
 * like the reference's blocking sparse_matrix::multiply
 * (matrix/sparse_matrix.cpp:465).  A context, and every matrix made from it,
 * is used by one host thread at a time (the reference's multiply is not
 * re-entrant either, matrix/sparse_matrix.cpp:436,471).
 *
 * Reference paths are relative to the FlashX source tree.
 */
#ifndef FMX_H
#define FMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FMX_OK = 0,
    FMX_EINVAL = 1,      /* bad argument                                     */
    FMX_ETYPE = 2,       /* unsupported or mismatched element type           */
    FMX_EDIM = 3,        /* dimension mismatch                               */
    FMX_EFORMAT = 4,     /* malformed tile image / header                    */
    FMX_ENOMEM = 5,      /* device or host allocation failed                 */
    FMX_EHIP = 6,        /* HIP runtime error                                */
    FMX_EUNSUPPORTED = 7 /* valid request this build does not implement      */
} fmx_status;

typedef enum {
    FMX_I32 = 0,  /* int      (P_INTEGER)  */
    FMX_I64 = 1,  /* long     (P_LONG)     */
    FMX_F32 = 2,  /* float    (P_FLOAT)    */
    FMX_F64 = 3,  /* double   (P_DOUBLE)   */
    FMX_BF16 = 4, /* bfloat16 (dense-op inputs only) */
    FMX_U64 = 5   /* size_t   (P_ULONG), SpMM only: spmm_creator<size_t, size_t>,
                     matrix/sparse_matrix.cpp:568-569 */
} fmx_dtype;

/* Edge normalisation flags for fmx_spm_build. */
#define FMX_BUILD_DEDUP      1u /* std::unique per row, fm_utils.cpp:37,460-464 */
#define FMX_BUILD_KEEP_SELF  2u /* keep self edges, fm_utils.cpp:39,450-452     */
#define FMX_BUILD_SYMMETRIZE 4u /* add (dst,src) for every (src,dst)            */

typedef struct fmx_ctx fmx_ctx;
typedef struct fmx_spm fmx_spm;

typedef struct {
    uint64_t nrows, ncols;     /* matrix_header nrows / ncols                  */
    uint64_t nnz;              /* non-zeros in the image                       */
    uint64_t block_h, block_w; /* tile height / width (powers of two)          */
    uint64_t nbr;              /* block rows                                   */
    uint64_t image_bytes;      /* header + tiles (offs[nbr])                   */
    uint64_t tile_bytes;       /* image_bytes - 4096                           */
    uint64_t entry_size;       /* 0 = binary                                   */
    uint64_t ntiles;           /* stored tiles (including empty stubs)         */
    uint64_t stream_chunks;    /* 0: image in HBM; else block-row chunks the   */
                               /* host-resident image is streamed in per SpMM  */
    uint64_t build_ns;         /* fmx_spm_build*: device time of the           */
                               /* construction kernels (0 for other matrices)  */
    uint64_t edges_in;         /* fmx_spm_build*: m, the edges given            */
    uint64_t hot_columns;      /* columns in the gather-order image's hot block */
                               /* (0: none)                                     */
} fmx_spm_info;

/* ---- context -------------------------------------------------------------
 * Replaces fm::init_flash_matrix / destroy_flash_matrix
 * (matrix/sparse_matrix.cpp:835-872): binds a device and owns a stream. */
fmx_status fmx_ctx_create(int device, fmx_ctx **out);
fmx_status fmx_ctx_destroy(fmx_ctx *ctx);
fmx_status fmx_ctx_stream(fmx_ctx *ctx, void **stream);
fmx_status fmx_ctx_sync(fmx_ctx *ctx);
const char *fmx_strerror(fmx_status st);
/* Message of the last failing call on this context ("" if none). */
const char *fmx_ctx_last_error(const fmx_ctx *ctx);

/* ---- sparse matrix ---------------------------------------------------------
 * fmx_spm_upload replaces sparse_matrix::create(SpM_2d_index::ptr,
 * SpM_2d_storage::ptr) (matrix/sparse_matrix.h:522, .cpp:679-683): `image` is
 * the host copy of a reference-format image (4096-byte header followed by the
 * block rows, matrix/sparse_matrix_format.cpp:308-334) of `bytes` bytes and
 * `offs` the nbr+1 block-row offsets of SpM_2d_index
 * (matrix/sparse_matrix_format.h:441-485).  The image is validated
 * (SpM_2d_index::verify, sparse_block_2d::verify,
 * matrix/sparse_matrix_format.cpp:36-60,128-136) and copied into HBM. */
fmx_status fmx_spm_upload(fmx_ctx *ctx, const void *image, size_t bytes,
                          const int64_t *offs, size_t nbr, fmx_spm **out);

/* fmx_spm_build replaces fm::create_2d_matrix(data_frame, block_2d_size, ...)
 * (matrix/fm_utils.h:37-39, matrix/fm_utils.cpp:617-715,726-841): builds the
 * tile image on the device from a DEVICE edge list src[m], dst[m] (u32 ids).
 * n == 0 takes n = max id + 1 as the reference does; a non-zero n must exceed
 * every id.  H = 2^log2_h, W = 2^log2_w, both <= 32768.  Bit-exact with the
 * reference writer. */
fmx_status fmx_spm_build(fmx_ctx *ctx, const uint32_t *src, const uint32_t *dst,
                         size_t m, uint64_t n, uint32_t log2_h, uint32_t log2_w,
                         uint32_t flags, fmx_spm **out);

/* Weighted form of fmx_spm_build (fm::create_2d_matrix with an attribute
 * column; attr_apply_operate, matrix/fm_utils.cpp:484-535): vals is a DEVICE
 * array of m values of dtype (I32, I64, F32, F64), written to each tile's
 * values region (entry size 4 or 8, prim_type P_INTEGER / P_LONG / P_FLOAT /
 * P_DOUBLE).  Among equal (row, column) edges input order is kept and
 * FMX_BUILD_DEDUP keeps the first one's value (the reference's std::sort
 * leaves that choice unspecified); SYMMETRIZE gives (dst, src) the same value.
 * Bit-exact with the oracle's weighted writer.  m < 2^31. */
fmx_status fmx_spm_build_w(fmx_ctx *ctx, const uint32_t *src, const uint32_t *dst, const void *vals,
                           fmx_dtype dtype, size_t m, uint64_t n, uint32_t log2_h, uint32_t log2_w,
                           uint32_t flags, fmx_spm **out);

fmx_status fmx_spm_info_get(const fmx_spm *spm, fmx_spm_info *info);

/* Copies the image (cap >= image_bytes) and the nbr+1 offsets to the host;
 * the counterpart of SpM_2d_storage::dump / SpM_2d_index::dump
 * (matrix/sparse_matrix_format.cpp:164-181,417-420). */
fmx_status fmx_spm_download(const fmx_spm *spm, void *image, size_t cap,
                            size_t *bytes, int64_t *offs);
fmx_status fmx_spm_destroy(fmx_spm *spm);

/* The .mat / .mat_idx pair on disk.  fmx_spm_load replaces
 * SpM_2d_index::load + SpM_2d_storage::load + sparse_matrix::create
 * (matrix/sparse_matrix_format.cpp:238-306, matrix/test/test-2d_multiply.cpp:
 * 147-161): the matrix file is the image (header page, then the block rows)
 * and the index file is the same header page followed by the nbr+1 int64
 * block-row offsets.  Both are validated as fmx_spm_upload does and the
 * headers must agree.  fmx_spm_dump writes the same two files
 * (SpM_2d_storage::dump, SpM_2d_index::dump, sparse_matrix_format.cpp:
 * 164-181,417-420).  Unreadable or unwritable files give FMX_EINVAL. */
fmx_status fmx_spm_load(fmx_ctx *ctx, const char *mat_file, const char *idx_file, fmx_spm **out);

/* Out-of-core matrices: the image stays in host memory and every SpMM
 * streams it through HBM.  fmx_spm_open replaces the SAFS-backed
 * sparse_matrix::create(SpM_2d_index::ptr, file_io_factory) (matrix/
 * sparse_matrix.h:534-535, .cpp:693-698) and its block-row I/O (matrix/
 * matrix_io.cpp:188-236, sparse_matrix.cpp:796-830): the .mat file is read
 * into pinned host memory (the index file as for fmx_spm_load), validated,
 * and cut into chunks of whole block rows of at most chunk_bytes (>= 4096;
 * a larger block row is a chunk of its own).  The tile directory and work
 * plans live in HBM, the image does not: fmx_spmm (and every caller of it)
 * then copies chunk c + 1 into one of two HBM staging buffers on the
 * matrix's own copy stream while the kernel of chunk c runs on the caller's
 * stream.  Results are bit-identical to the resident matrix's.  Calls on
 * one streamed matrix from several host threads at once are not supported
 * (they share the staging buffers).  fmx_spm_upload_streamed does the same
 * from an image in host memory (copied into pinned memory), and
 * fmx_spm_open_index from the .mat file plus an index already in memory: the
 * index's header page (its first 64 bytes must equal the file's) and its
 * nbr+1 block-row offsets. */
fmx_status fmx_spm_open(fmx_ctx *ctx, const char *mat_file, const char *idx_file, uint64_t chunk_bytes,
                        fmx_spm **out);
fmx_status fmx_spm_upload_streamed(fmx_ctx *ctx, const void *image, size_t bytes, const int64_t *offs, size_t nbr,
                                   uint64_t chunk_bytes, fmx_spm **out);
fmx_status fmx_spm_open_index(fmx_ctx *ctx, const char *mat_file, const void *header, const int64_t *offs,
                              size_t nbr, uint64_t chunk_bytes, fmx_spm **out);
fmx_status fmx_spm_dump(const fmx_spm *spm, const char *mat_file, const char *idx_file);

/* ---- SpMM ------------------------------------------------------------------
 * Y = A * X, replacing sparse_matrix::multiply(in, out)
 * (matrix/sparse_matrix.h:636-644, matrix/sparse_matrix.cpp:709-831) with
 * spmm_creator<T,T> (matrix/sparse_matrix.h:377-436).  X is ncols x k and Y
 * nrows x k, both row-major DEVICE buffers with leading dimensions ldx, ldy
 * (elements).  Y is overwritten; rows without non-zeros become 0
 * (matrix/sparse_matrix.cpp:339-345).  dtype is one of I32, I64, F32, F64;
 * a weighted image must carry values of the same type (sparse_matrix.cpp:
 * 754-766).  Integer results are exact; floating-point results are
 * deterministic from call to call. */
fmx_status fmx_spmm(fmx_ctx *ctx, const fmx_spm *spm, fmx_dtype dtype, int k,
                    const void *X, int64_t ldx, void *Y, int64_t ldy,
                    void *stream);

/* fmx_spmm with options (flags, OR-ed):
 *   FMX_SPMM_SEQ: floating-point sums in the reference's own order.  Each
 *     output element is the left-to-right sum, one rounding per addition, of
 *     its row's products in ascending column order starting from 0 -- what
 *     row_part_func / coo_func compute (`dest_row[j] += src_row[j] * data`,
 *     matrix/sparse_matrix.h:219-331) with the default sequential tile order
 *     (seq_exec_order, :161-176), for any tile shape.  FP32 and FP64 results
 *     are then bit-identical to the reference's in-memory CPU SpMM.  Without
 *     it (the default), FP32 rows are summed in FP64 and rounded once (the
 *     correctly rounded sum, which differs from the reference's sequential
 *     FP32 sum by that sum's own rounding error, about 1e-5 relative on
 *     power-law hub rows) and FP64 rows in a fixed order of the kernel's own;
 *     both are deterministic.  Integer results are exact either way.  The
 *     sequential mode does not split power-law rows over several waves, so
 *     it is slower on skewed matrices.
 *   FMX_SPMM_X_COLMAJOR: X is column-major, element (i, j) at X[j*ldx + i]
 *     (ldx >= ncols); it is converted to row-major in a context workspace
 *     first, as the reference converts a non-row-major input
 *     (matrix/sparse_matrix.cpp:720-736).
 *   FMX_SPMM_Y_COLMAJOR: Y is column-major, element (i, j) at Y[j*ldy + i]
 *     (ldy >= nrows); written directly (an L_COL output store, as the
 *     reference's eigen operators pass, matrix/test/test-eigen.cpp:115-118). */
#define FMX_SPMM_SEQ        1u
#define FMX_SPMM_X_COLMAJOR 2u
#define FMX_SPMM_Y_COLMAJOR 4u
fmx_status fmx_spmm_ex(fmx_ctx *ctx, const fmx_spm *spm, fmx_dtype dtype, int k,
                       const void *X, int64_t ldx, void *Y, int64_t ldy,
                       uint32_t flags, void *stream);

/* Y = diag(post) * A * diag(pre) * X: the SpMM of a custom task creator,
 * sparse_matrix::multiply(in, out, creator) (matrix/sparse_matrix.cpp:
 * 709-741, task_creator matrix/sparse_matrix.h:52-62) as the reference's
 * normalised-adjacency operator drives it (scale_rows by D^-1/2 on both
 * sides around spmm_creator<double, float>, matrix/test/test-eigen.cpp:
 * 112-127).  pre (ncols) and post (nrows) are DEVICE FP64 vectors or NULL
 * for none; dtype F32 or F64.  X * pre[i] is rounded to dtype once (staged
 * in a context workspace); post[r] multiplies row r's unrounded sum, which
 * is then rounded once (without FMX_SPMM_SEQ; with it, the sequential sum is
 * rounded, scaled, and rounded again).  A weighted image with FP32 values
 * multiplies FP64 dense data as spmm_creator<double, float> does: each
 * product is the FP64 product of the FP64 element and the FP32 value.
 * flags as for fmx_spmm_ex. */
fmx_status fmx_spmm_scaled(fmx_ctx *ctx, const fmx_spm *spm, fmx_dtype dtype, int k,
                           const double *pre, const double *post,
                           const void *X, int64_t ldx, void *Y, int64_t ldy,
                           uint32_t flags, void *stream);

/* Y = A * X with X (ncols x k) and Y (nrows x k) in HOST memory, for dense
 * operands larger than the HBM set aside for them: the column-panel branch
 * of sparse_matrix::multiply(dense_matrix, mem_size) (matrix/
 * sparse_matrix.cpp:773-831, panels at :796-830).  Columns go through HBM in
 * panels of kb = mem_bytes / ((ncols + nrows) * element size) columns (all
 * k if they fit): each panel of X is copied in, multiplied with fmx_spmm and
 * its Y panel copied back.  kb == 0 is FMX_ENOMEM ("The memory is too small
 * for the SpMM", :801-804).  Synchronous; leading dimensions in elements.
 * Pinned host buffers copy fastest.  Each column's result is the one
 * fmx_spmm gives for it. */
fmx_status fmx_spmm_host(fmx_ctx *ctx, const fmx_spm *spm, fmx_dtype dtype, int k, const void *X, int64_t ldx,
                         void *Y, int64_t ldy, uint64_t mem_bytes, void *stream);

/* The algorithmic bytes one fmx_spmm moves: tile bytes + 8*(nbr+1) +
 * ncols*k*s (X read once) + nrows*k*s (Y written once). */
fmx_status fmx_spmm_alg_bytes(const fmx_spm *spm, fmx_dtype dtype, int k,
                              uint64_t *bytes);

/* ---- tall-skinny dense operations (device operands, row-major) -------------- */
#define FMX_GRAM_F64  0  /* FP64 accumulation                              */
#define FMX_GRAM_BF16 1  /* BF16 MFMA products, FP32 accumulation (F32 in)  */

/* G[i][j] = sum_r A[r][i] * B[r][j]: G is ka x kb FP64 on the device.
 * Replaces MvTransMv / block_multi_vector::MvTransMv
 * (matrix/eigensolver/FM_MultiVector.h:499-558,
 * matrix/eigensolver/block_dense_matrix.cpp:1429-1513).  ka, kb <= 256. */
fmx_status fmx_gram(fmx_ctx *ctx, fmx_dtype dtype, int mode, uint64_t n, int ka, int kb,
                    const void *A, int64_t lda, const void *B, int64_t ldb, double *G,
                    void *stream);

/* No counterpart in the reference.  fmx_gram for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The per-block partial sums go to a fixed
 * 4 MiB device array that the library holds on every GPU it runs on until
 * the process ends, not to the context's grow-only workspace, and the block
 * count is capped so that they fit (a ka * kb * 8-byte output larger than
 * the array is FMX_EDIM; every ka, kb <= 256 fits).  Arguments, results and
 * errors are fmx_gram's; no device memory is allocated or freed.  All calls
 * on one GPU share that array: calls that overlap in time on one GPU (on
 * different streams or contexts) race, so order them. */
fmx_status fmx_gram_noalloc(fmx_ctx *ctx, fmx_dtype dtype, int mode, uint64_t n, int ka, int kb,
                            const void *A, int64_t lda, const void *B, int64_t ldb, double *G,
                            void *stream);

/* Y = alpha * A * C + beta * Y with A n x ka, C ka x kc (FP64, device,
 * row-major), Y n x kc; FP64 accumulation.  Replaces MvTimesMatAddMv /
 * gemm_op::run (FM_MultiVector.h:389-433, block_dense_matrix.cpp:822-914).
 * ka * kc <= 8192. */
fmx_status fmx_gemm_tall(fmx_ctx *ctx, fmx_dtype dtype, uint64_t n, int ka, int kc,
                         double alpha, const void *A, int64_t lda, const double *C,
                         double beta, void *Y, int64_t ldy, void *stream);

/* norms[j] = ||A[:, j]||_2 (FP64, device), MvNorm (FM_MultiVector.h:589-628). */
fmx_status fmx_colnorm(fmx_ctx *ctx, fmx_dtype dtype, uint64_t n, int k, const void *A,
                       int64_t lda, double *norms, void *stream);

/* Y[i][:] = d[i] * X[i][:] (d FP64 on the device): the diagonal scaling of
 * the normalised-adjacency operator (matrix/test/test-eigen.cpp:112-127). */
fmx_status fmx_row_scale(fmx_ctx *ctx, fmx_dtype dtype, uint64_t n, int k, const double *d,
                         const void *X, int64_t ldx, void *Y, int64_t ldy, void *stream);

/* ---- block Krylov eigensolver ---------------------------------------------- */
#define FMX_EIG_MIXED 0  /* FP32 SpMM, BF16-MFMA first Gram pass, FP64 reorth/RR/residual */
#define FMX_EIG_FP32  1  /* FP32 SpMM, FP64(-MFMA) Gram passes, FP64 RR/residual (default;
                            the BF16 pass reads the same FP32 operands and is no faster) */
#define FMX_EIG_FP64  2  /* everything FP64                                              */

/* eigen_options (matrix/eigensolver/eigensolver.h:59-73): which is "LA"
 * (largest algebraic), the solver is the block Krylov-Schur form. */
typedef struct {
    int nev;           /* wanted eigenpairs                                 */
    int block_size;    /* columns per Krylov block (<= 32)                  */
    int num_blocks;    /* blocks in the basis (block_size*num_blocks <= 128) */
    int max_restarts;
    double tol;        /* ||N x - theta x|| / |theta| (FP64) per pair       */
    int precision;     /* FMX_EIG_*                                          */
    int verbose;
    uint64_t seed;     /* random start block, uniform [-1, 1)               */
    double *history;   /* host array (or NULL): the largest relative        */
    int history_cap;   /* residual after each restart, estimated from the   */
                       /* Lanczos relation, or the FP64 one where checked   */
} fmx_eigen_opts;

typedef struct {
    int spmm_blocks;       /* block SpMMs (two per application of A^T A)     */
    int fp64_spmm_blocks;  /* FP64 block SpMMs of the residual checks        */
    int restarts;
    int converged;
    double max_resid;
    double seconds;
    int history_len;       /* entries written to opts->history               */
} fmx_eigen_stats;

fmx_status fmx_eigen_opts_init(fmx_eigen_opts *opts, int nev);

/* Top-nev eigenpairs of N = D^-1/2 A D^-1/2 for the symmetric binary
 * matrix A (NA_eigen_Operator, matrix/test/test-eigen.cpp:62-136, solved as
 * fm::eigen::compute_eigen does, eigensolver.cpp:118-308), with
 * D^-1/2 := 0 for isolated vertices.  X0 (device, n x block_size, FP32 or
 * FP64, leading dimension ldx0) is a warm start or NULL for a random one.
 * evals[nev] and resid[nev] are host arrays (descending), evecs a device
 * n x nev FP64 array (or NULL). */
fmx_status fmx_eigen_normalized(fmx_ctx *ctx, const fmx_spm *spm, const fmx_eigen_opts *opts,
                                const void *X0, int64_t ldx0, fmx_dtype x0_dtype,
                                double *evals, double *evecs, int64_t ldv, double *resid,
                                fmx_eigen_stats *stats);

/* The operators of the reference's eigensolver driver
 * (matrix/test/test-eigen.cpp:32-177), each solved as above: */
#define FMX_EIG_OP_NORMALIZED 0  /* D^-1/2 A D^-1/2, A symmetric (NA_eigen_Operator)  */
#define FMX_EIG_OP_PLAIN      1  /* A, symmetric (eigen_Operator, :32-58)               */
#define FMX_EIG_OP_SVD        2  /* A^T A, x -> A^T (A x) (SVD_Operator, :138-177)      */

/* compute_eigen(op, true, opts) (matrix/eigensolver/eigensolver.cpp:118-308)
 * for op = FMX_EIG_OP_*.  For the SVD operator A is nrows x ncols and At its
 * transpose (built from the swapped edge list, or loaded from the
 * reference's transpose image; NULL when A is symmetric, which is its own
 * transpose); evals are then the squared singular values of A, descending,
 * and evecs (ncols x nev) the right singular vectors.  The other operators
 * take At = NULL.  FP32 and mixed precision run the operator's SpMMs on FP32
 * data (a weighted matrix then needs FP32 values); the residual check runs
 * them on FP64 data. */
fmx_status fmx_eigen(fmx_ctx *ctx, int op, const fmx_spm *A, const fmx_spm *At, const fmx_eigen_opts *opts,
                     const void *X0, int64_t ldx0, fmx_dtype x0_dtype,
                     double *evals, double *evecs, int64_t ldv, double *resid,
                     fmx_eigen_stats *stats);

/* ---- SpMV callers ------------------------------------------------------------
 * PageRank as matrix/libmatrix-algs/PageRank.cpp:49-137 computes it (FP32):
 * out_deg = max(A*1, 1), pr = 1/n; each iteration in = pr/out_deg,
 * out = At*in, pr = damping*out + (1-damping)/n, until every vertex moved by
 * less than 0.001/n or max_iters.  At is the transpose of A (NULL: A is
 * symmetric).  pr is a device array of n floats; *iters the iterations run. */
fmx_status fmx_pagerank(fmx_ctx *ctx, const fmx_spm *A, const fmx_spm *At, int max_iters, float damping,
                        float *pr, int *iters, void *stream);

/* NMF of a binary A (n x m) by Lee-Seung multiplicative updates, as
 * fm::alg::NMF (matrix/libmatrix-algs/NMF.cpp:164-230,259-316) computes it:
 * each iteration updates H from A^T W and (W^T W) H, then W from A H^T and
 * W (H H^T), with X = max(X .* num, 1e-8) ./ (den + 1e-8) in FP32.  W (n x k)
 * and tH = H^T (m x k; the reference's column-major k x m H) are device FP32
 * row-major arrays holding the start on entry (the reference draws both
 * U[0,1)) and the result on return.  At is A's transpose (NULL: A is
 * symmetric).  dist (host, NULL to skip) receives ||A - W H||_F^2 after
 * iterations 0, 5, 10, ... (Fnorm, NMF.cpp:110-126): ceil(max_iters/5)
 * values.  1 <= k <= 64; a weighted A is FMX_EUNSUPPORTED as in the
 * reference. */
fmx_status fmx_nmf(fmx_ctx *ctx, const fmx_spm *A, const fmx_spm *At, int k, int max_iters, float *W, float *tH,
                   double *dist, void *stream);

/* ---- measurement -------------------------------------------------------------
 * No reference counterpart: the memory-system floor of the SpMM's dense-row
 * gathers.  Reads row cols[i] (clamped to x_rows - 1) of X for i < n with the
 * SpMM's load shape -- 16 bytes per lane, the lanes of a row side by side,
 * eight loads in flight per lane -- and nothing else.  row_bytes = k * element
 * size must be 16 times a power of two (at most 1024); X and ldx_bytes are
 * 16-byte aligned.  Timing it over a matrix's column sequence bounds what any
 * SpMM fetching the same rows in the same order can reach. */
fmx_status fmx_gather_floor(fmx_ctx *ctx, uint64_t row_bytes, const void *X, int64_t ldx_bytes, uint64_t x_rows,
                            const uint32_t *cols, uint64_t n, void *stream);

/* ---- seeded generators (device output) ------------------------------------
 * R-MAT edges with the project's SplitMix64 specification (the reference's
 * generator, flash-graph/tools/rmat-gen.cpp:43-48, is a time-seeded Boost
 * rmat_iterator); thr_* = floor(cumulative probability * 2^32). */
fmx_status fmx_rmat_edges(fmx_ctx *ctx, int scale, uint64_t m, uint64_t seed,
                          uint32_t thr_a, uint32_t thr_ab, uint32_t thr_abc,
                          int permute, uint32_t *src, uint32_t *dst,
                          void *stream);

/* Degree-corrected stochastic block model (no reference counterpart; the
 * closest is the plain SBM of flash-graph/utils/sbm.cpp:117-242).
 * fmx_dcsbm_params fills bstart[nblocks+1] (contiguous blocks, sizes
 * proportional to size_decay^-b) and wtab[1024] (quantiles of the degree
 * weight theta = (1-u)^(-1/(gamma-1)), capped at sqrt(n), in 1/1024 units)
 * on the host.  fmx_dcsbm_edges draws m edges on the device: u in proportion
 * to theta, then an intra-block partner with odds ratio:1 against an
 * inter-block one, both in proportion to theta; ids are relabelled with the
 * R-MAT permutation.  n must be a power of two. */
fmx_status fmx_dcsbm_params(uint64_t n, int nblocks, double gamma, double size_decay,
                            uint32_t *bstart, uint32_t *wtab);
fmx_status fmx_dcsbm_edges(fmx_ctx *ctx, uint64_t n, int nblocks, const uint32_t *bstart,
                           const uint32_t *wtab, uint32_t ratio, uint64_t m, uint64_t seed,
                           uint32_t *src, uint32_t *dst, void *stream);

/* ---- learned warm start (no reference counterpart) ----------------------
 * A 3-layer message-passing network over N = D^-1/2 A D^-1/2 predicting the
 * eigensolver's starting block: H_{l+1} = relu(H_l Ws_l + N H_l Wn_l + b_l),
 * Y = H_3 Wo; inputs X0 = [log(1+deg), 8 uniform features, 7 zeros] (n x 16),
 * hidden 64, output 16.  Parameters are one flat FP32 device array of
 * fmx_gnn_param_count() values; activations and work buffers are device
 * arrays of fmx_gnn_acts_floats(n) / fmx_gnn_work_floats(n) floats.
 * mode 0: BF16 MFMA products with FP32 accumulation; mode 1: FP32. */
uint64_t fmx_gnn_param_count(void);
uint64_t fmx_gnn_acts_floats(uint64_t n);
uint64_t fmx_gnn_work_floats(uint64_t n);
fmx_status fmx_gnn_features(fmx_ctx *ctx, const fmx_spm *spm, uint64_t seed, float *X0, double *dinv,
                            void *stream);
fmx_status fmx_gnn_forward(fmx_ctx *ctx, const fmx_spm *spm, const double *dinv, const float *params,
                           const float *X0, float *acts, float *Y, float *work, int mode, void *stream);
fmx_status fmx_gnn_backward(fmx_ctx *ctx, const fmx_spm *spm, const double *dinv, const float *params,
                            const float *X0, const float *acts, const float *dY, float *grads,
                            float *work, int mode, void *stream);
/* loss = k - tr((Y^T Y)^-1 (U^T Y)^T (U^T Y)) for orthonormal U (n x k), and
 * dY = dloss/dY (device, may be NULL). */
fmx_status fmx_subspace_loss(fmx_ctx *ctx, uint64_t n, int k, const float *Y, const float *U,
                             double *loss, float *dY, void *stream);
fmx_status fmx_adam(fmx_ctx *ctx, uint64_t np, float *params, const float *grads, float *m, float *v,
                    float lr, float beta1, float beta2, float eps, int step, void *stream);

#ifdef __cplusplus
}
#endif
#endif
