// dense.h -- internal entry points of the tall-skinny dense operations.
#pragma once

#include "fmx_internal.h"

namespace fmx {

fmx_status ctx_scratch(fmx_ctx *ctx, size_t bytes, void **out);
// Finds this device's copy of the fixed scratch of fmx_gram_noalloc (the
// context's device must be current) and records it on the context.
fmx_status noalloc_scratch_lookup(fmx_ctx *ctx);
// fixed_ws: the partial sums go to this buffer of fixed_bytes (the block
// count is capped to fit) instead of the context's grow-only workspace.
fmx_status gram(fmx_ctx *ctx, fmx_dtype dt, int mode, uint64_t n, int ka, int kb, const void *A,
                int64_t lda, const void *B, int64_t ldb, double *G, hipStream_t s, void *fixed_ws = nullptr,
                size_t fixed_bytes = 0);
fmx_status gemm_tall(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int ka, int kc, double alpha,
                     const void *A, int64_t lda, const double *C, double beta, void *Y, int64_t ldy,
                     hipStream_t s);
// G = A^T B with A already BF16 (n x ka, leading dimension lda, ka and lda
// multiples of 8, ka <= 128, 16-byte aligned) and B FP32 (kb <= 32, a
// multiple of 4): BF16 MFMA, FP32 accumulation, partials summed in FP64.
fmx_status gram_bf16_shadow(fmx_ctx *ctx, uint64_t n, int ka, int kb, const unsigned short *Ab, int64_t lda,
                            const float *B, int64_t ldb, double *G, hipStream_t s);
// gemm_tall (kc <= 32, ka <= 128) that also writes, for each stored value
// y of row r, y rounded to BF16 into Yb (FP32 only; or NULL) and y * ds[r]
// into Ys (or NULL), with leading dimensions ldyb and ldys.
fmx_status gemm_tall_copies(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int ka, int kc, double alpha, const void *A,
                            int64_t lda, const double *C, double beta, void *Y, int64_t ldy, unsigned short *Yb,
                            int64_t ldyb, const double *ds, void *Ys, int64_t ldys, hipStream_t s);
fmx_status gemm_tall_f32_f64(fmx_ctx *ctx, uint64_t n, int ka, int kc, double alpha, const float *A,
                             int64_t lda, const double *C, double beta, double *Y, int64_t ldy,
                             hipStream_t s);
fmx_status colnorm(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int k, const void *A, int64_t lda,
                   double *norms, int squared, hipStream_t s);
fmx_status row_scale(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int k, const double *d, const void *X,
                     int64_t ldx, void *Y, int64_t ldy, hipStream_t s);

}  // namespace fmx
