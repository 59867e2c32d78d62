// small.h -- small dense FP64 kernels of the block eigensolver.
#pragma once

#include <algorithm>

#include "fmx_internal.h"

namespace fmx {

// Eigen-decomposition of the symmetric m x m matrix H (device, row-major,
// m even, m <= 128): evals[m] unsorted, V[m x m] eigenvectors in columns;
// info[0] sweeps, info[1] final relative off-diagonal norm.  ws holds
// small_sym_eig_ws(m) bytes (the logged rotations of at most kJacobiSweeps
// sweeps).
constexpr int kJacobiSweeps = 40;
size_t small_sym_eig_ws(int m);
fmx_status small_sym_eig(fmx_ctx *ctx, const double *H, int m, double *evals, double *V,
                         double *info, void *ws, hipStream_t s);
// G (b x b) = R^T R, Rinv = R^-1 upper; info[0] = first bad pivot (1-based) or 0.
fmx_status small_chol_inv(fmx_ctx *ctx, const double *G, int b, double *Rinv, double *info,
                          hipStream_t s);
fmx_status h_update(fmx_ctx *ctx, double *H, int ld, int rows, int c0, int b, const double *C1,
                    const double *C2, hipStream_t s);
fmx_status cast_f32_f64(fmx_ctx *ctx, const float *X, int64_t ldx, uint64_t n, int k, double *Y,
                        int64_t ldy, hipStream_t s);
fmx_status cast_f64_f32(fmx_ctx *ctx, const double *X, int64_t ldx, uint64_t n, int k, float *Y,
                        int64_t ldy, hipStream_t s);
// out[j] = ||Y[:, j] - theta[j] X[:, j]||_2; partial holds nblk * k doubles.
fmx_status resid_norms(fmx_ctx *ctx, const double *Y, int64_t ldy, const double *X, int64_t ldx,
                       const double *theta, uint64_t n, int k, double *out, double *partial,
                       int nblk, hipStream_t s);
fmx_status rand_uniform(fmx_ctx *ctx, fmx_dtype dt, void *X, int64_t ldx, uint64_t n, int k, uint64_t seed,
                        hipStream_t s);
fmx_status degree_inv_sqrt(fmx_ctx *ctx, const float *deg, uint64_t n, double *dinv, hipStream_t s);

}  // namespace fmx
