// algs.hip -- callers of the SpMM path: PageRank (k = 1) and NMF (k columns).
//
// fmx_pagerank restates matrix/libmatrix-algs/PageRank.cpp:49-137 on the
// device, FP32 as the reference computes it:
//   out_deg = max(A * 1, 1)                       (get_out_degree, :49-60, :83)
//   pr = 1/n                                      (mat_init_operate, :31-43)
//   repeat: in = pr / out_deg; pr1 = in * out_deg (:101-105)
//           out = A^T * in                        (:110, the SpMM kernel, k = 1)
//           pr2 = d * out + (1 - d) / n           (:118-119)
//           converged = #{|pr1 - pr2| < 0.001/n}  (:97, :120-124); pr = pr2
//   until every vertex has converged or max_iters.
// The element-wise steps use explicitly rounded operations, so that no
// multiply-add is contracted and each step rounds as the reference's does;
// only the SpMV sum differs (FP64 accumulation, rounded once).
#include <vector>

#include "dense.h"
#include "fmx_internal.h"

namespace fmx {
namespace {

constexpr int kT = 256;

uint32_t grid_for_n(fmx_ctx *ctx, uint64_t n)
{
    uint64_t b = (n + kT - 1) / kT;
    uint64_t cap = (uint64_t)ctx->num_cus * 8;
    return (uint32_t)(b < 1 ? 1 : (b > cap ? cap : b));
}

__global__ void __launch_bounds__(kT) pr_fill(float *x, uint64_t n, float v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT) x[i] = v;
}

__global__ void __launch_bounds__(kT) pr_init(float *deg, float *pr, uint64_t n, float p0)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT) {
        deg[i] = deg[i] > 1.0f ? deg[i] : 1.0f;
        pr[i] = p0;
    }
}

__global__ void __launch_bounds__(kT) pr_prepare(const float *pr, const float *deg, float *in, float *pr1,
                                                 uint64_t n, unsigned long long *conv)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *conv = 0;  // the counter of this iteration
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT) {
        const float v = __fdiv_rn(pr[i], deg[i]);
        in[i] = v;
        pr1[i] = __fmul_rn(v, deg[i]);
    }
}

__global__ void __launch_bounds__(kT) pr_update(const float *out, const float *pr1, float *pr, uint64_t n,
                                                float damping, float add, float eps, unsigned long long *conv)
{
    unsigned long long c = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT) {
        const float p2 = __fadd_rn(__fmul_rn(out[i], damping), add);
        c += fabsf(pr1[i] - p2) < eps ? 1ull : 0ull;
        pr[i] = p2;
    }
    for (int d = 32; d; d >>= 1) c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(conv, c);
}

// X = max(X * N, eps) / (D + eps), element-wise in FP32 with each step
// rounded (fm.pmax2(X * N, eps) / (den + eps), NMF.cpp:180-192,206-217).
__global__ void __launch_bounds__(kT) nmf_mult_update(float *X, const float *N, const float *D, uint64_t len,
                                                      float eps)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < len; i += (uint64_t)gridDim.x * kT)
        X[i] = __fdiv_rn(fmaxf(__fmul_rn(X[i], N[i]), eps), __fadd_rn(D[i], eps));
}

}  // namespace
}  // namespace fmx

using namespace fmx;

extern "C" {

fmx_status fmx_pagerank(fmx_ctx *ctx, const fmx_spm *A, const fmx_spm *At, int max_iters, float damping,
                        float *pr, int *iters, void *stream)
{
    if (!ctx || !A || !pr) return fail(ctx, FMX_EINVAL, "null argument");
    if (!At) At = A;
    if (A->nrows != A->ncols || At->nrows != A->nrows || At->ncols != A->ncols)
        return fail(ctx, FMX_EDIM, "PageRank needs a square matrix and its transpose");
    if (A->esz || At->esz) return fail(ctx, FMX_EUNSUPPORTED, "PageRank runs on binary matrices");
    if (max_iters < 0) return fail(ctx, FMX_EINVAL, "max_iters must be >= 0");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t n = A->nrows;
    if (iters) *iters = 0;
    if (n == 0) return FMX_OK;
    float *buf = nullptr;
    unsigned long long *conv = nullptr;
    FMX_HIP(ctx, hipMalloc(&buf, 4 * n * sizeof(float)));
    if (hipMalloc(&conv, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipFree(buf);
        return fail(ctx, FMX_ENOMEM, "PageRank counter");
    }
    float *deg = buf, *in = buf + n, *out = buf + 2 * n, *pr1 = buf + 3 * n;
    const uint32_t g = grid_for_n(ctx, n);
    fmx_status st = FMX_OK;
    // out_deg = max(A * 1, 1); the ones vector lives in `in` for this step
    hipLaunchKernelGGL(pr_fill, dim3(g), dim3(kT), 0, s, in, n, 1.0f);
    st = fmx_spmm(ctx, A, FMX_F32, 1, in, 1, deg, 1, s);
    if (!st) {
        hipLaunchKernelGGL(pr_init, dim3(g), dim3(kT), 0, s, deg, pr, n, (float)(1.0 / (double)n));
        const float eps = (float)(0.001 / (double)n);
        const float add = (float)((1.0 - (double)damping) / (double)n);
        int it = 0;
        unsigned long long c = 0;
        while (c < n && it < max_iters) {
            hipLaunchKernelGGL(pr_prepare, dim3(g), dim3(kT), 0, s, pr, deg, in, pr1, n, conv);
            st = fmx_spmm(ctx, At, FMX_F32, 1, in, 1, out, 1, s);
            if (st) break;
            hipLaunchKernelGGL(pr_update, dim3(g), dim3(kT), 0, s, out, pr1, pr, n, damping, add, eps, conv);
            if (hipMemcpyAsync(&c, conv, sizeof(c), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                st = fail(ctx, FMX_EHIP, "PageRank convergence count");
                break;
            }
            it++;
        }
        if (iters) *iters = it;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    (void)hipFree(conv);
    return st;
}

// NMF by Lee-Seung multiplicative updates, matrix/libmatrix-algs/NMF.cpp:
// 164-230 (update_lee) and 259-316 (NMF), with H kept transposed (tH, m x k
// row-major: the reference's L_COL k x m H byte for byte).  Per iteration:
//   tD  = tH * (W^T W)                      (:168-173)
//   tH  = max(tH .* (A^T W), eps) ./ (tD + eps)   (:176-193)
//   D2  = W * (tH^T tH)                     (:201-208)
//   W   = max(W .* (A tH), eps) ./ (D2 + eps)     (:211-224)
//   WtW = W^T W                             (:227-228)
// and after iterations 0, 5, 10, ... the distance ||A - W H||_F^2 =
// nnz - 2 tr((A^T W)^T tH) + tr((tH WtW)^T tH) (Fnorm, :110-126).  The SpMMs
// are FP32 (FP64 sums rounded once), the Gram and k x k products accumulate
// in FP64 (the reference: FP32 BLAS), the update is FP32.
fmx_status fmx_nmf(fmx_ctx *ctx, const fmx_spm *A, const fmx_spm *At, int k, int max_iters, float *W, float *tH,
                   double *dist, void *stream)
{
    if (!ctx || !A || !W || !tH) return fail(ctx, FMX_EINVAL, "null argument");
    if (A->esz) return fail(ctx, FMX_EUNSUPPORTED, "can only handle binary sparse matrix");
    if (!At) {
        if (A->nrows != A->ncols) return fail(ctx, FMX_EDIM, "a non-square matrix needs its transpose");
        At = A;
    }
    if (At->esz) return fail(ctx, FMX_EUNSUPPORTED, "can only handle binary sparse matrix");
    if (At->nrows != A->ncols || At->ncols != A->nrows) return fail(ctx, FMX_EDIM, "At is not A's transpose");
    if (k < 1 || k > 64) return fail(ctx, FMX_EDIM, "NMF rank must be in [1, 64]");
    if (max_iters < 0) return fail(ctx, FMX_EINVAL, "max_iters must be >= 0");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t n = A->nrows, m = A->ncols;
    const float eps = 10e-9f;
    float *buf = nullptr;
    double *g = nullptr;
    const uint64_t big = std::max(n, m) * (uint64_t)k;
    FMX_HIP(ctx, hipMalloc(&buf, 2 * big * sizeof(float)));
    if (hipMalloc(&g, 2 * (size_t)k * k * sizeof(double)) != hipSuccess) {
        (void)hipFree(buf);
        return fail(ctx, FMX_ENOMEM, "NMF Gram buffers");
    }
    float *P = buf, *Q = buf + big;  // (m|n) x k products and denominators
    double *G = g, *G2 = g + (size_t)k * k;
    std::vector<double> h1((size_t)k * k), h2((size_t)k * k);
    auto upd = [&](float *X, const float *N, const float *D, uint64_t rows) {
        const uint64_t len = rows * (uint64_t)k;
        hipLaunchKernelGGL(nmf_mult_update, dim3(grid_for_n(ctx, len)), dim3(kT), 0, s, X, N, D, len, eps);
        return hipGetLastError() == hipSuccess ? FMX_OK : fail(ctx, FMX_EHIP, "NMF update launch");
    };
    fmx_status st = gram(ctx, FMX_F32, FMX_GRAM_F64, n, k, k, W, k, W, k, G, s);  // W^T W
    for (int it = 0; !st && it < max_iters; it++) {
        if (!st) st = gemm_tall(ctx, FMX_F32, m, k, k, 1.0, tH, k, G, 0.0, Q, k, s);
        if (!st) st = fmx_spmm(ctx, At, FMX_F32, k, W, k, P, k, s);
        if (!st) st = upd(tH, P, Q, m);
        if (!st) st = gram(ctx, FMX_F32, FMX_GRAM_F64, m, k, k, tH, k, tH, k, G2, s);  // H H^T
        if (!st) st = gemm_tall(ctx, FMX_F32, n, k, k, 1.0, W, k, G2, 0.0, Q, k, s);
        if (!st) st = fmx_spmm(ctx, A, FMX_F32, k, tH, k, P, k, s);
        if (!st) st = upd(W, P, Q, n);
        if (!st) st = gram(ctx, FMX_F32, FMX_GRAM_F64, n, k, k, W, k, W, k, G, s);
        if (!st && dist && it % 5 == 0) {
            st = fmx_spmm(ctx, At, FMX_F32, k, W, k, P, k, s);                            // A^T W
            if (!st) st = gram(ctx, FMX_F32, FMX_GRAM_F64, m, k, k, P, k, tH, k, G2, s);  // (A^T W)^T tH
            if (!st && hipMemcpyAsync(h1.data(), G2, h1.size() * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
                st = fail(ctx, FMX_EHIP, "NMF distance");
            if (!st) st = gemm_tall(ctx, FMX_F32, m, k, k, 1.0, tH, k, G, 0.0, Q, k, s);  // tH W^T W
            if (!st) st = gram(ctx, FMX_F32, FMX_GRAM_F64, m, k, k, Q, k, tH, k, G2, s);
            if (!st && (hipMemcpyAsync(h2.data(), G2, h2.size() * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess))
                st = fail(ctx, FMX_EHIP, "NMF distance");
            if (!st) {
                double t1 = 0, t2 = 0;
                for (int j = 0; j < k; j++) {
                    t1 += h1[(size_t)j * k + j];
                    t2 += h2[(size_t)j * k + j];
                }
                dist[it / 5] = (double)A->nnz - 2.0 * t1 + t2;
            }
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    (void)hipFree(g);
    return st;
}

}  // extern "C"
