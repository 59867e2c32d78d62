// algs.hip -- SpMV callers of the SpMM path: PageRank.
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

}  // extern "C"
