// dense.hip -- tall-skinny dense operations of the block eigensolver.
//
// Reference operations (all HBM-bound on an n x k operand, k <= 256):
//   Gram        B = A^T X      MvTransMv (FM_MultiVector.h:499-558,
//                              block_dense_matrix.cpp:1429-1513)
//   gemm_tall   C = a*A*B+b*C  MvTimesMatAddMv (FM_MultiVector.h:389-433,
//                              block_dense_matrix.cpp:822-914)
//   colnorm     ||a_j||_2      MvNorm (FM_MultiVector.h:589-628)
//   row_scale   diag(d)*X      NA_eigen_Operator's scale_rows
//                              (matrix/test/test-eigen.cpp:112-127)
// The Gram comes in two precisions: FP64 accumulation (the reorthogonalising
// pass) and BF16 MFMA with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
// Partial results per row block are reduced in a fixed order, so every
// result is deterministic.
#include "dense.h"

namespace fmx {
namespace {

constexpr int kT = 256;

template <typename T>
__device__ __forceinline__ double ld64(const T *p) { return (double)*p; }

// ---- FP64 Gram ----------------------------------------------------------------
// Block b accumulates G_b = sum over its rows of A[r][:]^T B[r][:] for the
// output slice [o0, o0 + 4096) of the ka*kb outputs, rows staged in LDS.
constexpr int kGRows = 32;

template <typename T>
__global__ void __launch_bounds__(kT) gram64_kernel(const T *A, int64_t lda, const T *B, int64_t ldb,
                                                    uint64_t n, int ka, int kb, uint64_t rows_per_blk,
                                                    double *partial)
{
    extern __shared__ double sm[];
    double *As = sm;                  // kGRows x ka
    double *Bs = sm + kGRows * ka;    // kGRows x kb
    const int nout = ka * kb;
    const int o0 = blockIdx.y * 4096;
    double acc[16];
#pragma unroll
    for (int t = 0; t < 16; t++) acc[t] = 0.0;
    const uint64_t r0 = blockIdx.x * rows_per_blk;
    const uint64_t r1 = min(n, r0 + rows_per_blk);
    for (uint64_t rb = r0; rb < r1; rb += kGRows) {
        const int nr = (int)min((uint64_t)kGRows, r1 - rb);
        for (int e = threadIdx.x; e < kGRows * ka; e += kT) {
            int r = e / ka, c = e - r * ka;
            As[e] = r < nr ? ld64(A + (rb + r) * lda + c) : 0.0;
        }
        for (int e = threadIdx.x; e < kGRows * kb; e += kT) {
            int r = e / kb, c = e - r * kb;
            Bs[e] = r < nr ? ld64(B + (rb + r) * ldb + c) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int o = o0 + t * kT + threadIdx.x;
            if (o < nout) {
                const int i = o / kb, j = o - i * kb;
                double s = acc[t];
                for (int r = 0; r < nr; r++) s = fma(As[r * ka + i], Bs[r * kb + j], s);
                acc[t] = s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const int o = o0 + t * kT + threadIdx.x;
        if (o < nout) partial[(uint64_t)blockIdx.x * nout + o] = acc[t];
    }
}

// out[o] = sum_b partial[b][o] in block order (deterministic)
template <typename TP, typename TO>
__global__ void __launch_bounds__(kT) reduce_partials(const TP *partial, int nblk, int nout,
                                                      TO *out, int sqrt_out)
{
    for (int o = blockIdx.x * kT + threadIdx.x; o < nout; o += gridDim.x * kT) {
        TP s = 0;
        for (int b = 0; b < nblk; b++) s += partial[(uint64_t)b * nout + o];
        double v = (double)s;
        out[o] = (TO)(sqrt_out ? sqrt(v) : v);
    }
}

// ---- BF16 MFMA Gram ------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f)
{
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (unsigned short)(u >> 16);  // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

constexpr int kBK = 64;          // rows per LDS stage (two MFMA k-steps)
constexpr int kBPad = kBK + 8;   // padded row length (bf16 elements)

// ka, kb multiples of 16 (caller pads); tiles of 16x16 spread over 4 waves.
__global__ void __launch_bounds__(kT) gram_bf16_kernel(const float *A, int64_t lda, const float *B,
                                                       int64_t ldb, uint64_t n, int ka, int kb,
                                                       int ka_real, int kb_real,
                                                       uint64_t rows_per_blk, float *partial)
{
    extern __shared__ unsigned short smb[];
    unsigned short *At = smb;                    // ka x kBPad  (A transposed)
    unsigned short *Bt = smb + ka * kBPad;       // kb x kBPad
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tiles_i = ka / 16, tiles_j = kb / 16, ntiles = tiles_i * tiles_j;
    constexpr int kMaxTiles = 16;                // per wave (ka*kb <= 64*256*16)
    floatx4 acc[kMaxTiles];
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    const uint64_t r0 = blockIdx.x * rows_per_blk;
    const uint64_t r1 = min(n, r0 + rows_per_blk);
    for (uint64_t rb = r0; rb < r1; rb += kBK) {
        const int nr = (int)min((uint64_t)kBK, r1 - rb);
        for (int e = threadIdx.x; e < kBK * ka; e += kT) {
            int r = e / ka, c = e - r * ka;
            float v = (r < nr && c < ka_real) ? A[(rb + r) * lda + c] : 0.f;
            At[c * kBPad + r] = f2bf(v);
        }
        for (int e = threadIdx.x; e < kBK * kb; e += kT) {
            int r = e / kb, c = e - r * kb;
            float v = (r < nr && c < kb_real) ? B[(rb + r) * ldb + c] : 0.f;
            Bt[c * kBPad + r] = f2bf(v);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kMaxTiles; t++) {
            const int tile = w + 4 * t;
            if (tile < ntiles) {
                const int ti = tile / tiles_j, tj = tile - ti * tiles_j;
#pragma unroll
                for (int ks = 0; ks < kBK / 32; ks++) {
                    const int k0 = ks * 32 + (lane >> 4) * 8;
                    bf16x8 a = *(const bf16x8 *)(At + (ti * 16 + (lane & 15)) * kBPad + k0);
                    bf16x8 b = *(const bf16x8 *)(Bt + (tj * 16 + (lane & 15)) * kBPad + k0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    const int nout = ka_real * kb_real;
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) {
        const int tile = w + 4 * t;
        if (tile < ntiles) {
            const int ti = tile / tiles_j, tj = tile - ti * tiles_j;
            const int j = tj * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = ti * 16 + (lane >> 4) * 4 + q;
                if (i < ka_real && j < kb_real)
                    partial[(uint64_t)blockIdx.x * nout + i * kb_real + j] = acc[t][q];
            }
        }
    }
}

// ---- gemm_tall -------------------------------------------------------------------
// Y[r][j] = alpha * sum_i A[r][i] C[i][j] + beta * Y[r][j], FP64 accumulation.
constexpr int kMRows = 32;

template <typename T, typename TY = T>
__global__ void __launch_bounds__(kT) gemm_tall_kernel(const T *A, int64_t lda, uint64_t n, int ka,
                                                       int kc, const double *C, double alpha,
                                                       double beta, TY *Y, int64_t ldy)
{
    extern __shared__ double sg[];
    double *Cs = sg;                 // ka x kc
    double *As = sg + ka * kc;       // kMRows x ka
    for (int e = threadIdx.x; e < ka * kc; e += kT) Cs[e] = C[e];
    for (uint64_t rb = blockIdx.x * (uint64_t)kMRows; rb < n; rb += (uint64_t)gridDim.x * kMRows) {
        const int nr = (int)min((uint64_t)kMRows, n - rb);
        __syncthreads();
        for (int e = threadIdx.x; e < kMRows * ka; e += kT) {
            int r = e / ka, c = e - r * ka;
            As[e] = r < nr ? ld64(A + (rb + r) * lda + c) : 0.0;
        }
        __syncthreads();
        for (int o = threadIdx.x; o < nr * kc; o += kT) {
            const int r = o / kc, j = o - r * kc;
            double s = 0.0;
            for (int i = 0; i < ka; i++) s = fma(As[r * ka + i], Cs[i * kc + j], s);
            TY *y = Y + (rb + r) * ldy + j;
            const double prev = (beta == 0.0) ? 0.0 : beta * (double)*y;
            *y = (TY)(prev + alpha * s);
        }
    }
}

// ---- column norms / sums of squares ----------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kT) colsq_kernel(const T *A, int64_t lda, uint64_t n, int k,
                                                   uint64_t rows_per_blk, double *partial)
{
    // thread t handles column t % k for rows t / k, t / k + kT / k, ...
    const int per = kT / k;
    const int c = threadIdx.x % k, rr = threadIdx.x / k;
    double s = 0.0;
    if (rr < per) {
        const uint64_t r0 = blockIdx.x * rows_per_blk;
        const uint64_t r1 = min(n, r0 + rows_per_blk);
        for (uint64_t r = r0 + rr; r < r1; r += per) {
            double v = ld64(A + r * lda + c);
            s = fma(v, v, s);
        }
    }
    __shared__ double red[kT];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < k) {
        double t = 0.0;
        for (int q = 0; q < per; q++) t += red[q * k + threadIdx.x];
        partial[(uint64_t)blockIdx.x * k + threadIdx.x] = t;
    }
}

template <typename T, typename TD>
__global__ void __launch_bounds__(kT) row_scale_kernel(uint64_t n, int k, const TD *d, const T *X,
                                                       int64_t ldx, T *Y, int64_t ldy)
{
    const uint64_t tot = n * (uint64_t)k;
    for (uint64_t e = blockIdx.x * (uint64_t)kT + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * kT) {
        const uint64_t r = e / k, c = e - r * k;
        Y[r * ldy + c] = (T)((TD)X[r * ldx + c] * d[r]);
    }
}

uint32_t blocks_for(fmx_ctx *ctx, uint64_t n, uint64_t rows_min)
{
    uint64_t b = (n + rows_min - 1) / rows_min;
    uint64_t cap = (uint64_t)ctx->num_cus * 4;
    if (b > cap) b = cap;
    return (uint32_t)(b ? b : 1);
}

}  // namespace

fmx_status ctx_scratch(fmx_ctx *ctx, size_t bytes, void **out)
{
    if (bytes > ctx->ws_bytes) {
        if (ctx->ws) FMX_HIP(ctx, hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
        FMX_HIP(ctx, hipMalloc(&ctx->ws, bytes));
        ctx->ws_bytes = bytes;
    }
    *out = ctx->ws;
    return FMX_OK;
}

fmx_status gram(fmx_ctx *ctx, fmx_dtype dt, int mode, uint64_t n, int ka, int kb, const void *A,
                int64_t lda, const void *B, int64_t ldb, double *G, hipStream_t s)
{
    if (ka <= 0 || kb <= 0 || ka > 256 || kb > 256) return fail(ctx, FMX_EDIM, "gram widths must be in [1, 256]");
    if (lda < ka || ldb < kb) return fail(ctx, FMX_EDIM, "leading dimension too small");
    const int nout = ka * kb;
    if (mode == FMX_GRAM_BF16) {
        if (dt != FMX_F32) return fail(ctx, FMX_ETYPE, "BF16 Gram takes FP32 operands");
        const int kap = (ka + 15) & ~15, kbp = (kb + 15) & ~15;
        if ((kap / 16) * (kbp / 16) > 64) return fail(ctx, FMX_EDIM, "BF16 Gram output too large");
        const uint32_t nblk = blocks_for(ctx, n, 1024);
        const uint64_t rpb = (n + nblk - 1) / nblk;
        void *ws;
        fmx_status st = ctx_scratch(ctx, (size_t)nblk * nout * sizeof(float), &ws);
        if (st) return st;
        const size_t lds = (size_t)(kap + kbp) * kBPad * sizeof(unsigned short);
        static bool attr = false;
        if (!attr) {
            FMX_HIP(ctx, hipFuncSetAttribute((const void *)gram_bf16_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr = true;
        }
        hipLaunchKernelGGL(gram_bf16_kernel, dim3(nblk), dim3(kT), lds, s, (const float *)A, lda,
                           (const float *)B, ldb, n, kap, kbp, ka, kb, rpb, (float *)ws);
        FMX_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL((reduce_partials<float, double>), dim3((nout + kT - 1) / kT), dim3(kT), 0, s,
                           (const float *)ws, (int)nblk, nout, G, 0);
        FMX_HIP(ctx, hipGetLastError());
        return FMX_OK;
    }
    if (mode != FMX_GRAM_F64) return fail(ctx, FMX_EINVAL, "unknown Gram mode");
    const uint32_t nblk = blocks_for(ctx, n, 512);
    const uint64_t rpb = (n + nblk - 1) / nblk;
    void *ws;
    fmx_status st = ctx_scratch(ctx, (size_t)nblk * nout * sizeof(double), &ws);
    if (st) return st;
    const size_t lds = (size_t)kGRows * (ka + kb) * sizeof(double);
    dim3 grid(nblk, (nout + 4095) / 4096);
    if (dt == FMX_F32)
        hipLaunchKernelGGL(gram64_kernel<float>, grid, dim3(kT), lds, s, (const float *)A, lda,
                           (const float *)B, ldb, n, ka, kb, rpb, (double *)ws);
    else if (dt == FMX_F64)
        hipLaunchKernelGGL(gram64_kernel<double>, grid, dim3(kT), lds, s, (const double *)A, lda,
                           (const double *)B, ldb, n, ka, kb, rpb, (double *)ws);
    else
        return fail(ctx, FMX_ETYPE, "Gram takes FP32 or FP64 operands");
    FMX_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL((reduce_partials<double, double>), dim3((nout + kT - 1) / kT), dim3(kT), 0, s,
                       (const double *)ws, (int)nblk, nout, G, 0);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status gemm_tall_f32_f64(fmx_ctx *ctx, uint64_t n, int ka, int kc, double alpha, const float *A,
                             int64_t lda, const double *C, double beta, double *Y, int64_t ldy,
                             hipStream_t s)
{
    if (ka <= 0 || kc <= 0 || (size_t)ka * kc > 8192 || ka > 512)
        return fail(ctx, FMX_EDIM, "gemm_tall needs ka*kc <= 8192");
    const size_t lds = ((size_t)ka * kc + (size_t)kMRows * ka) * sizeof(double);
    const uint32_t nblk = blocks_for(ctx, n, kMRows);
    static bool attr = false;
    if (!attr) {
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)gemm_tall_kernel<float, double>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    hipLaunchKernelGGL((gemm_tall_kernel<float, double>), dim3(nblk), dim3(kT), lds, s, A, lda, n, ka, kc,
                       C, alpha, beta, Y, ldy);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status gemm_tall(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int ka, int kc, double alpha,
                     const void *A, int64_t lda, const double *C, double beta, void *Y, int64_t ldy,
                     hipStream_t s)
{
    if (ka <= 0 || kc <= 0 || (size_t)ka * kc > 8192 || ka > 512)
        return fail(ctx, FMX_EDIM, "gemm_tall needs ka*kc <= 8192");
    if (lda < ka || ldy < kc) return fail(ctx, FMX_EDIM, "leading dimension too small");
    const size_t lds = ((size_t)ka * kc + (size_t)kMRows * ka) * sizeof(double);
    const uint32_t nblk = blocks_for(ctx, n, kMRows);
    static bool attr = false;
    if (!attr) {
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)gemm_tall_kernel<float>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)gemm_tall_kernel<double>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    if (dt == FMX_F32)
        hipLaunchKernelGGL(gemm_tall_kernel<float>, dim3(nblk), dim3(kT), lds, s, (const float *)A, lda, n,
                           ka, kc, C, alpha, beta, (float *)Y, ldy);
    else if (dt == FMX_F64)
        hipLaunchKernelGGL(gemm_tall_kernel<double>, dim3(nblk), dim3(kT), lds, s, (const double *)A, lda,
                           n, ka, kc, C, alpha, beta, (double *)Y, ldy);
    else
        return fail(ctx, FMX_ETYPE, "gemm_tall takes FP32 or FP64 operands");
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status colnorm(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int k, const void *A, int64_t lda,
                   double *norms, int squared, hipStream_t s)
{
    if (k <= 0 || k > kT) return fail(ctx, FMX_EDIM, "colnorm width must be in [1, 256]");
    if (lda < k) return fail(ctx, FMX_EDIM, "leading dimension too small");
    const uint32_t nblk = blocks_for(ctx, n, 2048);
    const uint64_t rpb = (n + nblk - 1) / nblk;
    void *ws;
    fmx_status st = ctx_scratch(ctx, (size_t)nblk * k * sizeof(double), &ws);
    if (st) return st;
    if (dt == FMX_F32)
        hipLaunchKernelGGL(colsq_kernel<float>, dim3(nblk), dim3(kT), 0, s, (const float *)A, lda, n, k,
                           rpb, (double *)ws);
    else if (dt == FMX_F64)
        hipLaunchKernelGGL(colsq_kernel<double>, dim3(nblk), dim3(kT), 0, s, (const double *)A, lda, n, k,
                           rpb, (double *)ws);
    else
        return fail(ctx, FMX_ETYPE, "colnorm takes FP32 or FP64 operands");
    FMX_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL((reduce_partials<double, double>), dim3(1), dim3(kT), 0, s, (const double *)ws,
                       (int)nblk, k, norms, squared ? 0 : 1);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status row_scale(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int k, const double *d, const void *X,
                     int64_t ldx, void *Y, int64_t ldy, hipStream_t s)
{
    if (k <= 0 || ldx < k || ldy < k) return fail(ctx, FMX_EDIM, "bad width or leading dimension");
    uint64_t tot = n * (uint64_t)k;
    uint64_t blocks = (tot + kT - 1) / kT;
    if (blocks > (uint64_t)ctx->num_cus * 16) blocks = (uint64_t)ctx->num_cus * 16;
    if (!blocks) return FMX_OK;
    if (dt == FMX_F32)
        hipLaunchKernelGGL((row_scale_kernel<float, double>), dim3((uint32_t)blocks), dim3(kT), 0, s, n, k,
                           d, (const float *)X, ldx, (float *)Y, ldy);
    else if (dt == FMX_F64)
        hipLaunchKernelGGL((row_scale_kernel<double, double>), dim3((uint32_t)blocks), dim3(kT), 0, s, n,
                           k, d, (const double *)X, ldx, (double *)Y, ldy);
    else
        return fail(ctx, FMX_ETYPE, "row_scale takes FP32 or FP64 operands");
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

}  // namespace fmx

using namespace fmx;

extern "C" {

fmx_status fmx_gram(fmx_ctx *ctx, fmx_dtype dtype, int mode, uint64_t n, int ka, int kb,
                    const void *A, int64_t lda, const void *B, int64_t ldb, double *G, void *stream)
{
    if (!ctx || (n && (!A || !B)) || !G) return fail(ctx, FMX_EINVAL, "null argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    return gram(ctx, dtype, mode, n, ka, kb, A, lda, B, ldb, G,
                (hipStream_t)stream);
}

fmx_status fmx_gemm_tall(fmx_ctx *ctx, fmx_dtype dtype, uint64_t n, int ka, int kc, double alpha,
                         const void *A, int64_t lda, const double *C, double beta, void *Y,
                         int64_t ldy, void *stream)
{
    if (!ctx || (n && (!A || !Y)) || !C) return fail(ctx, FMX_EINVAL, "null argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    return gemm_tall(ctx, dtype, n, ka, kc, alpha, A, lda, C, beta, Y, ldy,
                     (hipStream_t)stream);
}

fmx_status fmx_colnorm(fmx_ctx *ctx, fmx_dtype dtype, uint64_t n, int k, const void *A,
                       int64_t lda, double *norms, void *stream)
{
    if (!ctx || (n && !A) || !norms) return fail(ctx, FMX_EINVAL, "null argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    return colnorm(ctx, dtype, n, k, A, lda, norms, 0, (hipStream_t)stream);
}

fmx_status fmx_row_scale(fmx_ctx *ctx, fmx_dtype dtype, uint64_t n, int k, const double *d,
                         const void *X, int64_t ldx, void *Y, int64_t ldy, void *stream)
{
    if (!ctx || (n && (!X || !Y || !d))) return fail(ctx, FMX_EINVAL, "null argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    return row_scale(ctx, dtype, n, k, d, X, ldx, Y, ldy, (hipStream_t)stream);
}

}  // extern "C"
