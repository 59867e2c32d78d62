// gnn.hip -- the learned warm start: a 3-layer message-passing network that
// predicts the starting block of the block Krylov eigensolver.
//
// No reference counterpart (the reference starts Anasazi from a random block,
// matrix/eigensolver/FM_MultiVector.h:164-168).  Network, for the normalised
// adjacency N = D^-1/2 A D^-1/2:
//   H0 = X0 (n x 16: log(1+deg), 8 uniform features, 7 zeros)
//   Agg_l = N H_l                      (fmx_spmm + two row scalings)
//   H_{l+1} = relu(H_l Ws_l + Agg_l Wn_l + b_l),  l = 0, 1, 2  (hidden 64)
//   Y = H_3 Wo                         (n x 16)
// Dense products run as v_mfma_f32_16x16x32_bf16 with BF16 operands (weights
// and activations rounded to BF16) and FP32 accumulation; a plain-FP32 mode
// exists for gradient checking.  Weight gradients are Gram products
// (fmx_gram, BF16 MFMA or FP64), the bias gradient a column sum.
// Parameters, flat FP32: per layer Ws (din x 64), Wn (din x 64), b (64);
// then Wo (64 x 16); din = 16, 64, 64 (19,648 values).
#include <cmath>
#include <vector>

#include "dense.h"
#include "small.h"

namespace fmx {
namespace {

constexpr int kIn = 16, kHid = 64, kOut = 16, kLayers = 3;
constexpr int kT = 256;
constexpr int kRows = 64;       // rows per block tile
constexpr int kKMax = 128;      // max inner dimension
constexpr int kPad = 8;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (unsigned short)(u >> 16);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

int din_of(int l) { return l == 0 ? kIn : kHid; }

size_t layer_off(int l)
{
    size_t o = 0;
    for (int i = 0; i < l; i++) o += 2ull * din_of(i) * kHid + kHid;
    return o;
}

// out[r][c] = act( sum_k A[r][k] B[k][c] + bias[c] + add[r][c] ), A = [A1 | A2]
// (n x K), B = W (K x dout row-major) or W^T (W dout x K row-major, wtrans).
template <bool BF16>
__global__ void __launch_bounds__(kT) gnn_gemm_kernel(const float *A1, int64_t lda1, int d1, const float *A2,
                                                      int64_t lda2, int d2, const float *W, int wtrans,
                                                      int dout, const float *bias, const float *add,
                                                      int64_t ldadd, int relu, float *out, int64_t ldo,
                                                      uint64_t n)
{
    const int K = d1 + d2;
    const int KP = (K + 31) & ~31;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if constexpr (BF16) {
        __shared__ unsigned short As[kRows * (kKMax + kPad)];
        __shared__ unsigned short Bt[kHid * (kKMax + kPad)];
        const int ld = KP + kPad;
        for (int e = threadIdx.x; e < dout * KP; e += kT) {
            const int c = e / KP, k = e - c * KP;
            float v = 0.f;
            if (k < K) v = wtrans ? W[(size_t)c * K + k] : W[(size_t)k * dout + c];
            Bt[c * ld + k] = f2bf(v);
        }
        const int nct = dout / 16;
        for (uint64_t r0 = blockIdx.x * (uint64_t)kRows; r0 < n; r0 += (uint64_t)gridDim.x * kRows) {
            __syncthreads();
            for (int e = threadIdx.x; e < kRows * KP; e += kT) {
                const int r = e / KP, k = e - r * KP;
                const uint64_t row = r0 + r;
                float v = 0.f;
                if (row < n && k < K) v = k < d1 ? A1[row * lda1 + k] : A2[row * lda2 + (k - d1)];
                As[r * ld + k] = f2bf(v);
            }
            __syncthreads();
            floatx4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
            for (int ks = 0; ks < KP; ks += 32) {
                const int k0 = ks + (lane >> 4) * 8;
                const bf16x8 a = *(const bf16x8 *)(As + (w * 16 + (lane & 15)) * ld + k0);
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    if (t < nct) {
                        const bf16x8 b = *(const bf16x8 *)(Bt + (t * 16 + (lane & 15)) * ld + k0);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (t >= nct) continue;
                const int c = t * 16 + (lane & 15);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint64_t row = r0 + w * 16 + (lane >> 4) * 4 + q;
                    if (row >= n) continue;
                    float v = acc[t][q];
                    if (bias) v += bias[c];
                    if (add) v += add[row * ldadd + c];
                    if (relu) v = v > 0.f ? v : 0.f;
                    out[row * ldo + c] = v;
                }
            }
        }
    } else {
        constexpr int kR = kRows / 2;  // 32-row tiles keep the FP32 operands in 48 KiB
        __shared__ float As[kR * (kKMax + 1)];
        __shared__ float Bs[kKMax * kHid];
        for (int e = threadIdx.x; e < K * dout; e += kT) {
            const int k = e / dout, c = e - k * dout;
            Bs[e] = wtrans ? W[(size_t)c * K + k] : W[(size_t)k * dout + c];
        }
        for (uint64_t r0 = blockIdx.x * (uint64_t)kR; r0 < n; r0 += (uint64_t)gridDim.x * kR) {
            __syncthreads();
            for (int e = threadIdx.x; e < kR * K; e += kT) {
                const int r = e / K, k = e - r * K;
                const uint64_t row = r0 + r;
                float v = 0.f;
                if (row < n) v = k < d1 ? A1[row * lda1 + k] : A2[row * lda2 + (k - d1)];
                As[r * (kKMax + 1) + k] = v;
            }
            __syncthreads();
            for (int o = threadIdx.x; o < kR * dout; o += kT) {
                const int r = o / dout, c = o - r * dout;
                const uint64_t row = r0 + r;
                if (row >= n) continue;
                float v = 0.f;
                for (int k = 0; k < K; k++) v = fmaf(As[r * (kKMax + 1) + k], Bs[k * dout + c], v);
                if (bias) v += bias[c];
                if (add) v += add[row * ldadd + c];
                if (relu) v = v > 0.f ? v : 0.f;
                out[row * ldo + c] = v;
            }
        }
    }
}

// dZ = dH * (H > 0), in place on dH
__global__ void __launch_bounds__(kT) relu_back_kernel(const float *H, float *dH, uint64_t tot)
{
    for (uint64_t e = blockIdx.x * (uint64_t)kT + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * kT)
        if (!(H[e] > 0.f)) dH[e] = 0.f;
}

__global__ void __launch_bounds__(kT) colsum_kernel(const float *A, uint64_t n, int k, uint64_t rpb,
                                                    double *partial)
{
    const int per = kT / k;
    const int c = threadIdx.x % k, rr = threadIdx.x / k;
    double s = 0.0;
    if (rr < per) {
        const uint64_t r0 = blockIdx.x * rpb, r1 = min(n, r0 + rpb);
        for (uint64_t r = r0 + rr; r < r1; r += per) s += (double)A[r * k + c];
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

__global__ void __launch_bounds__(kT) colsum_finish(const double *partial, int nblk, int k, float *out)
{
    for (int j = threadIdx.x; j < k; j += kT) {
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += partial[(uint64_t)b * k + j];
        out[j] = (float)s;
    }
}

__global__ void __launch_bounds__(kT) f64_to_f32(const double *a, float *b, int n)
{
    for (int i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) b[i] = (float)a[i];
}

__global__ void __launch_bounds__(kT) adam_kernel(uint64_t np, float *p, const float *g, float *m, float *v,
                                                  float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < np; i += (uint64_t)gridDim.x * kT) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// X0[v] = [log(1 + deg_v), 8 x uniform[-1, 1), 7 x 0]
__global__ void __launch_bounds__(kT) features_kernel(const float *deg, uint64_t n, uint64_t seed, float *X0)
{
    for (uint64_t e = blockIdx.x * (uint64_t)kT + threadIdx.x; e < n * kIn; e += (uint64_t)gridDim.x * kT) {
        const uint64_t v = e / kIn;
        const int j = (int)(e - v * kIn);
        float x = 0.f;
        if (j == 0) {
            x = logf(1.f + deg[v]);
        } else if (j <= 8) {
            uint64_t z = seed + (v * 8 + j) * 0x9E3779B97F4A7C15ULL;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            z ^= z >> 31;
            x = (float)((double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
        }
        X0[e] = x;
    }
}

uint32_t grid_n(fmx_ctx *ctx, uint64_t work)
{
    uint64_t b = (work + kT - 1) / kT;
    uint64_t cap = (uint64_t)ctx->num_cus * 16;
    return (uint32_t)(b < 1 ? 1 : (b > cap ? cap : b));
}

fmx_status gemm(fmx_ctx *ctx, int mode, const float *A1, int64_t lda1, int d1, const float *A2, int64_t lda2,
                int d2, const float *W, int wtrans, int dout, const float *bias, const float *add,
                int64_t ldadd, int relu, float *out, int64_t ldo, uint64_t n, hipStream_t s)
{
    const uint64_t tiles = (n + kRows - 1) / kRows;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(tiles ? tiles : 1, (uint64_t)ctx->num_cus * 8);
    if (mode == 0)
        hipLaunchKernelGGL(gnn_gemm_kernel<true>, dim3(grid), dim3(kT), 0, s, A1, lda1, d1, A2, lda2, d2, W,
                           wtrans, dout, bias, add, ldadd, relu, out, ldo, n);
    else
        hipLaunchKernelGGL(gnn_gemm_kernel<false>, dim3(grid), dim3(kT), 0, s, A1, lda1, d1, A2, lda2, d2, W,
                           wtrans, dout, bias, add, ldadd, relu, out, ldo, n);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

// out = N X for d columns: row scale, SpMM, row scale (T is n x d scratch)
fmx_status norm_agg(fmx_ctx *ctx, const fmx_spm *spm, const double *dinv, const float *X, int d, float *T,
                    float *out, hipStream_t s)
{
    const uint64_t n = spm->nrows;
    fmx_status st = row_scale(ctx, FMX_F32, n, d, dinv, X, d, T, d, s);
    if (st) return st;
    st = fmx_spmm(ctx, spm, FMX_F32, d, T, d, out, d, s);
    if (st) return st;
    return row_scale(ctx, FMX_F32, n, d, dinv, out, d, out, d, s);
}

// grads (FP32) <- Gram(A, B) through an FP64 scratch
fmx_status gram_grad(fmx_ctx *ctx, int mode, uint64_t n, int ka, int kb, const float *A, const float *B,
                     double *g64, float *out, hipStream_t s)
{
    fmx_status st = gram(ctx, FMX_F32, mode == 0 ? FMX_GRAM_BF16 : FMX_GRAM_F64, n, ka, kb, A, ka, B, kb, g64, s);
    if (st) return st;
    hipLaunchKernelGGL(f64_to_f32, dim3(16), dim3(kT), 0, s, g64, out, ka * kb);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

}  // namespace
}  // namespace fmx

using namespace fmx;

extern "C" {

uint64_t fmx_gnn_param_count(void) { return layer_off(kLayers) + (uint64_t)kHid * kOut; }
uint64_t fmx_gnn_acts_floats(uint64_t n) { return n * (uint64_t)(kIn + 5 * kHid); }
uint64_t fmx_gnn_work_floats(uint64_t n) { return n * (uint64_t)(4 * kHid); }

fmx_status fmx_gnn_features(fmx_ctx *ctx, const fmx_spm *spm, uint64_t seed, float *X0, double *dinv,
                            void *stream)
{
    if (!ctx || !spm || !X0 || !dinv) return fail(ctx, FMX_EINVAL, "null argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t n = spm->nrows;
    float *ones = nullptr, *deg = nullptr;
    FMX_HIP(ctx, hipMalloc(&ones, n * sizeof(float)));
    FMX_HIP(ctx, hipMalloc(&deg, n * sizeof(float)));
    std::vector<float> h1(n, 1.0f);
    fmx_status st = FMX_OK;
    if (hipMemcpyAsync(ones, h1.data(), n * 4, hipMemcpyHostToDevice, s) != hipSuccess) st = FMX_EHIP;
    if (!st) st = fmx_spmm(ctx, spm, FMX_F32, 1, ones, 1, deg, 1, s);
    if (!st) st = degree_inv_sqrt(ctx, deg, n, dinv, s);
    if (!st) {
        hipLaunchKernelGGL(features_kernel, dim3(grid_n(ctx, n * kIn)), dim3(kT), 0, s, deg, n, seed, X0);
        if (hipGetLastError() != hipSuccess) st = FMX_EHIP;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(ones);
    (void)hipFree(deg);
    return st;
}

fmx_status fmx_gnn_forward(fmx_ctx *ctx, const fmx_spm *spm, const double *dinv, const float *P,
                           const float *X0, float *acts, float *Y, float *work, int mode, void *stream)
{
    if (!ctx || !spm || !dinv || !P || !X0 || !acts || !Y || !work) return fail(ctx, FMX_EINVAL, "null argument");
    if (mode != 0 && mode != 1) return fail(ctx, FMX_EINVAL, "mode is 0 (BF16 MFMA) or 1 (FP32)");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t n = spm->nrows;
    const float *H = X0;
    float *a = acts;
    for (int l = 0; l < kLayers; l++) {
        const int d = din_of(l);
        float *Agg = a;
        a += n * d;
        float *Hn = a;
        a += n * kHid;
        fmx_status st = norm_agg(ctx, spm, dinv, H, d, work, Agg, s);
        if (st) return st;
        const float *Wl = P + layer_off(l);
        st = gemm(ctx, mode, H, d, d, Agg, d, d, Wl, 0, kHid, Wl + 2 * (size_t)d * kHid, nullptr, 0, 1, Hn, kHid, n, s);
        if (st) return st;
        H = Hn;
    }
    return gemm(ctx, mode, H, kHid, kHid, nullptr, 0, 0, P + layer_off(kLayers), 0, kOut, nullptr, nullptr, 0, 0, Y,
                kOut, n, s);
}

fmx_status fmx_gnn_backward(fmx_ctx *ctx, const fmx_spm *spm, const double *dinv, const float *P,
                            const float *X0, const float *acts, const float *dY, float *G, float *work, int mode,
                            void *stream)
{
    if (!ctx || !spm || !dinv || !P || !X0 || !acts || !dY || !G || !work) return fail(ctx, FMX_EINVAL, "null argument");
    if (mode != 0 && mode != 1) return fail(ctx, FMX_EINVAL, "mode is 0 (BF16 MFMA) or 1 (FP32)");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t n = spm->nrows;
    // activations: per layer Agg_l (n x din), H_{l+1} (n x 64)
    const float *Agg[kLayers], *Hs[kLayers + 1];
    Hs[0] = X0;
    const float *a = acts;
    for (int l = 0; l < kLayers; l++) {
        Agg[l] = a;
        a += n * din_of(l);
        Hs[l + 1] = a;
        a += n * kHid;
    }
    float *dH = work, *T = work + n * kHid, *T2 = work + 2 * n * kHid, *T3 = work + 3 * n * kHid;
    double *g64 = nullptr, *part = nullptr;
    FMX_HIP(ctx, hipMalloc(&g64, (size_t)kHid * kHid * sizeof(double)));
    const uint32_t nblk = 512;
    FMX_HIP(ctx, hipMalloc(&part, (size_t)nblk * kHid * sizeof(double)));
    fmx_status st = FMX_OK;
    const float *Wo = P + layer_off(kLayers);
    float *dWo = G + layer_off(kLayers);
    st = gram_grad(ctx, mode, n, kHid, kOut, Hs[kLayers], dY, g64, dWo, s);
    // dH3 = dY Wo^T  (Wo is 64 x 16 = dout x K)
    if (!st) st = gemm(ctx, mode, dY, kOut, kOut, nullptr, 0, 0, Wo, 1, kHid, nullptr, nullptr, 0, 0, dH, kHid, n, s);
    for (int l = kLayers - 1; l >= 0 && !st; l--) {
        const int d = din_of(l);
        const float *Wl = P + layer_off(l);
        float *Gl = G + layer_off(l);
        hipLaunchKernelGGL(relu_back_kernel, dim3(grid_n(ctx, n * kHid)), dim3(kT), 0, s, Hs[l + 1], dH, n * kHid);
        if (hipGetLastError() != hipSuccess) { st = FMX_EHIP; break; }
        st = gram_grad(ctx, mode, n, d, kHid, Hs[l], dH, g64, Gl, s);
        if (!st) st = gram_grad(ctx, mode, n, d, kHid, Agg[l], dH, g64, Gl + (size_t)d * kHid, s);
        if (!st) {
            const uint64_t rpb = (n + nblk - 1) / nblk;
            hipLaunchKernelGGL(colsum_kernel, dim3(nblk), dim3(kT), 0, s, dH, n, kHid, rpb, part);
            hipLaunchKernelGGL(colsum_finish, dim3(1), dim3(kT), 0, s, part, (int)nblk, kHid,
                               Gl + 2 * (size_t)d * kHid);
            if (hipGetLastError() != hipSuccess) st = FMX_EHIP;
        }
        if (st || l == 0) break;
        // dH_l = dZ Ws^T + N (dZ Wn^T); Ws, Wn are d x 64 = dout x K
        st = gemm(ctx, mode, dH, kHid, kHid, nullptr, 0, 0, Wl + (size_t)d * kHid, 1, d, nullptr, nullptr, 0, 0, T,
                  d, n, s);
        if (!st) st = norm_agg(ctx, spm, dinv, T, d, T2, T3, s);
        if (!st) st = gemm(ctx, mode, dH, kHid, kHid, nullptr, 0, 0, Wl, 1, d, nullptr, T3, d, 0, T, d, n, s);
        if (!st && hipMemcpyAsync(dH, T, n * d * sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess)
            st = FMX_EHIP;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(g64);
    (void)hipFree(part);
    return st ? fail(ctx, st, "gnn backward failed") : FMX_OK;
}

fmx_status fmx_subspace_loss(fmx_ctx *ctx, uint64_t n, int k, const float *Y, const float *U, double *loss,
                             float *dY, void *stream)
{
    if (!ctx || !Y || !U || !loss || k <= 0 || k > 64) return fail(ctx, FMX_EINVAL, "bad argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    double *dG = nullptr, *dC = nullptr, *dM = nullptr;
    FMX_HIP(ctx, hipMalloc(&dG, (size_t)k * k * sizeof(double)));
    FMX_HIP(ctx, hipMalloc(&dC, (size_t)k * k * sizeof(double)));
    FMX_HIP(ctx, hipMalloc(&dM, 2 * (size_t)k * k * sizeof(double)));
    std::vector<double> G(k * k), C(k * k), Gi(k * k), B(k * k), M1(k * k), M2(k * k), T(k * k);
    fmx_status st = gram(ctx, FMX_F32, FMX_GRAM_F64, n, k, k, Y, k, Y, k, dG, s);
    if (!st) st = gram(ctx, FMX_F32, FMX_GRAM_F64, n, k, k, U, k, Y, k, dC, s);
    if (!st && hipMemcpyAsync(G.data(), dG, k * k * 8, hipMemcpyDeviceToHost, s) != hipSuccess) st = FMX_EHIP;
    if (!st && hipMemcpyAsync(C.data(), dC, k * k * 8, hipMemcpyDeviceToHost, s) != hipSuccess) st = FMX_EHIP;
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = FMX_EHIP;
    if (!st) {
        // Gi = G^-1 (Gauss-Jordan, partial pivoting; k x k on the host)
        std::vector<double> M(2 * k * k);
        for (int i = 0; i < k; i++)
            for (int j = 0; j < 2 * k; j++) M[i * 2 * k + j] = j < k ? G[i * k + j] : (j - k == i);
        for (int c = 0; c < k && !st; c++) {
            int p = c;
            for (int r = c + 1; r < k; r++) if (fabs(M[r * 2 * k + c]) > fabs(M[p * 2 * k + c])) p = r;
            if (M[p * 2 * k + c] == 0.0) { st = fail(ctx, FMX_EINVAL, "Y has rank < k"); break; }
            for (int j = 0; j < 2 * k; j++) std::swap(M[c * 2 * k + j], M[p * 2 * k + j]);
            const double dv = M[c * 2 * k + c];
            for (int j = 0; j < 2 * k; j++) M[c * 2 * k + j] /= dv;
            for (int r = 0; r < k; r++) {
                if (r == c) continue;
                const double f = M[r * 2 * k + c];
                if (f != 0.0) for (int j = 0; j < 2 * k; j++) M[r * 2 * k + j] -= f * M[c * 2 * k + j];
            }
        }
        if (!st) {
            for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) Gi[i * k + j] = M[i * 2 * k + k + j];
            for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) {
                double v = 0; for (int t = 0; t < k; t++) v += C[t * k + i] * C[t * k + j];
                B[i * k + j] = v;
            }
            double tr = 0;
            for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) tr += Gi[i * k + j] * B[j * k + i];
            *loss = k - tr;
            for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) {
                double a1 = 0, a2 = 0;
                for (int t = 0; t < k; t++) { a1 += C[i * k + t] * Gi[t * k + j]; a2 += Gi[i * k + t] * B[t * k + j]; }
                M1[i * k + j] = a1;
                T[i * k + j] = a2;
            }
            for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) {
                double v = 0; for (int t = 0; t < k; t++) v += T[i * k + t] * Gi[t * k + j];
                M2[i * k + j] = v;
            }
        }
    }
    if (!st && dY) {
        if (hipMemcpyAsync(dM, M1.data(), k * k * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(dM + k * k, M2.data(), k * k * 8, hipMemcpyHostToDevice, s) != hipSuccess)
            st = FMX_EHIP;
        if (!st) st = gemm_tall(ctx, FMX_F32, n, k, k, -2.0, U, k, dM, 0.0, dY, k, s);
        if (!st) st = gemm_tall(ctx, FMX_F32, n, k, k, 2.0, Y, k, dM + k * k, 1.0, dY, k, s);
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(dG);
    (void)hipFree(dC);
    (void)hipFree(dM);
    return st;
}

fmx_status fmx_adam(fmx_ctx *ctx, uint64_t np, float *params, const float *grads, float *m, float *v, float lr,
                    float beta1, float beta2, float eps, int step, void *stream)
{
    if (!ctx || !params || !grads || !m || !v || step < 1) return fail(ctx, FMX_EINVAL, "bad argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_n(ctx, np)), dim3(kT), 0, (hipStream_t)stream, np, params, grads, m,
                       v, lr, beta1, beta2, eps, c1, c2);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

}  // extern "C"
