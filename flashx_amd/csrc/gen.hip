// gen.hip -- seeded graph generators on the device.
//
// R-MAT: the project's counter-based specification (SplitMix64 output
// function, Steele/Lea/Flood 2014), one thread per edge.  The reference
// generator (flash-graph/tools/rmat-gen.cpp:43-48) draws from a time-seeded
// Boost rmat_iterator, so only the distribution is shared with it:
// quadrant probabilities (a, b, c, d) per level and a vertex relabelling.
#include <algorithm>
#include <cmath>

#include "fmx_internal.h"

namespace fmx {
namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ULL;
constexpr uint64_t kPermSalt = 0x7065726D75746531ULL;  // "permute1"

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t key, uint64_t ctr)
{
    uint64_t z = key + (ctr + 1) * kGamma;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct PermKeys {
    uint64_t add[3];
    uint64_t mul[3];
};

__device__ __forceinline__ uint32_t perm(uint32_t v, int scale, const PermKeys &pk)
{
    const uint64_t mask = (scale >= 64) ? ~0ULL : ((1ULL << scale) - 1);
    const int sh = (scale + 1) / 2;
    uint64_t x = v;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        x = (x + pk.add[i]) & mask;
        x = (x * pk.mul[i]) & mask;
        x ^= x >> sh;
    }
    return (uint32_t)x;
}

__global__ void __launch_bounds__(256) rmat_kernel(int scale, uint64_t m, uint64_t seed,
                                                    uint32_t ta, uint32_t tab, uint32_t tabc,
                                                    int permute, PermKeys pk,
                                                    uint32_t *src, uint32_t *dst)
{
    const uint64_t ndraw = (uint64_t)(scale + 1) / 2;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t s = 0, d = 0;
        for (uint64_t j = 0; j < ndraw; j++) {
            const uint64_t r = splitmix64(seed, e * ndraw + j);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int l = (int)(2 * j) + h;
                if (l < scale) {
                    const uint32_t u = h ? (uint32_t)(r >> 32) : (uint32_t)r;
                    const uint32_t q = (u < ta) ? 0u : (u < tab) ? 1u : (u < tabc) ? 2u : 3u;
                    s |= (q >> 1) << (scale - 1 - l);
                    d |= (q & 1u) << (scale - 1 - l);
                }
            }
        }
        if (permute) {
            s = perm(s, scale, pk);
            d = perm(d, scale, pk);
        }
        src[e] = s;
        dst[e] = d;
    }
}

// ---- degree-corrected SBM ----------------------------------------------------
constexpr uint64_t kSbmSaltW = 0x73626D5745494748ULL;
constexpr uint64_t kSbmSaltE = 0x73626D4544474553ULL;

__global__ void __launch_bounds__(256) sbm_weights(uint64_t n, uint64_t seed, const uint32_t *wtab,
                                                    uint64_t *w)
{
    for (uint64_t v = blockIdx.x * 256ull + threadIdx.x; v < n; v += (uint64_t)gridDim.x * 256)
        w[v] = wtab[splitmix64(seed ^ kSbmSaltW, v) >> 54];
    if (blockIdx.x == 0 && threadIdx.x == 0) w[n] = 0;
}

__device__ __forceinline__ uint64_t sbm_find(const uint64_t *C, uint64_t lo, uint64_t hi, uint64_t t)
{
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (C[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256) sbm_edges(uint64_t n, int scale, int nblocks, const uint32_t *bstart,
                                                  uint32_t ratio, uint64_t m, uint64_t seed, PermKeys pk,
                                                  const uint64_t *C, uint32_t *src, uint32_t *dst)
{
    __shared__ uint32_t bs[257];
    for (int i = threadIdx.x; i <= nblocks; i += 256) bs[i] = bstart[i];
    __syncthreads();
    const uint64_t vtot = C[n];
    for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < m; e += (uint64_t)gridDim.x * 256) {
        const uint64_t r1 = splitmix64(seed ^ kSbmSaltE, 3 * e);
        const uint64_t r2 = splitmix64(seed ^ kSbmSaltE, 3 * e + 1);
        const uint64_t r3 = splitmix64(seed ^ kSbmSaltE, 3 * e + 2);
        const uint64_t u = sbm_find(C, 0, n, r1 % vtot);
        int lo = 0, hi = nblocks;
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (bs[mid] <= u) lo = mid;
            else hi = mid;
        }
        const uint64_t c0 = C[bs[lo]], vbf = C[bs[lo + 1]] - c0;
        const uint64_t vb = vbf >> 8, vt = vtot >> 8;
        const uint64_t num = (uint64_t)ratio * vb, den = num + vt - vb;
        uint64_t v;
        if ((r2 >> 40) * den < (num << 24)) {
            v = sbm_find(C, bs[lo], bs[lo + 1], c0 + r3 % vbf);
        } else {
            uint64_t t = r3 % (vtot - vbf);
            if (t >= c0) t += vbf;
            v = sbm_find(C, 0, n, t);
        }
        src[e] = perm((uint32_t)u, scale, pk);
        dst[e] = perm((uint32_t)v, scale, pk);
    }
}

}  // namespace

fmx_status scan_u64_exclusive(fmx_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total,
                              hipStream_t s);

}  // namespace fmx

using namespace fmx;

extern "C" fmx_status fmx_dcsbm_params(uint64_t n, int nblocks, double gamma, double size_decay,
                                       uint32_t *bstart, uint32_t *wtab)
{
    if (!bstart || !wtab || nblocks < 1 || nblocks > 256 || (uint64_t)nblocks > n || gamma <= 1.0 ||
        size_decay <= 0.0)
        return FMX_EINVAL;
    // block b holds a share proportional to size_decay^-b of the vertices
    double tot = 0.0;
    for (int b = 0; b < nblocks; b++) tot += pow(size_decay, -b);
    uint64_t acc = 0;
    bstart[0] = 0;
    for (int b = 0; b < nblocks; b++) {
        uint64_t sz = (uint64_t)floor((double)n * pow(size_decay, -b) / tot);
        if (sz < 1) sz = 1;
        acc += sz;
        bstart[b + 1] = (uint32_t)(b + 1 == nblocks ? n : (acc < n ? acc : n));
    }
    // weights: quantiles of theta = (1-u)^(-1/(gamma-1)), theta >= 1, in 1/1024 units
    const double cap = sqrt((double)n);
    for (int i = 0; i < 1024; i++) {
        double th = pow(1.0 - (i + 0.5) / 1024.0, -1.0 / (gamma - 1.0));
        if (th > cap) th = cap;
        wtab[i] = (uint32_t)llround(th * 1024.0);
    }
    return FMX_OK;
}

extern "C" fmx_status fmx_dcsbm_edges(fmx_ctx *ctx, uint64_t n, int nblocks, const uint32_t *bstart,
                                      const uint32_t *wtab, uint32_t ratio, uint64_t m, uint64_t seed,
                                      uint32_t *src, uint32_t *dst, void *stream)
{
    if (!ctx || !bstart || !wtab) return FMX_EINVAL;
    int scale = 0;
    while ((1ull << scale) < n) scale++;
    if (n < 2 || (1ull << scale) != n || scale > 32) return fail(ctx, FMX_EINVAL, "n must be a power of two");
    if (nblocks < 1 || nblocks > 256 || bstart[0] != 0 || bstart[nblocks] != n)
        return fail(ctx, FMX_EINVAL, "bad block boundaries");
    for (int b = 0; b < nblocks; b++)
        if (bstart[b + 1] <= bstart[b]) return fail(ctx, FMX_EINVAL, "empty block");
    if (m && (!src || !dst)) return fail(ctx, FMX_EINVAL, "null edge buffer");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    uint64_t *C = nullptr;
    uint32_t *dwt = nullptr, *dbs = nullptr;
    FMX_HIP(ctx, hipMalloc(&C, (n + 1) * sizeof(uint64_t)));
    FMX_HIP(ctx, hipMalloc(&dwt, 1024 * sizeof(uint32_t)));
    FMX_HIP(ctx, hipMalloc(&dbs, (nblocks + 1) * sizeof(uint32_t)));
    FMX_HIP(ctx, hipMemcpyAsync(dwt, wtab, 1024 * 4, hipMemcpyHostToDevice, s));
    FMX_HIP(ctx, hipMemcpyAsync(dbs, bstart, (nblocks + 1) * 4, hipMemcpyHostToDevice, s));
    PermKeys pk;
    for (int i = 0; i < 3; i++) {
        pk.add[i] = splitmix64(seed ^ kPermSalt, 2 * i);
        pk.mul[i] = splitmix64(seed ^ kPermSalt, 2 * i + 1) | 1ULL;
    }
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)ctx->num_cus * 16);
    hipLaunchKernelGGL(sbm_weights, dim3(grid), dim3(256), 0, s, n, seed, dwt, C);
    fmx_status st = scan_u64_exclusive(ctx, C, C, n + 1, nullptr, s);
    if (st == FMX_OK && m) {
        const uint32_t g2 = (uint32_t)std::min<uint64_t>((m + 255) / 256, (uint64_t)ctx->num_cus * 16);
        hipLaunchKernelGGL(sbm_edges, dim3(g2), dim3(256), 0, s, n, scale, nblocks, dbs, ratio, m, seed, pk, C,
                           src, dst);
        if (hipGetLastError() != hipSuccess) st = fail(ctx, FMX_EHIP, "sbm_edges launch");
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(C);
    (void)hipFree(dwt);
    (void)hipFree(dbs);
    return st;
}

extern "C" fmx_status fmx_rmat_edges(fmx_ctx *ctx, int scale, uint64_t m, uint64_t seed,
                                     uint32_t thr_a, uint32_t thr_ab, uint32_t thr_abc,
                                     int permute, uint32_t *src, uint32_t *dst, void *stream)
{
    if (!ctx) return FMX_EINVAL;
    if (scale < 1 || scale > 32) return fail(ctx, FMX_EINVAL, "scale must be in [1, 32]");
    if (m && (!src || !dst)) return fail(ctx, FMX_EINVAL, "null edge buffer");
    if (!(thr_a <= thr_ab && thr_ab <= thr_abc)) return fail(ctx, FMX_EINVAL, "thresholds must be cumulative");
    if (m == 0) return FMX_OK;
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    PermKeys pk;
    for (int i = 0; i < 3; i++) {
        pk.add[i] = splitmix64(seed ^ kPermSalt, 2 * i);
        pk.mul[i] = splitmix64(seed ^ kPermSalt, 2 * i + 1) | 1ULL;
    }
    hipStream_t s = (hipStream_t)stream;
    uint64_t blocks = (m + 255) / 256;
    if (blocks > (uint64_t)ctx->num_cus * 64) blocks = (uint64_t)ctx->num_cus * 64;
    hipLaunchKernelGGL(rmat_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, scale, m, seed, thr_a,
                       thr_ab, thr_abc, permute, pk, src, dst);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}
