// gen.hip -- seeded graph generators on the device.
//
// R-MAT: the project's counter-based specification (SplitMix64 output
// function, Steele/Lea/Flood 2014), one thread per edge.  The reference
// generator (flash-graph/tools/rmat-gen.cpp:43-48) draws from a time-seeded
// Boost rmat_iterator, so only the distribution is shared with it:
// quadrant probabilities (a, b, c, d) per level and a vertex relabelling.
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

}  // namespace
}  // namespace fmx

using namespace fmx;

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
