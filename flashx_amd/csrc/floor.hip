// floor.hip -- the memory-system floor of the SpMM's dense-row gathers.
//
// fmx_gather_floor reads the rows of X named by a column sequence (the
// order in which an SpMM visits them) with the SpMM's load shape -- 16 bytes
// per lane, the lanes of one row side by side, eight loads in flight per
// lane -- and does nothing else.  Its time is what any kernel that fetches
// the same rows in the same order needs for the fetches alone; the SpMM's
// time over it is the cost of decoding the tiles and summing.
#include "dense.h"
#include "fmx_internal.h"

namespace fmx {
namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kFloorT = 256;
constexpr uint32_t kPerWave = 2048;

template <int L>
__global__ void __launch_bounds__(kFloorT) gather_floor(const unsigned char *X, uint64_t ldx_bytes,
                                                        uint64_t x_rows, const uint32_t *cols, uint64_t n,
                                                        uint32_t *sink)
{
    constexpr int NZP = 64 / L;  // rows per load instruction
    const uint32_t lane = threadIdx.x & 63, q = lane % L, g = lane / L;
    const uint64_t wave = (blockIdx.x * (uint64_t)kFloorT + threadIdx.x) >> 6;
    const uint64_t b0 = wave * kPerWave;
    if (b0 >= n) return;
    const uint64_t last = x_rows - 1;
    u32x4 acc = {0, 0, 0, 0};
    for (uint32_t i = 0; i < kPerWave; i += NZP * 8) {
        u32x4 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            uint64_t e = b0 + i + u * NZP + g;
            e = e < n ? e : n - 1;
            uint64_t c = cols[e];
            c = c < last ? c : last;
            v[u] = *(const u32x4 *)(X + c * ldx_bytes + q * 16);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc ^= v[u];
    }
    const uint32_t x = acc.x ^ acc.y ^ acc.z ^ acc.w;
    if (x == 0x9E3779B9u) *sink = x;  // keeps the loads; harmless when taken
}

}  // namespace
}  // namespace fmx

using namespace fmx;

extern "C" {

fmx_status fmx_gather_floor(fmx_ctx *ctx, uint64_t row_bytes, const void *X, int64_t ldx_bytes, uint64_t x_rows,
                            const uint32_t *cols, uint64_t n, void *stream)
{
    if (!ctx || (n && (!X || !cols))) return fail(ctx, FMX_EINVAL, "null argument");
    const uint64_t L = row_bytes / 16;
    if (row_bytes % 16 || L < 1 || L > 64 || (64 % L))
        return fail(ctx, FMX_EINVAL, "row_bytes must be 16 times a power of two, at most 1024");
    if ((uintptr_t)X % 16 || ldx_bytes < (int64_t)row_bytes || ldx_bytes % 16)
        return fail(ctx, FMX_EINVAL, "X rows must be 16-byte aligned");
    if (n && x_rows == 0) return fail(ctx, FMX_EDIM, "X has no rows");
    if (n == 0) return FMX_OK;
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    void *ws;
    fmx_status st = ctx_scratch(ctx, 16, &ws);
    if (st) return st;
    const uint64_t waves = (n + kPerWave - 1) / kPerWave;
    const dim3 grid((uint32_t)((waves * 64 + kFloorT - 1) / kFloorT));
    hipStream_t s = (hipStream_t)stream;
    const unsigned char *x = (const unsigned char *)X;
    uint32_t *sink = (uint32_t *)ws;
    switch (L) {
    case 1: hipLaunchKernelGGL(gather_floor<1>, grid, dim3(kFloorT), 0, s, x, ldx_bytes, x_rows, cols, n, sink); break;
    case 2: hipLaunchKernelGGL(gather_floor<2>, grid, dim3(kFloorT), 0, s, x, ldx_bytes, x_rows, cols, n, sink); break;
    case 4: hipLaunchKernelGGL(gather_floor<4>, grid, dim3(kFloorT), 0, s, x, ldx_bytes, x_rows, cols, n, sink); break;
    case 8: hipLaunchKernelGGL(gather_floor<8>, grid, dim3(kFloorT), 0, s, x, ldx_bytes, x_rows, cols, n, sink); break;
    case 16: hipLaunchKernelGGL(gather_floor<16>, grid, dim3(kFloorT), 0, s, x, ldx_bytes, x_rows, cols, n, sink); break;
    case 32: hipLaunchKernelGGL(gather_floor<32>, grid, dim3(kFloorT), 0, s, x, ldx_bytes, x_rows, cols, n, sink); break;
    default: hipLaunchKernelGGL(gather_floor<64>, grid, dim3(kFloorT), 0, s, x, ldx_bytes, x_rows, cols, n, sink); break;
    }
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

}  // extern "C"
