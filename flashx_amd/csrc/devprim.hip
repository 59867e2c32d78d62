// devprim.hip -- hand-written device primitives for tile construction.
//
//   * scan_exclusive: tile reduce -> recursive scan of tile sums -> tile scan
//     (4096 values per 256-thread block, wave scans by DPP-free shuffles);
//   * radix_sort_u64: stable LSD radix sort, 8-bit digits; per pass a digit
//     histogram per 4096-key tile, an exclusive scan of the digit-major
//     histogram, and a scatter that ranks equal digits inside each wave with
//     ballots (so equal keys keep their order: the sort is stable);
//   * max_u32_pair: max of two id arrays.
#include "devprim.h"

namespace fmx {
namespace {

constexpr int kT = 256;          // threads per block
constexpr int kI = 16;           // items per thread
constexpr int kTile = kT * kI;   // 4096

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, int d)
{
    if constexpr (sizeof(T) == 8) {
        uint32_t lo = (uint32_t)v, hi = (uint32_t)((uint64_t)v >> 32);
        lo = __shfl_up(lo, d, 64);
        hi = __shfl_up(hi, d, 64);
        return (T)(((uint64_t)hi << 32) | lo);
    } else {
        return (T)__shfl_up((uint32_t)v, d, 64);
    }
}

// Exclusive scan of one value per thread over the block; returns the block total.
template <typename T>
__device__ __forceinline__ T block_exclusive(T v, T *lds_waves, T &total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = shfl_up_t(inc, d);
        if (lane >= d) inc += o;
    }
    if (lane == 63) lds_waves[w] = inc;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kT / 64; i++) {
        T x = lds_waves[i];
        if (i < w) off += x;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

template <typename T>
__global__ void __launch_bounds__(kT) scan_reduce(const T *in, uint64_t n, T *partials)
{
    __shared__ T wl[kT / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    T s = 0;
#pragma unroll
    for (int i = 0; i < kI; i++) {
        uint64_t idx = base + (uint64_t)i * kT + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    T tot;
    (void)block_exclusive(s, wl, tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

template <typename T>
__global__ void __launch_bounds__(kT) scan_tile(const T *in, T *out, uint64_t n, const T *carries,
                                                 T *d_total)
{
    __shared__ T wl[kT / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kI;
    T v[kI];
    T s = 0;
#pragma unroll
    for (int i = 0; i < kI; i++) {
        uint64_t idx = base + i;
        v[i] = idx < n ? in[idx] : T(0);
        s += v[i];
    }
    T tot;
    T run = block_exclusive(s, wl, tot);
    if (carries) run += carries[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kI; i++) {
        uint64_t idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
    if (d_total && threadIdx.x == 0) *d_total = tot;
}

template <typename T>
fmx_status scan_rec(fmx_ctx *ctx, const T *in, T *out, uint64_t n, T *d_total, hipStream_t s)
{
    const uint64_t nb = (n + kTile - 1) / kTile;
    if (nb <= 1) {
        hipLaunchKernelGGL(scan_tile<T>, dim3(1), dim3(kT), 0, s, in, out, n, (const T *)nullptr,
                           d_total);
        FMX_HIP(ctx, hipGetLastError());
        return FMX_OK;
    }
    DevBuf<T> partials;
    FMX_HIP(ctx, partials.alloc(nb));
    hipLaunchKernelGGL(scan_reduce<T>, dim3((uint32_t)nb), dim3(kT), 0, s, in, n, partials.p);
    FMX_HIP(ctx, hipGetLastError());
    fmx_status st = scan_rec<T>(ctx, partials.p, partials.p, nb, d_total, s);
    if (st) return st;
    hipLaunchKernelGGL(scan_tile<T>, dim3((uint32_t)nb), dim3(kT), 0, s, in, out, n,
                       (const T *)partials.p, (T *)nullptr);
    FMX_HIP(ctx, hipGetLastError());
    // partials must outlive the kernels that read them
    FMX_HIP(ctx, hipStreamSynchronize(s));
    return FMX_OK;
}

// ---- radix sort -------------------------------------------------------------
constexpr int kWaves = kT / 64;
constexpr int kChunks = kTile / kT;  // 64-key chunks per wave

__global__ void __launch_bounds__(kT) rs_count(const uint64_t *keys, uint64_t n, int shift,
                                               uint64_t *hist, uint32_t nb)
{
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
#pragma unroll 4
    for (int i = 0; i < kI; i++) {
        uint64_t idx = base + (uint64_t)i * kT + threadIdx.x;
        if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

// Scatter of one pass; the optional u32 payload (vin -> vout) moves with its
// key.  The tile's keys are first ranked into digit order in LDS (equal
// digits ranked inside each wave with ballots, waves and 64-key chunks in
// order, so equal keys keep their order: the sort is stable), then written
// out in that order, so consecutive threads write consecutive slots of each
// digit's run in the output (coalesced).
template <bool KV>
__global__ void __launch_bounds__(kT) rs_scatter_t(const uint64_t *in, uint64_t *out, const uint32_t *vin,
                                                   uint32_t *vout, uint64_t n, int shift, const uint64_t *offs,
                                                   uint32_t nb)
{
    __shared__ uint32_t wcnt[kWaves][256];   // keys of each digit per wave
    __shared__ uint32_t lbase[kWaves][256];  // next tile slot of each digit per wave
    __shared__ uint32_t dstart[256];         // first tile slot of each digit
    __shared__ uint64_t gstart[256];         // first output slot of this tile's keys of each digit
    __shared__ uint32_t wl[kWaves];
    __shared__ uint64_t sk[kTile];           // the tile's keys in digit order
    __shared__ uint32_t sv[KV ? kTile : 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * 256; i += kT) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    // wave w owns keys [w*1024, (w+1)*1024) of the tile, 64 at a time, in order
    const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)w * (kChunks * 64);
    uint64_t k[kChunks];
    uint32_t pv[KV ? kChunks : 1];
#pragma unroll
    for (int c = 0; c < kChunks; c++) {
        uint64_t idx = base + (uint64_t)c * 64 + lane;
        k[c] = idx < n ? in[idx] : 0;
        if constexpr (KV) pv[c] = idx < n ? vin[idx] : 0u;
        if (idx < n) atomicAdd(&wcnt[w][(k[c] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    {
        const int d = threadIdx.x;  // one digit per thread
        uint32_t tot = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ww++) tot += wcnt[ww][d];
        uint32_t all;
        uint32_t run = block_exclusive(tot, wl, all);
        dstart[d] = run;
#pragma unroll
        for (int ww = 0; ww < kWaves; ww++) {
            lbase[ww][d] = run;
            run += wcnt[ww][d];
        }
        gstart[d] = offs[(uint64_t)d * nb + blockIdx.x];
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int c = 0; c < kChunks; c++) {
        const uint64_t idx = base + (uint64_t)c * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = (uint32_t)(k[c] >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bm = __ballot(valid && bit);
            peers &= bit ? bm : ~bm;
        }
        const uint32_t lpos = lbase[w][d] + __popcll(peers & lt);
        if (valid) {
            sk[lpos] = k[c];
            if constexpr (KV) sv[lpos] = pv[c];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers & lt) == 0) lbase[w][d] += __popcll(peers);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
    const uint32_t nt = (uint32_t)min((uint64_t)kTile, n - t0);
    for (uint32_t j = threadIdx.x; j < nt; j += kT) {
        const uint64_t key = sk[j];
        const uint32_t d = (uint32_t)(key >> shift) & 0xFFu;
        const uint64_t pos = gstart[d] + (j - dstart[d]);
        out[pos] = key;
        if constexpr (KV) vout[pos] = sv[j];
    }
}

__global__ void __launch_bounds__(kT) max_pair_kernel(const uint32_t *a, const uint32_t *b,
                                                      uint64_t n, uint32_t *out)
{
    uint32_t m = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kT)
        m = max(m, max(a[i], b[i]));
    for (int d = 32; d; d >>= 1) m = max(m, (uint32_t)__shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

}  // namespace

template <typename T>
fmx_status scan_exclusive(fmx_ctx *ctx, const T *in, T *out, uint64_t n, T *total, hipStream_t s)
{
    DevBuf<T> d_total;
    FMX_HIP(ctx, d_total.alloc(1));
    if (n == 0) {
        if (total) *total = 0;
        return FMX_OK;
    }
    fmx_status st = scan_rec<T>(ctx, in, out, n, d_total.p, s);
    if (st) return st;
    if (total) {
        FMX_HIP(ctx, hipMemcpyAsync(total, d_total.p, sizeof(T), hipMemcpyDeviceToHost, s));
    }
    FMX_HIP(ctx, hipStreamSynchronize(s));
    return FMX_OK;
}

template fmx_status scan_exclusive<uint32_t>(fmx_ctx *, const uint32_t *, uint32_t *, uint64_t,
                                             uint32_t *, hipStream_t);
template fmx_status scan_exclusive<uint64_t>(fmx_ctx *, const uint64_t *, uint64_t *, uint64_t,
                                             uint64_t *, hipStream_t);

fmx_status scan_u64_exclusive(fmx_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total,
                              hipStream_t s)
{
    return scan_exclusive<uint64_t>(ctx, in, out, n, total, s);
}

fmx_status radix_sort_u64(fmx_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, int bits,
                          uint64_t **sorted, hipStream_t s)
{
    uint32_t *dummy = nullptr;
    return radix_sort_u64_kv(ctx, keys, tmp, nullptr, nullptr, n, bits, sorted, &dummy, s);
}

fmx_status radix_sort_u64_kv(fmx_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint32_t *vals, uint32_t *vtmp,
                             uint64_t n, int bits, uint64_t **sorted, uint32_t **sorted_vals, hipStream_t s)
{
    *sorted = keys;
    *sorted_vals = vals;
    if (n < 2 || bits <= 0) return FMX_OK;
    const uint64_t nb64 = (n + kTile - 1) / kTile;
    if (nb64 >= (1ull << 32) / 256) return fail(ctx, FMX_EUNSUPPORTED, "too many keys to sort");
    const uint32_t nb = (uint32_t)nb64;
    DevBuf<uint64_t> hist;
    FMX_HIP(ctx, hist.alloc((uint64_t)nb * 256));
    uint64_t *a = keys, *b = tmp;
    uint32_t *va = vals, *vb = vtmp;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(rs_count, dim3(nb), dim3(kT), 0, s, a, n, shift, hist.p, nb);
        FMX_HIP(ctx, hipGetLastError());
        fmx_status st = scan_exclusive<uint64_t>(ctx, hist.p, hist.p, (uint64_t)nb * 256, nullptr, s);
        if (st) return st;
        if (vals)
            hipLaunchKernelGGL(rs_scatter_t<true>, dim3(nb), dim3(kT), 0, s, a, b, va, vb, n, shift, hist.p, nb);
        else
            hipLaunchKernelGGL(rs_scatter_t<false>, dim3(nb), dim3(kT), 0, s, a, b, nullptr, nullptr, n, shift,
                               hist.p, nb);
        FMX_HIP(ctx, hipGetLastError());
        uint64_t *t = a;
        a = b;
        b = t;
        uint32_t *vt = va;
        va = vb;
        vb = vt;
    }
    FMX_HIP(ctx, hipStreamSynchronize(s));
    *sorted = a;
    *sorted_vals = va;
    return FMX_OK;
}

fmx_status max_u32_pair(fmx_ctx *ctx, const uint32_t *a, const uint32_t *b, uint64_t n,
                        uint32_t *out, hipStream_t s)
{
    DevBuf<uint32_t> d;
    FMX_HIP(ctx, d.alloc(1));
    FMX_HIP(ctx, hipMemsetAsync(d.p, 0, sizeof(uint32_t), s));
    uint64_t blocks = (n + kT - 1) / kT;
    if (blocks > (uint64_t)ctx->num_cus * 8) blocks = (uint64_t)ctx->num_cus * 8;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(max_pair_kernel, dim3((uint32_t)blocks), dim3(kT), 0, s, a, b, n, d.p);
    FMX_HIP(ctx, hipGetLastError());
    FMX_HIP(ctx, hipMemcpyAsync(out, d.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FMX_HIP(ctx, hipStreamSynchronize(s));
    return FMX_OK;
}

}  // namespace fmx
