// devprim.h -- device-wide primitives used by tile construction:
// stable LSD radix sort of 64-bit keys, exclusive scans and a max reduction.
#pragma once

#include "fmx_internal.h"

namespace fmx {

// Exclusive prefix sum of n values (in may equal out).  Returns the total
// through *total (host) when total != nullptr (synchronises the stream).
template <typename T>
fmx_status scan_exclusive(fmx_ctx *ctx, const T *in, T *out, uint64_t n, T *total,
                          hipStream_t s);

// Stable sort of n keys by bits [0, bits).  `keys` and `tmp` have n entries;
// on return *sorted points at whichever buffer holds the result.
fmx_status radix_sort_u64(fmx_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, int bits,
                          uint64_t **sorted, hipStream_t s);

// The same with a u32 payload per key (vals, vtmp: n entries each; NULL vals
// sorts keys only); *sorted_vals points at the payload in key order.
fmx_status radix_sort_u64_kv(fmx_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint32_t *vals, uint32_t *vtmp,
                             uint64_t n, int bits, uint64_t **sorted, uint32_t **sorted_vals, hipStream_t s);

// max over a[0..n) and b[0..n) (n > 0), returned on the host.
fmx_status max_u32_pair(fmx_ctx *ctx, const uint32_t *a, const uint32_t *b, uint64_t n,
                        uint32_t *out, hipStream_t s);

// RAII device buffer.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(uint64_t n) { return hipMalloc(&p, (n ? n : 1) * sizeof(T)); }
    T *release() { T *q = p; p = nullptr; return q; }
};

}  // namespace fmx
