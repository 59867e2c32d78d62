// spmm.hip -- gfx950 SpMM over the reference's 2D tile image.
//
// Reference semantics: block_spmm_task_impl::run_on_block with row_part_func
// and coo_func (matrix/sparse_matrix.h:219-375): for every non-zero (r, c) of
// a tile, Y[r][0..k) += X[c][0..k) (binary matrix, value 1).
//
// Execution model (one wave = one work item, no cross-wave sharing):
//   * an item is a contiguous word range of one block row (tile headers and
//     all), plus a band of <= Hs output rows kept in LDS;
//   * the wave streams the range 64*U words at a time with coalesced u16
//     loads, classifies every word (tile header / row head / column) with
//     ballots against a 64-entry window of the tile directory, resolves each
//     column's row with a ballot prefix (the nearest head to its left, or the
//     row carried over), and appends (band row, global column) to an LDS queue;
//   * the queue is drained with vector gathers of X rows (16 bytes per lane,
//     L lanes per non-zero) issued P at a time, accumulated with LDS adds into
//     the band;
//   * the band is written to Y once (rows without non-zeros are 0), or to a
//     partial slot when a power-law band was split over several items; split
//     bands are summed slot by slot, in stream order, by spmm_reduce.
// Each band is owned by exactly one wave, so results do not depend on timing.
// FP32 data accumulate in an FP64 band and are rounded once on write-out.
#include <cstdlib>

#include "fmx_internal.h"

namespace fmx {
namespace {

constexpr int kWPB = 4;  // waves per workgroup (independent items)

struct SpmmArgs {
    const uint16_t *img;
    const uint2 *dir;
    const Item *items;
    const void *X;
    void *Y;
    void *partial;
    int64_t ldx;
    int64_t ldy;
    uint32_t nitems;
    uint32_t hs;
    uint32_t wave_bytes;
    uint32_t band_bytes;
    uint32_t band_stride;
    int log2w;
    int k;
    uint64_t ncols;
};

template <typename T, int E>
struct alignas(sizeof(T) * E) Vec {
    T v[E];
};

__device__ __forceinline__ uint64_t mask_le(int lane) { return (2ull << lane) - 1ull; }
__device__ __forceinline__ uint64_t mask_lt(int lane) { return (1ull << lane) - 1ull; }

__device__ __forceinline__ uint32_t rdlane(uint32_t v, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

__device__ __forceinline__ uint32_t uni(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

template <typename T>
__device__ __forceinline__ void lds_add(T *p, T v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// LDS accumulator type: FP32 data accumulate in FP64 (ds_add_f64 is fast on
// gfx950 where ds_add_f32 is not, and the sum is then rounded once).
template <typename T> struct AccOf { using type = T; };
template <> struct AccOf<float> { using type = double; };

template <typename T, int K>
struct Lanes {
    static constexpr int S = sizeof(T);
    static constexpr int VB = (K * S < 16) ? K * S : 16;  // bytes per lane
    static constexpr int E = VB / S;                       // elements per lane
    static constexpr int L = K / E;                        // lanes per non-zero
    static constexpr int NZP = 64 / L;                     // non-zeros per pass
};

// Drains queue entries [0, qn) (qn >= 1) into the band.  K > 0: compile-time
// width.  Loads are unconditional (indices clamped to the last entry) so the
// P gathers of a batch are all in flight before the first LDS add.
template <typename T, int K, int P, typename A>
__device__ __forceinline__ void drain(const SpmmArgs &a, A *band, uint32_t S, const uint32_t *qcol,
                                      const uint16_t *qrow, uint32_t qn, int lane)
{
    using C = Lanes<T, K>;
    const T *X = (const T *)a.X;
    const int q = lane % C::L;
    const int g = lane / C::L;
    const uint32_t last = qn - 1;
    for (uint32_t base = 0; base < qn; base += C::NZP * P) {
        Vec<T, C::E> v[P];
        uint32_t rr[P];
#pragma unroll
        for (int pp = 0; pp < P; pp++) {
            const uint32_t e = base + pp * C::NZP + g;
            const uint32_t ec = e < last ? e : last;
            const uint32_t gc = qcol[ec];
            rr[pp] = e <= last ? (uint32_t)qrow[ec] : 0xFFFFFFFFu;
            v[pp] = *(const Vec<T, C::E> *)(X + (uint64_t)gc * a.ldx + q * C::E);
        }
#pragma unroll
        for (int pp = 0; pp < P; pp++) {
            if (rr[pp] != 0xFFFFFFFFu) {
                A *dst = band + rr[pp] * S + q * C::E;
#pragma unroll
                for (int i = 0; i < C::E; i++) lds_add(dst + i, (A)v[pp].v[i]);
            }
        }
    }
}

// Runtime width: flat (entry, element) pairs, one element per lane.
template <typename T, int P, typename A>
__device__ __forceinline__ void drain_generic(const SpmmArgs &a, A *band, uint32_t S, const uint32_t *qcol,
                                              const uint16_t *qrow, uint32_t qn, int lane)
{
    const T *X = (const T *)a.X;
    const uint32_t k = (uint32_t)a.k;
    const uint32_t tot = qn * k;
    for (uint32_t base = 0; base < tot; base += 64 * P) {
        T v[P];
        uint32_t dst[P];
#pragma unroll
        for (int pp = 0; pp < P; pp++) {
            const uint32_t f = base + pp * 64 + lane;
            const uint32_t fc = f < tot ? f : tot - 1;
            const uint32_t e = fc / k, j = fc - e * k;
            v[pp] = X[(uint64_t)qcol[e] * a.ldx + j];
            dst[pp] = f < tot ? (uint32_t)qrow[e] * S + j : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int pp = 0; pp < P; pp++)
            if (dst[pp] != 0xFFFFFFFFu) lds_add(band + dst[pp], (A)v[pp]);
    }
}

template <typename T, int K, int U, int P>
__global__ void __launch_bounds__(64 * kWPB) spmm_stream(SpmmArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int QCAP = 64 * U;
    const int lane = threadIdx.x & 63;
    const uint32_t wv = uni(threadIdx.x >> 6);
    const uint32_t item_id = blockIdx.x * kWPB + wv;
    if (item_id >= a.nitems) return;
    const uint32_t k = K ? K : (uint32_t)a.k;

    unsigned char *wbase = smem + wv * a.wave_bytes;
    using A = typename AccOf<T>::type;
    A *band = (A *)wbase;
    uint32_t *qcol = (uint32_t *)(wbase + a.band_bytes);
    uint16_t *qrow = (uint16_t *)(qcol + QCAP);

    const Item &it = a.items[item_id];
    const uint64_t brow_word = it.brow_word;
    const uint32_t begin = it.begin, nwords = it.nwords;
    const uint32_t dir_end = it.dir_end;
    const uint64_t row0 = ((uint64_t)it.row0_hi << 32) | it.row0_lo;
    const uint32_t sb_base = it.sb_base, rows = it.rows;
    const uint32_t slot = it.slot;
    int cur_row = (it.carry == kNone) ? -1 : (int)it.carry;
    uint32_t ct = it.dir_begin;

    const uint32_t S = a.band_stride;  // padded band row (elements)
    for (uint32_t i = lane; i < rows * S; i += 64) band[i] = A(0);

    const uint16_t *wp = a.img + brow_word;
    const uint32_t pend = begin + nwords;  // nwords >= 1
    const uint32_t plast = pend - 1;
    uint2 cur = a.dir[ct];
    cur.x = uni(cur.x);
    cur.y = uni(cur.y);
    const uint32_t dlast = dir_end - 1;
    // directory window (lane j holds tile ct+1+j) and words of the first chunk
    uint2 dj;
    uint32_t wrd[U];
    {
        const uint32_t wj = ct + 1 + lane;
        dj = a.dir[wj < dlast ? wj : dlast];
        if (wj > dlast) dj = make_uint2(0xFFFFFFFFu, 0u);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t pos = begin + u * 64 + lane;
            wrd[u] = wp[pos < plast ? pos : plast];
        }
    }
    uint32_t p = begin;
    while (p < pend) {
        uint32_t qn = 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t p0 = p + u * 64;
            if (p0 >= pend) break;
            const uint32_t pos = p0 + lane;
            const bool valid = pos < pend;
            // tiles starting inside (p0, p0+64)
            const int A = __popcll(__ballot(dj.x <= p0));
            const int A2 = __popcll(__ballot(dj.x < p0 + 64));
            uint64_t B = 0;
            for (int j = A; j < A2; j++) B |= 1ull << (rdlane(dj.x, j) - p0);
            uint32_t bstart = cur.x, bbc = cur.y;
            if (A) { bstart = rdlane(dj.x, A - 1); bbc = rdlane(dj.y, A - 1); }
            const int c = __popcll(B & mask_le(lane));
            const int src = c ? (A + c - 1) : 0;
            const uint32_t sx = __shfl(dj.x, src), sy = __shfl(dj.y, src);
            const uint32_t tstart = c ? sx : bstart, tbc = c ? sy : bbc;
            const uint32_t w = wrd[u];
            const bool hdr = (pos - tstart) < 8u;
            const bool head = valid && !hdr && (w & 0x8000u);
            const bool col = valid && !hdr && !(w & 0x8000u);
            const uint64_t hm = __ballot(head);
            const uint64_t m = hm & mask_le(lane);
            const int hl = m ? 63 - __clzll(m) : 0;
            const int hrow = __shfl((int)(w & 0x7FFFu), hl);
            const int row = m ? hrow : cur_row;
            if (hm) cur_row = (int)rdlane(w & 0x7FFFu, 63 - __clzll(hm));
            const uint32_t brow = (uint32_t)(row - (int)sb_base);
            const uint32_t gcol = (tbc << a.log2w) | w;
            const bool keep = col && row >= 0 && brow < rows && gcol < a.ncols;
            const uint64_t cm = __ballot(keep);
            if (keep) {
                const uint32_t r = qn + __popcll(cm & mask_lt(lane));
                qcol[r] = gcol;
                qrow[r] = (uint16_t)brow;
            }
            qn += __popcll(cm);
        }
        // tile holding the next chunk's first word
        const uint32_t pn = p + 64 * U;
        const int An = __popcll(__ballot(dj.x <= pn));
        if (An) {
            cur.x = rdlane(dj.x, An - 1);
            cur.y = rdlane(dj.y, An - 1);
            ct += An;
        }
        // prefetch the next chunk (clamped, so always in bounds) before the
        // gathers of this one
        {
            const uint32_t wj = ct + 1 + lane;
            dj = a.dir[wj < dlast ? wj : dlast];
            if (wj > dlast) dj = make_uint2(0xFFFFFFFFu, 0u);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t pos = pn + u * 64 + lane;
                wrd[u] = wp[pos < plast ? pos : plast];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (qn) {
            if (K) drain<T, (K ? K : 1), P, A>(a, band, S, qcol, qrow, qn, lane);
            else drain_generic<T, P, A>(a, band, S, qcol, qrow, qn, lane);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        p = pn;
    }

    // write the band (rounded once to T), or the partial slot in A
    if (slot == kNone) {
        T *dst = (T *)a.Y + row0 * (uint64_t)a.ldy;
        for (uint32_t i = lane; i < rows * k; i += 64) {
            const uint32_t r = i / k, cc = i - r * k;
            dst[r * (uint64_t)a.ldy + cc] = (T)band[r * S + cc];
        }
    } else {
        A *dst = (A *)a.partial + (uint64_t)slot * a.hs * k;
        for (uint32_t i = lane; i < rows * k; i += 64) {
            const uint32_t r = i / k, cc = i - r * k;
            dst[i] = band[r * S + cc];
        }
    }
}

// Y rows of a split band = partial slots summed in stream order.
template <typename T, typename A>
__global__ void __launch_bounds__(256) spmm_reduce(const Split *splits, const A *partial,
                                                    uint32_t hs, uint32_t k, T *Y, int64_t ldy)
{
    const Split sp = splits[blockIdx.x];
    const uint32_t tot = sp.rows * k;
    for (uint32_t i = threadIdx.x; i < tot; i += blockDim.x) {
        A acc = partial[(uint64_t)sp.slot_begin * hs * k + i];
        for (uint32_t s = 1; s < sp.nslots; s++)
            acc += partial[(uint64_t)(sp.slot_begin + s) * hs * k + i];
        uint32_t r = i / k, c = i - r * k;
        Y[(sp.row0 + r) * (uint64_t)ldy + c] = (T)acc;
    }
}

template <typename T, int K>
constexpr int passes_for()
{
    if constexpr (K == 0) {
        return 8;
    } else {
        constexpr int NZP = Lanes<T, K>::NZP;
        constexpr int E = Lanes<T, K>::E;
        constexpr int vgprs_per_pass = E * (int)sizeof(T) / 4;
        constexpr int p = 128 / vgprs_per_pass;          // ~128 VGPRs of loads in flight
        constexpr int pmax = (64 * 8) / NZP;              // no more than the queue
        return p < pmax ? (p < 1 ? 1 : p) : pmax;
    }
}

template <typename T, int K>
fmx_status launch_k(fmx_ctx *ctx, SpmmArgs a, const Plan *plan, hipStream_t stream)
{
    constexpr int U = 8;
    constexpr int P = passes_for<T, K>();
    auto kern = spmm_stream<T, K, U, P>;
    const uint32_t lds = a.wave_bytes * kWPB;
    static bool attr_set = false;
    if (!attr_set) {
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
    }
    const uint32_t blocks = (plan->nitems + kWPB - 1) / kWPB;
    if (blocks) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * kWPB), lds, stream, a);
        FMX_HIP(ctx, hipGetLastError());
    }
    return FMX_OK;
}

template <typename T>
fmx_status launch_t(fmx_ctx *ctx, const fmx_spm *spm, const Plan *plan, int k,
                    const void *X, int64_t ldx, void *Y, int64_t ldy, void *partial,
                    hipStream_t stream)
{
    constexpr int U = 8;
    SpmmArgs a;
    a.img = (const uint16_t *)spm->d_image;
    a.dir = spm->d_dir;
    a.items = plan->d_items;
    a.X = X;
    a.Y = Y;
    a.partial = partial;
    a.ldx = ldx;
    a.ldy = ldy;
    a.nitems = plan->nitems;
    a.hs = plan->hs;
    // pad band rows by one element to spread the LDS adds over the banks
    a.band_stride = (uint32_t)k + ((k >= 4 && getenv("FMX_NO_PAD") == nullptr) ? 1u : 0u);
    using A = typename AccOf<T>::type;
    a.band_bytes = (uint32_t)(((uint64_t)plan->hs * a.band_stride * sizeof(A) + 15) & ~15ull);
    a.wave_bytes = a.band_bytes + 64 * U * 6;
    a.wave_bytes = (a.wave_bytes + 15) & ~15u;
    a.log2w = spm->log2w;
    a.k = k;
    a.ncols = spm->ncols;
    if ((uint64_t)a.wave_bytes * kWPB > ctx->lds_per_block)
        return fail(ctx, FMX_EUNSUPPORTED, "band does not fit in LDS");
    // 16-byte vector gathers need aligned rows; otherwise the runtime-width path.
    const int vb = (k * (int)sizeof(T) < 16) ? k * (int)sizeof(T) : 16;
    const bool aligned = ((uintptr_t)X % vb == 0) && ((ldx * (int64_t)sizeof(T)) % vb == 0);
    fmx_status st;
    switch (aligned ? k : 0) {
    case 1: st = launch_k<T, 1>(ctx, a, plan, stream); break;
    case 2: st = launch_k<T, 2>(ctx, a, plan, stream); break;
    case 4: st = launch_k<T, 4>(ctx, a, plan, stream); break;
    case 8: st = launch_k<T, 8>(ctx, a, plan, stream); break;
    case 16: st = launch_k<T, 16>(ctx, a, plan, stream); break;
    case 32: st = launch_k<T, 32>(ctx, a, plan, stream); break;
    case 64: st = launch_k<T, 64>(ctx, a, plan, stream); break;
    default: st = launch_k<T, 0>(ctx, a, plan, stream); break;
    }
    if (st) return st;
    if (plan->nsplits) {
        hipLaunchKernelGGL((spmm_reduce<T, A>), dim3(plan->nsplits), dim3(256), 0, stream,
                           plan->d_splits, (const A *)partial, plan->hs, (uint32_t)k, (T *)Y, ldy);
        FMX_HIP(ctx, hipGetLastError());
    }
    return FMX_OK;
}

}  // namespace

fmx_status launch_spmm(fmx_ctx *ctx, const fmx_spm *spm, const Plan *plan, fmx_dtype dt,
                       int k, const void *X, int64_t ldx, void *Y, int64_t ldy,
                       void *partial, hipStream_t stream)
{
    switch (dt) {
    case FMX_I32: return launch_t<int32_t>(ctx, spm, plan, k, X, ldx, Y, ldy, partial, stream);
    case FMX_I64: return launch_t<int64_t>(ctx, spm, plan, k, X, ldx, Y, ldy, partial, stream);
    case FMX_F32: return launch_t<float>(ctx, spm, plan, k, X, ldx, Y, ldy, partial, stream);
    case FMX_F64: return launch_t<double>(ctx, spm, plan, k, X, ldx, Y, ldy, partial, stream);
    default: return fail(ctx, FMX_ETYPE, "wrong matrix type in spmm");
    }
}

}  // namespace fmx
