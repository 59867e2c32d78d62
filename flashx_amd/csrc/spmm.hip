// spmm.hip -- gfx950 SpMM over the reference's 2D tile image.
//
// Reference semantics: block_spmm_task_impl::run_on_block with row_part_func
// and coo_func (matrix/sparse_matrix.h:219-375): for every non-zero (r, c) of
// a tile, Y[r][0..k) += X[c][0..k) (binary matrix, value 1).
//
// Execution model (one wave = one work item, no cross-wave sharing):
//   * an item is a contiguous word range of one block row (tile headers and
//     all) plus a band of <= Hs output rows kept in LDS;
//   * the range is decoded 64*R words at a time (a "chunk", 16-byte aligned):
//     lane l owns R consecutive words, finds the tile holding its first word
//     by a binary search of an LDS window of the tile directory, and
//     classifies its words (tile header / row head / column) with per-lane
//     state.  The row in effect at a lane's first word is the last row head
//     of the nearest lane below it (ballot + shuffle), and a wave prefix sum
//     of per-lane column counts writes the (global column, band row) entries
//     in stream order to an LDS queue;
//   * the queue is drained with vector gathers of X rows (16 bytes per lane,
//     L lanes per non-zero).  Lane group g takes P consecutive entries, so a
//     run of one row (a row part of a tile) is summed in registers and reaches
//     the band with one LDS add per element at each row change;
//   * the next chunk's words are requested before the drain; a wave's own
//     decode and gathers alternate, and the other resident waves (up to four
//     per SIMD, as LDS allows) keep gathers in flight meanwhile;
//   * the band is written to Y once (rows without non-zeros are 0), or to a
//     partial slot when a power-law band was split over several items; split
//     bands are summed slot by slot, in stream order, by spmm_reduce.
// Each band is owned by exactly one wave, so results do not depend on timing.
// FP32 data accumulate in an FP64 band and are rounded once on write-out.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "fmx_internal.h"

namespace fmx {
namespace {

constexpr int kWPB = 1;  // waves per workgroup: one item, so LDS packs per wave

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
    const uint32_t *vdir;  // weighted images: values-region word of each tile
    int vwords;            // weighted images: u16 words per value (2 or 4)
    const double *rscale;  // floating-point data: Y row r is scaled by rscale[r] (or null)
    int ycol;              // Y column-major: element (r, j) at Y[j * ldy + r]
    const void *Xhot;      // gather-order images: rows of columns < nhot (leading dimension k)
    uint64_t nhot;         // columns c >= nhot are row c - nhot of X
};

// Row gcol of the gathered operand: with HOT (a gather-order image) the
// dense hot block for the first nhot columns and X for the others, else X.
template <typename T, bool HOT>
__device__ __forceinline__ const T *xrow(const SpmmArgs &a, uint32_t gcol)
{
    if constexpr (!HOT) return (const T *)a.X + (uint64_t)gcol * a.ldx;
    return gcol < a.nhot ? (const T *)a.Xhot + (uint64_t)gcol * (uint32_t)a.k
                         : (const T *)a.X + (uint64_t)(gcol - a.nhot) * a.ldx;
}

template <typename T, int E>
struct alignas(sizeof(T) * E) Vec {
    T v[E];
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

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

// Orders this wave's LDS writes before its later LDS reads.  LDS executes one
// wave's instructions in order, so only compiler reordering must be stopped;
// a fence would also wait for every outstanding global load.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
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

// Band column of output column c for compile-time widths: lane q of a
// non-zero's L lanes holds columns q*E .. q*E+E-1 and adds its i-th one at
// band column i*L+q, so the L lanes of one add touch consecutive words.
template <typename T, int K>
__device__ __forceinline__ uint32_t band_col(uint32_t c)
{
    using C = Lanes<T, K>;
    return (c % C::E) * C::L + c / C::E;
}

// Queue entry e lives at slot e + e/P: the P-entry segments of the lane
// groups then start in different LDS banks.
template <int P>
__device__ __forceinline__ uint32_t qslot(uint32_t e) { return e + e / P; }

// ---------------------------------------------------------------------------
// Lane-sequential decode
template <int R>
struct LaneCfg {
    static constexpr int CW = 64 * R;                                // words per chunk
    static constexpr int NB = 1 + (R + 10) / 11;                     // tiles a lane can touch
    static constexpr int DWIN = ((CW - 1) / 11 + 2 + 63) / 64 * 64;  // window entries
    static constexpr int DPL = DWIN / 64;
};
// A stored tile is >= 11 words (16-byte header, one entry, end marker; a
// 16-byte stub only for an empty block row, which is an item by itself), so
// a chunk meets at most (CW-1)/11+1 tile starts after the tile holding its
// first word, and R words at most NB-1.

struct ItemCtx {
    uint64_t bw;          // absolute word index of the block row
    int32_t begin, pend;  // item words [begin, pend), relative to the block row
    uint32_t sb_base, rows;
    int log2w;
    uint64_t ncols;
};

// Decodes the chunk starting at absolute word c (8-word aligned) from this
// lane's R words in wreg and the directory window win (DWIN entries and a
// sentinel).  Writes the chunk's entries in stream order at queue slots
// qslot<SEGP>(e) -- band row 0xFFFF for an entry outside the band -- and
// returns their number; carry becomes the row in effect after the chunk.
template <int R, int SEGP>
__device__ __forceinline__ uint32_t decode_chunk(const ItemCtx &ic, uint64_t c, const uint32_t *wreg,
                                                 const uint2 *win, int &carry, uint32_t *qcol, uint16_t *qrow,
                                                 int lane)
{
    using LC = LaneCfg<R>;
    constexpr int NB = LC::NB, DWIN = LC::DWIN;
    const int32_t p0 = (int32_t)((int64_t)c - (int64_t)ic.bw) + lane * R;
    const int32_t q = p0 > ic.begin ? p0 : ic.begin;
    // last window tile starting at or before q (window entry 0 holds the
    // chunk's first item word, sentinels compare greater)
    uint32_t j = 0;
#pragma unroll
    for (int st = DWIN / 2; st >= 1; st >>= 1)
        if (win[j + st].x <= (uint32_t)q) j += st;
    uint32_t ts0[NB], tb0[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const uint32_t jj = j + b < DWIN ? j + b : DWIN;
        const uint2 e = win[jj];
        ts0[b] = e.x;
        tb0[b] = e.y;
    }
    // pass 1: column count and last row head of this lane
    int cnt = 0, lrow = 0;
    bool hh = false;
    {
        uint32_t ts[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) ts[b] = ts0[b];
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int32_t pos = p0 + i;
            const uint32_t w = (wreg[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu;
            const bool valid = pos >= ic.begin && pos < ic.pend;
            if (NB > 1 && valid && (uint32_t)pos >= ts[NB > 1 ? 1 : 0]) {
#pragma unroll
                for (int b = 0; b + 1 < NB; b++) ts[b] = ts[b + 1];
                ts[NB - 1] = 0xFFFFFFFFu;
            }
            const bool body = valid && ((uint32_t)pos - ts[0]) >= 8u;
            if (body && (w & 0x8000u)) {
                lrow = (int)(w & 0x7FFFu);
                hh = true;
            }
            cnt += (body && !(w & 0x8000u)) ? 1 : 0;
        }
    }
    // row in effect at this lane's first word; row after the chunk
    const uint64_t hm = __ballot(hh);
    const uint64_t below = hm & mask_lt(lane);
    const int src = below ? 63 - __clzll(below) : 0;
    const int prow = __shfl(lrow, src);
    const int row_in = below ? prow : carry;
    if (hm) carry = (int)rdlane((uint32_t)lrow, 63 - __clzll(hm));
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    const uint32_t qn = rdlane((uint32_t)incl, 63);
    // pass 2: entries in stream order
    {
        uint32_t ts[NB], tb[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            ts[b] = ts0[b];
            tb[b] = tb0[b];
        }
        int r = row_in;
        uint32_t o = (uint32_t)(incl - cnt);
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int32_t pos = p0 + i;
            const uint32_t w = (wreg[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu;
            const bool valid = pos >= ic.begin && pos < ic.pend;
            if (NB > 1 && valid && (uint32_t)pos >= ts[NB > 1 ? 1 : 0]) {
#pragma unroll
                for (int b = 0; b + 1 < NB; b++) {
                    ts[b] = ts[b + 1];
                    tb[b] = tb[b + 1];
                }
                ts[NB - 1] = 0xFFFFFFFFu;
            }
            const bool body = valid && ((uint32_t)pos - ts[0]) >= 8u;
            if (body && (w & 0x8000u)) r = (int)(w & 0x7FFFu);
            if (body && !(w & 0x8000u)) {
                const uint32_t br = (uint32_t)(r - (int)ic.sb_base);
                const uint32_t gcol = (tb[0] << ic.log2w) | w;
                const bool keep = r >= 0 && br < ic.rows && gcol < ic.ncols;
                const uint32_t sl = qslot<SEGP>(o);
                qcol[sl] = keep ? gcol : 0u;
                qrow[sl] = keep ? (uint16_t)br : (uint16_t)0xFFFFu;
                o++;
            }
        }
    }
    return qn;
}

// This lane's words of chunk c, clamped to the aligned 8-word group holding
// the item's last word lastw (always inside the image allocation, whose size
// is a multiple of 4096 bytes).  The image is streamed once: non-temporal
// loads, so it does not displace X in L2.
template <int R>
__device__ __forceinline__ void load_words(const uint16_t *img, uint64_t c, uint64_t lastw, int lane,
                                           uint32_t *wreg)
{
    static_assert(R % 8 == 0, "lane word count");
    const uint64_t last8 = lastw & ~7ull;
#pragma unroll
    for (int h = 0; h < R / 8; h++) {
        uint64_t g = c + (uint64_t)lane * R + h * 8;
        g = g < last8 ? g : last8;
        const u32x4 v = __builtin_nontemporal_load((const u32x4 *)(img + g));
        wreg[h * 4 + 0] = v.x;
        wreg[h * 4 + 1] = v.y;
        wreg[h * 4 + 2] = v.z;
        wreg[h * 4 + 3] = v.w;
    }
}

// Directory entries ct + h*64 + lane (clamped); `live` marks those inside the
// block row, the others become sentinels when the window is stored.
template <int DPL>
__device__ __forceinline__ void load_window(const uint2 *dir, uint32_t ct, uint32_t dlast, int lane, uint2 *dreg,
                                            bool *live)
{
#pragma unroll
    for (int h = 0; h < DPL; h++) {
        const uint32_t t = ct + h * 64 + lane;
        live[h] = t <= dlast;
        dreg[h] = dir[t < dlast ? t : dlast];
    }
}

template <int DPL>
__device__ __forceinline__ void store_window(uint2 *win, const uint2 *dreg, const bool *live, int lane)
{
#pragma unroll
    for (int h = 0; h < DPL; h++) win[h * 64 + lane] = live[h] ? dreg[h] : make_uint2(0xFFFFFFFFu, 0u);
}

// Start of the next chunk's window: entries of this window starting at or
// before its first word (the window is sorted, sentinels compare greater).
template <int DPL>
__device__ __forceinline__ uint32_t window_advance(const uint2 *dreg, const bool *live, uint32_t pn)
{
    int n = 0;
#pragma unroll
    for (int h = 0; h < DPL; h++) n += __popcll(__ballot(live[h] && dreg[h].x <= pn));
    return (uint32_t)n;
}

// ---------------------------------------------------------------------------
// Drains (compile-time width K): gathers of P consecutive entries per lane
// group, summed in registers per row run, added to the band at row changes.
template <typename T, int K, int P, typename A, bool HOT>
__device__ __forceinline__ void gather_batch(const SpmmArgs &a, const uint32_t *qcol, const uint16_t *qrow,
                                             uint32_t base, uint32_t qn, int lane, Vec<T, Lanes<T, K>::E> *v,
                                             uint32_t *rr)
{
    using C = Lanes<T, K>;
    const int q = lane % C::L;
    const int g = lane / C::L;
    uint32_t gcs[P];
#pragma unroll
    for (int pp = 0; pp < P; pp++) {
        const uint32_t e = base + g * P + pp;
        const uint32_t sl = qslot<P>(e);
        const bool ok = e < qn;
        const uint32_t qr = qrow[sl];
        gcs[pp] = ok ? qcol[sl] : 0u;
        rr[pp] = (ok && qr != 0xFFFFu) ? qr : 0xFFFFFFFFu;
    }
    // every queue read before the first gather, every gather before the first use
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int pp = 0; pp < P; pp++) v[pp] = *(const Vec<T, C::E> *)(xrow<T, HOT>(a, gcs[pp]) + q * C::E);
    __builtin_amdgcn_sched_barrier(0);
}

template <typename T, int K, int P, typename A>
__device__ __forceinline__ void accumulate_batch(A *band, uint32_t S, int lane, const Vec<T, Lanes<T, K>::E> *v,
                                                 const uint32_t *rr)
{
    using C = Lanes<T, K>;
    const int q = lane % C::L;
    if constexpr (sizeof(T) == 8) {
        // 8-byte data: each entry adds straight into the band.  Run sums
        // save few LDS adds (rows average 2.4 entries per tile on R-MAT)
        // and their divergent branches cost more than that for E <= 2
        // (FP64 k = 16: 1.17 ms against 1.34 ms with run sums)
#pragma unroll
        for (int pp = 0; pp < P; pp++) {
            if (rr[pp] != 0xFFFFFFFFu) {
                A *dst = band + rr[pp] * S + q;
#pragma unroll
                for (int i = 0; i < C::E; i++) lds_add(dst + i * C::L, (A)v[pp].v[i]);
            }
        }
        return;
    }
    A acc[C::E];
    uint32_t cur = rr[0];
#pragma unroll
    for (int i = 0; i < C::E; i++) acc[i] = (A)v[0].v[i];
#pragma unroll
    for (int pp = 1; pp < P; pp++) {
        if (rr[pp] != cur) {
            if (cur != 0xFFFFFFFFu) {
                A *dst = band + cur * S + q;
#pragma unroll
                for (int i = 0; i < C::E; i++) lds_add(dst + i * C::L, acc[i]);
            }
            cur = rr[pp];
#pragma unroll
            for (int i = 0; i < C::E; i++) acc[i] = (A)v[pp].v[i];
        } else {
#pragma unroll
            for (int i = 0; i < C::E; i++) acc[i] += (A)v[pp].v[i];
        }
    }
    if (cur != 0xFFFFFFFFu) {
        A *dst = band + cur * S + q;
#pragma unroll
        for (int i = 0; i < C::E; i++) lds_add(dst + i * C::L, acc[i]);
    }
}

// ---------------------------------------------------------------------------
// Sequential mode: every output element is the left-to-right sum of its
// row's products in stream order, one rounding per addition, starting from
// the zeroed band -- the reference's own order (row_part_func / coo_func,
// matrix/sparse_matrix.h:219-331, `dest_row[j] += src_row[j] * data` per
// non-zero, tiles in ascending block column).  Within a tile a row's columns
// ascend and tiles ascend, so this is each row's ascending-column sum, the
// same for every tile shape.  Lane groups hold consecutive queue entries, so
// they add in group order: group g reads the band after group g-1 wrote it
// (one wave's LDS accesses execute in program order).  A run of one row
// inside a group is chained in registers, which is the same sequence of
// roundings.
template <typename T, int K, int P>
__device__ __forceinline__ void accumulate_batch_seq(T *band, uint32_t S, int lane, const Vec<T, Lanes<T, K>::E> *v,
                                                     const uint32_t *rr)
{
#pragma clang fp contract(off)
    using C = Lanes<T, K>;
    const int q = lane % C::L;
    const int g = lane / C::L;
    for (int gg = 0; gg < C::NZP; gg++) {
        if (g == gg) {
            T acc[C::E];
            uint32_t cur = 0xFFFFFFFFu;
#pragma unroll
            for (int i = 0; i < C::E; i++) acc[i] = T(0);
#pragma unroll
            for (int pp = 0; pp < P; pp++) {
                if (rr[pp] != cur) {
                    if (cur != 0xFFFFFFFFu) {
                        T *dst = band + cur * S + q;
#pragma unroll
                        for (int i = 0; i < C::E; i++) dst[i * C::L] = acc[i];
                    }
                    cur = rr[pp];
                    if (cur != 0xFFFFFFFFu) {
                        const T *src = band + cur * S + q;
#pragma unroll
                        for (int i = 0; i < C::E; i++) acc[i] = src[i * C::L];
                    }
                }
                if (cur != 0xFFFFFFFFu) {
#pragma unroll
                    for (int i = 0; i < C::E; i++) acc[i] = acc[i] + v[pp].v[i];
                }
            }
            if (cur != 0xFFFFFFFFu) {
                T *dst = band + cur * S + q;
#pragma unroll
                for (int i = 0; i < C::E; i++) dst[i * C::L] = acc[i];
            }
        }
        wave_sync();
    }
}

// Sequential mode at runtime widths: lane j owns output column j (of each
// 64-column slice) and adds the queue's entries in order.  Weighted images
// (W) multiply each gathered element by the entry's value of type V first,
// rounded in T, as `src_row[j] * data` is.
template <typename T, typename V, int P, bool W>
__device__ __forceinline__ void drain_generic_seq(const SpmmArgs &a, T *band, uint32_t S, const uint32_t *qcol,
                                                  const uint16_t *qrow, const uint32_t *qval, const uint16_t *vimg,
                                                  uint32_t qn, int lane);

// Runtime width: flat (entry, element) pairs, one element per lane.
template <typename T, int P, typename A, bool HOT>
__device__ __forceinline__ void drain_generic(const SpmmArgs &a, A *band, uint32_t S, const uint32_t *qcol,
                                              const uint16_t *qrow, uint32_t qn, int lane)
{
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
            const uint32_t qr = qrow[e];
            v[pp] = xrow<T, HOT>(a, qcol[e])[j];
            dst[pp] = (f < tot && qr != 0xFFFFu) ? qr * S + j : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int pp = 0; pp < P; pp++)
            if (dst[pp] != 0xFFFFFFFFu) lds_add(band + dst[pp], (A)v[pp]);
    }
}

// One value of type T from the u16-aligned values region (u16 loads, so no
// alignment beyond 2 bytes is assumed).
template <typename T>
__device__ __forceinline__ T load_value(const uint16_t *p)
{
    if constexpr (sizeof(T) == 4) {
        const uint32_t u = (uint32_t)p[0] | ((uint32_t)p[1] << 16);
        return __builtin_bit_cast(T, u);
    } else {
        const uint64_t u = (uint64_t)p[0] | ((uint64_t)p[1] << 16) | ((uint64_t)p[2] << 32) |
                           ((uint64_t)p[3] << 48);
        return __builtin_bit_cast(T, u);
    }
}

template <typename T, typename V, int P, bool W>
__device__ __forceinline__ void drain_generic_seq(const SpmmArgs &a, T *band, uint32_t S, const uint32_t *qcol,
                                                  const uint16_t *qrow, const uint32_t *qval, const uint16_t *vimg,
                                                  uint32_t qn, int lane)
{
#pragma clang fp contract(off)
    const uint32_t k = (uint32_t)a.k;
    for (uint32_t j0 = 0; j0 < k; j0 += 64) {
        const uint32_t j = j0 + (uint32_t)lane;
        const bool act = j < k;
        for (uint32_t e0 = 0; e0 < qn; e0 += P) {
            T v[P];
            uint32_t dst[P];
#pragma unroll
            for (int pp = 0; pp < P; pp++) {
                const uint32_t e = e0 + pp;
                const uint32_t ec = e < qn ? e : qn - 1;
                const uint32_t qr = qrow[ec];
                v[pp] = xrow<T, false>(a, qcol[ec])[act ? j : 0u];
                if constexpr (W) v[pp] = v[pp] * (T)load_value<V>(vimg + qval[ec]);
                dst[pp] = (e < qn && act && qr != 0xFFFFu) ? qr * S + j : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int pp = 0; pp < P; pp++)
                if (dst[pp] != 0xFFFFFFFFu) band[dst[pp]] = band[dst[pp]] + v[pp];
        }
    }
}

// The output scales of a band's first 128 rows, requested when the item
// starts so that their latency is hidden by the item's work.
struct RowScale {
    double v[2];
};

template <typename T>
__device__ __forceinline__ RowScale prefetch_rscale(const SpmmArgs &a, uint64_t row0, uint32_t rows, int lane)
{
    RowScale rs{{1.0, 1.0}};
    if constexpr (std::is_floating_point<T>::value) {
        if (a.rscale) {
#pragma unroll
            for (int h = 0; h < 2; h++)
                if ((uint32_t)(h * 64 + lane) < rows) rs.v[h] = a.rscale[row0 + h * 64 + lane];
        }
    }
    return rs;
}

// Writes the band to Y, or to its partial slot.  With row scales, the
// scales of bands of at most 128 rows are staged in scratch LDS (the drained
// queue), so the write loop waits on no global load.
template <typename T, int K, typename A>
__device__ __forceinline__ void write_band(const SpmmArgs &a, const A *band, uint32_t S, uint64_t row0,
                                           uint32_t rows, uint32_t slot, int lane, const RowScale &rs,
                                           double *scratch)
{
    const uint32_t k = K ? K : (uint32_t)a.k;
    auto bcol = [&](uint32_t cc) -> uint32_t {
        if constexpr (K > 0) return band_col<T, (K ? K : 1)>(cc);
        else return cc;
    };
    if (slot == kNone) {
        T *dst = (T *)a.Y + row0 * (uint64_t)a.ldy;
        bool staged = false;
        if constexpr (std::is_floating_point<T>::value) {
            if (a.rscale && rows <= 128) {
                wave_sync();
#pragma unroll
                for (int h = 0; h < 2; h++)
                    if ((uint32_t)(h * 64 + lane) < rows) scratch[h * 64 + lane] = rs.v[h];
                wave_sync();
                staged = true;
            }
        }
        if (a.ycol) {
            // column-major Y: lanes run down a column, so the stores of one
            // column are consecutive
            T *dc = (T *)a.Y + row0;
            for (uint32_t i = lane; i < rows * k; i += 64) {
                const uint32_t cc = i / rows, r = i - cc * rows;
                A v = band[r * S + bcol(cc)];
                if constexpr (std::is_floating_point<T>::value)
                    if (a.rscale) v = (A)((double)v * (staged ? scratch[r] : a.rscale[row0 + r]));
                __builtin_nontemporal_store((T)v, &dc[cc * (uint64_t)a.ldy + r]);
            }
            return;
        }
        for (uint32_t i = lane; i < rows * k; i += 64) {
            const uint32_t r = i / k, cc = i - r * k;
            A v = band[r * S + bcol(cc)];
            if constexpr (std::is_floating_point<T>::value)
                if (a.rscale) v = (A)((double)v * (staged ? scratch[r] : a.rscale[row0 + r]));  // one rounding
            // written once and not read here: non-temporal
            __builtin_nontemporal_store((T)v, &dst[r * (uint64_t)a.ldy + cc]);
        }
    } else {
        A *dst = (A *)a.partial + (uint64_t)slot * a.hs * k;
        for (uint32_t i = lane; i < rows * k; i += 64) {
            const uint32_t r = i / k, cc = i - r * k;
            dst[i] = band[r * S + bcol(cc)];
        }
    }
}

struct ItemState {
    ItemCtx ic;
    uint32_t dlast, slot, ct;
    uint64_t row0;
    int carry;
    int32_t begin, end, rot;  // the whole item, and where processing starts
    uint32_t dir_begin, rot_dir;
    int item_carry;
};

// Segment 0 is [rot, end) and segment 1 is [begin, rot) (empty without a
// rotation).  A rotated start is a tile start, where no row is carried in.
__device__ __forceinline__ bool select_segment(ItemState &st, int seg)
{
    if (seg == 0) {
        st.ic.begin = st.rot;
        st.ic.pend = st.end;
        st.ct = st.rot_dir;
        st.carry = st.rot == st.begin ? st.item_carry : -1;
        return true;
    }
    if (st.rot == st.begin) return false;
    st.ic.begin = st.begin;
    st.ic.pend = st.rot;
    st.ct = st.dir_begin;
    st.carry = st.item_carry;
    return true;
}

__device__ __forceinline__ ItemState item_state(const SpmmArgs &a, const Item &it)
{
    ItemState s;
    s.ic.bw = it.brow_word;
    s.ic.begin = (int32_t)it.begin;
    s.ic.pend = (int32_t)(it.begin + it.nwords);  // nwords >= 1
    s.ic.sb_base = it.sb_base;
    s.ic.rows = it.rows;
    s.ic.log2w = a.log2w;
    s.ic.ncols = a.ncols;
    s.dlast = it.dir_end - 1;
    s.row0 = ((uint64_t)it.row0_hi << 32) | it.row0_lo;
    s.slot = it.slot;
    s.carry = (it.carry == kNone) ? -1 : (int)it.carry;
    s.ct = it.dir_begin;
    s.begin = (int32_t)it.begin;
    s.end = (int32_t)(it.begin + it.nwords);
    s.rot = (int32_t)it.rot_begin;
    s.dir_begin = it.dir_begin;
    s.rot_dir = it.rot_dir;
    s.item_carry = s.carry;
    return s;
}

// ---------------------------------------------------------------------------
// The SpMM kernel: chunks of 64*R words, each decoded and then drained
// batch by batch.  The next chunk's words and window are requested behind
// the first batch's gathers, so their latency overlaps the gathers'; the
// gathers of one wave are not overlapped with its own decode, other resident
// waves fill that gap (four per SIMD where LDS allows).
// SEQ: the sequential mode (accumulate_batch_seq): the band holds T and
// every element is summed in the reference's order; the plan of such a call
// neither splits nor rotates items.
template <typename T, int K, int R, int P, bool SEQ, bool HOT>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(5, 5)))
spmm_lane(SpmmArgs a)
{
    using LC = LaneCfg<R>;
    constexpr int CW = LC::CW, DWIN = LC::DWIN, DPL = LC::DPL;
    constexpr int SEGP = K ? P : (1 << 30);
    constexpr int QS = CW + CW / (K ? P : CW) + 1;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const uint32_t wv = uni(threadIdx.x >> 6);
    const uint32_t item_id = blockIdx.x * kWPB + wv;
    if (item_id >= a.nitems) return;

    unsigned char *wbase = smem + wv * a.wave_bytes;
    using A = typename std::conditional<SEQ, T, typename AccOf<T>::type>::type;
    A *band = (A *)wbase;
    uint32_t *qcol = (uint32_t *)(wbase + a.band_bytes);
    uint16_t *qrow = (uint16_t *)(qcol + QS);
    uint2 *win = (uint2 *)(wbase + ((a.band_bytes + QS * 6 + 7) & ~7u));

    ItemState st = item_state(a, a.items[item_id]);
    const RowScale rs = prefetch_rscale<T>(a, st.row0, st.ic.rows, lane);
    const uint32_t S = a.band_stride;
    for (uint32_t i = lane; i < st.ic.rows * S; i += 64) band[i] = A(0);
    if (lane == 0) win[DWIN] = make_uint2(0xFFFFFFFFu, 0u);

    for (int seg = 0; seg < 2; seg++) {
    if (!select_segment(st, seg)) break;
    const uint64_t abs_end = st.ic.bw + (uint64_t)st.ic.pend;
    const uint64_t lastw = abs_end - 1;
    uint64_t c = (st.ic.bw + (uint64_t)st.ic.begin) & ~7ull;
    uint32_t wreg[R / 2];
    uint2 dreg[DPL];
    bool live[DPL];
    load_words<R>(a.img, c, lastw, lane, wreg);
    load_window<DPL>(a.dir, st.ct, st.dlast, lane, dreg, live);

    while (c < abs_end) {
        store_window<DPL>(win, dreg, live, lane);
        wave_sync();
        const uint32_t qn = decode_chunk<R, SEGP>(st.ic, c, wreg, win, st.carry, qcol, qrow, lane);
        const uint64_t cn = c + CW;
        // the next chunk's words and window (this chunk's are in LDS now)
        auto fetch_next = [&]() {
            if (cn < abs_end) {
                st.ct += window_advance<DPL>(dreg, live, (uint32_t)((int64_t)cn - (int64_t)st.ic.bw)) - 1;
                load_words<R>(a.img, cn, lastw, lane, wreg);
                load_window<DPL>(a.dir, st.ct, st.dlast, lane, dreg, live);
            }
        };
        if constexpr (K == 0) fetch_next();
        wave_sync();
        if constexpr (K > 0) {
            // requested behind the first batch's gathers: a wait for the
            // gathers need not wait for these, and their latency overlaps
            // the gathers' instead of preceding them
            using C = Lanes<T, K>;
            bool fetched = false;
            for (uint32_t base = 0; base < qn; base += C::NZP * P) {
                Vec<T, C::E> v[P];
                uint32_t rr[P];
                gather_batch<T, K, P, A, HOT>(a, qcol, qrow, base, qn, lane, v, rr);
                if (!fetched) {
                    fetch_next();
                    fetched = true;
                }
                if constexpr (SEQ) accumulate_batch_seq<T, K, P>(band, S, lane, v, rr);
                else accumulate_batch<T, K, P, A>(band, S, lane, v, rr);
            }
            if (!fetched) fetch_next();
        } else if (qn) {
            if constexpr (SEQ) drain_generic_seq<T, T, P, false>(a, band, S, qcol, qrow, nullptr, nullptr, qn, lane);
            else drain_generic<T, P, A, HOT>(a, band, S, qcol, qrow, qn, lane);
        }
        wave_sync();
        c = cn;
    }
    }
    write_band<T, K, A>(a, band, S, st.row0, st.ic.rows, st.slot, lane, rs, (double *)qcol);
}


// ---------------------------------------------------------------------------
// Weighted images (entry_size 4 or 8): each tile's values follow its COO
// region, one per non-zero in stream order (SCSR values, then COO values,
// sparse_matrix_format.h:233-241).  The decode is decode_chunk's plus two
// facts per entry: words at or past the tile's values region are not
// entries, and the entry's value is the tile's (o - wq[j])-th, where o is
// the entry's index in the chunk and wq[j] the chunk index at which window
// tile j starts (minus the columns of tile 0 seen in earlier chunks, for
// j = 0).  Work items start at tile starts (rotation and splits), so no
// value position is carried into an item.
template <int R, int SEGP>
__device__ __forceinline__ uint32_t decode_chunk_w(const ItemCtx &ic, uint64_t c, const uint32_t *wreg,
                                                   const uint2 *win, const uint32_t *wvb, int32_t *wq, int c0,
                                                   int vwords, int &carry, uint32_t *qcol, uint16_t *qrow,
                                                   uint32_t *qval, int lane)
{
    using LC = LaneCfg<R>;
    constexpr int NB = LC::NB, DWIN = LC::DWIN;
    const int32_t p0 = (int32_t)((int64_t)c - (int64_t)ic.bw) + lane * R;
    const int32_t q = p0 > ic.begin ? p0 : ic.begin;
    uint32_t j = 0;
#pragma unroll
    for (int st = DWIN / 2; st >= 1; st >>= 1)
        if (win[j + st].x <= (uint32_t)q) j += st;
    uint32_t ts0[NB], tb0[NB], tv0[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const uint32_t jj = j + b < DWIN ? j + b : DWIN;
        const uint2 e = win[jj];
        ts0[b] = e.x;
        tb0[b] = e.y;
        tv0[b] = wvb[jj];
    }
    // pass 1: counts, last row head, and the column count at each tile
    // start inside this lane's words (a tile starting exactly at the lane's
    // first item word counts too, unless it is window tile 0)
    int cnt = 0, lrow = 0;
    bool hh = false;
    int nx = 0;
    int xcnt[NB];
    uint32_t xj[NB];
    {
        uint32_t ts[NB], tv[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            ts[b] = ts0[b];
            tv[b] = tv0[b];
            xcnt[b] = 0;
            xj[b] = 0;
        }
        if (j > 0 && ts0[0] == (uint32_t)q && q < ic.pend) {
            xcnt[0] = 0;
            xj[0] = j;
            nx = 1;
        }
        uint32_t jc = j;
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int32_t pos = p0 + i;
            const uint32_t w = (wreg[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu;
            const bool valid = pos >= ic.begin && pos < ic.pend;
            if (NB > 1 && valid && (uint32_t)pos >= ts[NB > 1 ? 1 : 0]) {
#pragma unroll
                for (int b = 0; b + 1 < NB; b++) {
                    ts[b] = ts[b + 1];
                    tv[b] = tv[b + 1];
                }
                ts[NB - 1] = 0xFFFFFFFFu;
                tv[NB - 1] = 0xFFFFFFFFu;
                jc++;
                if (nx < NB) {
                    xcnt[nx] = cnt;
                    xj[nx] = jc;
                    nx++;
                }
            }
            const bool body = valid && ((uint32_t)pos - ts[0]) >= 8u && (uint32_t)pos < tv[0];
            if (body && (w & 0x8000u)) {
                lrow = (int)(w & 0x7FFFu);
                hh = true;
            }
            cnt += (body && !(w & 0x8000u)) ? 1 : 0;
        }
    }
    const uint64_t hm = __ballot(hh);
    const uint64_t below = hm & mask_lt(lane);
    const int src = below ? 63 - __clzll(below) : 0;
    const int prow = __shfl(lrow, src);
    const int row_in = below ? prow : carry;
    if (hm) carry = (int)rdlane((uint32_t)lrow, 63 - __clzll(hm));
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    const uint32_t qn = rdlane((uint32_t)incl, 63);
    const int base = incl - cnt;
    if (lane == 0) wq[0] = -c0;
#pragma unroll
    for (int b = 0; b < NB; b++)
        if (b < nx) wq[xj[b]] = base + xcnt[b];
    wave_sync();
    // pass 2: entries with their value words
    {
        uint32_t ts[NB], tb[NB], tv[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            ts[b] = ts0[b];
            tb[b] = tb0[b];
            tv[b] = tv0[b];
        }
        int qt = wq[j];
        int r = row_in;
        uint32_t o = (uint32_t)base;
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int32_t pos = p0 + i;
            const uint32_t w = (wreg[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu;
            const bool valid = pos >= ic.begin && pos < ic.pend;
            if (NB > 1 && valid && (uint32_t)pos >= ts[NB > 1 ? 1 : 0]) {
#pragma unroll
                for (int b = 0; b + 1 < NB; b++) {
                    ts[b] = ts[b + 1];
                    tb[b] = tb[b + 1];
                    tv[b] = tv[b + 1];
                }
                ts[NB - 1] = 0xFFFFFFFFu;
                tv[NB - 1] = 0xFFFFFFFFu;
                qt = (int)o;
            }
            const bool body = valid && ((uint32_t)pos - ts[0]) >= 8u && (uint32_t)pos < tv[0];
            if (body && (w & 0x8000u)) r = (int)(w & 0x7FFFu);
            if (body && !(w & 0x8000u)) {
                const uint32_t br = (uint32_t)(r - (int)ic.sb_base);
                const uint32_t gcol = (tb[0] << ic.log2w) | w;
                const bool keep = r >= 0 && br < ic.rows && gcol < ic.ncols;
                const uint32_t sl = qslot<SEGP>(o);
                qcol[sl] = keep ? gcol : 0u;
                qrow[sl] = keep ? (uint16_t)br : (uint16_t)0xFFFFu;
                qval[sl] = tv[0] + (uint32_t)((int)o - qt) * (uint32_t)vwords;
                o++;
            }
        }
    }
    return qn;
}

template <typename T, typename V, int K, int R, int P, bool SEQ, bool HOT>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(4, 4)))
spmm_lane_w(SpmmArgs a)
{
    using LC = LaneCfg<R>;
    constexpr int CW = LC::CW, DWIN = LC::DWIN, DPL = LC::DPL;
    constexpr int SEGP = K ? P : (1 << 30);
    constexpr int QS = CW + CW / (K ? P : CW) + 1;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const uint32_t wv = uni(threadIdx.x >> 6);
    const uint32_t item_id = blockIdx.x * kWPB + wv;
    if (item_id >= a.nitems) return;

    unsigned char *wbase = smem + wv * a.wave_bytes;
    using A = typename std::conditional<SEQ, T, typename AccOf<T>::type>::type;
    A *band = (A *)wbase;
    uint32_t *qcol = (uint32_t *)(wbase + a.band_bytes);
    uint32_t *qval = qcol + QS;
    uint16_t *qrow = (uint16_t *)(qval + QS);
    const uint32_t win_off = (a.band_bytes + QS * 10 + 7) & ~7u;
    uint2 *win = (uint2 *)(wbase + win_off);
    uint32_t *wvb = (uint32_t *)(wbase + win_off + (DWIN + 1) * 8);
    int32_t *wq = (int32_t *)(wbase + win_off + (DWIN + 1) * 12);

    ItemState st = item_state(a, a.items[item_id]);
    const RowScale rs = prefetch_rscale<T>(a, st.row0, st.ic.rows, lane);
    const uint32_t S = a.band_stride;
    for (uint32_t i = lane; i < st.ic.rows * S; i += 64) band[i] = A(0);
    if (lane == 0) {
        win[DWIN] = make_uint2(0xFFFFFFFFu, 0u);
        wvb[DWIN] = 0xFFFFFFFFu;
    }
    const uint16_t *vimg = a.img + st.ic.bw;

    for (int seg = 0; seg < 2; seg++) {
    if (!select_segment(st, seg)) break;
    const uint64_t abs_end = st.ic.bw + (uint64_t)st.ic.pend;
    const uint64_t lastw = abs_end - 1;
    uint64_t c = (st.ic.bw + (uint64_t)st.ic.begin) & ~7ull;
    int c0 = 0;  // columns of window tile 0 decoded in earlier chunks
    uint32_t wreg[R / 2];
    uint2 dreg[DPL];
    uint32_t vreg[DPL];
    bool live[DPL];
    load_words<R>(a.img, c, lastw, lane, wreg);
    load_window<DPL>(a.dir, st.ct, st.dlast, lane, dreg, live);
#pragma unroll
    for (int h = 0; h < DPL; h++) {
        const uint32_t t = st.ct + h * 64 + lane;
        vreg[h] = a.vdir[t < st.dlast ? t : st.dlast];
    }

    while (c < abs_end) {
        store_window<DPL>(win, dreg, live, lane);
#pragma unroll
        for (int h = 0; h < DPL; h++) wvb[h * 64 + lane] = live[h] ? vreg[h] : 0xFFFFFFFFu;
        wave_sync();
        const uint32_t qn = decode_chunk_w<R, SEGP>(st.ic, c, wreg, win, wvb, wq, c0, a.vwords, st.carry,
                                                    qcol, qrow, qval, lane);
        const uint64_t cn = c + CW;
        wave_sync();
        // the next chunk's words, window and value-directory entries (this
        // chunk's are in LDS now)
        auto fetch_next = [&]() {
            if (cn < abs_end) {
                const uint32_t pn = (uint32_t)((int64_t)cn - (int64_t)st.ic.bw);
                const uint32_t adv = window_advance<DPL>(dreg, live, pn);
                // the next chunk's window tile 0 is window tile adv-1; one
                // that starts exactly at the next chunk's first word has no
                // columns in this chunk, and its wq slot was not written by
                // this decode
                c0 = win[adv - 1].x == pn ? 0 : (int)qn - wq[adv - 1];
                st.ct += adv - 1;
                load_words<R>(a.img, cn, lastw, lane, wreg);
                load_window<DPL>(a.dir, st.ct, st.dlast, lane, dreg, live);
#pragma unroll
                for (int h = 0; h < DPL; h++) {
                    const uint32_t t = st.ct + h * 64 + lane;
                    vreg[h] = a.vdir[t < st.dlast ? t : st.dlast];
                }
            }
        };
        if constexpr (K == 0) fetch_next();
        if constexpr (K > 0) {
            // requested behind the first batch's gathers and value loads (as
            // in spmm_lane)
            using C = Lanes<T, K>;
            const int q = lane % C::L;
            const int g = lane / C::L;
            bool fetched = false;
            for (uint32_t base = 0; base < qn; base += C::NZP * P) {
                Vec<T, C::E> v[P];
                uint32_t rr[P], gcs[P];
                V vals[P];
#pragma unroll
                for (int pp = 0; pp < P; pp++) {
                    const uint32_t e = base + g * P + pp;
                    const uint32_t sl = qslot<P>(e);
                    const bool ok = e < qn;
                    const uint32_t qr = qrow[sl];
                    gcs[pp] = ok ? qcol[sl] : 0u;
                    rr[pp] = (ok && qr != 0xFFFFu) ? qr : 0xFFFFFFFFu;
                    const uint32_t vo = ok ? qval[sl] : 0u;  // a safe address when unused
                    vals[pp] = load_value<V>(vimg + vo);
                }
#pragma unroll
                for (int pp = 0; pp < P; pp++)
                    v[pp] = *(const Vec<T, C::E> *)(xrow<T, HOT>(a, gcs[pp]) + q * C::E);
                if (!fetched) {
                    fetch_next();
                    fetched = true;
                }
                if constexpr (SEQ) {
                    // products rounded in T, then added in order
#pragma clang fp contract(off)
#pragma unroll
                    for (int pp = 0; pp < P; pp++)
#pragma unroll
                        for (int i = 0; i < C::E; i++) v[pp].v[i] = v[pp].v[i] * (T)vals[pp];
                    accumulate_batch_seq<T, K, P>(band, S, lane, v, rr);
                    continue;
                }
                A acc[C::E];
                uint32_t cur = 0xFFFFFFFFu;
#pragma unroll
                for (int i = 0; i < C::E; i++) acc[i] = A(0);
#pragma unroll
                for (int pp = 0; pp < P; pp++) {
                    if (rr[pp] != cur) {
                        if (cur != 0xFFFFFFFFu) {
                            A *dst = band + cur * S + q;
#pragma unroll
                            for (int i = 0; i < C::E; i++) lds_add(dst + i * C::L, acc[i]);
                        }
                        cur = rr[pp];
#pragma unroll
                        for (int i = 0; i < C::E; i++) acc[i] = (A)v[pp].v[i] * (A)vals[pp];
                    } else {
#pragma unroll
                        for (int i = 0; i < C::E; i++) acc[i] += (A)v[pp].v[i] * (A)vals[pp];
                    }
                }
                if (cur != 0xFFFFFFFFu) {
                    A *dst = band + cur * S + q;
#pragma unroll
                    for (int i = 0; i < C::E; i++) lds_add(dst + i * C::L, acc[i]);
                }
            }
            if (!fetched) fetch_next();
        } else if (SEQ) {
            if (qn) drain_generic_seq<T, V, P, true>(a, (T *)band, S, qcol, qrow, qval, vimg, qn, lane);
        } else {
            const uint32_t k = (uint32_t)a.k;
            const uint32_t tot = qn * k;
            for (uint32_t base = 0; base < tot; base += 64 * P) {
                T xv[P];
                V vv[P];
                uint32_t dst[P];
#pragma unroll
                for (int pp = 0; pp < P; pp++) {
                    const uint32_t f = base + pp * 64 + lane;
                    const uint32_t fc = f < tot ? f : tot - 1;
                    const uint32_t e = fc / k, jj = fc - e * k;
                    const uint32_t qr = qrow[e];
                    xv[pp] = xrow<T, HOT>(a, qcol[e])[jj];
                    vv[pp] = load_value<V>(vimg + qval[e]);
                    dst[pp] = (f < tot && qr != 0xFFFFu) ? qr * S + jj : 0xFFFFFFFFu;
                }
#pragma unroll
                for (int pp = 0; pp < P; pp++)
                    if (dst[pp] != 0xFFFFFFFFu) lds_add(band + dst[pp], (A)xv[pp] * (A)vv[pp]);
            }
        }
        wave_sync();
        c = cn;
    }
    }
    write_band<T, K, A>(a, band, S, st.row0, st.ic.rows, st.slot, lane, rs, (double *)qcol);
}

// Y rows of a split band = partial slots summed in stream order.
template <typename T, typename A>
__global__ void __launch_bounds__(256) spmm_reduce(const Split *splits, const A *partial,
                                                    uint32_t hs, uint32_t k, T *Y, int64_t ldy,
                                                    const double *rscale, int ycol)
{
    const Split sp = splits[blockIdx.x];
    const uint32_t tot = sp.rows * k;
    for (uint32_t i = threadIdx.x; i < tot; i += blockDim.x) {
        A acc = partial[(uint64_t)sp.slot_begin * hs * k + i];
        for (uint32_t s = 1; s < sp.nslots; s++)
            acc += partial[(uint64_t)(sp.slot_begin + s) * hs * k + i];
        uint32_t r = i / k, c = i - r * k;
        if constexpr (std::is_floating_point<T>::value)
            if (rscale) acc = (A)((double)acc * rscale[sp.row0 + r]);
        if (ycol) Y[c * (uint64_t)ldy + sp.row0 + r] = (T)acc;
        else Y[(sp.row0 + r) * (uint64_t)ldy + c] = (T)acc;
    }
}

// Gathers in flight per lane group.  Eight 16-byte gathers per lane keep
// spmm_lane under 80 VGPRs, so the LDS of its band and queue, not registers,
// bounds the waves per CU.  More resident waves with short batches put more
// gathers in flight than fewer waves with long ones: on R-MAT scale 22, FP32
// k = 16, 64x32768 tiles, 0.93 ms against 1.30 ms for 512-entry batches
// overlapped with the next decode at two waves per SIMD (the best earlier
// schedule), and 57 ms against 92 ms on scale 27, FP32 k = 32.
template <typename T, int K>
constexpr int passes_for()
{
    if constexpr (K == 0) return 8;
    else return 512 / Lanes<T, K>::NZP < 8 ? 512 / Lanes<T, K>::NZP : 8;
}

constexpr int kLaneR = 8;  // words per lane: 512-word chunks

// Raises a kernel's dynamic LDS limit once per process.
template <void (*KERN)(SpmmArgs)>
fmx_status allow_lds(fmx_ctx *ctx)
{
    static bool done = false;
    if (!done) {
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)KERN, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         160 * 1024));
        done = true;
    }
    return FMX_OK;
}

template <void (*KERN)(SpmmArgs)>
fmx_status launch_items(fmx_ctx *ctx, const SpmmArgs &a, hipStream_t stream)
{
    const uint32_t lds = a.wave_bytes * kWPB;
    if (lds > ctx->lds_per_block) return fail(ctx, FMX_EUNSUPPORTED, "band does not fit in LDS");
    fmx_status st = allow_lds<KERN>(ctx);
    if (st) return st;
    const uint32_t blocks = (a.nitems + kWPB - 1) / kWPB;
    if (blocks) {
        hipLaunchKernelGGL(KERN, dim3(blocks), dim3(64 * kWPB), lds, stream, a);
        FMX_HIP(ctx, hipGetLastError());
    }
    return FMX_OK;
}

// Weighted images; V is the type of the stored values (T, or float values
// with FP64 dense data: spmm_creator<double, float>).
template <typename T, typename V, int K, bool SEQ, bool HOT>
fmx_status launch_kw(fmx_ctx *ctx, SpmmArgs a, hipStream_t stream)
{
    constexpr int P = passes_for<T, K>();
    constexpr int R = kLaneR;
    constexpr int CW = 64 * R;
    constexpr int QS = CW + CW / (K ? P : CW) + 1;
    const uint32_t win_off = (a.band_bytes + QS * 10 + 7) & ~7u;
    a.wave_bytes = (win_off + (LaneCfg<R>::DWIN + 1) * 16 + 15) & ~15u;
    return launch_items<spmm_lane_w<T, V, K, R, P, SEQ, HOT>>(ctx, a, stream);
}

template <typename T, int K, bool SEQ, bool HOT>
fmx_status launch_k(fmx_ctx *ctx, SpmmArgs a, int vprim, hipStream_t stream)
{
    if (a.vdir) {
        if constexpr (std::is_same<T, double>::value)
            if (vprim == 8) return launch_kw<double, float, K, SEQ, HOT>(ctx, a, stream);
        return launch_kw<T, T, K, SEQ, HOT>(ctx, a, stream);
    }
    constexpr int P = passes_for<T, K>();
    constexpr int CW = 64 * kLaneR;
    constexpr int QS = CW + CW / (K ? P : CW) + 1;
    const uint32_t win_off = (a.band_bytes + QS * 6 + 7) & ~7u;
    a.wave_bytes = (win_off + (LaneCfg<kLaneR>::DWIN + 1) * 8 + 15) & ~15u;
    return launch_items<spmm_lane<T, K, kLaneR, P, SEQ, HOT>>(ctx, a, stream);
}

// HOT (a gather-order image, rows narrower than a 128-byte line) has kernels
// for the widths it is used with; other widths take the runtime-width form.
template <typename T, bool SEQ, bool HOT>
fmx_status launch_w(fmx_ctx *ctx, SpmmArgs a, int k, int vprim, hipStream_t stream)
{
    constexpr int KMAX = HOT ? (int)(64 / sizeof(T)) : 64;
    switch (k) {
    case 1: return launch_k<T, 1, SEQ, HOT>(ctx, a, vprim, stream);
    case 2: return launch_k<T, 2, SEQ, HOT>(ctx, a, vprim, stream);
    case 4: return launch_k<T, 4, SEQ, HOT>(ctx, a, vprim, stream);
    case 8: return launch_k<T, 8, SEQ, HOT>(ctx, a, vprim, stream);
    case 16: return launch_k<T, (16 <= KMAX ? 16 : 0), SEQ, HOT>(ctx, a, vprim, stream);
    case 32: return launch_k<T, (32 <= KMAX ? 32 : 0), SEQ, HOT>(ctx, a, vprim, stream);
    case 64: return launch_k<T, (64 <= KMAX ? 64 : 0), SEQ, HOT>(ctx, a, vprim, stream);
    default: return launch_k<T, 0, SEQ, HOT>(ctx, a, vprim, stream);
    }
}

template <typename T>
fmx_status launch_t(fmx_ctx *ctx, const fmx_spm *spm, const Plan *plan, int k,
                    const void *X, int64_t ldx, void *Y, int64_t ldy, void *partial,
                    hipStream_t stream, const uint16_t *img, uint32_t item_begin, uint32_t item_end,
                    bool reduce, const double *rscale, uint32_t flags, const void *Xhot, uint64_t nhot)
{
    // integer sums are exact in any order: the sequential mode is the default one
    const bool seq = (flags & FMX_SPMM_SEQ) && std::is_floating_point<T>::value;
    SpmmArgs a;
    a.img = img;
    a.dir = spm->d_dir;
    a.items = plan->d_items + item_begin;
    a.X = X;
    a.Y = Y;
    a.partial = partial;
    a.ldx = ldx;
    a.ldy = ldy;
    a.nitems = item_end - item_begin;
    a.hs = plan->hs;
    // pad band rows by one element to spread the LDS adds over the banks
    static const bool no_pad = getenv("FMX_NO_PAD") != nullptr;  // read once per process
    a.band_stride = (uint32_t)k + ((k >= 4 && !no_pad) ? 1u : 0u);
    const uint64_t acc = seq ? sizeof(T) : sizeof(typename AccOf<T>::type);
    a.band_bytes = (uint32_t)(((uint64_t)plan->hs * a.band_stride * acc + 15) & ~15ull);
    a.wave_bytes = 0;
    a.log2w = spm->log2w;
    a.k = k;
    a.ncols = spm->ncols;
    a.vdir = spm->esz ? spm->d_vdir : nullptr;
    a.vwords = (int)(spm->esz / 2);
    a.rscale = rscale;
    a.ycol = (flags & FMX_SPMM_Y_COLMAJOR) ? 1 : 0;
    a.Xhot = Xhot;
    a.nhot = Xhot ? nhot : 0;
    // 16-byte vector gathers need aligned rows; otherwise the runtime-width path.
    const int vb = (k * (int)sizeof(T) < 16) ? k * (int)sizeof(T) : 16;
    const bool aligned = ((uintptr_t)X % vb == 0) && ((ldx * (int64_t)sizeof(T)) % vb == 0) &&
                         ((uintptr_t)Xhot % vb == 0) && ((k * (int)sizeof(T)) % vb == 0);
    const int kk = aligned ? k : 0;
    fmx_status st;
    if (a.nhot)
        st = launch_w<T, false, true>(ctx, a, kk, spm->prim, stream);  // never with the sequential mode
    else if constexpr (std::is_floating_point<T>::value)
        st = seq ? launch_w<T, true, false>(ctx, a, kk, spm->prim, stream)
                 : launch_w<T, false, false>(ctx, a, kk, spm->prim, stream);
    else
        st = launch_w<T, false, false>(ctx, a, kk, spm->prim, stream);
    if (st) return st;
    if (reduce && plan->nsplits) {
        using A = typename AccOf<T>::type;
        hipLaunchKernelGGL((spmm_reduce<T, A>), dim3(plan->nsplits), dim3(256), 0, stream,
                           plan->d_splits, (const A *)partial, plan->hs, (uint32_t)k, (T *)Y, ldy, rscale, a.ycol);
        FMX_HIP(ctx, hipGetLastError());
    }
    return FMX_OK;
}

}  // namespace

// Row-major copy of a dense operand, optionally scaled by pre[i] per row
// (one rounding): out[i * k + j] = X(i, j) * pre[i].  A column-major X goes
// through a 32 x 64 LDS tile so that both the reads (down columns) and the
// writes (along rows) are consecutive.
template <typename T>
__global__ void __launch_bounds__(256) stage_dense(const T *X, int64_t ldx, int xcol, const double *pre,
                                                    uint64_t n, uint32_t k, T *out)
{
    __shared__ T tile[32][65];
    const uint64_t r0 = (uint64_t)blockIdx.x * 64;
    const uint32_t c0 = blockIdx.y * 32;
    const uint32_t tid = threadIdx.x;
    auto put = [&](uint64_t r, uint32_t c, T v) {
        if constexpr (std::is_floating_point<T>::value)
            if (pre) v = (T)((double)v * pre[r]);
        out[r * k + c] = v;
    };
    if (xcol) {
        for (uint32_t cc = tid / 64; cc < 32; cc += 4) {
            const uint64_t r = r0 + tid % 64;
            const uint32_t c = c0 + cc;
            if (r < n && c < k) tile[cc][tid % 64] = X[c * (uint64_t)ldx + r];
        }
        __syncthreads();
        for (uint32_t rr = tid / 32; rr < 64; rr += 8) {
            const uint64_t r = r0 + rr;
            const uint32_t c = c0 + tid % 32;
            if (r < n && c < k) put(r, c, tile[tid % 32][rr]);
        }
    } else {
        for (uint32_t i = tid; i < 64 * 32; i += 256) {
            const uint64_t r = r0 + i / 32;
            const uint32_t c = c0 + i % 32;
            if (r < n && c < k) put(r, c, X[r * (uint64_t)ldx + c]);
        }
    }
}

fmx_status stage_dense_rows(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int k, const void *X, int64_t ldx, bool xcol,
                            const double *pre, void *out, hipStream_t stream)
{
    if (n == 0) return FMX_OK;
    const dim3 grid((unsigned)((n + 63) / 64), (unsigned)((k + 31) / 32));
    switch (dt) {
    case FMX_I32:
        hipLaunchKernelGGL(stage_dense<int32_t>, grid, dim3(256), 0, stream, (const int32_t *)X, ldx, (int)xcol,
                           nullptr, n, (uint32_t)k, (int32_t *)out);
        break;
    case FMX_I64:
    case FMX_U64:
        hipLaunchKernelGGL(stage_dense<int64_t>, grid, dim3(256), 0, stream, (const int64_t *)X, ldx, (int)xcol,
                           nullptr, n, (uint32_t)k, (int64_t *)out);
        break;
    case FMX_F32:
        hipLaunchKernelGGL(stage_dense<float>, grid, dim3(256), 0, stream, (const float *)X, ldx, (int)xcol, pre,
                           n, (uint32_t)k, (float *)out);
        break;
    case FMX_F64:
        hipLaunchKernelGGL(stage_dense<double>, grid, dim3(256), 0, stream, (const double *)X, ldx, (int)xcol, pre,
                           n, (uint32_t)k, (double *)out);
        break;
    default: return fail(ctx, FMX_ETYPE, "wrong matrix type in spmm");
    }
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

// out[i][0..k) = X[hot[i]][0..k): rows of V-sized pieces, a row's pieces on
// consecutive lanes.
template <typename V>
__global__ void __launch_bounds__(256) stage_hot(const V *X, uint64_t ldx_v, const uint32_t *hot, uint64_t nhot,
                                                  uint32_t kv, V *out)
{
    const uint64_t tot = nhot * kv;
    for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * 256) {
        const uint64_t i = e / kv, j = e - i * kv;
        out[e] = X[(uint64_t)hot[i] * ldx_v + j];
    }
}

fmx_status stage_hot_rows(fmx_ctx *ctx, fmx_dtype dt, uint64_t nhot, int k, const void *X, int64_t ldx,
                          const uint32_t *hot, void *out, hipStream_t stream)
{
    if (nhot == 0) return FMX_OK;
    const uint64_t rb = (uint64_t)k * elem_size(dt), lb = (uint64_t)ldx * elem_size(dt);
    const uint64_t blocks = std::min<uint64_t>((nhot * k + 255) / 256, (uint64_t)ctx->num_cus * 32);
    // the widest piece that divides the row, the stride and the base
    if (rb % 16 == 0 && lb % 16 == 0 && (uintptr_t)X % 16 == 0)
        hipLaunchKernelGGL(stage_hot<uint4>, dim3((unsigned)blocks), dim3(256), 0, stream, (const uint4 *)X, lb / 16,
                           hot, nhot, (uint32_t)(rb / 16), (uint4 *)out);
    else if (rb % 8 == 0 && lb % 8 == 0 && (uintptr_t)X % 8 == 0)
        hipLaunchKernelGGL(stage_hot<uint2>, dim3((unsigned)blocks), dim3(256), 0, stream, (const uint2 *)X, lb / 8,
                           hot, nhot, (uint32_t)(rb / 8), (uint2 *)out);
    else
        hipLaunchKernelGGL(stage_hot<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, stream, (const uint32_t *)X,
                           lb / 4, hot, nhot, (uint32_t)(rb / 4), (uint32_t *)out);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status launch_spmm(fmx_ctx *ctx, const fmx_spm *spm, const Plan *plan, fmx_dtype dt,
                       int k, const void *X, int64_t ldx, void *Y, int64_t ldy,
                       void *partial, hipStream_t stream, const uint16_t *img,
                       uint32_t ib, uint32_t ie, bool reduce, const double *rscale, uint32_t flags,
                       const void *Xhot, uint64_t nhot)
{
    if (rscale && dt != FMX_F32 && dt != FMX_F64) return fail(ctx, FMX_ETYPE, "row scaling needs FP data");
    switch (dt) {
    case FMX_I32: return launch_t<int32_t>(ctx, spm, plan, k, X, ldx, Y, ldy, partial, stream, img, ib, ie, reduce, rscale, flags, Xhot, nhot);
    // size_t sums wrap exactly as long sums do: one kernel serves both
    case FMX_I64:
    case FMX_U64: return launch_t<int64_t>(ctx, spm, plan, k, X, ldx, Y, ldy, partial, stream, img, ib, ie, reduce, rscale, flags, Xhot, nhot);
    case FMX_F32: return launch_t<float>(ctx, spm, plan, k, X, ldx, Y, ldy, partial, stream, img, ib, ie, reduce, rscale, flags, Xhot, nhot);
    case FMX_F64: return launch_t<double>(ctx, spm, plan, k, X, ldx, Y, ldy, partial, stream, img, ib, ie, reduce, rscale, flags, Xhot, nhot);
    default: return fail(ctx, FMX_ETYPE, "wrong matrix type in spmm");
    }
}

}  // namespace fmx
