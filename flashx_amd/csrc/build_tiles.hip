// build_tiles.hip -- on-device construction of the reference's 2D tile image
// from an edge list (fm::create_2d_matrix, matrix/fm_utils.cpp:617-715,
// 726-841), bit-exact with the reference writer.
//
// The reference sorts the edge list by source, groups rows into block rows,
// sorts each row's neighbours, optionally de-duplicates, and then for every
// block row sizes (collect_block_info, fm_utils.cpp:267-327) and serialises
// (run_blocks/add_nz, fm_utils.cpp:228-265,329-433) its tiles.  Here the same
// image comes out of data-parallel steps:
//   1. keys = (block row, block col, row in tile, col in tile) packed in a u64,
//      self edges dropped unless kept (fm_utils.cpp:450-452), both directions
//      when symmetrising;
//   2. stable LSD radix sort of the keys (tile order = image order), then
//      std::unique-style compaction when de-duplicating (fm_utils.cpp:460-464);
//   3. runs = maximal groups of one row inside one tile; a run of one non-zero
//      is a COO entry, a longer run an SCSR row part (add_nz, fm_utils.cpp:
//      235-264);
//   4. per tile: nnz, nrow, num_coo, SCSR words -> tile size
//      (sparse_block_2d::get_size, sparse_matrix_format.h:294-299); per block
//      row: size or a 16-byte empty tile (fm_utils.cpp:341-351); offsets by
//      exclusive scans (fm_utils.cpp:701-712);
//   5. every key writes its own u16 words; every tile writes its header and
//      end marker (sparse_block_2d::finalize, sparse_matrix_format.cpp:119-126).
#include <algorithm>
#include <cstring>

#include "devprim.h"

namespace fmx {
namespace {

constexpr int kB = 256;

struct KeyFmt {
    int lh, lw, bbc;   // log2 H, log2 W, bits of the block-column field
    uint32_t H, W;
    __host__ __device__ uint64_t key(uint32_t s, uint32_t d) const
    {
        return ((uint64_t)(s >> lh) << (bbc + lh + lw)) | ((uint64_t)(d >> lw) << (lh + lw)) |
               ((uint64_t)(s & (H - 1)) << lw) | (uint64_t)(d & (W - 1));
    }
    __host__ __device__ uint32_t br(uint64_t k) const { return (uint32_t)(k >> (bbc + lh + lw)); }
    __host__ __device__ uint32_t bc(uint64_t k) const
    {
        return (uint32_t)((k >> (lh + lw)) & ((1ull << bbc) - 1));
    }
    __host__ __device__ uint32_t rrel(uint64_t k) const { return (uint32_t)((k >> lw) & (H - 1)); }
    __host__ __device__ uint32_t crel(uint64_t k) const { return (uint32_t)(k & (W - 1)); }
};

__device__ __forceinline__ uint64_t gtid() { return blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

// One reservation per block and 8 edges per thread keep the global counter
// cold; the order of keys is irrelevant (they are sorted next).
constexpr int kEPT = 8;

__global__ void __launch_bounds__(kB) make_keys(const uint32_t *src, const uint32_t *dst, uint64_t m,
                                                 KeyFmt f, int keep_self, int sym, uint64_t *keys,
                                                 unsigned long long *counter)
{
    __shared__ uint32_t wsum[kB / 64];
    __shared__ unsigned long long bbase;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint64_t base = blockIdx.x * (uint64_t)kB * kEPT; base < m;
         base += (uint64_t)gridDim.x * kB * kEPT) {
        uint32_t s[kEPT], d[kEPT];
        bool ok[kEPT];
        uint32_t cnt = 0;
#pragma unroll
        for (int t = 0; t < kEPT; t++) {
            const uint64_t i = base + (uint64_t)t * kB + threadIdx.x;
            s[t] = 0;
            d[t] = 0;
            ok[t] = false;
            if (i < m) {
                s[t] = src[i];
                d[t] = dst[i];
                ok[t] = keep_self || s[t] != d[t];
            }
            cnt += ok[t] ? ((sym && s[t] != d[t]) ? 2u : 1u) : 0u;
        }
        uint32_t inc = cnt;
#pragma unroll
        for (int dd = 1; dd < 64; dd <<= 1) {
            uint32_t o = __shfl_up(inc, dd, 64);
            if (lane >= dd) inc += o;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t woff = 0, btot = 0;
#pragma unroll
        for (int q = 0; q < kB / 64; q++) {
            if (q < w) woff += wsum[q];
            btot += wsum[q];
        }
        if (threadIdx.x == 0) bbase = btot ? atomicAdd(counter, (unsigned long long)btot) : 0ull;
        __syncthreads();
        uint64_t pos = bbase + woff + inc - cnt;
#pragma unroll
        for (int t = 0; t < kEPT; t++) {
            if (!ok[t]) continue;
            keys[pos++] = f.key(s[t], d[t]);
            if (sym && s[t] != d[t]) keys[pos++] = f.key(d[t], s[t]);
        }
        __syncthreads();
    }
}

// flag[i] = 1 when keys[i] starts a new group under `shift` (keys >> shift)
__global__ void __launch_bounds__(kB) group_flags(const uint64_t *k, uint64_t n, int shift, uint32_t *flag)
{
    for (uint64_t i = gtid(); i < n; i += gstride())
        flag[i] = (i == 0 || (k[i] >> shift) != (k[i - 1] >> shift)) ? 1u : 0u;
}

__global__ void __launch_bounds__(kB) compact_unique(const uint64_t *k, uint64_t n, const uint32_t *ex,
                                                     uint64_t *out)
{
    for (uint64_t i = gtid(); i < n; i += gstride())
        if (i == 0 || k[i] != k[i - 1]) out[ex[i]] = k[i];
}

// Weighted builds keep input order among equal keys (the first value of a
// duplicate is kept), so keys are emitted in edge order: cnt[i] = keys of
// edge i, then after an exclusive scan each edge writes its keys at its
// offset with the edge index as payload.
__global__ void __launch_bounds__(kB) count_keys_w(const uint32_t *src, const uint32_t *dst, uint64_t m,
                                                   int keep_self, int sym, uint32_t *cnt)
{
    for (uint64_t i = gtid(); i < m; i += gstride()) {
        const uint32_t a = src[i], b = dst[i];
        const bool ok = keep_self || a != b;
        cnt[i] = ok ? ((sym && a != b) ? 2u : 1u) : 0u;
    }
}

__global__ void __launch_bounds__(kB) emit_keys_w(const uint32_t *src, const uint32_t *dst, uint64_t m,
                                                  KeyFmt f, int keep_self, int sym, const uint32_t *pos,
                                                  uint64_t *keys, uint32_t *eidx)
{
    for (uint64_t i = gtid(); i < m; i += gstride()) {
        const uint32_t a = src[i], b = dst[i];
        if (!keep_self && a == b) continue;
        const uint32_t p = pos[i];
        keys[p] = f.key(a, b);
        eidx[p] = (uint32_t)i;
        if (sym && a != b) {
            keys[p + 1] = f.key(b, a);
            eidx[p + 1] = (uint32_t)i;
        }
    }
}

__global__ void __launch_bounds__(kB) compact_unique_kv(const uint64_t *k, const uint32_t *v, uint64_t n,
                                                        const uint32_t *ex, uint64_t *out, uint32_t *vout)
{
    for (uint64_t i = gtid(); i < n; i += gstride())
        if (i == 0 || k[i] != k[i - 1]) {
            out[ex[i]] = k[i];
            vout[ex[i]] = v[i];
        }
}

// run_start[r] = first key of run r; run_start[R] = nk
__global__ void __launch_bounds__(kB) fill_run_start(const uint64_t *k, uint64_t nk, int lw,
                                                     const uint32_t *ex, uint32_t *run_start,
                                                     uint32_t R)
{
    for (uint64_t i = gtid(); i < nk; i += gstride())
        if (i == 0 || (k[i] >> lw) != (k[i - 1] >> lw)) run_start[ex[i]] = (uint32_t)i;
    if (gtid() == 0) run_start[R] = (uint32_t)nk;
}

// per run: tile-start flag, SCSR words, COO flag
__global__ void __launch_bounds__(kB) run_props(const uint64_t *k, const uint32_t *run_start, uint32_t R,
                                                int tshift, uint32_t *tflag, uint64_t *scsr_w,
                                                uint32_t *coo_f)
{
    for (uint64_t r = gtid(); r < R; r += gstride()) {
        const uint32_t s0 = run_start[r], len = run_start[r + 1] - s0;
        tflag[r] = (r == 0 || (k[s0] >> tshift) != (k[run_start[r - 1]] >> tshift)) ? 1u : 0u;
        scsr_w[r] = len >= 2 ? (uint64_t)len + 1 : 0;
        coo_f[r] = len == 1 ? 1u : 0u;
    }
    if (gtid() == 0) {
        scsr_w[R] = 0;
        coo_f[R] = 0;
    }
}

// tile_first_run[t], tile_of_run[r]; tile_first_run[T] = R
__global__ void __launch_bounds__(kB) fill_tiles(const uint64_t *k, const uint32_t *run_start, uint32_t R,
                                                 int tshift, const uint32_t *tex,
                                                 uint32_t *tile_first_run, uint32_t *tile_of_run,
                                                 uint32_t T)
{
    for (uint64_t r = gtid(); r < R; r += gstride()) {
        const bool first = r == 0 || (k[run_start[r]] >> tshift) != (k[run_start[r - 1]] >> tshift);
        const uint32_t t = tex[r] + (first ? 1u : 0u) - 1u;
        tile_of_run[r] = t;
        if (first) tile_first_run[t] = (uint32_t)r;
    }
    if (gtid() == 0) tile_first_run[T] = R;
}

__global__ void __launch_bounds__(kB) tile_props(const uint32_t *run_start, const uint32_t *tile_first_run,
                                                 uint32_t T, uint64_t *tsize, uint32_t esz)
{
    for (uint64_t t = gtid(); t < T; t += gstride()) {
        const uint32_t fr = tile_first_run[t], lr = tile_first_run[t + 1];
        const uint64_t nnz = run_start[lr] - run_start[fr];
        const uint64_t nrow = lr - fr;
        tsize[t] = 16 + 2 * (nrow + nnz + 1) + (uint64_t)esz * nnz;
    }
    if (gtid() == 0) tsize[T] = 0;
}

// Per block row with tiles (tiles are in block-row order): its bytes, from
// the scanned tile sizes, and its directory entries, written by the row's
// last tile.
__global__ void __launch_bounds__(kB) row_totals(const uint64_t *k, KeyFmt f, const uint32_t *run_start,
                                                 const uint32_t *tile_first_run, uint32_t T,
                                                 const uint64_t *size_ex, const uint32_t *br_first,
                                                 unsigned long long *brsize, uint32_t *ndir)
{
    for (uint64_t t = gtid(); t < T; t += gstride()) {
        const uint32_t b = f.br(k[run_start[tile_first_run[t]]]);
        if (t + 1 < T && f.br(k[run_start[tile_first_run[t + 1]]]) == b) continue;
        brsize[b] = size_ex[t + 1] - size_ex[br_first[b]];
        ndir[b] = (uint32_t)(t + 1 - br_first[b]);
    }
}

// empty block rows take a 16-byte empty tile (one directory entry); keep a
// copy of "had tiles"
__global__ void __launch_bounds__(kB) fix_empty_rows(unsigned long long *brsize, uint64_t nbr,
                                                     uint8_t *had, uint32_t *ndir)
{
    for (uint64_t b = gtid(); b < nbr; b += gstride()) {
        had[b] = brsize[b] != 0;
        if (!brsize[b]) {
            brsize[b] = 16;
            ndir[b] = 1;
        }
    }
    if (gtid() == 0) {
        brsize[nbr] = 0;
        ndir[nbr] = 0;
    }
}

__global__ void __launch_bounds__(kB) first_tile_of_row(const uint64_t *k, KeyFmt f,
                                                        const uint32_t *run_start,
                                                        const uint32_t *tile_first_run, uint32_t T,
                                                        uint32_t *br_first)
{
    for (uint64_t t = gtid(); t < T; t += gstride()) {
        const uint32_t b = f.br(k[run_start[tile_first_run[t]]]);
        if (t == 0 || f.br(k[run_start[tile_first_run[t - 1]]]) != b) br_first[b] = (uint32_t)t;
    }
}

// The tile directory the SpMM plans and kernels read, straight from the
// construction (a resident image built here is never copied to the host):
// per block row its first directory index (bdir, exclusive counts: its
// tiles, or 1 for the empty tile of an empty block row), per directory entry
// the tile's word offset inside its block row and its block column, and for
// weighted images the word where the tile's values start.
__global__ void __launch_bounds__(kB) fill_dir(const uint64_t *k, KeyFmt f, const uint32_t *run_start,
                                               const uint32_t *tile_first_run, uint32_t T, const uint64_t *offs_ex,
                                               const uint32_t *br_first, const uint64_t *tile_off,
                                               const uint32_t *bdir, const uint8_t *had, uint64_t nbr, int weighted,
                                               uint2 *dir, uint32_t *vdir)
{
    for (uint64_t t = gtid(); t < T; t += gstride()) {
        const uint32_t fr = tile_first_run[t], lr = tile_first_run[t + 1];
        const uint64_t key0 = k[run_start[fr]];
        const uint32_t b = f.br(key0);
        const uint32_t x = (uint32_t)((tile_off[t] - 4096 - offs_ex[b]) / 2);
        const uint32_t i = bdir[b] + (uint32_t)(t - br_first[b]);
        dir[i] = make_uint2(x, f.bc(key0));
        if (weighted) vdir[i] = x + 8 + (lr - fr) + (run_start[lr] - run_start[fr]) + 1;
    }
    for (uint64_t b = gtid(); b < nbr; b += gstride())
        if (!had[b]) {
            dir[bdir[b]] = make_uint2(0, 0);
            if (weighted) vdir[bdir[b]] = 9;  // the empty tile: 8 header words, no rows, no entries, end marker
        }
}

// The directory's invariants, per block row (the SpMM plans and kernels
// trust them): it has tiles, the first at word 0, word offsets strictly
// increasing and leaving room for a tile header inside the block row, block
// columns increasing and in range, value words (weighted) past their tile's
// header (one past the block row for the empty tile of an empty block row).
__global__ void __launch_bounds__(kB) check_dir(const uint2 *dir, const uint32_t *vdir, const uint32_t *bdir,
                                                const uint64_t *offs_ex, uint64_t nbr, uint64_t nbc, uint64_t ndir,
                                                int weighted, unsigned int *bad)
{
    for (uint64_t b = gtid(); b < nbr; b += gstride()) {
        const uint32_t lo = bdir[b], hi = bdir[b + 1];
        const int64_t words = (int64_t)(offs_ex[b + 1] - offs_ex[b]) / 2;
        bool ok = hi > lo && hi <= ndir && words >= 8 && dir[lo].x == 0;
        for (uint32_t i = lo; ok && i < hi; i++) {
            const uint2 d = dir[i];
            ok = (int64_t)d.x + 8 <= words && (uint64_t)d.y < (nbc ? nbc : 1) &&
                 (i == lo || (d.x > dir[i - 1].x && d.y > dir[i - 1].y)) &&
                 (!weighted || (vdir[i] >= d.x + 8 && (int64_t)vdir[i] <= words + 1));
        }
        if (!ok) atomicOr(bad, 1u);
    }
}

__device__ __forceinline__ void put16(uint8_t *img, uint64_t off, uint32_t v)
{
    *(uint16_t *)(img + off) = (uint16_t)v;
}

__global__ void __launch_bounds__(kB) write_tiles(const uint64_t *k, KeyFmt f, const uint32_t *run_start,
                                                  const uint32_t *tile_first_run, uint32_t T,
                                                  const uint64_t *scsr_ex, const uint32_t *coo_ex,
                                                  const uint64_t *offs_ex, const uint64_t *size_ex,
                                                  const uint32_t *br_first, uint64_t *tile_off,
                                                  uint8_t *img)
{
    for (uint64_t t = gtid(); t < T; t += gstride()) {
        const uint32_t fr = tile_first_run[t], lr = tile_first_run[t + 1];
        const uint64_t key0 = k[run_start[fr]];
        const uint32_t b = f.br(key0), c = f.bc(key0);
        const uint64_t off = 4096 + offs_ex[b] + size_ex[t] - size_ex[br_first[b]];
        tile_off[t] = off;
        const uint32_t nnz = run_start[lr] - run_start[fr];
        const uint32_t nrow = lr - fr;
        const uint32_t ncoo = coo_ex[lr] - coo_ex[fr];
        const uint64_t sw = scsr_ex[lr] - scsr_ex[fr];
        put16(img, off + 0, b & 0xFFFF);
        put16(img, off + 2, b >> 16);
        put16(img, off + 4, c & 0xFFFF);
        put16(img, off + 6, c >> 16);
        put16(img, off + 8, nnz & 0xFFFF);
        put16(img, off + 10, nnz >> 16);
        put16(img, off + 12, nrow);
        put16(img, off + 14, ncoo);
        put16(img, off + 16 + 2 * sw, 0xFFFF);
    }
}

__global__ void __launch_bounds__(kB) write_stubs(const uint8_t *had, uint64_t nbr, const uint64_t *offs_ex,
                                                  uint8_t *img)
{
    for (uint64_t b = gtid(); b < nbr; b += gstride()) {
        if (had[b]) continue;
        const uint64_t off = 4096 + offs_ex[b];
        put16(img, off + 0, (uint32_t)b & 0xFFFF);
        put16(img, off + 2, (uint32_t)(b >> 16));
        for (int i = 4; i < 16; i += 2) put16(img, off + i, 0);
    }
}

__global__ void __launch_bounds__(kB) write_nz(const uint64_t *k, uint64_t nk, KeyFmt f,
                                               const uint32_t *rex, const uint32_t *run_start,
                                               const uint32_t *tile_of_run,
                                               const uint32_t *tile_first_run,
                                               const uint64_t *scsr_ex, const uint32_t *coo_ex,
                                               const uint64_t *tile_off, uint8_t *img, const uint32_t *eidx,
                                               const uint8_t *vals, uint32_t esz)
{
    for (uint64_t i = gtid(); i < nk; i += gstride()) {
        const uint64_t key = k[i];
        const bool first = i == 0 || (key >> f.lw) != (k[i - 1] >> f.lw);
        const uint32_t r = rex[i] + (first ? 1u : 0u) - 1u;
        const uint32_t t = tile_of_run[r];
        const uint32_t fr = tile_first_run[t], lr = tile_first_run[t + 1];
        const uint32_t s0 = run_start[r], len = run_start[r + 1] - s0;
        const uint64_t off = tile_off[t];
        const uint64_t sw = scsr_ex[lr] - scsr_ex[fr];  // SCSR words (heads and columns) of the tile
        uint64_t vidx;                                   // the non-zero's index in the values region
        if (len == 1) {
            const uint64_t pos = off + 16 + 2 * (sw + 1) + 4ull * (coo_ex[r] - coo_ex[fr]);
            put16(img, pos, 0x8000u | f.rrel(key));
            put16(img, pos + 2, f.crel(key));
            const uint64_t scsr_runs = (uint64_t)(lr - fr) - (coo_ex[lr] - coo_ex[fr]);
            vidx = (sw - scsr_runs) + (coo_ex[r] - coo_ex[fr]);
        } else {
            const uint64_t sb = off + 16 + 2 * (scsr_ex[r] - scsr_ex[fr]);
            if (i == s0) put16(img, sb, 0x8000u | f.rrel(key));
            put16(img, sb + 2 * (1 + (i - s0)), f.crel(key));
            const uint64_t runs_before = (uint64_t)(r - fr) - (coo_ex[r] - coo_ex[fr]);
            vidx = (scsr_ex[r] - scsr_ex[fr]) - runs_before + (i - s0);
        }
        if (esz) {
            // values follow the COO region: SCSR values in order, then COO values
            const uint64_t ncoo = coo_ex[lr] - coo_ex[fr];
            const uint64_t vb = off + 16 + 2 * (sw + 1) + 4 * ncoo + vidx * esz;
            const uint8_t *v = vals + (uint64_t)eidx[i] * esz;
            for (uint32_t b = 0; b < esz; b += 2) put16(img, vb + b, (uint32_t)v[b] | ((uint32_t)v[b + 1] << 8));
        }
    }
}

int bits_for(uint64_t n)
{
    int b = 0;
    while (b < 64 && (1ull << b) < n) b++;
    return b;
}

uint32_t grid_for(fmx_ctx *ctx, uint64_t n)
{
    uint64_t b = (n + kB - 1) / kB;
    uint64_t cap = (uint64_t)ctx->num_cus * 16;
    if (b > cap) b = cap;
    return (uint32_t)(b ? b : 1);
}

}  // namespace


}  // namespace fmx

using namespace fmx;

namespace fmx {

// vals (device, esz bytes per edge) != nullptr builds a weighted image.
// ncols != 0 builds an nrows = n by ncols matrix (ids below max(n, ncols);
// the library's own work images); internal skips the work images of the
// result.
fmx_status build_impl(fmx_ctx *ctx, const uint32_t *src, const uint32_t *dst, size_t m, uint64_t n,
                      uint32_t log2_h, uint32_t log2_w, uint32_t flags, const void *vals, uint32_t esz,
                      int32_t prim_type, fmx_spm **out, uint64_t ncols, bool internal)
{
    if (!ctx || !out) return fail(ctx, FMX_EINVAL, "null argument");
    *out = nullptr;
    if (m && (!src || !dst)) return fail(ctx, FMX_EINVAL, "null edge list");
    if (log2_h > 15 || log2_w > 15) return fail(ctx, FMX_EINVAL, "tile dimensions must be <= 32768");
    if (flags & ~7u) return fail(ctx, FMX_EINVAL, "unknown build flags");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    // the edge list may come from any stream: construction is synchronous
    FMX_HIP(ctx, hipDeviceSynchronize());
    hipStream_t s = ctx->stream;
    // device time of the construction, from the first kernel to the last
    // write (host waits for totals in between are included)
    struct EvPair {
        hipEvent_t a = nullptr, b = nullptr;
        ~EvPair()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    FMX_HIP(ctx, hipEventCreate(&ev.a));
    FMX_HIP(ctx, hipEventCreate(&ev.b));
    FMX_HIP(ctx, hipEventRecord(ev.a, s));

    // n = max id + 1 (fm_utils.cpp:730-739), or a caller-given n above every id
    if (m) {
        uint32_t mx = 0;
        fmx_status st = max_u32_pair(ctx, src, dst, m, &mx, s);
        if (st) return st;
        if (n == 0) n = (uint64_t)mx + 1;
        else if ((uint64_t)mx >= std::max(n, ncols)) return fail(ctx, FMX_EDIM, "vertex id >= n");
    } else if (n == 0) {
        return fail(ctx, FMX_EINVAL, "empty edge list needs an explicit n");
    }
    if (ncols == 0) ncols = n;
    if (n > (1ull << 32) || ncols > (1ull << 32)) return fail(ctx, FMX_EINVAL, "n exceeds 2^32");
    const uint64_t H = 1ull << log2_h, W = 1ull << log2_w;
    const uint64_t nbr = (n + H - 1) / H, nbc = (ncols + W - 1) / W;
    KeyFmt f;
    f.lh = (int)log2_h;
    f.lw = (int)log2_w;
    f.bbc = bits_for(nbc);
    f.H = (uint32_t)H;
    f.W = (uint32_t)W;
    const int keybits = bits_for(nbr) + f.bbc + f.lh + f.lw;
    const int sym = (flags & FMX_BUILD_SYMMETRIZE) != 0;
    const uint64_t cap = sym ? 2 * (uint64_t)m : (uint64_t)m;

    // 1. keys (weighted: in edge order, with the edge index as payload)
    const int keep_self = (flags & FMX_BUILD_KEEP_SELF) ? 1 : 0;
    DevBuf<uint64_t> keys, tmp;
    DevBuf<uint32_t> eidx, etmp;
    DevBuf<unsigned long long> counter;
    FMX_HIP(ctx, keys.alloc(cap));
    FMX_HIP(ctx, tmp.alloc(cap));
    FMX_HIP(ctx, counter.alloc(1));
    FMX_HIP(ctx, hipMemsetAsync(counter.p, 0, sizeof(unsigned long long), s));
    unsigned long long nk64 = 0;
    fmx_status st = FMX_OK;
    if (vals) {
        if (m >= (1ull << 31)) return fail(ctx, FMX_EUNSUPPORTED, "weighted builds take fewer than 2^31 edges");
        FMX_HIP(ctx, eidx.alloc(cap));
        FMX_HIP(ctx, etmp.alloc(cap));
        DevBuf<uint32_t> pos;
        FMX_HIP(ctx, pos.alloc((uint64_t)m + 1));
        uint32_t total = 0;
        if (m) {
            hipLaunchKernelGGL(count_keys_w, dim3(grid_for(ctx, m)), dim3(kB), 0, s, src, dst, (uint64_t)m,
                               keep_self, sym, pos.p);
            FMX_HIP(ctx, hipGetLastError());
            st = scan_exclusive<uint32_t>(ctx, pos.p, pos.p, m, &total, s);
            if (st) return st;
            hipLaunchKernelGGL(emit_keys_w, dim3(grid_for(ctx, m)), dim3(kB), 0, s, src, dst, (uint64_t)m, f,
                               keep_self, sym, pos.p, keys.p, eidx.p);
            FMX_HIP(ctx, hipGetLastError());
        }
        nk64 = total;
    } else {
        if (m) {
            hipLaunchKernelGGL(make_keys, dim3(grid_for(ctx, m)), dim3(kB), 0, s, src, dst, (uint64_t)m, f,
                               keep_self, sym, keys.p, counter.p);
            FMX_HIP(ctx, hipGetLastError());
        }
        FMX_HIP(ctx, hipMemcpyAsync(&nk64, counter.p, sizeof(nk64), hipMemcpyDeviceToHost, s));
        FMX_HIP(ctx, hipStreamSynchronize(s));
    }
    if (nk64 >= (1ull << 32)) return fail(ctx, FMX_EUNSUPPORTED, "more than 2^32 - 1 non-zeros");

    // 2. sort + unique (a stable sort: equal keys keep edge order)
    uint64_t *K = keys.p;
    uint32_t *V = eidx.p;
    st = radix_sort_u64_kv(ctx, keys.p, tmp.p, vals ? eidx.p : nullptr, vals ? etmp.p : nullptr, nk64, keybits,
                           &K, &V, s);
    if (st) return st;
    uint64_t *other = (K == keys.p) ? tmp.p : keys.p;
    uint32_t *vother = (V == eidx.p) ? etmp.p : eidx.p;
    uint64_t nk = nk64;
    DevBuf<uint32_t> flag;
    FMX_HIP(ctx, flag.alloc(nk + 1));
    if ((flags & FMX_BUILD_DEDUP) && nk) {
        hipLaunchKernelGGL(group_flags, dim3(grid_for(ctx, nk)), dim3(kB), 0, s, K, nk, 0, flag.p);
        FMX_HIP(ctx, hipGetLastError());
        uint32_t nu = 0;
        st = scan_exclusive<uint32_t>(ctx, flag.p, flag.p, nk, &nu, s);
        if (st) return st;
        if (vals) {
            hipLaunchKernelGGL(compact_unique_kv, dim3(grid_for(ctx, nk)), dim3(kB), 0, s, K, V, nk, flag.p,
                               other, vother);
            V = vother;
        } else {
            hipLaunchKernelGGL(compact_unique, dim3(grid_for(ctx, nk)), dim3(kB), 0, s, K, nk, flag.p, other);
        }
        FMX_HIP(ctx, hipGetLastError());
        K = other;
        nk = nu;
    }
    // the key (and payload) buffer that does not hold the result is not read
    // again: its memory goes back before the run and tile arrays are made
    // (17 GB at 2^31 edges)
    FMX_HIP(ctx, hipStreamSynchronize(s));
    for (DevBuf<uint64_t> *b : {&keys, &tmp})
        if (b->p && b->p != K) {
            FMX_HIP(ctx, hipFree(b->p));
            b->p = nullptr;
        }
    for (DevBuf<uint32_t> *b : {&eidx, &etmp})
        if (b->p && b->p != V) {
            FMX_HIP(ctx, hipFree(b->p));
            b->p = nullptr;
        }

    // 3. runs
    uint32_t R = 0;
    if (nk) {
        hipLaunchKernelGGL(group_flags, dim3(grid_for(ctx, nk)), dim3(kB), 0, s, K, nk, f.lw, flag.p);
        FMX_HIP(ctx, hipGetLastError());
        st = scan_exclusive<uint32_t>(ctx, flag.p, flag.p, nk, &R, s);
        if (st) return st;
    }
    DevBuf<uint32_t> run_start, tflag, coo_ex, tile_of_run;
    DevBuf<uint64_t> scsr_ex;
    FMX_HIP(ctx, run_start.alloc((uint64_t)R + 1));
    FMX_HIP(ctx, tflag.alloc((uint64_t)R + 1));
    FMX_HIP(ctx, coo_ex.alloc((uint64_t)R + 1));
    FMX_HIP(ctx, scsr_ex.alloc((uint64_t)R + 1));
    FMX_HIP(ctx, tile_of_run.alloc((uint64_t)R + 1));
    const int tshift = f.lh + f.lw;
    uint32_t T = 0;
    if (nk) {
        hipLaunchKernelGGL(fill_run_start, dim3(grid_for(ctx, nk)), dim3(kB), 0, s, K, nk, f.lw, flag.p,
                           run_start.p, R);
        FMX_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(run_props, dim3(grid_for(ctx, R)), dim3(kB), 0, s, K, run_start.p, R, tshift,
                           tflag.p, scsr_ex.p, coo_ex.p);
        FMX_HIP(ctx, hipGetLastError());
        st = scan_exclusive<uint32_t>(ctx, tflag.p, tflag.p, R, &T, s);
        if (st) return st;
        st = scan_exclusive<uint64_t>(ctx, scsr_ex.p, scsr_ex.p, (uint64_t)R + 1, nullptr, s);
        if (st) return st;
        st = scan_exclusive<uint32_t>(ctx, coo_ex.p, coo_ex.p, (uint64_t)R + 1, nullptr, s);
        if (st) return st;
    }

    // 4. tiles and block rows
    DevBuf<uint32_t> tile_first_run, br_first;
    DevBuf<uint64_t> tsize, tile_off, offs_ex;
    DevBuf<unsigned long long> brsize;
    DevBuf<uint8_t> had;
    FMX_HIP(ctx, tile_first_run.alloc((uint64_t)T + 1));
    FMX_HIP(ctx, tsize.alloc((uint64_t)T + 1));
    FMX_HIP(ctx, tile_off.alloc((uint64_t)T + 1));
    FMX_HIP(ctx, brsize.alloc(nbr + 1));
    FMX_HIP(ctx, offs_ex.alloc(nbr + 1));
    FMX_HIP(ctx, br_first.alloc(nbr + 1));
    FMX_HIP(ctx, had.alloc(nbr + 1));
    DevBuf<uint32_t> bdir;  // directory entries per block row, then their exclusive scan
    FMX_HIP(ctx, bdir.alloc(nbr + 1));
    FMX_HIP(ctx, hipMemsetAsync(brsize.p, 0, (nbr + 1) * sizeof(unsigned long long), s));
    if (T) {
        hipLaunchKernelGGL(fill_tiles, dim3(grid_for(ctx, R)), dim3(kB), 0, s, K, run_start.p, R, tshift,
                           tflag.p, tile_first_run.p, tile_of_run.p, T);
        FMX_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(tile_props, dim3(grid_for(ctx, T)), dim3(kB), 0, s, run_start.p, tile_first_run.p, T,
                           tsize.p, vals ? esz : 0u);
        FMX_HIP(ctx, hipGetLastError());
        st = scan_exclusive<uint64_t>(ctx, tsize.p, tsize.p, (uint64_t)T + 1, nullptr, s);
        if (st) return st;
        hipLaunchKernelGGL(first_tile_of_row, dim3(grid_for(ctx, T)), dim3(kB), 0, s, K, f, run_start.p,
                           tile_first_run.p, T, br_first.p);
        FMX_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(row_totals, dim3(grid_for(ctx, T)), dim3(kB), 0, s, K, f, run_start.p,
                           tile_first_run.p, T, tsize.p, br_first.p, brsize.p, bdir.p);
        FMX_HIP(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(fix_empty_rows, dim3(grid_for(ctx, nbr)), dim3(kB), 0, s, brsize.p, nbr, had.p, bdir.p);
    FMX_HIP(ctx, hipGetLastError());
    uint64_t tiles_total = 0;
    st = scan_exclusive<uint64_t>(ctx, (const uint64_t *)brsize.p, offs_ex.p, nbr + 1, &tiles_total, s);
    if (st) return st;

    // 5. serialise
    std::unique_ptr<fmx_spm> spm(new fmx_spm());
    spm->ctx = ctx;
    const uint64_t bytes = 4096 + tiles_total;
    const uint64_t padded = (bytes + 4095) & ~4095ull;
    FMX_HIP(ctx, hipMalloc(&spm->d_image, padded));
    FMX_HIP(ctx, hipMemsetAsync(spm->d_image + bytes, 0, padded - bytes, s));
    {
        uint8_t hdr[4096];
        memset(hdr, 0, sizeof(hdr));
        const int64_t magic = 0x0123456789FFFFFFLL;
        const int32_t ver = 2, type = 2, layout = 2, prim = vals ? prim_type : 0;
        const uint64_t hesz = vals ? esz : 0;
        memcpy(hdr + 0, &magic, 8);
        memcpy(hdr + 8, &ver, 4);
        memcpy(hdr + 12, &type, 4);
        memcpy(hdr + 16, &hesz, 8);
        memcpy(hdr + 24, &n, 8);
        memcpy(hdr + 32, &ncols, 8);
        memcpy(hdr + 40, &layout, 4);
        memcpy(hdr + 44, &prim, 4);
        memcpy(hdr + 48, &H, 8);
        memcpy(hdr + 56, &W, 8);
        FMX_HIP(ctx, hipMemcpyAsync(spm->d_image, hdr, 4096, hipMemcpyHostToDevice, s));
        FMX_HIP(ctx, hipStreamSynchronize(s));
    }
    if (T) {
        hipLaunchKernelGGL(write_tiles, dim3(grid_for(ctx, T)), dim3(kB), 0, s, K, f, run_start.p,
                           tile_first_run.p, T, scsr_ex.p, coo_ex.p, offs_ex.p, tsize.p, br_first.p,
                           tile_off.p, spm->d_image);
        FMX_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(write_nz, dim3(grid_for(ctx, nk)), dim3(kB), 0, s, K, nk, f, flag.p,
                           run_start.p, tile_of_run.p, tile_first_run.p, scsr_ex.p, coo_ex.p,
                           tile_off.p, spm->d_image, V, (const uint8_t *)vals, vals ? esz : 0u);
        FMX_HIP(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(write_stubs, dim3(grid_for(ctx, nbr)), dim3(kB), 0, s, had.p, nbr, offs_ex.p,
                       spm->d_image);
    FMX_HIP(ctx, hipGetLastError());

    // 6. the directory, in HBM where the SpMM reads it (the host keeps each
    // block row's first directory index)
    uint32_t ndir = 0;
    st = scan_exclusive<uint32_t>(ctx, bdir.p, bdir.p, nbr + 1, &ndir, s);
    if (st) return st;
    FMX_HIP(ctx, hipMalloc(&spm->d_dir, std::max<uint64_t>(1, ndir) * sizeof(uint2)));
    if (vals) FMX_HIP(ctx, hipMalloc(&spm->d_vdir, std::max<uint64_t>(1, ndir) * sizeof(uint32_t)));
    hipLaunchKernelGGL(fill_dir, dim3(grid_for(ctx, std::max<uint64_t>(T, nbr))), dim3(kB), 0, s, K, f,
                       run_start.p, tile_first_run.p, T, offs_ex.p, br_first.p, tile_off.p, bdir.p, had.p, nbr,
                       vals ? 1 : 0, spm->d_dir, spm->d_vdir);
    FMX_HIP(ctx, hipGetLastError());
    DevBuf<unsigned int> bad;
    FMX_HIP(ctx, bad.alloc(1));
    FMX_HIP(ctx, hipMemsetAsync(bad.p, 0, sizeof(unsigned int), s));
    hipLaunchKernelGGL(check_dir, dim3(grid_for(ctx, nbr)), dim3(kB), 0, s, spm->d_dir, spm->d_vdir, bdir.p,
                       offs_ex.p, nbr, nbc, (uint64_t)ndir, vals ? 1 : 0, bad.p);
    FMX_HIP(ctx, hipGetLastError());
    FMX_HIP(ctx, hipEventRecord(ev.b, s));
    // the host keeps the block-row offsets and first directory indices; the
    // directory itself stays in HBM (plans copy the few block rows they split)
    std::vector<uint64_t> oex(nbr + 1);
    unsigned int hbad = 0;
    spm->brow_dir.resize(nbr + 1);
    FMX_HIP(ctx, hipMemcpyAsync(oex.data(), offs_ex.p, (nbr + 1) * 8, hipMemcpyDeviceToHost, s));
    FMX_HIP(ctx, hipMemcpyAsync(spm->brow_dir.data(), bdir.p, (nbr + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FMX_HIP(ctx, hipMemcpyAsync(&hbad, bad.p, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
    FMX_HIP(ctx, hipStreamSynchronize(s));
    if (hbad) return fail(ctx, FMX_EINVAL, "internal error: inconsistent device-built directory");
    float build_ms = 0.f;
    FMX_HIP(ctx, hipEventElapsedTime(&build_ms, ev.a, ev.b));
    spm->build_ns = (uint64_t)((double)build_ms * 1e6);
    spm->edges_in = m;
    spm->offs.resize(nbr + 1);
    for (uint64_t b = 0; b <= nbr; b++) spm->offs[b] = (int64_t)(4096 + oex[b]);
    spm->bytes = bytes;
    spm->nrows = n;
    spm->ncols = ncols;
    spm->nnz = nk;
    spm->H = H;
    spm->W = W;
    spm->log2h = (int)log2_h;
    spm->log2w = (int)log2_w;
    spm->nbr = nbr;
    spm->nbc = nbc;
    spm->esz = vals ? esz : 0;
    spm->prim = vals ? prim_type : 0;
    spm->ntiles = ndir;
    if (!internal) {
        st = prepare_work(ctx, spm.get());
        if (!st) st = prepare_fast(ctx, spm.get());
        if (st) return st;
    }
    *out = spm.release();
    return FMX_OK;
}

}  // namespace fmx

extern "C" fmx_status fmx_spm_build(fmx_ctx *ctx, const uint32_t *src, const uint32_t *dst,
                                    size_t m, uint64_t n, uint32_t log2_h, uint32_t log2_w,
                                    uint32_t flags, fmx_spm **out)
{
    return fmx::build_impl(ctx, src, dst, m, n, log2_h, log2_w, flags, nullptr, 0, 0, out, 0, false);
}

extern "C" fmx_status fmx_spm_build_w(fmx_ctx *ctx, const uint32_t *src, const uint32_t *dst, const void *vals,
                                      fmx_dtype dtype, size_t m, uint64_t n, uint32_t log2_h, uint32_t log2_w,
                                      uint32_t flags, fmx_spm **out)
{
    if (!ctx) return FMX_EINVAL;
    if (m && !vals) return fmx::fail(ctx, FMX_EINVAL, "null values");
    // prim_type of the header (generic_type.h:57-70): P_INTEGER, P_LONG, P_ULONG, P_FLOAT, P_DOUBLE
    int32_t prim;
    uint32_t esz;
    switch (dtype) {
    case FMX_I32: prim = 4; esz = 4; break;
    case FMX_I64: prim = 6; esz = 8; break;
    case FMX_U64: prim = 7; esz = 8; break;
    case FMX_F32: prim = 8; esz = 4; break;
    case FMX_F64: prim = 9; esz = 8; break;
    default: return fmx::fail(ctx, FMX_ETYPE, "values must be int32, int64, uint64, float32 or float64");
    }
    static const uint8_t none = 0;
    return fmx::build_impl(ctx, src, dst, m, n, log2_h, log2_w, flags, vals ? vals : &none, esz, prim, out, 0,
                           false);
}
