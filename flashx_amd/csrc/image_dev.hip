// image_dev.hip -- walks of an image resident in HBM, on the device.  A
// matrix built on the device has no host copy of its image; what the host
// walks of an uploaded or streamed image compute (hotcols.cpp, retile.cpp,
// the carries of get_plan) is computed here for it.
//
// decode_lists: every non-zero as (row, column[, value]) in stream order
// (block rows ascending; in each block row its tiles in directory order; in
// each tile the SCSR rows, then the COO entries):
//   tile_origin   each tile's block row (one thread per block row);
//   tile_count    each tile's non-zero count, from its header;
//   scan          each tile's first output position;
//   decode        one thread per tile walks its words and writes its
//                 non-zeros (and counts each column's uses, atomically, when
//                 asked).
//
// gather_order_lists: the same lists with the columns relabelled so that the
// nhot most used columns become 0..nhot-1 in descending order of use (ties by
// column) and every other column c becomes nhot + c, and the list of the hot
// columns in their new order:
//   rank          sort keys ((2^31 - uses) << 32 | column): descending use,
//                 ascending column among equals;
//   cut           the first rank at which the running sum of uses reaches
//                 the wanted share (or the uses reach 0) is nhot;
//   relabel       new ids, then every column of the list.
//
// plan_carries: for a word position inside a block row, the row of the SCSR
// row part in force there (the word is a column index) or none (it is a
// header), which a work item that starts there carries in (one wave per
// position, searching back 64 words at a time).
#include "devprim.h"

namespace fmx {
namespace {

constexpr int kT = 256;
constexpr uint64_t kMaxUses = 1ull << 31;  // non-zeros of a gather-order image stay below 2^31

// image words are 2-byte aligned (tiles start at even byte offsets)
__device__ __forceinline__ uint32_t ld16(const uint8_t *p) { return *(const uint16_t *)p; }
__device__ __forceinline__ uint32_t ld32(const uint8_t *p) { return ld16(p) | (ld16(p + 2) << 16); }

__global__ void __launch_bounds__(kT) tile_origin(const uint32_t *brow_dir, uint64_t nbr, uint32_t *tbr)
{
    for (uint64_t br = blockIdx.x * (uint64_t)kT + threadIdx.x; br < nbr; br += (uint64_t)gridDim.x * kT)
        for (uint32_t t = brow_dir[br]; t < brow_dir[br + 1]; t++) tbr[t] = (uint32_t)br;
}

__global__ void __launch_bounds__(kT) tile_count(const uint8_t *img, const int64_t *offs, const uint2 *dir,
                                                 const uint32_t *tbr, uint64_t ntiles, uint32_t *tnnz)
{
    for (uint64_t t = blockIdx.x * (uint64_t)kT + threadIdx.x; t < ntiles; t += (uint64_t)gridDim.x * kT)
        tnnz[t] = ld32(img + offs[tbr[t]] + 2ull * dir[t].x + 8);
}

__global__ void __launch_bounds__(kT) decode(const uint8_t *img, const int64_t *offs, const uint2 *dir,
                                             const uint32_t *tbr, const uint32_t *tpos, uint64_t ntiles,
                                             uint64_t H, uint64_t W, uint32_t esz, uint32_t *rows,
                                             uint32_t *cols, uint8_t *vals, uint32_t *uses)
{
    for (uint64_t t = blockIdx.x * (uint64_t)kT + threadIdx.x; t < ntiles; t += (uint64_t)gridDim.x * kT) {
        const uint64_t br = tbr[t];
        const uint8_t *p = img + offs[br] + 2ull * dir[t].x;
        const uint32_t nnz = ld32(p + 8);
        if (!nnz) continue;
        const uint32_t nrow = ld16(p + 12), ncoo = ld16(p + 14);
        const uint32_t scsr_words = (nrow - ncoo) + (nnz - ncoo);
        const uint8_t *w = p + 16, *coo = w + 2ull * scsr_words + 2, *vp = coo + 4ull * ncoo;
        const uint32_t r0 = (uint32_t)(br * H), c0 = (uint32_t)((uint64_t)dir[t].y * W);
        uint64_t o = tpos[t];
        uint32_t e = 0, r = r0;
        auto put = [&](uint32_t row, uint32_t col) {
            rows[o] = row;
            cols[o] = col;
            if (uses) atomicAdd(&uses[col], 1u);
            for (uint32_t b = 0; b < esz; b++) vals[o * esz + b] = vp[(uint64_t)e * esz + b];
            o++;
            e++;
        };
        for (uint32_t i = 0; i < scsr_words; i++) {
            const uint32_t v = ld16(w + 2ull * i);
            if (v & 0x8000u) r = r0 + (v & 0x7FFFu);
            else put(r, c0 + v);
        }
        for (uint32_t i = 0; i < ncoo; i++) put(r0 + (ld16(coo + 4ull * i) & 0x7FFFu), c0 + ld16(coo + 4ull * i + 2));
    }
}

__global__ void __launch_bounds__(kT) rank_keys(const uint32_t *uses, uint64_t ncols, uint64_t *keys)
{
    for (uint64_t c = blockIdx.x * (uint64_t)kT + threadIdx.x; c < ncols; c += (uint64_t)gridDim.x * kT)
        keys[c] = ((kMaxUses - uses[c]) << 32) | c;
}

__global__ void __launch_bounds__(kT) ranked_uses(const uint64_t *keys, uint64_t ncols, uint64_t *u)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < ncols; i += (uint64_t)gridDim.x * kT)
        u[i] = kMaxUses - (keys[i] >> 32);
}

// Rank i is hot while the uses before it fall short of the share and it has
// uses itself; both conditions hold on a prefix of the ranks, so nhot is the
// first rank that fails (ncols if none does): the one rank that fails while
// its predecessor passes writes it.
__global__ void __launch_bounds__(kT) cut(const uint64_t *keys, const uint64_t *before, uint64_t ncols,
                                          uint64_t target_permille, uint64_t *nhot)
{
    auto hot = [&](uint64_t i) { return before[i] * 1000 < target_permille && (keys[i] >> 32) < kMaxUses; };
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < ncols; i += (uint64_t)gridDim.x * kT)
        if (!hot(i) && (i == 0 || hot(i - 1))) *nhot = i;
}

__global__ void __launch_bounds__(kT) new_ids(uint64_t ncols, uint64_t nhot, uint32_t *newid)
{
    for (uint64_t c = blockIdx.x * (uint64_t)kT + threadIdx.x; c < ncols; c += (uint64_t)gridDim.x * kT)
        newid[c] = (uint32_t)(nhot + c);
}

__global__ void __launch_bounds__(kT) hot_ids(const uint64_t *keys, uint64_t nhot, uint32_t *newid, uint32_t *hot)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kT + threadIdx.x; i < nhot; i += (uint64_t)gridDim.x * kT) {
        const uint32_t c = (uint32_t)(keys[i] & 0xFFFFFFFFu);
        newid[c] = (uint32_t)i;
        hot[i] = c;
    }
}

__global__ void __launch_bounds__(kT) relabel(uint32_t *cols, uint64_t nnz, const uint32_t *newid)
{
    for (uint64_t e = blockIdx.x * (uint64_t)kT + threadIdx.x; e < nnz; e += (uint64_t)gridDim.x * kT)
        cols[e] = newid[cols[e]];
}

// carry[i] for word p[i] of the block row starting at word base[i] (p[i]
// lies past its tile's 8-word header): kNone if the word is a header (row
// part or COO entry), else the row of the nearest row-part header before it.
// One wave per cut: the lanes read 64 words back at a time and a ballot
// finds the nearest header (a hub row's part can be thousands of words long;
// the word after the tile header always has the header bit, so the search
// ends inside the tile).
__global__ void __launch_bounds__(kT) carries(const uint16_t *img, const uint64_t *base, const uint32_t *p,
                                              uint64_t n, uint32_t *carry)
{
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * kT) >> 6;
    for (uint64_t i = (blockIdx.x * (uint64_t)kT + threadIdx.x) >> 6; i < n; i += nw) {
        const uint16_t *w = img + base[i];
        const int64_t q = p[i];
        if (w[q] & 0x8000u) {
            if (lane == 0) carry[i] = kNone;
            continue;
        }
        for (int64_t hi = q - 1;; hi -= 64) {
            const int64_t at = hi - lane;
            const uint32_t v = at >= 0 ? w[at] : 0x8000u;
            const uint64_t m = __ballot((v & 0x8000u) != 0);
            if (m) {
                const int l0 = __ffsll((unsigned long long)m) - 1;  // nearest to q
                const uint32_t head = __shfl(v, l0, 64);
                if (lane == 0) carry[i] = head & 0x7FFFu;
                break;
            }
        }
    }
}

uint32_t grid_for(fmx_ctx *ctx, uint64_t n)
{
    const uint64_t want = (n + kT - 1) / kT, cap = (uint64_t)ctx->num_cus * 32;
    return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

}  // namespace

fmx_status decode_lists(fmx_ctx *ctx, const fmx_spm *src, uint32_t *rows, uint32_t *cols, void *vals,
                        uint32_t *uses)
{
    hipStream_t s = ctx->stream;
    if (!src->d_image) return fail(ctx, FMX_EINVAL, "internal error: the image is not in HBM");
    if (src->nnz >= (1ull << 32)) return fail(ctx, FMX_EUNSUPPORTED, "more than 2^32 - 1 non-zeros to list");
    const uint64_t nbr = src->nbr, ntiles = src->brow_dir[nbr];
    DevBuf<int64_t> offs;
    DevBuf<uint32_t> bdir, tbr, tnnz;
    FMX_HIP(ctx, offs.alloc(nbr + 1));
    FMX_HIP(ctx, bdir.alloc(nbr + 1));
    FMX_HIP(ctx, tbr.alloc(ntiles));
    FMX_HIP(ctx, tnnz.alloc(ntiles));
    FMX_HIP(ctx, hipMemcpyAsync(offs.p, src->offs.data(), (nbr + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    FMX_HIP(ctx, hipMemcpyAsync(bdir.p, src->brow_dir.data(), (nbr + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(tile_origin, dim3(grid_for(ctx, nbr)), dim3(kT), 0, s, bdir.p, nbr, tbr.p);
    hipLaunchKernelGGL(tile_count, dim3(grid_for(ctx, ntiles)), dim3(kT), 0, s, src->d_image, offs.p, src->d_dir,
                       tbr.p, ntiles, tnnz.p);
    FMX_HIP(ctx, hipGetLastError());
    uint32_t total = 0;
    fmx_status st = scan_exclusive<uint32_t>(ctx, tnnz.p, tnnz.p, ntiles, &total, s);
    if (st) return st;
    if (total != src->nnz) return fail(ctx, FMX_EINVAL, "internal error: tile counts disagree with the image");
    hipLaunchKernelGGL(decode, dim3(grid_for(ctx, ntiles)), dim3(kT), 0, s, src->d_image, offs.p, src->d_dir, tbr.p,
                       tnnz.p, ntiles, src->H, src->W, (uint32_t)src->esz, rows, cols, (uint8_t *)vals, uses);
    FMX_HIP(ctx, hipGetLastError());
    FMX_HIP(ctx, hipStreamSynchronize(s));  // the directory copies above read host memory
    return FMX_OK;
}

fmx_status gather_order_lists(fmx_ctx *ctx, const fmx_spm *src, uint64_t ncols, uint64_t permille,
                              uint32_t *rows, uint32_t *cols, void *vals, uint32_t **hot_out, uint64_t *nhot_out)
{
    *hot_out = nullptr;
    *nhot_out = 0;
    hipStream_t s = ctx->stream;
    DevBuf<uint32_t> uses, newid;
    FMX_HIP(ctx, uses.alloc(ncols));
    FMX_HIP(ctx, hipMemsetAsync(uses.p, 0, ncols * sizeof(uint32_t), s));
    fmx_status st = decode_lists(ctx, src, rows, cols, vals, uses.p);
    if (st) return st;
    DevBuf<uint64_t> keys, ktmp, before, nhot;
    FMX_HIP(ctx, keys.alloc(ncols));
    FMX_HIP(ctx, ktmp.alloc(ncols));
    FMX_HIP(ctx, before.alloc(ncols));
    FMX_HIP(ctx, nhot.alloc(1));
    hipLaunchKernelGGL(rank_keys, dim3(grid_for(ctx, ncols)), dim3(kT), 0, s, uses.p, ncols, keys.p);
    FMX_HIP(ctx, hipGetLastError());
    uint64_t *sorted = nullptr;
    st = radix_sort_u64(ctx, keys.p, ktmp.p, ncols, 64, &sorted, s);
    if (st) return st;
    hipLaunchKernelGGL(ranked_uses, dim3(grid_for(ctx, ncols)), dim3(kT), 0, s, sorted, ncols, before.p);
    FMX_HIP(ctx, hipGetLastError());
    st = scan_exclusive<uint64_t>(ctx, before.p, before.p, ncols, nullptr, s);
    if (st) return st;
    FMX_HIP(ctx, hipMemcpyAsync(nhot.p, &ncols, sizeof(uint64_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(cut, dim3(grid_for(ctx, ncols)), dim3(kT), 0, s, sorted, before.p, ncols,
                       src->nnz * permille, nhot.p);
    FMX_HIP(ctx, hipGetLastError());
    uint64_t nh = 0;
    FMX_HIP(ctx, hipMemcpyAsync(&nh, nhot.p, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    FMX_HIP(ctx, hipStreamSynchronize(s));
    if (nh == 0 || nh + ncols > (1ull << 32)) return FMX_OK;  // no hot columns, or ids past 32 bits
    DevBuf<uint32_t> hot;
    FMX_HIP(ctx, newid.alloc(ncols));
    FMX_HIP(ctx, hot.alloc(nh));
    hipLaunchKernelGGL(new_ids, dim3(grid_for(ctx, ncols)), dim3(kT), 0, s, ncols, nh, newid.p);
    hipLaunchKernelGGL(hot_ids, dim3(grid_for(ctx, nh)), dim3(kT), 0, s, sorted, nh, newid.p, hot.p);
    hipLaunchKernelGGL(relabel, dim3(grid_for(ctx, src->nnz)), dim3(kT), 0, s, cols, src->nnz, newid.p);
    FMX_HIP(ctx, hipGetLastError());
    FMX_HIP(ctx, hipStreamSynchronize(s));
    *hot_out = hot.release();
    *nhot_out = nh;
    return FMX_OK;
}

fmx_status plan_carries(fmx_ctx *ctx, const fmx_spm *s, const std::vector<uint64_t> &base,
                        const std::vector<uint32_t> &p, std::vector<uint32_t> &carry)
{
    const uint64_t n = p.size();
    carry.assign(n, kNone);
    if (!n) return FMX_OK;
    if (!s->d_image) return fail(ctx, FMX_EINVAL, "internal error: the image is not in HBM");
    hipStream_t st = ctx->stream;
    DevBuf<uint64_t> db;
    DevBuf<uint32_t> dp, dc;
    FMX_HIP(ctx, db.alloc(n));
    FMX_HIP(ctx, dp.alloc(n));
    FMX_HIP(ctx, dc.alloc(n));
    FMX_HIP(ctx, hipMemcpyAsync(db.p, base.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    FMX_HIP(ctx, hipMemcpyAsync(dp.p, p.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(carries, dim3(grid_for(ctx, n * 64)), dim3(kT), 0, st, (const uint16_t *)s->d_image, db.p, dp.p, n,
                       dc.p);
    FMX_HIP(ctx, hipGetLastError());
    FMX_HIP(ctx, hipMemcpyAsync(carry.data(), dc.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    FMX_HIP(ctx, hipStreamSynchronize(st));
    return FMX_OK;
}

}  // namespace fmx
