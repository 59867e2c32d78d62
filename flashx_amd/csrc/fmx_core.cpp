// fmx_core.cpp -- context, matrix upload/validation, work plans and the
// fmx_spmm entry point of libfmx.
//
// Reference paths are relative to the FlashX source tree.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>

#include "dense.h"
#include "fmx_internal.h"

namespace fmx {

Plan::~Plan()
{
    if (d_items) (void)hipFree(d_items);
    if (d_splits) (void)hipFree(d_splits);
}

Scratch::~Scratch()
{
    if (ptr) (void)hipFree(ptr);
}

fmx_status fail(fmx_ctx *ctx, fmx_status st, const std::string &msg)
{
    if (ctx) ctx->last_error = msg;
    return st;
}

fmx_status hip_fail(fmx_ctx *ctx, hipError_t e, const char *what)
{
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    return fail(ctx, e == hipErrorOutOfMemory ? FMX_ENOMEM : FMX_EHIP, msg);
}

int elem_size(fmx_dtype dt)
{
    switch (dt) {
    case FMX_I32: case FMX_F32: return 4;
    case FMX_I64: case FMX_U64: case FMX_F64: return 8;
    case FMX_BF16: return 2;
    }
    return 0;
}

static uint16_t rd16(const uint8_t *p) { uint16_t v; memcpy(&v, p, 2); return v; }
static uint32_t rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t rd64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }

static int log2_exact(uint64_t v)
{
    for (int b = 0; b <= 15; b++)
        if ((1ULL << b) == v) return b;
    return -1;
}

// matrix_header::verify (matrix/matrix_header.cpp:52-60), block_2d_size
// (matrix/matrix_header.h:42-91), SpM_2d_index::verify
// (matrix/sparse_matrix_format.cpp:128-136) and sparse_block_2d::verify
// (matrix/sparse_matrix_format.cpp:36-60), plus every structural fact the
// reference writer guarantees, so that the device kernels can index without
// further checks.
fmx_status analyse_image(fmx_ctx *ctx, fmx_spm *s)
{
    const uint8_t *img = s->host_image();
    const uint64_t bytes = s->host_bytes();
    if (bytes < kHeaderBytes) return fail(ctx, FMX_EFORMAT, "image shorter than its 4096-byte header");
    if (rd64(img) != 0x0123456789FFFFFFULL) return fail(ctx, FMX_EFORMAT, "wrong magic number");
    if (rd32(img + 8) != 2) return fail(ctx, FMX_EFORMAT, "wrong version number");
    if (rd32(img + 12) != 2) return fail(ctx, FMX_EFORMAT, "not a sparse matrix");
    if (rd32(img + 40) != 2) return fail(ctx, FMX_EFORMAT, "layout is not L_ROW_2D");
    s->esz = rd64(img + 16);
    s->nrows = rd64(img + 24);
    s->ncols = rd64(img + 32);
    s->prim = (int)rd32(img + 44);
    s->H = rd64(img + 48);
    s->W = rd64(img + 56);
    s->log2h = log2_exact(s->H);
    s->log2w = log2_exact(s->W);
    if (s->log2h < 0 || s->log2w < 0)
        return fail(ctx, FMX_EFORMAT, "tile dimensions must be powers of two <= 32768");
    if (s->nrows > (1ULL << 32) || s->ncols > (1ULL << 32))
        return fail(ctx, FMX_EFORMAT, "more than 2^32 rows or columns");
    // binary, or values of the header's primitive type: P_INTEGER / P_FLOAT
    // (4 bytes), P_LONG / P_ULONG / P_DOUBLE (8 bytes) (generic_type.h:57-70)
    if (s->esz != 0) {
        const bool ok4 = s->esz == 4 && (s->prim == 4 || s->prim == 8);
        const bool ok8 = s->esz == 8 && (s->prim == 6 || s->prim == 7 || s->prim == 9);
        if (!ok4 && !ok8)
            return fail(ctx, FMX_EUNSUPPORTED,
                        "weighted images need 4-byte int/float or 8-byte long/size_t/double values");
    }
    const uint64_t nbr = (s->nrows + s->H - 1) / s->H;
    if (s->offs.size() != nbr + 1)
        return fail(ctx, FMX_EFORMAT, "there are an incorrect number of offsets");
    s->nbr = nbr;
    s->nbc = (s->ncols + s->W - 1) / s->W;
    if (s->offs[0] != (int64_t)kHeaderBytes) return fail(ctx, FMX_EFORMAT, "offs[0] must be 4096");
    if ((uint64_t)s->offs[nbr] > bytes) return fail(ctx, FMX_EFORMAT, "offsets run past the image");
    for (uint64_t i = 1; i <= nbr; i++)
        if (s->offs[i] <= s->offs[i - 1] || (s->offs[i] & 1))
            return fail(ctx, FMX_EFORMAT, "block-row offsets must be even and strictly increasing");

    // Each block row is validated on its own (threads over block-row
    // stripes); the first failing block row's message is reported.  A
    // second pass fills the directory at each block row's prefix position.
    std::vector<uint32_t> ntile(nbr, 0);
    std::vector<uint64_t> rnnz(nbr, 0);
    std::vector<std::string> err(nbr);
    auto check_row = [&](uint64_t br) -> std::string {
        const uint8_t *base = img + s->offs[br];
        const uint8_t *p = base, *end = img + s->offs[br + 1];
        if ((uint64_t)(end - base) / 2 >= (1ULL << 31)) return "block row larger than 2^31 words";
        const uint64_t rows_here = std::min<uint64_t>(s->H, s->nrows - br * s->H);
        long prev_bc = -1;
        uint64_t empty_tiles = 0, nnz = 0;
        uint32_t tiles = 0;
        while (p < end) {
            if (end - p < 16) return "truncated tile header";
            uint32_t tbr = rd32(p), tbc = rd32(p + 4), tnnz = rd32(p + 8);
            uint32_t tnrow = rd16(p + 12), tncoo = rd16(p + 14);
            if (tbr != br) return "tile in the wrong block row";
            uint64_t size = 16;
            if (!tnnz) empty_tiles++;
            if (tnnz) {
                if ((long)tbc <= prev_bc) return "tiles not in ascending block column";
                if ((uint64_t)tbc >= s->nbc) return "block column out of range";
                prev_bc = tbc;
                if (tncoo > tnrow || tncoo > tnnz || tnrow > rows_here) return "inconsistent tile counts";
                size = 16 + 2ULL * tnrow + 2ULL * tnnz + 2 + s->esz * tnnz;
                if ((uint64_t)(end - p) < size) return "tile runs past its block row";
                const uint64_t cols_here = std::min<uint64_t>(s->W, s->ncols - (uint64_t)tbc * s->W);
                const uint8_t *w = p + 16;
                const uint8_t *rend = p + 16 + 2ULL * (tnrow - tncoo) + 2ULL * (tnnz - tncoo);
                long prev_row = -1;
                uint64_t nz = 0, nr = 0;
                while (w < rend) {
                    uint16_t h = rd16(w);
                    // 0xFFFF is row 32767 of a 32768-row tile; the region
                    // end is positional (sparse_matrix_format.h:339-347).
                    if (!(h & 0x8000)) return "bad row-part header";
                    long r = h & 0x7FFF;
                    if (r <= prev_row || (uint64_t)r >= rows_here) return "row-part rows out of order or range";
                    prev_row = r;
                    w += 2;
                    uint64_t len = 0;
                    while (w < rend && !(rd16(w) & 0x8000)) {
                        if (rd16(w) >= cols_here) return "column out of range";
                        w += 2;
                        len++;
                    }
                    if (len == 0) return "empty row part";
                    nz += len;
                    nr++;
                }
                if (rd16(rend) != 0xFFFF) return "missing end-of-row-parts marker";
                const uint8_t *coo = rend + 2;
                prev_row = -1;
                for (uint32_t i = 0; i < tncoo; i++) {
                    uint16_t h = rd16(coo + 4ULL * i), c = rd16(coo + 4ULL * i + 2);
                    if (!(h & 0x8000) || (c & 0x8000)) return "bad COO entry";
                    long r = h & 0x7FFF;
                    if (r <= prev_row || (uint64_t)r >= rows_here || c >= cols_here)
                        return "COO rows out of order or range";
                    prev_row = r;
                }
                if (nz + tncoo != tnnz || nr + tncoo != tnrow) return "tile counts disagree with its contents";
                nnz += tnnz;
            }
            tiles++;
            p += size;
        }
        if (p != end) return "block row size disagrees with its tiles";
        if (tiles == 0) return "block row without tiles";
        // an empty tile only stands for an empty block row (fm_utils.cpp:
        // 317,340-350); the SpMM decoder relies on every other tile being at
        // least 11 words long
        if (empty_tiles && tiles > 1) return "empty tile in a non-empty block row";
        ntile[br] = tiles;
        rnnz[br] = nnz;
        return std::string();
    };
    const unsigned nth = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>({16, nbr / 64 + 1,
                                                                               std::thread::hardware_concurrency()}));
    auto for_rows = [&](auto fn) {
        std::vector<std::thread> th;
        for (unsigned t = 1; t < nth; t++)
            th.emplace_back([&, t] {
                for (uint64_t br = t; br < nbr; br += nth) fn(br);
            });
        for (uint64_t br = 0; br < nbr; br += nth) fn(br);
        for (auto &x : th) x.join();
    };
    for_rows([&](uint64_t br) { err[br] = check_row(br); });
    for (uint64_t br = 0; br < nbr; br++)
        if (!err[br].empty()) return fail(ctx, FMX_EFORMAT, err[br]);
    s->brow_dir.assign(nbr + 1, 0);
    for (uint64_t br = 0; br < nbr; br++) s->brow_dir[br + 1] = s->brow_dir[br] + ntile[br];
    if ((uint64_t)s->brow_dir[nbr] != std::accumulate(ntile.begin(), ntile.end(), (uint64_t)0))
        return fail(ctx, FMX_EFORMAT, "more than 2^32 tiles");
    s->dir.assign(s->brow_dir[nbr], make_uint2(0, 0));
    s->vdir.assign(s->esz ? s->brow_dir[nbr] : 0, 0);
    for_rows([&](uint64_t br) {
        const uint8_t *base = img + s->offs[br], *p = base, *end = img + s->offs[br + 1];
        uint32_t t = s->brow_dir[br];
        while (p < end) {
            const uint32_t tnnz = rd32(p + 8), tnrow = rd16(p + 12);
            s->dir[t] = make_uint2((uint32_t)((p - base) / 2), rd32(p + 4));
            if (s->esz) s->vdir[t] = (uint32_t)((p - base) / 2 + 8 + tnrow + tnnz + 1);
            p += tnnz ? 16 + 2ULL * tnrow + 2ULL * tnnz + 2 + s->esz * tnnz : 16;
            t++;
        }
    });
    const uint64_t nnz = std::accumulate(rnnz.begin(), rnnz.end(), (uint64_t)0);
    s->brow_dir[nbr] = (uint32_t)s->dir.size();
    s->ntiles = s->dir.size();
    s->nnz = nnz;
    s->bytes = (uint64_t)s->offs[nbr];
    return FMX_OK;
}

uint32_t band_rows(const fmx_spm *spm, int k, int s)
{
    const char *e = getenv("FMX_BAND_BYTES");
    uint64_t budget = e ? strtoull(e, nullptr, 10) : 16384;
    if (budget < 256) budget = 256;
    uint64_t rows = budget / ((uint64_t)k * s);
    uint32_t hs = 1;
    while ((uint64_t)hs * 2 <= rows && hs < 32768) hs *= 2;
    return (uint32_t)std::min<uint64_t>(hs, spm->H);
}

// The directory entries of one block row at a time: from the host copy, or
// (an image built on the device keeps its directory in HBM only) copied
// from HBM when a plan first needs them, which is only for the block rows
// whose bands are split.
class DirRows {
public:
    DirRows(fmx_ctx *ctx, const fmx_spm *s) : ctx_(ctx), s_(s) {}
    // entries brow_dir[br] .. brow_dir[br + 1] - 1, or null after a failed copy
    const uint2 *row(uint64_t br)
    {
        const uint32_t lo = s_->brow_dir[br], hi = s_->brow_dir[br + 1];
        if (!s_->dir.empty()) return s_->dir.data() + lo;
        auto it = cache_.find(br);
        if (it != cache_.end()) return it->second.data();
        std::vector<uint2> &v = cache_[br];
        v.resize(hi - lo);
        if (hipMemcpy(v.data(), s_->d_dir + lo, (size_t)(hi - lo) * sizeof(uint2), hipMemcpyDeviceToHost) !=
            hipSuccess) {
            (void)hip_fail(ctx_, hipGetLastError(), "directory copy");
            return nullptr;
        }
        return v.data();
    }

private:
    fmx_ctx *ctx_;
    const fmx_spm *s_;
    std::map<uint64_t, std::vector<uint2>> cache_;
};

// The last tile of block row br (entries d, local indices 0..n-1) starting
// at or before word p, as a directory index.
static uint32_t tile_at(const fmx_spm *s, uint64_t br, const uint2 *d, uint32_t p)
{
    const uint32_t n = s->brow_dir[br + 1] - s->brow_dir[br];
    const uint32_t i = (uint32_t)(std::upper_bound(d, d + n, p, [](uint32_t v, const uint2 &e) { return v < e.x; }) -
                                  d) - 1;
    return s->brow_dir[br] + i;
}

// Tile-relative row in force at word `p` of the block row whose first word
// is `base` in a host image (kNone if `p` is a row-part or COO header);
// plan_carries computes the same on the device for an image without a host
// copy.
static uint32_t carry_at(const fmx_spm *s, uint64_t base, uint32_t p)
{
    const uint16_t *w = reinterpret_cast<const uint16_t *>(s->host_image()) + base;
    if (w[p] & 0x8000) return kNone;
    uint32_t q = p;
    while (!(w[q - 1] & 0x8000)) q--;
    return w[q - 1] & 0x7FFF;
}

fmx_status get_plan(fmx_ctx *ctx, fmx_spm *s, uint32_t hs, bool ordered, Plan **out)
{
    const uint32_t key = hs | (ordered ? 0x80000000u : 0u);
    auto it = s->plans.find(key);
    if (it != s->plans.end()) { *out = it->second.get(); return FMX_OK; }
    std::unique_ptr<Plan> plan(new Plan());
    plan->hs = hs;

    struct Band { uint64_t br; uint64_t row0; uint32_t sb_base, rows; uint32_t words; };
    std::vector<Band> bands;
    uint64_t total = 0;
    for (uint64_t br = 0; br < s->nbr; br++) {
        uint64_t rows_here = std::min<uint64_t>(s->H, s->nrows - br * s->H);
        uint32_t words = (uint32_t)((s->offs[br + 1] - s->offs[br]) / 2);
        for (uint64_t sb = 0; sb * hs < rows_here; sb++) {
            uint32_t rows = (uint32_t)std::min<uint64_t>(hs, rows_here - sb * hs);
            bands.push_back({br, br * s->H + sb * hs, (uint32_t)(sb * hs), rows, words});
            total += words;
        }
    }
    // Work cap: a band whose stream exceeds it is split into near-equal word
    // ranges reduced in order afterwards (power-law hub rows).
    uint64_t avg = bands.empty() ? 0 : total / bands.size();
    uint64_t cap = std::max<uint64_t>(8192, 4 * avg);
    const char *e = getenv("FMX_ITEM_CAP");
    if (e) cap = std::max<uint64_t>(64, strtoull(e, nullptr, 10));
    if (ordered) cap = UINT64_MAX;  // one item per band: its rows are summed in stream order

    std::vector<Split> splits;
    uint32_t nslots = 0;
    // carries of binary split items, resolved after the loop: item index,
    // block row's first word, cut word
    std::vector<size_t> carry_item;
    std::vector<uint64_t> carry_base;
    std::vector<uint32_t> carry_word;
    DirRows dir_rows(ctx, s);
    for (const Band &b : bands) {
        uint32_t pieces = (uint32_t)(b.words / cap + (b.words % cap != 0));
        if (pieces < 1) pieces = 1;
        Item base{};
        base.brow_word = (uint64_t)s->offs[b.br] / 2;
        base.dir_end = s->brow_dir[b.br + 1];
        base.row0_lo = (uint32_t)b.row0;
        base.row0_hi = (uint32_t)(b.row0 >> 32);
        base.sb_base = b.sb_base;
        base.rows = b.rows;
        if (pieces == 1) {
            Item t = base;
            t.begin = 0;
            t.nwords = b.words;
            t.dir_begin = s->brow_dir[b.br];
            t.carry = kNone;
            t.slot = kNone;
            plan->items.push_back(t);
            continue;
        }
        std::vector<uint32_t> cut(pieces + 1), dirs(pieces + 1), carries(pieces + 1);
        cut[0] = 0;
        dirs[0] = s->brow_dir[b.br];
        carries[0] = kNone;
        const uint2 *d = dir_rows.row(b.br);
        if (!d) return FMX_EHIP;
        for (uint32_t i = 1; i < pieces; i++) {
            uint32_t p = (uint32_t)((uint64_t)b.words * i / pieces);
            if (p <= cut[i - 1]) p = cut[i - 1] + 1;
            const uint32_t t = tile_at(s, b.br, d, p);
            const uint32_t tx = d[t - s->brow_dir[b.br]].x;
            if (s->esz) {
                // weighted: cut at the tile start at or before p, so no row
                // and no value position is carried into a piece
                p = std::max(tx, cut[i - 1]);
            } else if (p < tx + 8) {
                p = tx + 8;  // out of the tile header
            }
            dirs[i] = t;
            carries[i] = kNone;  // binary: resolved below
            cut[i] = p;
        }
        cut[pieces] = b.words;
        // drop pieces emptied by header adjustment
        std::vector<uint32_t> keep;
        for (uint32_t i = 0; i < pieces; i++)
            if (cut[i] < cut[i + 1]) keep.push_back(i);
        Split sp{b.row0, b.rows, nslots, (uint32_t)keep.size(), 0};
        for (uint32_t i : keep) {
            Item t = base;
            t.begin = cut[i];
            t.nwords = cut[i + 1] - cut[i];
            t.dir_begin = dirs[i];
            t.carry = carries[i];
            t.slot = nslots++;
            if (i > 0 && !s->esz) {
                carry_item.push_back(plan->items.size());
                carry_base.push_back((uint64_t)s->offs[b.br] / 2);
                carry_word.push_back(cut[i]);
            }
            plan->items.push_back(t);
        }
        splits.push_back(sp);
    }
    if (!carry_item.empty()) {
        std::vector<uint32_t> c;
        if (s->host_bytes()) {
            c.resize(carry_item.size());
            for (size_t j = 0; j < c.size(); j++) c[j] = carry_at(s, carry_base[j], carry_word[j]);
        } else {
            fmx_status st = plan_carries(ctx, s, carry_base, carry_word, c);
            if (st) return st;
        }
        for (size_t j = 0; j < c.size(); j++) plan->items[carry_item[j]].carry = c[j];
    }
    // Longest items first, so the hardware dispatcher approximates LPT.
    std::stable_sort(plan->items.begin(), plan->items.end(),
                     [](const Item &a, const Item &b) { return a.nwords > b.nwords; });
    // Column rotation (off unless FMX_ROT_SLOTS is set): item i (in dispatch
    // order) starts at the first of its tiles in block column
    // >= (i * nbc / slots) mod nbc and wraps around, so that items resident
    // together gather from the same tile columns of X.  The start depends
    // only on the item's index, so results stay deterministic.  Measured on
    // R-MAT scale 22 it does not pay: the gathers are not bound by X's
    // cache footprint (time per non-zero is the same with 1/8 of X), and
    // every rotation rate tried was 5-10% slower than none.
    const char *re = getenv("FMX_ROT_SLOTS");
    const uint64_t slots = (re && !ordered && !s->dir.empty()) ? strtoull(re, nullptr, 10) : 0;
    for (size_t i = 0; i < plan->items.size(); i++) {
        Item &t = plan->items[i];
        t.rot_begin = t.begin;
        t.rot_dir = t.dir_begin;
        if (slots == 0 || s->nbc < 2) continue;
        const uint32_t col = (uint32_t)(((uint64_t)i * s->nbc / slots) % s->nbc);
        // first tile after dir_begin with block column >= col, inside the item
        uint32_t lo = t.dir_begin + 1, hi = t.dir_end;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (s->dir[mid].y < col) lo = mid + 1;
            else hi = mid;
        }
        if (lo < t.dir_end && s->dir[lo].x < t.begin + t.nwords && s->dir[lo].x > t.begin) {
            t.rot_begin = s->dir[lo].x;
            t.rot_dir = lo;
        }
    }
    // The kernel trusts these invariants; check them before anything runs.
    for (const Item &t : plan->items) {
        // the first tile of the item starts at or before the item: the block
        // row's first tile at word 0, or the tile a cut fell in
        const uint64_t br = (((uint64_t)t.row0_hi << 32) | t.row0_lo) / s->H;
        bool first_ok = br < s->nbr;
        if (first_ok && t.begin == 0) {
            first_ok = t.dir_begin == s->brow_dir[br];
        } else if (first_ok) {
            const uint2 *d = dir_rows.row(br);
            if (!d) return FMX_EHIP;
            first_ok = t.dir_begin >= s->brow_dir[br] && t.dir_begin < s->brow_dir[br + 1] &&
                       d[t.dir_begin - s->brow_dir[br]].x <= t.begin;
        }
        const bool ok = first_ok && t.nwords >= 1 && t.dir_begin < t.dir_end && t.dir_end <= s->ntiles &&
                        t.rot_begin >= t.begin &&
                        t.rot_begin < t.begin + t.nwords && t.rot_dir >= t.dir_begin && t.rot_dir < t.dir_end &&
                        (t.rot_begin == t.begin ? t.rot_dir == t.dir_begin : s->dir[t.rot_dir].x == t.rot_begin) &&
                        t.rows >= 1 && t.rows <= hs;
        if (!ok) return fail(ctx, FMX_EINVAL, "internal error: inconsistent SpMM work item");
    }
    if (s->streamed()) {
        // streamed: the items of each chunk together, longest first within it
        const std::vector<uint64_t> &cb = s->chunk_br;
        auto chunk_of = [&](const Item &t) {
            const uint64_t br = (uint64_t)(((uint64_t)t.row0_hi << 32 | t.row0_lo) / s->H);
            return (size_t)(std::upper_bound(cb.begin(), cb.end(), br) - cb.begin()) - 1;
        };
        std::stable_sort(plan->items.begin(), plan->items.end(),
                         [&](const Item &a, const Item &b) { return chunk_of(a) < chunk_of(b); });
        const size_t nch = cb.size() - 1;
        plan->chunk_items.assign(nch + 1, 0);
        for (const Item &t : plan->items) plan->chunk_items[chunk_of(t) + 1]++;
        for (size_t c = 0; c < nch; c++) plan->chunk_items[c + 1] += plan->chunk_items[c];
        // a chunk's items stay inside the bytes staged for it
        for (size_t c = 0; c < nch; c++)
            for (uint32_t i = plan->chunk_items[c]; i < plan->chunk_items[c + 1]; i++) {
                const Item &t = plan->items[i];
                const uint64_t lo = 2 * ((t.brow_word + t.begin) & ~7ull);
                const uint64_t hi = 2 * (((t.brow_word + t.begin + t.nwords - 1) | 7ull) + 1);
                if (lo < s->chunk_lo[c] || hi > s->chunk_hi[c])
                    return fail(ctx, FMX_EINVAL, "internal error: work item outside its chunk");
            }
    }
    plan->nitems = (uint32_t)plan->items.size();
    plan->nslots = nslots;
    plan->nsplits = (uint32_t)splits.size();
    if (plan->nitems) {
        FMX_HIP(ctx, hipMalloc(&plan->d_items, plan->nitems * sizeof(Item)));
        FMX_HIP(ctx, hipMemcpy(plan->d_items, plan->items.data(), plan->nitems * sizeof(Item), hipMemcpyHostToDevice));
    }
    if (plan->nsplits) {
        FMX_HIP(ctx, hipMalloc(&plan->d_splits, plan->nsplits * sizeof(Split)));
        FMX_HIP(ctx, hipMemcpy(plan->d_splits, splits.data(), plan->nsplits * sizeof(Split), hipMemcpyHostToDevice));
    }
    *out = plan.get();
    s->plans[key] = std::move(plan);
    return FMX_OK;
}

// A device-built image: bring the host copy, validate it like an uploaded
// one and publish the tile directory.
// Tile directory (and, for weighted images, the values directory) to HBM.
fmx_status upload_dirs(fmx_ctx *ctx, fmx_spm *s)
{
    FMX_HIP(ctx, hipMalloc(&s->d_dir, std::max<size_t>(1, s->dir.size()) * sizeof(uint2)));
    FMX_HIP(ctx, hipMemcpy(s->d_dir, s->dir.data(), s->dir.size() * sizeof(uint2), hipMemcpyHostToDevice));
    if (s->esz) {
        FMX_HIP(ctx, hipMalloc(&s->d_vdir, std::max<size_t>(1, s->vdir.size()) * sizeof(uint32_t)));
        FMX_HIP(ctx, hipMemcpy(s->d_vdir, s->vdir.data(), s->vdir.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    return FMX_OK;
}

fmx_status prepare_work(fmx_ctx *ctx, fmx_spm *s)
{
    const uint32_t h2 = work_rows_for(s);
    if (!h2) return FMX_OK;
    if (!s->host_bytes() && s->d_image) {
        // built on the device, without a host copy: the non-zeros listed in
        // stream order and built again in h2-row tiles (every entry as it
        // is, so the result is the re-cut of the same image)
        struct Lists {
            uint32_t *rows = nullptr, *cols = nullptr;
            void *vals = nullptr;
            ~Lists()
            {
                if (rows) (void)hipFree(rows);
                if (cols) (void)hipFree(cols);
                if (vals) (void)hipFree(vals);
            }
        } l;
        const uint64_t nz = s->nnz ? s->nnz : 1;
        if (hipMalloc(&l.rows, nz * 4) != hipSuccess || hipMalloc(&l.cols, nz * 4) != hipSuccess ||
            (s->esz && hipMalloc(&l.vals, nz * s->esz) != hipSuccess))
            return fail(ctx, FMX_ENOMEM, "re-cut work image");
        fmx_status st = decode_lists(ctx, s, l.rows, l.cols, l.vals, nullptr);
        if (st) return st;
        int lh2 = 0;
        while ((1u << lh2) < h2) lh2++;
        fmx_spm *w = nullptr;
        st = build_impl(ctx, l.rows, l.cols, s->nnz, s->nrows, (uint32_t)lh2, (uint32_t)s->log2w, FMX_BUILD_KEEP_SELF,
                        s->esz ? l.vals : nullptr, (uint32_t)s->esz, s->prim, &w, s->ncols, true);
        if (st) return st;
        std::unique_ptr<fmx_spm> wp(w);
        if (wp->nnz != s->nnz) return fail(ctx, FMX_EINVAL, "internal error: the re-cut work image lost non-zeros");
        s->work = std::move(wp);
        return FMX_OK;
    }
    std::unique_ptr<fmx_spm> w(new fmx_spm());
    w->ctx = ctx;
    retile_image(s, h2, w->h_image, w->offs);
    fmx_status st = analyse_image(ctx, w.get());
    if (st) return fail(ctx, FMX_EINVAL, "internal error: the re-cut work image is invalid: " + ctx->last_error);
    if (w->nnz != s->nnz) return fail(ctx, FMX_EINVAL, "internal error: the re-cut work image lost non-zeros");
    const size_t padded = (w->bytes + 4095) & ~size_t(4095);
    FMX_HIP(ctx, hipMalloc(&w->d_image, padded));
    FMX_HIP(ctx, hipMemcpy(w->d_image, w->h_image.data(), w->bytes, hipMemcpyHostToDevice));
    if (padded > w->bytes) FMX_HIP(ctx, hipMemset(w->d_image + w->bytes, 0, padded - w->bytes));
    st = upload_dirs(ctx, w.get());
    if (st) return st;
    s->work = std::move(w);
    return FMX_OK;
}

// Out-of-core SpMM (the SAFS-streamed multiply of sparse_matrix.cpp:796-830
// and matrix_io.cpp:188-236, here host memory -> HBM): chunk c is copied into
// staging buffer c % 2 on the matrix's copy stream while the kernel of chunk
// c - 1 runs on the caller's stream.  A buffer is refilled only after the
// kernel that last read it has finished (ev_free, also across calls and
// streams), and the copies wait for the caller's earlier work on its stream.
// Split bands are reduced once, after the last chunk.
fmx_status spmm_streamed(fmx_ctx *ctx, fmx_spm *spm, Plan *plan, fmx_dtype dtype, int k, const void *X,
                         int64_t ldx, void *Y, int64_t ldy, hipStream_t s, const double *rscale, uint32_t flags,
                         const void *Xhot = nullptr, uint64_t nhot = 0)
{
    const size_t nch = spm->chunk_br.size() - 1;
    FMX_HIP(ctx, hipEventRecord(spm->ev_start, s));
    FMX_HIP(ctx, hipStreamWaitEvent(spm->copy_stream, spm->ev_start, 0));
    for (size_t c = 0; c < nch; c++) {
        const int b = (int)(c & 1);
        const uint64_t lo = spm->chunk_lo[c], hi = spm->chunk_hi[c];
        // the kernel that last read buffer b (this call's, or an earlier
        // call's on any stream) has finished
        if (spm->free_recorded[b]) FMX_HIP(ctx, hipStreamWaitEvent(spm->copy_stream, spm->ev_free[b], 0));
        FMX_HIP(ctx, hipMemcpyAsync(spm->d_stage[b], spm->h_pinned + lo, hi - lo, hipMemcpyHostToDevice,
                                    spm->copy_stream));
        FMX_HIP(ctx, hipEventRecord(spm->ev_copied[b], spm->copy_stream));
        FMX_HIP(ctx, hipStreamWaitEvent(s, spm->ev_copied[b], 0));
        // absolute word w of the image is at staging byte 2w - lo
        const uint16_t *img = (const uint16_t *)(spm->d_stage[b] - lo);
        fmx_status st = launch_spmm(ctx, spm, plan, dtype, k, X, ldx, Y, ldy, spm->partial.ptr, s, img,
                                    plan->chunk_items[c], plan->chunk_items[c + 1], c + 1 == nch, rscale, flags,
                                    Xhot, nhot);
        if (st) return st;
        FMX_HIP(ctx, hipEventRecord(spm->ev_free[b], s));
        spm->free_recorded[b] = true;
    }
    return FMX_OK;
}

// Pinned host image (already holding h_pinned_len bytes) -> validated,
// chunked streamed matrix.
fmx_status finish_streamed(fmx_ctx *ctx, fmx_spm *s, uint64_t chunk_bytes, bool whole_file)
{
    fmx_status st = analyse_image(ctx, s);
    if (st) return st;
    if (whole_file && s->bytes != s->h_pinned_len)
        return fail(ctx, FMX_EFORMAT, "matrix file is longer than its block rows");
    if (chunk_bytes < 4096) return fail(ctx, FMX_EINVAL, "chunk_bytes must be at least 4096");
    st = finish_streamed_chunks(ctx, s, chunk_bytes);
    if (st) return st;
    // the gather-order image, when its construction fits on the device (an
    // image larger than HBM streams in the reference's column order only)
    st = prepare_fast(ctx, s, chunk_bytes);
    if (st == FMX_ENOMEM) {
        s->fast.reset();
        st = FMX_OK;
    }
    return st;
}

fmx_status finish_streamed_chunks(fmx_ctx *ctx, fmx_spm *s, uint64_t chunk_bytes)
{
    fmx_status st = FMX_OK;
    if (const uint32_t h2 = work_rows_for(s)) {
        // tiles taller than an SpMM band: the re-cut work image (in pinned
        // host memory too) is what every SpMM streams, as a resident matrix
        // runs on its re-cut image; the given image stays for downloads
        std::vector<uint8_t> img;
        std::vector<int64_t> offs;
        retile_image(s, h2, img, offs);
        std::unique_ptr<fmx_spm> w(new fmx_spm());
        w->ctx = ctx;
        st = alloc_pinned(ctx, w.get(), img.size());
        if (st) return st;
        memcpy(w->h_pinned, img.data(), img.size());
        std::vector<uint8_t>().swap(img);
        w->offs = std::move(offs);
        st = analyse_image(ctx, w.get());
        if (!st) st = finish_streamed_chunks(ctx, w.get(), chunk_bytes);  // its tiles are short: no re-cut
        if (st) return fail(ctx, FMX_EINVAL, "internal error: the re-cut work image: " + ctx->last_error);
        if (w->nnz != s->nnz) return fail(ctx, FMX_EINVAL, "internal error: the re-cut work image lost non-zeros");
        s->work = std::move(w);
        return FMX_OK;
    }
    // chunks of whole block rows, each at most chunk_bytes unless one block
    // row alone is larger
    s->chunk_br.clear();
    s->chunk_lo.clear();
    s->chunk_hi.clear();
    uint64_t br = 0;
    while (br < s->nbr) {
        const uint64_t lo = (uint64_t)s->offs[br] & ~15ull;
        uint64_t e = br + 1;
        while (e < s->nbr && ((((uint64_t)s->offs[e + 1] + 15) & ~15ull) - lo) <= chunk_bytes) e++;
        s->chunk_br.push_back(br);
        s->chunk_lo.push_back(lo);
        s->chunk_hi.push_back(((uint64_t)s->offs[e] + 15) & ~15ull);
        s->stage_bytes = std::max(s->stage_bytes, s->chunk_hi.back() - lo);
        br = e;
    }
    s->chunk_br.push_back(s->nbr);
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    for (int b = 0; b < 2; b++) {
        FMX_HIP(ctx, hipMalloc(&s->d_stage[b], std::max<uint64_t>(s->stage_bytes, 16)));
        FMX_HIP(ctx, hipEventCreateWithFlags(&s->ev_copied[b], hipEventDisableTiming));
        FMX_HIP(ctx, hipEventCreateWithFlags(&s->ev_free[b], hipEventDisableTiming));
    }
    FMX_HIP(ctx, hipEventCreateWithFlags(&s->ev_start, hipEventDisableTiming));
    FMX_HIP(ctx, hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    return upload_dirs(ctx, s);
}

// Pinned buffer of `len` bytes padded with zeros to 4 KiB.
fmx_status alloc_pinned(fmx_ctx *ctx, fmx_spm *s, uint64_t len)
{
    const uint64_t padded = std::max<uint64_t>((len + 4095) & ~4095ull, 4096);
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    FMX_HIP(ctx, hipHostMalloc((void **)&s->h_pinned, padded, hipHostMallocDefault));
    memset(s->h_pinned + len, 0, padded - len);
    s->h_pinned_len = len;
    return FMX_OK;
}

// Prim type a weighted image must carry for dense data of dtype
// (sparse_matrix.cpp:754-766): P_INTEGER, P_LONG, P_ULONG, P_FLOAT, P_DOUBLE.
static int value_prim(fmx_dtype dt)
{
    switch (dt) {
    case FMX_I32: return 4;
    case FMX_I64: return 6;
    case FMX_U64: return 7;
    case FMX_F32: return 8;
    case FMX_F64: return 9;
    default: return -1;
    }
}

// The SpMM entries (fmx_spmm, fmx_spmm_ex, fmx_spmm_scaled).
fmx_status spmm_exec(fmx_ctx *ctx, const fmx_spm *cspm, fmx_dtype dtype, int k, const double *pre,
                     const double *post, const void *X, int64_t ldx, void *Y, int64_t ldy, uint32_t flags,
                     bool mixed, hipStream_t s)
{
    if (!ctx || !cspm) return fail(ctx, FMX_EINVAL, "null context or matrix");
    fmx_spm *spm = const_cast<fmx_spm *>(cspm);
    if (spm->ctx != ctx) return fail(ctx, FMX_EINVAL, "matrix belongs to another context");
    if (value_prim(dtype) < 0) return fail(ctx, FMX_ETYPE, "wrong matrix type in spmm");
    if (flags & ~(FMX_SPMM_SEQ | FMX_SPMM_X_COLMAJOR | FMX_SPMM_Y_COLMAJOR))
        return fail(ctx, FMX_EINVAL, "unknown SpMM flags");
    if ((pre || post) && dtype != FMX_F32 && dtype != FMX_F64)
        return fail(ctx, FMX_ETYPE, "row scaling needs FP32 or FP64 data");
    // a weighted matrix multiplies dense data of its own value type
    // (sparse_matrix.cpp:754-766); a custom creator may pair FP64 dense data
    // with FP32 values (spmm_creator<double, float>)
    if (spm->esz && spm->prim != value_prim(dtype) && !(mixed && dtype == FMX_F64 && spm->prim == 8))
        return fail(ctx, FMX_ETYPE, "wrong matrix type in spmm");
    if (k <= 0) return fail(ctx, FMX_EDIM, "k must be positive");
    const bool xcol = flags & FMX_SPMM_X_COLMAJOR, ycol = flags & FMX_SPMM_Y_COLMAJOR;
    if ((xcol ? (uint64_t)ldx < spm->ncols : ldx < k) || (ycol ? (uint64_t)ldy < spm->nrows : ldy < k))
        return fail(ctx, FMX_EDIM, "leading dimension smaller than the operand's rows (column-major) or k");
    const int es = elem_size(dtype);
    const bool seq = flags & FMX_SPMM_SEQ;
    // LDS accumulator bytes: FP32 sums in FP64, except in the sequential mode
    const int acc = (dtype == FMX_F32 && !seq) ? 8 : es;
    if ((uint64_t)k * acc > 32768) return fail(ctx, FMX_EUNSUPPORTED, "dense rows wider than 32 KiB");
    if ((!X && spm->ncols) || (!Y && spm->nrows)) return fail(ctx, FMX_EINVAL, "null dense operand");
    if (spm->nrows == 0) return FMX_OK;
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    // the image the kernels read: in the default mode the gather-order one;
    // in the sequential mode the reference's column order, re-cut if its
    // tiles are taller than a band
    const fmx_spm *orig = spm;
    // the gather-order image pays where X's rows share 128-byte lines and X
    // is larger than the XCDs' L2s together (32 MB): a row of a full line
    // gains nothing from being next to another hot row, and a small X stays
    // in L2 anyway, so only the restaging would remain
    // (FMX_HOT_MIN_BYTES, read at every call, moves the X-size threshold)
    const char *hm = getenv("FMX_HOT_MIN_BYTES");
    const uint64_t min_x = hm ? strtoull(hm, nullptr, 10) : (32ull << 20);
    const bool hot = !seq && spm->fast && (uint64_t)k * es < 128 && spm->ncols * (uint64_t)k * es > min_x;
    if (hot) spm = spm->fast.get();
    else if (spm->work) spm = spm->work.get();
    Plan *plan = nullptr;
    fmx_status st = get_plan(ctx, spm, band_rows(spm, k, acc), seq, &plan);
    if (st) return st;
    size_t need = (size_t)plan->nslots * plan->hs * k * (dtype == FMX_F32 ? 8 : es);
    if (need > spm->partial.bytes) {
        if (spm->partial.ptr) FMX_HIP(ctx, hipFree(spm->partial.ptr));
        spm->partial.ptr = nullptr;
        spm->partial.bytes = 0;
        FMX_HIP(ctx, hipMalloc(&spm->partial.ptr, need));
        spm->partial.bytes = need;
    }
    // a column-major or pre-scaled X is restaged row-major first (the
    // reference converts a non-row-major input, sparse_matrix.cpp:720-736)
    if ((xcol || pre) && spm->ncols) {
        const size_t xb = spm->ncols * (size_t)k * es;
        if (xb > ctx->xstage.bytes) {
            if (ctx->xstage.ptr) FMX_HIP(ctx, hipFree(ctx->xstage.ptr));
            ctx->xstage.ptr = nullptr;
            ctx->xstage.bytes = 0;
            FMX_HIP(ctx, hipMalloc(&ctx->xstage.ptr, xb));
            ctx->xstage.bytes = xb;
        }
        st = stage_dense_rows(ctx, dtype, spm->ncols, k, X, ldx, xcol, pre, ctx->xstage.ptr, s);
        if (st) return st;
        X = ctx->xstage.ptr;
        ldx = k;
    }
    const uint32_t kflags = flags & (FMX_SPMM_SEQ | FMX_SPMM_Y_COLMAJOR);
    if (hot) {
        // X's hot rows, densely, in the gather-order image's column order
        const size_t hb = orig->nhot * (size_t)k * es;
        if (hb > ctx->xhot.bytes) {
            if (ctx->xhot.ptr) FMX_HIP(ctx, hipFree(ctx->xhot.ptr));
            ctx->xhot.ptr = nullptr;
            ctx->xhot.bytes = 0;
            FMX_HIP(ctx, hipMalloc(&ctx->xhot.ptr, hb));
            ctx->xhot.bytes = hb;
        }
        st = stage_hot_rows(ctx, dtype, orig->nhot, k, X, ldx, orig->d_hot, ctx->xhot.ptr, s);
        if (st) return st;
        if (spm->streamed())
            return spmm_streamed(ctx, spm, plan, dtype, k, X, ldx, Y, ldy, s, post, kflags, ctx->xhot.ptr,
                                 orig->nhot);
        return launch_spmm(ctx, spm, plan, dtype, k, X, ldx, Y, ldy, spm->partial.ptr, s,
                           (const uint16_t *)spm->d_image, 0, plan->nitems, true, post, kflags, ctx->xhot.ptr,
                           orig->nhot);
    }
    if (!spm->streamed())
        return launch_spmm(ctx, spm, plan, dtype, k, X, ldx, Y, ldy, spm->partial.ptr, s,
                           (const uint16_t *)spm->d_image, 0, plan->nitems, true, post, kflags);
    return spmm_streamed(ctx, spm, plan, dtype, k, X, ldx, Y, ldy, s, post, kflags);
}

fmx_status spmm_scaled(fmx_ctx *ctx, const fmx_spm *spm, fmx_dtype dtype, int k, const void *X, int64_t ldx,
                       void *Y, int64_t ldy, const double *rscale, hipStream_t s)
{
    return spmm_exec(ctx, spm, dtype, k, nullptr, rscale, X, ldx, Y, ldy, 0, false, s);
}

}  // namespace fmx

fmx_spm::~fmx_spm()
{
    // no copy or kernel may still use the staging buffers
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    for (int b = 0; b < 2; b++)
        if (free_recorded[b]) (void)hipEventSynchronize(ev_free[b]);
    if (d_image) (void)hipFree(d_image);
    if (d_dir) (void)hipFree(d_dir);
    if (d_vdir) (void)hipFree(d_vdir);
    if (d_hot) (void)hipFree(d_hot);
    for (int b = 0; b < 2; b++) {
        if (d_stage[b]) (void)hipFree(d_stage[b]);
        if (ev_copied[b]) (void)hipEventDestroy(ev_copied[b]);
        if (ev_free[b]) (void)hipEventDestroy(ev_free[b]);
    }
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (h_pinned) (void)hipHostFree(h_pinned);
}

using namespace fmx;

extern "C" {

const char *fmx_strerror(fmx_status st)
{
    switch (st) {
    case FMX_OK: return "success";
    case FMX_EINVAL: return "invalid argument";
    case FMX_ETYPE: return "unsupported or mismatched element type";
    case FMX_EDIM: return "dimension mismatch";
    case FMX_EFORMAT: return "malformed tile image";
    case FMX_ENOMEM: return "out of memory";
    case FMX_EHIP: return "HIP runtime error";
    case FMX_EUNSUPPORTED: return "not supported by this build";
    }
    return "unknown status";
}

fmx_status fmx_ctx_create(int device, fmx_ctx **out)
{
    if (!out) return FMX_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FMX_EHIP;
    if (device < 0 || device >= ndev) return FMX_EINVAL;
    std::unique_ptr<fmx_ctx> c(new fmx_ctx());
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) return FMX_EHIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return FMX_EHIP;
    c->num_cus = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return FMX_EHIP;
    // the fixed scratch of fmx_gram_noalloc, looked up with the device current
    if (noalloc_scratch_lookup(c.get()) != FMX_OK) {
        (void)hipStreamDestroy(c->stream);
        return FMX_EHIP;
    }
    *out = c.release();
    return FMX_OK;
}

fmx_status fmx_ctx_destroy(fmx_ctx *ctx)
{
    if (!ctx) return FMX_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    if (ctx->ws) (void)hipFree(ctx->ws);
    delete ctx;
    return FMX_OK;
}

fmx_status fmx_ctx_stream(fmx_ctx *ctx, void **stream)
{
    if (!ctx || !stream) return FMX_EINVAL;
    *stream = ctx->stream;
    return FMX_OK;
}

fmx_status fmx_ctx_sync(fmx_ctx *ctx)
{
    if (!ctx) return FMX_EINVAL;
    FMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FMX_OK;
}

const char *fmx_ctx_last_error(const fmx_ctx *ctx)
{
    return ctx ? ctx->last_error.c_str() : "";
}

fmx_status fmx_spm_upload(fmx_ctx *ctx, const void *image, size_t bytes,
                          const int64_t *offs, size_t nbr, fmx_spm **out)
{
    if (!ctx || !image || !offs || !out) return fail(ctx, FMX_EINVAL, "null argument");
    *out = nullptr;
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<fmx_spm> s(new fmx_spm());
    s->ctx = ctx;
    s->h_image.assign((const uint8_t *)image, (const uint8_t *)image + bytes);
    s->offs.assign(offs, offs + nbr + 1);
    fmx_status st = analyse_image(ctx, s.get());
    if (st) return st;
    size_t padded = (s->bytes + 4095) & ~size_t(4095);
    FMX_HIP(ctx, hipMalloc(&s->d_image, padded));
    FMX_HIP(ctx, hipMemcpy(s->d_image, s->h_image.data(), s->bytes, hipMemcpyHostToDevice));
    if (padded > s->bytes) FMX_HIP(ctx, hipMemset(s->d_image + s->bytes, 0, padded - s->bytes));
    st = upload_dirs(ctx, s.get());
    if (st) return st;
    st = prepare_work(ctx, s.get());
    if (st) return st;
    st = prepare_fast(ctx, s.get());
    if (st) return st;
    *out = s.release();
    return FMX_OK;
}

fmx_status fmx_spm_info_get(const fmx_spm *spm, fmx_spm_info *info)
{
    if (!spm || !info) return FMX_EINVAL;
    info->nrows = spm->nrows;
    info->ncols = spm->ncols;
    info->nnz = spm->nnz;
    info->block_h = spm->H;
    info->block_w = spm->W;
    info->nbr = spm->nbr;
    info->image_bytes = spm->bytes;
    info->tile_bytes = spm->bytes - kHeaderBytes;
    info->entry_size = spm->esz;
    info->ntiles = spm->ntiles;
    const fmx_spm *run = spm->work ? spm->work.get() : spm;  // the image every SpMM reads
    info->stream_chunks = spm->streamed() ? run->chunk_br.size() - 1 : 0;
    info->build_ns = spm->build_ns;
    info->edges_in = spm->edges_in;
    info->hot_columns = spm->fast ? spm->nhot : 0;
    return FMX_OK;
}

fmx_status fmx_spm_download(const fmx_spm *spm, void *image, size_t cap,
                            size_t *bytes, int64_t *offs)
{
    if (!spm) return FMX_EINVAL;
    fmx_ctx *ctx = spm->ctx;
    if (bytes) *bytes = spm->bytes;
    if (image) {
        if (cap < spm->bytes) return fail(ctx, FMX_EINVAL, "image buffer too small");
        if (spm->streamed()) {
            memcpy(image, spm->h_pinned, spm->bytes);
        } else {
            FMX_HIP(ctx, hipSetDevice(ctx->device));
            FMX_HIP(ctx, hipMemcpy(image, spm->d_image, spm->bytes, hipMemcpyDeviceToHost));
        }
    }
    if (offs) memcpy(offs, spm->offs.data(), spm->offs.size() * sizeof(int64_t));
    return FMX_OK;
}

namespace {

bool read_file(const char *path, std::vector<uint8_t> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    bool ok = fseek(f, 0, SEEK_END) == 0;
    long n = ok ? ftell(f) : -1;
    ok = ok && n >= 0 && fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        buf.resize((size_t)n);
        ok = n == 0 || fread(buf.data(), (size_t)n, 1, f) == 1;
    }
    fclose(f);
    return ok;
}

bool write_file(const char *path, const void *a, size_t na, const void *b, size_t nb)
{
    FILE *f = fopen(path, "wb");
    if (!f) return false;
    bool ok = (na == 0 || fwrite(a, na, 1, f) == 1) && (nb == 0 || fwrite(b, nb, 1, f) == 1);
    ok = (fclose(f) == 0) && ok;
    return ok;
}

}  // namespace

fmx_status fmx_spm_load(fmx_ctx *ctx, const char *mat_file, const char *idx_file, fmx_spm **out)
{
    if (!ctx || !mat_file || !idx_file || !out) return fail(ctx, FMX_EINVAL, "null argument");
    *out = nullptr;
    std::vector<uint8_t> idx, img;
    if (!read_file(idx_file, idx)) return fail(ctx, FMX_EINVAL, std::string("can't read ") + idx_file);
    if (!read_file(mat_file, img)) return fail(ctx, FMX_EINVAL, std::string("can't read ") + mat_file);
    // SpM_2d_index: the matrix header page, then nbr+1 block-row offsets
    if (idx.size() < kHeaderBytes + 8 || (idx.size() - kHeaderBytes) % 8)
        return fail(ctx, FMX_EFORMAT, "index file is not a header plus 8-byte offsets");
    if (img.size() < kHeaderBytes) return fail(ctx, FMX_EFORMAT, "matrix file shorter than its header");
    // both files describe the same matrix: type, entry size, shape, layout, tiles
    if (memcmp(idx.data(), img.data(), 64) != 0)
        return fail(ctx, FMX_EFORMAT, "index and matrix headers differ");
    const size_t nent = (idx.size() - kHeaderBytes) / 8;
    std::vector<int64_t> offs(nent);
    memcpy(offs.data(), idx.data() + kHeaderBytes, nent * 8);
    fmx_status st = fmx_spm_upload(ctx, img.data(), img.size(), offs.data(), nent - 1, out);
    if (st) return st;
    if ((*out)->bytes != img.size()) {
        fmx_spm_destroy(*out);
        *out = nullptr;
        return fail(ctx, FMX_EFORMAT, "matrix file is longer than its block rows");
    }
    return FMX_OK;
}

fmx_status fmx_spm_upload_streamed(fmx_ctx *ctx, const void *image, size_t bytes, const int64_t *offs, size_t nbr,
                                   uint64_t chunk_bytes, fmx_spm **out)
{
    if (!ctx || !image || !offs || !out) return fail(ctx, FMX_EINVAL, "null argument");
    *out = nullptr;
    std::unique_ptr<fmx_spm> s(new fmx_spm());
    s->ctx = ctx;
    fmx_status st = alloc_pinned(ctx, s.get(), bytes);
    if (st) return st;
    memcpy(s->h_pinned, image, bytes);
    s->offs.assign(offs, offs + nbr + 1);
    st = finish_streamed(ctx, s.get(), chunk_bytes, false);
    if (st) return st;
    *out = s.release();
    return FMX_OK;
}

fmx_status fmx_spm_open_index(fmx_ctx *ctx, const char *mat_file, const void *header, const int64_t *offs,
                              size_t nbr, uint64_t chunk_bytes, fmx_spm **out)
{
    if (!ctx || !mat_file || !header || !offs || !out) return fail(ctx, FMX_EINVAL, "null argument");
    *out = nullptr;
    FILE *f = fopen(mat_file, "rb");
    if (!f) return fail(ctx, FMX_EINVAL, std::string("can't read ") + mat_file);
    long n = -1;
    if (fseek(f, 0, SEEK_END) == 0) n = ftell(f);
    if (n < 0 || fseek(f, 0, SEEK_SET) != 0) {
        fclose(f);
        return fail(ctx, FMX_EINVAL, std::string("can't read ") + mat_file);
    }
    std::unique_ptr<fmx_spm> s(new fmx_spm());
    s->ctx = ctx;
    fmx_status st = alloc_pinned(ctx, s.get(), (uint64_t)n);
    if (st) {
        fclose(f);
        return st;
    }
    const bool ok = n == 0 || fread(s->h_pinned, (size_t)n, 1, f) == 1;
    fclose(f);
    if (!ok) return fail(ctx, FMX_EINVAL, std::string("can't read ") + mat_file);
    if ((uint64_t)n < kHeaderBytes) return fail(ctx, FMX_EFORMAT, "matrix file shorter than its header");
    // both describe the same matrix: type, entry size, shape, layout, tiles
    if (memcmp(header, s->h_pinned, 64) != 0) return fail(ctx, FMX_EFORMAT, "index and matrix headers differ");
    s->offs.assign(offs, offs + nbr + 1);
    st = finish_streamed(ctx, s.get(), chunk_bytes, true);
    if (st) return st;
    *out = s.release();
    return FMX_OK;
}

fmx_status fmx_spm_open(fmx_ctx *ctx, const char *mat_file, const char *idx_file, uint64_t chunk_bytes,
                        fmx_spm **out)
{
    if (!ctx || !mat_file || !idx_file || !out) return fail(ctx, FMX_EINVAL, "null argument");
    *out = nullptr;
    std::vector<uint8_t> idx;
    if (!read_file(idx_file, idx)) return fail(ctx, FMX_EINVAL, std::string("can't read ") + idx_file);
    if (idx.size() < kHeaderBytes + 8 || (idx.size() - kHeaderBytes) % 8)
        return fail(ctx, FMX_EFORMAT, "index file is not a header plus 8-byte offsets");
    const size_t nent = (idx.size() - kHeaderBytes) / 8;
    std::vector<int64_t> offs(nent);
    memcpy(offs.data(), idx.data() + kHeaderBytes, nent * 8);
    return fmx_spm_open_index(ctx, mat_file, idx.data(), offs.data(), nent - 1, chunk_bytes, out);
}

fmx_status fmx_spm_dump(const fmx_spm *spm, const char *mat_file, const char *idx_file)
{
    if (!spm || !mat_file || !idx_file) return FMX_EINVAL;
    fmx_ctx *ctx = spm->ctx;
    std::vector<uint8_t> img(spm->bytes);
    fmx_status st = fmx_spm_download(spm, img.data(), img.size(), nullptr, nullptr);
    if (st) return st;
    if (!write_file(mat_file, img.data(), img.size(), nullptr, 0))
        return fail(ctx, FMX_EINVAL, std::string("can't write ") + mat_file);
    if (!write_file(idx_file, img.data(), kHeaderBytes, spm->offs.data(), spm->offs.size() * sizeof(int64_t)))
        return fail(ctx, FMX_EINVAL, std::string("can't write ") + idx_file);
    return FMX_OK;
}

fmx_status fmx_spm_destroy(fmx_spm *spm)
{
    if (!spm) return FMX_EINVAL;
    (void)hipSetDevice(spm->ctx->device);
    delete spm;
    return FMX_OK;
}

fmx_status fmx_spmm_alg_bytes(const fmx_spm *spm, fmx_dtype dtype, int k, uint64_t *bytes)
{
    if (!spm || !bytes || k <= 0) return FMX_EINVAL;
    int es = elem_size(dtype);
    if (!es || dtype == FMX_BF16) return FMX_ETYPE;
    *bytes = (spm->bytes - kHeaderBytes) + 8 * (spm->nbr + 1) +
             spm->ncols * (uint64_t)k * es + spm->nrows * (uint64_t)k * es;
    return FMX_OK;
}

fmx_status fmx_spmm(fmx_ctx *ctx, const fmx_spm *cspm, fmx_dtype dtype, int k,
                    const void *X, int64_t ldx, void *Y, int64_t ldy, void *stream)
{
    return fmx::spmm_exec(ctx, cspm, dtype, k, nullptr, nullptr, X, ldx, Y, ldy, 0, false, (hipStream_t)stream);
}

fmx_status fmx_spmm_ex(fmx_ctx *ctx, const fmx_spm *cspm, fmx_dtype dtype, int k, const void *X, int64_t ldx,
                       void *Y, int64_t ldy, uint32_t flags, void *stream)
{
    return fmx::spmm_exec(ctx, cspm, dtype, k, nullptr, nullptr, X, ldx, Y, ldy, flags, false,
                          (hipStream_t)stream);
}

fmx_status fmx_spmm_scaled(fmx_ctx *ctx, const fmx_spm *cspm, fmx_dtype dtype, int k, const double *pre,
                           const double *post, const void *X, int64_t ldx, void *Y, int64_t ldy, uint32_t flags,
                           void *stream)
{
    return fmx::spmm_exec(ctx, cspm, dtype, k, pre, post, X, ldx, Y, ldy, flags, true, (hipStream_t)stream);
}

fmx_status fmx_spmm_host(fmx_ctx *ctx, const fmx_spm *spm, fmx_dtype dtype, int k, const void *X, int64_t ldx,
                         void *Y, int64_t ldy, uint64_t mem_bytes, void *stream)
{
    if (!ctx || !spm) return fail(ctx, FMX_EINVAL, "null context or matrix");
    const int es = elem_size(dtype);
    if (!es || dtype == FMX_BF16) return fail(ctx, FMX_ETYPE, "wrong matrix type in spmm");
    if (k <= 0) return fail(ctx, FMX_EDIM, "k must be positive");
    if (ldx < k || ldy < k) return fail(ctx, FMX_EDIM, "leading dimension smaller than k");
    if ((!X && spm->ncols) || (!Y && spm->nrows)) return fail(ctx, FMX_EINVAL, "null dense operand");
    if (spm->nrows == 0) return FMX_OK;
    const uint64_t per_col = (spm->ncols + spm->nrows) * (uint64_t)es;
    const uint64_t kb = std::min<uint64_t>((uint64_t)k, mem_bytes / per_col);
    if (kb == 0) return fail(ctx, FMX_ENOMEM, "The memory is too small for the SpMM");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    void *dx = nullptr, *dy = nullptr;
    FMX_HIP(ctx, hipMalloc(&dx, std::max<uint64_t>(spm->ncols * kb * es, 16)));
    if (hipMalloc(&dy, spm->nrows * kb * es) != hipSuccess) {
        (void)hipFree(dx);
        return fail(ctx, FMX_ENOMEM, "dense panel");
    }
    fmx_status st = FMX_OK;
    for (uint64_t j0 = 0; j0 < (uint64_t)k && !st; j0 += kb) {
        const uint64_t w = std::min<uint64_t>(kb, (uint64_t)k - j0);
        if (spm->ncols &&
            hipMemcpy2DAsync(dx, w * es, (const char *)X + j0 * es, (size_t)ldx * es, w * es, spm->ncols,
                             hipMemcpyHostToDevice, s) != hipSuccess) {
            st = fail(ctx, FMX_EHIP, "dense panel copy");
            break;
        }
        st = fmx_spmm(ctx, spm, dtype, (int)w, dx, (int64_t)w, dy, (int64_t)w, s);
        if (!st && hipMemcpy2DAsync((char *)Y + j0 * es, (size_t)ldy * es, dy, w * es, w * es, spm->nrows,
                                    hipMemcpyDeviceToHost, s) != hipSuccess)
            st = fail(ctx, FMX_EHIP, "dense panel copy");
    }
    if (hipStreamSynchronize(s) != hipSuccess && !st) st = fail(ctx, FMX_EHIP, "dense panel copy");
    (void)hipFree(dx);
    (void)hipFree(dy);
    return st;
}

}  // extern "C"
