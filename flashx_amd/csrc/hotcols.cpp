// hotcols.cpp -- the gather-order image of a matrix.
//
// Every non-zero (r, c) of the SpMM gathers row c of X, and on a power-law
// graph a few columns take most of the gathers (R-MAT scale 22: the 65,536
// most used columns hold 63% of the non-zeros, the 600,000 most used 95%).
// In the reference's column order those rows lie scattered over X, each
// sharing its 128-byte cache line with an unrelated row, so the L2 of an XCD
// holds far fewer of them than it could.  The gather-order image relabels
// the columns instead: the nhot most used columns (the fewest whose
// non-zeros make up 95% of the total) become columns 0..nhot-1 in
// descending order of use, and every other column c becomes nhot + c.  Each
// default-mode SpMM first copies X's hot rows into a dense nhot x k block
// (stage_hot_rows; 2 * nhot * k * s bytes, a few percent of the SpMM's
// traffic) and gathers from it, and from X for the rest.  The rows of A are
// not moved, so Y is written in place and the result is the same sum of the
// same products; only the FP64 accumulation order inside a row changes, and
// FP32 data are still rounded once from the FP64 sum.  The sequential mode
// sums in the reference's column order and keeps using the image in that
// order.
//
// The image is built once, when the matrix is created: the column degrees
// and the relabelled (row, column[, value]) list come from the image in HBM
// (image_dev.hip) or, for a matrix streamed from host memory, from its
// host copy (below); the device builder (build_tiles.hip) then writes the
// new image in tiles of at most 64 rows by 32768 columns.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>

#include "fmx_internal.h"

namespace fmx {
namespace {

uint16_t rd16(const uint8_t *p)
{
    uint16_t v;
    memcpy(&v, p, 2);
    return v;
}
uint32_t rd32(const uint8_t *p)
{
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

// Calls fn(row, col, value pointer) for every non-zero of block row br, in
// stream order (SCSR row parts, then COO entries, tile by tile).
template <typename F>
void for_nonzeros(const fmx_spm *s, uint64_t br, F fn)
{
    const uint8_t *base = s->host_image() + s->offs[br];
    for (uint32_t t = s->brow_dir[br]; t < s->brow_dir[br + 1]; t++) {
        const uint8_t *p = base + 2ull * s->dir[t].x;
        const uint32_t nnz = rd32(p + 8);
        if (!nnz) continue;
        const uint32_t nrow = rd16(p + 12), ncoo = rd16(p + 14);
        const uint32_t scsr_words = (nrow - ncoo) + (nnz - ncoo);
        const uint8_t *w = p + 16, *coo = w + 2ull * scsr_words + 2, *vals = coo + 4ull * ncoo;
        const uint64_t r0 = br * s->H, c0 = (uint64_t)s->dir[t].y * s->W;
        uint32_t e = 0;
        uint64_t r = 0;
        for (uint32_t i = 0; i < scsr_words; i++) {
            const uint16_t v = rd16(w + 2ull * i);
            if (v & 0x8000u) {
                r = r0 + (v & 0x7FFFu);
                continue;
            }
            fn(r, c0 + v, vals + s->esz * e);
            e++;
        }
        for (uint32_t i = 0; i < ncoo; i++, e++)
            fn(r0 + (rd16(coo + 4ull * i) & 0x7FFFu), c0 + rd16(coo + 4ull * i + 2), vals + s->esz * e);
    }
}

unsigned host_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

// The host form of gather_order_lists (image_dev.hip) for an image that
// is not in HBM: the same lists, from the host copy, uploaded into rows,
// cols and vals; *hot_out (device) and *nhot_out as there.
fmx_status host_lists(fmx_ctx *ctx, const fmx_spm *src, uint64_t ncols, uint64_t permille, uint32_t *d_rows,
                      uint32_t *d_cols, void *d_vals, uint32_t **hot_out, uint64_t *nhot_out)
{
    *hot_out = nullptr;
    *nhot_out = 0;
    const uint64_t nnz = src->nnz, esz = src->esz;
    const unsigned nth = host_threads();
    // column degrees: per-thread counts over block-row stripes, then summed
    std::vector<std::vector<uint32_t>> part(nth, std::vector<uint32_t>(ncols, 0));
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; t++)
            th.emplace_back([&, t] {
                uint32_t *d = part[t].data();
                for (uint64_t br = t; br < src->nbr; br += nth)
                    for_nonzeros(src, br, [&](uint64_t, uint64_t c, const uint8_t *) { d[c]++; });
            });
        for (auto &x : th) x.join();
    }
    std::vector<uint32_t> deg(ncols, 0);
    for (unsigned t = 0; t < nth; t++) {
        for (uint64_t c = 0; c < ncols; c++) deg[c] += part[t][c];
        std::vector<uint32_t>().swap(part[t]);
    }
    // the most used columns, descending, ties by column: the fewest whose
    // non-zeros make up the wanted share of the total
    std::vector<uint32_t> order(ncols);
    {
        // counting sort by descending degree, ascending column among equals
        const uint32_t maxd = *std::max_element(deg.begin(), deg.end());
        std::vector<uint64_t> pos((size_t)maxd + 2, 0);
        for (uint64_t c = 0; c < ncols; c++) pos[maxd - deg[c] + 1]++;
        for (size_t i = 1; i < pos.size(); i++) pos[i] += pos[i - 1];
        for (uint64_t c = 0; c < ncols; c++) order[pos[maxd - deg[c]]++] = (uint32_t)c;
    }
    uint64_t nhot = 0, acc = 0;
    while (nhot < ncols && acc * 1000 < nnz * permille && deg[order[nhot]] > 0) acc += deg[order[nhot++]];
    if (nhot == 0 || nhot + ncols > (1ull << 32)) return FMX_OK;  // no hot columns, or ids past 32 bits
    std::vector<uint32_t> newid(ncols);
    for (uint64_t c = 0; c < ncols; c++) newid[c] = (uint32_t)(nhot + c);
    for (uint64_t i = 0; i < nhot; i++) newid[order[i]] = (uint32_t)i;
    // the relabelled non-zeros in stream order, each block row's at its
    // prefix-sum position
    std::vector<uint64_t> nzoff(src->nbr + 1, 0);
    for (uint64_t br = 0; br < src->nbr; br++) {
        uint64_t cnt = 0;
        for (uint32_t t = src->brow_dir[br]; t < src->brow_dir[br + 1]; t++)
            cnt += rd32(src->host_image() + src->offs[br] + 2ull * src->dir[t].x + 8);
        nzoff[br + 1] = nzoff[br] + cnt;
    }
    // pinned, so that the uploads run at the host link's rate
    struct Pinned {
        void *p = nullptr;
        ~Pinned() { if (p) (void)hipHostFree(p); }
    } prow, pcol, pval;
    FMX_HIP(ctx, hipHostMalloc(&prow.p, nnz * 4, hipHostMallocDefault));
    FMX_HIP(ctx, hipHostMalloc(&pcol.p, nnz * 4, hipHostMallocDefault));
    if (esz) FMX_HIP(ctx, hipHostMalloc(&pval.p, nnz * esz, hipHostMallocDefault));
    uint32_t *rows = (uint32_t *)prow.p, *cols = (uint32_t *)pcol.p;
    uint8_t *vals = (uint8_t *)pval.p;
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; t++)
            th.emplace_back([&, t] {
                for (uint64_t br = t; br < src->nbr; br += nth) {
                    uint64_t i = nzoff[br];
                    for_nonzeros(src, br, [&](uint64_t r, uint64_t c, const uint8_t *v) {
                        rows[i] = (uint32_t)r;
                        cols[i] = newid[c];
                        if (esz) memcpy(vals + esz * i, v, esz);
                        i++;
                    });
                }
            });
        for (auto &x : th) x.join();
    }
    if (hipMemcpy(d_rows, rows, nnz * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_cols, cols, nnz * 4, hipMemcpyHostToDevice) != hipSuccess ||
        (esz && hipMemcpy(d_vals, vals, nnz * esz, hipMemcpyHostToDevice) != hipSuccess))
        return fail(ctx, FMX_EHIP, "gather-order image upload");
    uint32_t *hot = nullptr;
    FMX_HIP(ctx, hipMalloc(&hot, nhot * 4));
    if (hipMemcpy(hot, order.data(), nhot * 4, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(hot);
        return fail(ctx, FMX_EHIP, "gather-order image upload");
    }
    *hot_out = hot;
    *nhot_out = nhot;
    return FMX_OK;
}

}  // namespace

fmx_status prepare_fast(fmx_ctx *ctx, fmx_spm *s, uint64_t stream_chunk)
{
    const char *e = getenv("FMX_HOT_COLS");  // "0" turns the gather-order image off
    if ((e && !strcmp(e, "0")) || s->nnz == 0 || s->nnz >= (1ull << 31)) return FMX_OK;
    const fmx_spm *src = s->work ? s->work.get() : s;  // either walks the same non-zeros
    // the share of the non-zeros the hot columns cover (FMX_HOT_PERMILLE)
    const char *pm = getenv("FMX_HOT_PERMILLE");
    const uint64_t permille = std::min<uint64_t>(1000, pm ? strtoull(pm, nullptr, 10) : 950);
    struct Lists {
        uint32_t *rows = nullptr, *cols = nullptr, *hot = nullptr;
        void *vals = nullptr;
        ~Lists()
        {
            if (rows) (void)hipFree(rows);
            if (cols) (void)hipFree(cols);
            if (vals) (void)hipFree(vals);
            if (hot) (void)hipFree(hot);
        }
    } l;
    if (hipMalloc(&l.rows, s->nnz * 4) != hipSuccess || hipMalloc(&l.cols, s->nnz * 4) != hipSuccess ||
        (s->esz && hipMalloc(&l.vals, s->nnz * s->esz) != hipSuccess))
        return fail(ctx, FMX_ENOMEM, "gather-order image");
    uint64_t nhot = 0;
    fmx_status st = src->d_image
                        ? gather_order_lists(ctx, src, s->ncols, permille, l.rows, l.cols, l.vals, &l.hot, &nhot)
                        : host_lists(ctx, src, s->ncols, permille, l.rows, l.cols, l.vals, &l.hot, &nhot);
    if (st || nhot == 0) return st;
    fmx_spm *f = nullptr;
    {
        const uint32_t lh = (uint32_t)std::min(s->log2h, 6), lw = 15;
        // every non-zero as it is: self edges kept, duplicates kept (in
        // stream order, so a weighted row keeps its value order)
        st = build_impl(ctx, l.rows, l.cols, s->nnz, s->nrows, lh, lw, FMX_BUILD_KEEP_SELF,
                        s->esz ? l.vals : nullptr, (uint32_t)s->esz, s->prim, &f, nhot + s->ncols, true);
    }
    if (st) return st;
    std::unique_ptr<fmx_spm> fast(f);
    if (fast->nnz != s->nnz) return fail(ctx, FMX_EINVAL, "internal error: the gather-order image lost non-zeros");
    if (stream_chunk) {
        // a streamed matrix streams its gather-order image from pinned host
        // memory as well, chunked like the image it was made from
        fmx_spm *d = fast.get();
        std::unique_ptr<fmx_spm> h(new fmx_spm());
        h->ctx = ctx;
        fmx_status st2 = alloc_pinned(ctx, h.get(), d->bytes);
        if (st2) return st2;
        FMX_HIP(ctx, hipMemcpy(h->h_pinned, d->d_image, d->bytes, hipMemcpyDeviceToHost));
        h->offs = d->offs;
        fast.reset();
        st2 = analyse_image(ctx, h.get());
        if (!st2) st2 = finish_streamed_chunks(ctx, h.get(), stream_chunk);
        if (st2) return st2;
        fast = std::move(h);
    }
    s->d_hot = l.hot;
    l.hot = nullptr;
    s->nhot = nhot;
    s->fast = std::move(fast);
    return FMX_OK;
}

}  // namespace fmx
