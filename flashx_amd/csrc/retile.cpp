// retile.cpp -- the work image of a matrix whose tiles are taller than an
// SpMM band.
//
// The SpMM kernel gives each work item one band of rows of one block row and
// walks that block row's words; when a tile is taller than the band (the
// reference's default 32768 x 32768 tiles, flash-graph/utils/fg2fm.cpp:
// 155-156, against bands of 64-2048 rows), every band walks the whole block
// row to pick out its own rows.  Such an image is therefore re-cut once, when
// the matrix is created, into a work image of 64-row tiles (work_rows_for)
// with the same tile width: every row part and single-entry row of a tile
// stays what
// it is (its columns, its values, SCSR or COO), and moves to the tile of its
// new block row, in the same block column.  Row r of the matrix then
// accumulates the same non-zeros in the same ascending-column order, so the
// sequential mode is bit-identical on both images.  Resident and streamed
// matrices are re-cut alike (a streamed one keeps its work image in pinned
// host memory), so the two give the same bits in every mode.  The image
// given by the caller is kept unchanged: downloads and dumps return it, and
// the algorithmic bytes are counted on it.
#include <algorithm>
#include <cstdlib>
#include <cstring>
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
void wr16(uint8_t *p, uint16_t v) { memcpy(p, &v, 2); }
void wr32(uint8_t *p, uint32_t v) { memcpy(p, &v, 4); }

// One sub-tile: the part of an original tile in one new block row.
struct Part {
    uint32_t sub;           // new block row, relative to the original's first
    uint32_t w0, w1;        // SCSR words [w0, w1) (row heads and columns), relative to the tile
    uint32_t nrow_scsr, nnz_scsr;
    uint32_t v0;            // first SCSR value index
    uint32_t c0, c1;        // COO entries [c0, c1)
};

// The parts of one validated tile at p (analyse_image checked its layout),
// in ascending new block row.
void tile_parts(const uint8_t *p, uint32_t log2h2, std::vector<Part> &out)
{
    out.clear();
    const uint32_t nnz = rd32(p + 8), nrow = rd16(p + 12), ncoo = rd16(p + 14);
    const uint32_t scsr_words = (nrow - ncoo) + (nnz - ncoo);
    const uint8_t *w = p + 16;
    const uint8_t *coo = w + 2ull * scsr_words + 2;
    // SCSR row parts, ascending rows: a part is its head and the columns up
    // to the next head (the region end is positional)
    uint32_t i = 0, vi = 0;
    while (i < scsr_words) {
        const uint32_t r = rd16(w + 2ull * i) & 0x7FFFu, sub = r >> log2h2;
        uint32_t j = i + 1;
        while (j < scsr_words && !(rd16(w + 2ull * j) & 0x8000u)) j++;
        if (out.empty() || out.back().sub != sub) out.push_back({sub, i, i, 0, 0, vi, 0, 0});
        Part &pt = out.back();
        pt.w1 = j;
        pt.nrow_scsr++;
        pt.nnz_scsr += j - i - 1;
        vi += j - i - 1;
        i = j;
    }
    // COO entries, ascending rows, merged into the parts of their rows
    size_t k = 0;
    for (uint32_t c = 0; c < ncoo;) {
        const uint32_t sub = (rd16(coo + 4ull * c) & 0x7FFFu) >> log2h2;
        uint32_t e = c + 1;
        while (e < ncoo && ((rd16(coo + 4ull * e) & 0x7FFFu) >> log2h2) == sub) e++;
        while (k < out.size() && out[k].sub < sub) k++;
        if (k == out.size() || out[k].sub != sub) {
            out.insert(out.begin() + k, Part{sub, 0, 0, 0, 0, 0, 0, 0});
        }
        out[k].c0 = c;
        out[k].c1 = e;
        c = e;
    }
}

uint64_t part_bytes(const Part &pt, uint64_t esz)
{
    const uint64_t ncoo = pt.c1 - pt.c0, nnz = pt.nnz_scsr + ncoo, nrow = pt.nrow_scsr + ncoo;
    return 16 + 2 * nrow + 2 * nnz + 2 + esz * nnz;
}

}  // namespace

// Builds the work image of `s` (H > h2) in img / offs.
void retile_image(const fmx_spm *s, uint32_t h2, std::vector<uint8_t> &img, std::vector<int64_t> &offs)
{
    const uint8_t *src = s->host_image();
    uint32_t log2h2 = 0;
    while ((1u << log2h2) < h2) log2h2++;
    const uint64_t per = s->H >> log2h2;  // new block rows per original one
    const uint64_t nbr2 = (s->nrows + h2 - 1) >> log2h2;
    const uint64_t esz = s->esz;
    std::vector<uint64_t> size2(nbr2, 0);
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto for_rows = [&](auto fn) {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; t++)
            th.emplace_back([&, t] {
                std::vector<Part> parts;
                for (uint64_t br = t; br < s->nbr; br += nth) fn(br, parts);
            });
        for (auto &x : th) x.join();
    };
    // pass 1: bytes of every new block row
    for_rows([&](uint64_t br, std::vector<Part> &parts) {
        for (uint32_t t = s->brow_dir[br]; t < s->brow_dir[br + 1]; t++) {
            const uint8_t *p = src + s->offs[br] + 2ull * s->dir[t].x;
            if (!rd32(p + 8)) continue;  // the stub of an empty block row
            tile_parts(p, log2h2, parts);
            for (const Part &pt : parts) size2[br * per + pt.sub] += part_bytes(pt, esz);
        }
    });
    offs.assign(nbr2 + 1, 0);
    offs[0] = (int64_t)kHeaderBytes;
    for (uint64_t b = 0; b < nbr2; b++) {
        if (!size2[b]) size2[b] = 16;  // an empty block row is one empty tile
        offs[b + 1] = offs[b] + (int64_t)size2[b];
    }
    img.assign((size_t)offs[nbr2], 0);
    memcpy(img.data(), src, kHeaderBytes);
    const uint64_t h2v = h2;
    memcpy(img.data() + 48, &h2v, 8);
    // pass 2: the tiles, each new block row's in ascending block column
    for_rows([&](uint64_t br, std::vector<Part> &parts) {
        std::vector<uint64_t> cur(per);
        for (uint64_t q = 0; q < per && br * per + q < nbr2; q++) cur[q] = (uint64_t)offs[br * per + q];
        for (uint32_t t = s->brow_dir[br]; t < s->brow_dir[br + 1]; t++) {
            const uint8_t *p = src + s->offs[br] + 2ull * s->dir[t].x;
            const uint32_t nnz = rd32(p + 8);
            if (!nnz) continue;
            const uint32_t nrow = rd16(p + 12), ncoo = rd16(p + 14);
            const uint32_t scsr_words = (nrow - ncoo) + (nnz - ncoo);
            const uint8_t *w = p + 16, *coo = w + 2ull * scsr_words + 2, *vals = coo + 4ull * ncoo;
            const uint32_t nnz_scsr = nnz - ncoo;
            tile_parts(p, log2h2, parts);
            for (const Part &pt : parts) {
                uint8_t *o = img.data() + cur[pt.sub];
                const uint32_t pcoo = pt.c1 - pt.c0, pnnz = pt.nnz_scsr + pcoo;
                wr32(o, (uint32_t)(br * per + pt.sub));
                wr32(o + 4, s->dir[t].y);
                wr32(o + 8, pnnz);
                wr16(o + 12, (uint16_t)(pt.nrow_scsr + pcoo));
                wr16(o + 14, (uint16_t)pcoo);
                uint8_t *q = o + 16;
                const uint32_t rbase = pt.sub << log2h2;
                for (uint32_t i = pt.w0; i < pt.w1; i++, q += 2) {
                    const uint16_t v = rd16(w + 2ull * i);
                    wr16(q, (v & 0x8000u) ? (uint16_t)(0x8000u | ((v & 0x7FFFu) - rbase)) : v);
                }
                wr16(q, 0xFFFF);
                q += 2;
                for (uint32_t c = pt.c0; c < pt.c1; c++, q += 4) {
                    wr16(q, (uint16_t)(0x8000u | ((rd16(coo + 4ull * c) & 0x7FFFu) - rbase)));
                    wr16(q + 2, rd16(coo + 4ull * c + 2));
                }
                if (esz) {
                    memcpy(q, vals + esz * pt.v0, esz * pt.nnz_scsr);
                    q += esz * pt.nnz_scsr;
                    memcpy(q, vals + esz * (nnz_scsr + pt.c0), esz * pcoo);
                    q += esz * pcoo;
                }
                cur[pt.sub] = (uint64_t)(q - img.data());
            }
        }
        // empty new block rows: one empty tile {br, 0, 0, 0, 0}
        for (uint64_t q = 0; q < per && br * per + q < nbr2; q++)
            if (cur[q] == (uint64_t)offs[br * per + q]) wr32(img.data() + cur[q], (uint32_t)(br * per + q));
    });
}

// Tile height of the work image (0: none); FMX_WORK_ROWS overrides it (0
// turns re-cutting off), read at every matrix creation.
uint32_t work_rows_for(const fmx_spm *s)
{
    uint64_t h2 = 64;
    if (const char *e = getenv("FMX_WORK_ROWS")) h2 = strtoull(e, nullptr, 10);
    if (h2 == 0 || (h2 & (h2 - 1)) || s->H <= h2) return 0;
    return (uint32_t)h2;
}

}  // namespace fmx
