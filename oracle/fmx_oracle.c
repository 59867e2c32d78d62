/*
 * fmx_oracle.c -- CPU restatement of FlashMatrix's 2D-tiled SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline).  Loaded by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py; never by the
 * product library.  Reference citations are file:line in the FlashX tree.
 *
 * What is restated here:
 *   - edge normalisation: fm_utils.cpp:726-782 (n = max id + 1, every id is a
 *     row), fm_utils.cpp:435-482 (drop self edges unless disabled, sort each
 *     row's neighbours ascending, optional std::unique);
 *   - tile writer: fm_utils.cpp:228-433 (collect_block_info two-pass sizing,
 *     add_nz: runs of length 1 -> COO, >= 2 -> SCSR row part), empty block
 *     rows as one 16-byte empty tile (fm_utils.cpp:341-351), header on block
 *     row 0 (fm_utils.cpp:355-358), offsets (fm_utils.cpp:701-712);
 *   - byte layout: sparse_matrix_format.h:130-240, matrix_header.h:97-121;
 *   - SpMM: sparse_matrix.h:219-375 (row_part_func, coo_func, run_on_block),
 *     sparse_matrix.cpp:251-393 (super-block walk, zeroed band, copy-out),
 *     matrix_io.cpp:188-236 (block-row task hand-out), sparse_matrix.h:450-463
 *     (super-block size).
 *
 * The R-MAT generator is this project's own seeded specification (the
 * reference's flash-graph/tools/rmat-gen.cpp:43-48 draws from a time-seeded
 * Boost rmat_iterator and is not reproducible); its parity is therefore
 * pinned by invariants, not by reference vectors.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* RNG: output function of SplitMix64 (Steele, Lea, Flood, OOPSLA 2014).     */
/* The ctr-th output of a SplitMix64 stream whose state starts at `key`.     */
/* ------------------------------------------------------------------------ */
#define SM_GAMMA 0x9E3779B97F4A7C15ULL

uint64_t orc_splitmix64(uint64_t key, uint64_t ctr)
{
    uint64_t z = key + (ctr + 1) * SM_GAMMA;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Seed salt for the vertex permutation ("permute1" in ASCII). */
#define PERM_SALT 0x7065726D75746531ULL

/*
 * Vertex relabelling: a bijection on [0, 2^scale) made of three rounds of
 * (add key, multiply by an odd key, xor-shift right), all modulo 2^scale.
 * Each step is invertible modulo a power of two, so the map is a permutation.
 */
uint32_t orc_rmat_perm(uint32_t v, int scale, uint64_t seed)
{
    uint64_t mask = (scale >= 64) ? ~0ULL : ((1ULL << scale) - 1);
    int sh = (scale + 1) / 2;
    uint64_t x = v;
    for (int i = 0; i < 3; i++) {
        uint64_t add = orc_splitmix64(seed ^ PERM_SALT, 2 * i);
        uint64_t mul = orc_splitmix64(seed ^ PERM_SALT, 2 * i + 1) | 1ULL;
        x = (x + add) & mask;
        x = (x * mul) & mask;
        x ^= x >> sh;
    }
    return (uint32_t)x;
}

/*
 * R-MAT edge e: ceil(scale/2) SplitMix64 draws r_j = splitmix64(seed,
 * e*ndraw + j); level 2j uses the low 32 bits of r_j, level 2j+1 the high 32
 * bits.  Level l (0 = most significant bit) picks quadrant
 *   u < thr_a -> (0,0),  u < thr_ab -> (0,1),  u < thr_abc -> (1,0),  else (1,1)
 * with thr_* = floor(cumulative probability * 2^32) computed by the caller.
 * The (row bit, column bit) of the quadrant is bit (scale-1-l) of (src, dst).
 */
void orc_rmat_edges(int scale, uint64_t m, uint64_t seed,
                    uint32_t thr_a, uint32_t thr_ab, uint32_t thr_abc,
                    int permute, uint32_t *src, uint32_t *dst)
{
    uint64_t ndraw = (uint64_t)(scale + 1) / 2;
    for (uint64_t e = 0; e < m; e++) {
        uint32_t s = 0, d = 0;
        for (int l = 0; l < scale; l++) {
            uint64_t r = orc_splitmix64(seed, e * ndraw + (uint64_t)(l >> 1));
            uint32_t u = (l & 1) ? (uint32_t)(r >> 32) : (uint32_t)r;
            uint32_t q = (u < thr_a) ? 0u : (u < thr_ab) ? 1u : (u < thr_abc) ? 2u : 3u;
            s |= (q >> 1) << (scale - 1 - l);
            d |= (q & 1u) << (scale - 1 - l);
        }
        if (permute) {
            s = orc_rmat_perm(s, scale, seed);
            d = orc_rmat_perm(d, scale, seed);
        }
        src[e] = s;
        dst[e] = d;
    }
}

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* Degree-corrected stochastic block model (the project's own seeded spec;   */
/* no reference counterpart: its nearest relative, flash-graph/utils/sbm.cpp */
/* :117-242, is a plain SBM).  Integer arithmetic throughout so that the     */
/* device generator reproduces it bit for bit:                               */
/*   weight w_v = wtab[splitmix64(seed^W, v) >> 54] (a 1024-quantile table   */
/*   of a power law computed by the caller), C = exclusive prefix of w,      */
/*   blocks are contiguous id ranges bstart[0..B];                           */
/*   edge e: u = vertex of rank (r1 % Vtot) in C, b = block(u);              */
/*   intra-block iff (r2 >> 40) * den < (num << 24) with num = ratio*vb,     */
/*   den = num + vt - vb (vb, vt = volumes >> 8); the partner is drawn in    */
/*   proportion to w inside b, or outside b by skipping b's volume range;    */
/*   finally both ends go through the R-MAT relabelling orc_rmat_perm.       */
/* ------------------------------------------------------------------------ */
#define SBM_SALT_W 0x73626D5745494748ULL
#define SBM_SALT_E 0x73626D4544474553ULL

static uint64_t sbm_find(const uint64_t *C, uint64_t lo, uint64_t hi, uint64_t t)
{
    /* largest v in [lo, hi) with C[v] <= t */
    while (hi - lo > 1) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (C[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

int orc_dcsbm_edges(uint64_t n, int scale, int nblocks, const uint32_t *bstart,
                    const uint32_t *wtab, uint32_t ratio, uint64_t m, uint64_t seed,
                    uint32_t *src, uint32_t *dst)
{
    uint64_t *C = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    if (!C) return ORC_ENOMEM;
    C[0] = 0;
    for (uint64_t v = 0; v < n; v++)
        C[v + 1] = C[v] + wtab[orc_splitmix64(seed ^ SBM_SALT_W, v) >> 54];
    const uint64_t vtot = C[n];
    for (uint64_t e = 0; e < m; e++) {
        uint64_t r1 = orc_splitmix64(seed ^ SBM_SALT_E, 3 * e);
        uint64_t r2 = orc_splitmix64(seed ^ SBM_SALT_E, 3 * e + 1);
        uint64_t r3 = orc_splitmix64(seed ^ SBM_SALT_E, 3 * e + 2);
        uint64_t u = sbm_find(C, 0, n, r1 % vtot);
        int lo = 0, hi = nblocks;
        while (hi - lo > 1) { int mid = (lo + hi) / 2; if (bstart[mid] <= u) lo = mid; else hi = mid; }
        const int b = lo;
        const uint64_t c0 = C[bstart[b]], vb_full = C[bstart[b + 1]] - c0;
        const uint64_t vb = vb_full >> 8, vt = vtot >> 8;
        const uint64_t num = (uint64_t)ratio * vb, den = num + vt - vb;
        uint64_t v;
        if ((r2 >> 40) * den < (num << 24)) {
            v = sbm_find(C, bstart[b], bstart[b + 1], c0 + r3 % vb_full);
        } else {
            uint64_t t = r3 % (vtot - vb_full);
            if (t >= c0) t += vb_full;
            v = sbm_find(C, 0, n, t);
        }
        src[e] = orc_rmat_perm((uint32_t)u, scale, seed);
        dst[e] = orc_rmat_perm((uint32_t)v, scale, seed);
    }
    free(C);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Edge normalisation.                                                       */
/* ------------------------------------------------------------------------ */
static int bits_for(uint64_t n)
{
    int b = 0;
    while (b < 64 && (1ULL << b) < n) b++;
    return b;
}

/* LSD radix sort of 64-bit keys over the bit ranges [0,bd) and [32,32+bs). */
static int radix_sort_keys(uint64_t *keys, uint64_t cnt, int bd, int bs)
{
    if (cnt < 2) return ORC_OK;
    uint64_t *tmp = (uint64_t *)malloc(cnt * sizeof(uint64_t));
    if (!tmp) return ORC_ENOMEM;
    enum { DB = 11, NB = 1 << DB };
    uint64_t *count = (uint64_t *)malloc(NB * sizeof(uint64_t));
    if (!count) { free(tmp); return ORC_ENOMEM; }
    int ranges[2][2] = {{0, bd}, {32, 32 + bs}};
    uint64_t *a = keys, *b = tmp;
    for (int rr = 0; rr < 2; rr++) {
        for (int lo = ranges[rr][0]; lo < ranges[rr][1]; lo += DB) {
            int w = ranges[rr][1] - lo < DB ? ranges[rr][1] - lo : DB;
            uint64_t dmask = (1ULL << w) - 1;
            memset(count, 0, NB * sizeof(uint64_t));
            for (uint64_t i = 0; i < cnt; i++) count[(a[i] >> lo) & dmask]++;
            uint64_t sum = 0;
            for (int d = 0; d < NB; d++) { uint64_t c = count[d]; count[d] = sum; sum += c; }
            for (uint64_t i = 0; i < cnt; i++) b[count[(a[i] >> lo) & dmask]++] = a[i];
            uint64_t *t = a; a = b; b = t;
        }
    }
    if (a != keys) memcpy(keys, a, cnt * sizeof(uint64_t));
    free(tmp);
    free(count);
    return ORC_OK;
}

/*
 * Normalised edge list as sorted keys (src << 32 | dst).
 * n == 0 means n = max id + 1 (fm_utils.cpp:730-739,788); otherwise n must
 * exceed every id.  Self edges are dropped unless ORC_KEEP_SELF
 * (fm_utils.cpp:39,450-452); neighbours are sorted ascending
 * (fm_utils.cpp:459) and de-duplicated under ORC_DEDUP (fm_utils.cpp:460-464).
 */
int orc_normalize(const uint32_t *src, const uint32_t *dst, uint64_t m,
                  uint64_t n, unsigned flags, uint64_t **keys_out,
                  uint64_t *nkeys, uint64_t *n_out)
{
    uint64_t maxid = 0;
    for (uint64_t i = 0; i < m; i++) {
        if (src[i] > maxid) maxid = src[i];
        if (dst[i] > maxid) maxid = dst[i];
    }
    if (n == 0) {
        if (m == 0) return ORC_EINVAL;
        n = maxid + 1;
    } else if (m > 0 && maxid >= n) {
        return ORC_EDIM;
    }
    int sym = (flags & ORC_SYMMETRIZE) != 0;
    int keep_self = (flags & ORC_KEEP_SELF) != 0;
    uint64_t cap = sym ? 2 * m : m;
    uint64_t *keys = (uint64_t *)malloc((cap ? cap : 1) * sizeof(uint64_t));
    if (!keys) return ORC_ENOMEM;
    uint64_t c = 0;
    for (uint64_t i = 0; i < m; i++) {
        uint32_t s = src[i], d = dst[i];
        if (s == d && !keep_self) continue;
        keys[c++] = ((uint64_t)s << 32) | d;
        if (sym && s != d) keys[c++] = ((uint64_t)d << 32) | s;
    }
    int bits = bits_for(n);
    int rc = radix_sort_keys(keys, c, bits, bits);
    if (rc) { free(keys); return rc; }
    if (flags & ORC_DEDUP) {
        uint64_t o = 0;
        for (uint64_t i = 0; i < c; i++)
            if (o == 0 || keys[o - 1] != keys[i]) keys[o++] = keys[i];
        c = o;
    }
    *keys_out = keys;
    *nkeys = c;
    *n_out = n;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Tile image writer.                                                        */
/* ------------------------------------------------------------------------ */
#define HDR_BYTES 4096
#define TILE_HDR 16
#define MAGIC 0x0123456789FFFFFFLL

static void put_u16(uint8_t *p, uint16_t v) { memcpy(p, &v, 2); }
static void put_u32(uint8_t *p, uint32_t v) { memcpy(p, &v, 4); }
static void put_u64(uint8_t *p, uint64_t v) { memcpy(p, &v, 8); }
static uint16_t get_u16(const uint8_t *p) { uint16_t v; memcpy(&v, p, 2); return v; }
static uint32_t get_u32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t get_u64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }

/* matrix_header.cpp:28-49 + matrix_header.h:102-121: 4096 zero-filled bytes. */
static void write_header(uint8_t *h, uint64_t nrows, uint64_t ncols,
                         uint64_t hgt, uint64_t wid)
{
    memset(h, 0, HDR_BYTES);
    put_u64(h + 0, (uint64_t)MAGIC);
    put_u32(h + 8, 2);      /* version */
    put_u32(h + 12, 2);     /* matrix_type SPARSE */
    put_u64(h + 16, 0);     /* entry_size: binary */
    put_u64(h + 24, nrows);
    put_u64(h + 32, ncols);
    put_u32(h + 40, 2);     /* L_ROW_2D */
    put_u32(h + 44, 0);     /* P_BOOL */
    put_u64(h + 48, hgt);
    put_u64(h + 56, wid);
}

typedef struct { uint32_t nnz; uint32_t nrow; uint32_t ncoo; } binfo_t;

/* sparse_block_2d::get_size, sparse_matrix_format.h:233-241,294-299 */
static uint64_t tile_bytes(const binfo_t *b, uint64_t esz)
{
    return TILE_HDR + 2ULL * b->nrow + 2ULL * b->nnz + 2 + esz * b->nnz;
}

/* Writes the image of the sorted, normalised keys (src << 32 | dst); vals,
 * when esz > 0, holds esz bytes per key in the same order and becomes each
 * tile's values region (SCSR values in SCSR order, then COO values,
 * sparse_matrix_format.h:233-241,294-299).  Takes ownership of keys/vals. */
static int build_2d_impl(uint64_t *keys, uint64_t nk, uint8_t *vals, uint64_t esz, int prim,
                         uint64_t n, int log2h, int log2w,
                         uint8_t **image_out, uint64_t *image_bytes,
                         int64_t **offs_out, uint64_t *nbr_out, uint64_t *n_out,
                         uint64_t *nnz_out)
{
    int rc = ORC_OK;
    const uint64_t H = 1ULL << log2h, W = 1ULL << log2w;
    const uint64_t nbr = (n + H - 1) / H, nbc = (n + W - 1) / W;

    binfo_t *bi = (binfo_t *)calloc(nbc, sizeof(binfo_t));
    int64_t *offs = (int64_t *)malloc((nbr + 1) * sizeof(int64_t));
    uint64_t *brow_lo = (uint64_t *)malloc((nbr + 1) * sizeof(uint64_t));
    if (!bi || !offs || !brow_lo) { free(bi); free(offs); free(brow_lo); free(keys); free(vals); return ORC_ENOMEM; }

    /* Key range of each block row (keys are sorted by src). */
    {
        uint64_t p = 0;
        for (uint64_t br = 0; br <= nbr; br++) {
            uint64_t row0 = br * H;
            while (p < nk && (keys[p] >> 32) < row0) p++;
            brow_lo[br] = p;
        }
        brow_lo[nbr] = nk;
    }

    /* Pass 1 over every block row: sizes (collect_block_info). */
    uint64_t total = HDR_BYTES;
    for (uint64_t br = 0; br < nbr; br++) {
        memset(bi, 0, nbc * sizeof(binfo_t));
        uint64_t lo = brow_lo[br], hi = brow_lo[br + 1];
        uint64_t i = lo;
        while (i < hi) {
            uint64_t j = i + 1;
            uint64_t row = keys[i] >> 32, bc = (uint32_t)keys[i] >> log2w;
            while (j < hi && (keys[j] >> 32) == row && ((uint32_t)keys[j] >> log2w) == bc) j++;
            uint64_t len = j - i;
            if (len > W) { rc = ORC_EFORMAT; goto fail; } /* fm_utils.cpp:403-406 */
            bi[bc].nrow++;
            bi[bc].nnz += (uint32_t)len;
            if (len == 1) bi[bc].ncoo++;
            i = j;
        }
        uint64_t bytes = 0;
        for (uint64_t c = 0; c < nbc; c++)
            if (bi[c].nnz) bytes += tile_bytes(&bi[c], esz);
        if (bytes == 0) bytes = TILE_HDR; /* empty block row, fm_utils.cpp:341-351 */
        offs[br] = (int64_t)total;
        total += bytes;
    }
    offs[nbr] = (int64_t)total;

    uint8_t *img = (uint8_t *)malloc(total);
    if (!img) { rc = ORC_ENOMEM; goto fail; }
    write_header(img, n, n, H, W);
    put_u64(img + 16, esz);
    put_u32(img + 44, (uint32_t)prim);

    /* Pass 2: serialise tiles (run_blocks + add_nz + finalize). */
    {
        uint64_t *scsr_p = (uint64_t *)malloc(nbc * sizeof(uint64_t));
        uint64_t *coo_p = (uint64_t *)malloc(nbc * sizeof(uint64_t));
        uint64_t *scsr_v = (uint64_t *)malloc(nbc * sizeof(uint64_t));
        uint64_t *coo_v = (uint64_t *)malloc(nbc * sizeof(uint64_t));
        if (!scsr_p || !coo_p || !scsr_v || !coo_v) {
            free(scsr_p); free(coo_p); free(scsr_v); free(coo_v); free(img); rc = ORC_ENOMEM; goto fail;
        }
        for (uint64_t br = 0; br < nbr; br++) {
            uint8_t *base = img + offs[br];
            uint64_t lo = brow_lo[br], hi = brow_lo[br + 1];
            if (lo == hi) {
                memset(base, 0, TILE_HDR);
                put_u32(base, (uint32_t)br);
                continue;
            }
            memset(bi, 0, nbc * sizeof(binfo_t));
            uint64_t i = lo;
            while (i < hi) {
                uint64_t j = i + 1;
                uint64_t row = keys[i] >> 32, bc = (uint32_t)keys[i] >> log2w;
                while (j < hi && (keys[j] >> 32) == row && ((uint32_t)keys[j] >> log2w) == bc) j++;
                bi[bc].nrow++;
                bi[bc].nnz += (uint32_t)(j - i);
                if (j - i == 1) bi[bc].ncoo++;
                i = j;
            }
            uint64_t cur = 0;
            for (uint64_t c = 0; c < nbc; c++) {
                if (!bi[c].nnz) continue;
                uint8_t *t = base + cur;
                put_u32(t + 0, (uint32_t)br);
                put_u32(t + 4, (uint32_t)c);
                put_u32(t + 8, bi[c].nnz);
                put_u16(t + 12, (uint16_t)bi[c].nrow);
                put_u16(t + 14, (uint16_t)bi[c].ncoo);
                uint64_t rheader = 2ULL * (bi[c].nrow - bi[c].ncoo) + 2ULL * (bi[c].nnz - bi[c].ncoo) + 2;
                scsr_p[c] = offs[br] + cur + TILE_HDR;
                coo_p[c] = offs[br] + cur + TILE_HDR + rheader;
                scsr_v[c] = offs[br] + cur + TILE_HDR + 2ULL * bi[c].nrow + 2ULL * bi[c].nnz + 2;
                coo_v[c] = scsr_v[c] + esz * (bi[c].nnz - bi[c].ncoo);
                cur += tile_bytes(&bi[c], esz);
            }
            i = lo;
            while (i < hi) {
                uint64_t j = i + 1;
                uint64_t row = keys[i] >> 32, bc = (uint32_t)keys[i] >> log2w;
                while (j < hi && (keys[j] >> 32) == row && ((uint32_t)keys[j] >> log2w) == bc) j++;
                uint16_t rrel = (uint16_t)(row & (H - 1));
                if (j - i == 1) {
                    put_u16(img + coo_p[bc], (uint16_t)(rrel | 0x8000));
                    put_u16(img + coo_p[bc] + 2, (uint16_t)((uint32_t)keys[i] & (W - 1)));
                    coo_p[bc] += 4;
                    if (esz) { memcpy(img + coo_v[bc], vals + i * esz, esz); coo_v[bc] += esz; }
                } else {
                    put_u16(img + scsr_p[bc], (uint16_t)(rrel | 0x8000));
                    scsr_p[bc] += 2;
                    for (uint64_t q = i; q < j; q++) {
                        put_u16(img + scsr_p[bc], (uint16_t)((uint32_t)keys[q] & (W - 1)));
                        scsr_p[bc] += 2;
                        if (esz) { memcpy(img + scsr_v[bc], vals + q * esz, esz); scsr_v[bc] += esz; }
                    }
                }
                i = j;
            }
            for (uint64_t c = 0; c < nbc; c++)
                if (bi[c].nnz) put_u16(img + scsr_p[c], 0xFFFF); /* end marker */
        }
        free(scsr_p);
        free(coo_p);
        free(scsr_v);
        free(coo_v);
    }

    free(bi);
    free(brow_lo);
    free(keys);
    free(vals);
    *image_out = img;
    *image_bytes = total;
    *offs_out = offs;
    *nbr_out = nbr;
    *n_out = n;
    *nnz_out = nk;
    return ORC_OK;
fail:
    free(bi);
    free(offs);
    free(brow_lo);
    free(keys);
    free(vals);
    return rc;
}

int orc_build_2d(const uint32_t *src, const uint32_t *dst, uint64_t m,
                 uint64_t n, int log2h, int log2w, unsigned flags,
                 uint8_t **image_out, uint64_t *image_bytes,
                 int64_t **offs_out, uint64_t *nbr_out, uint64_t *n_out,
                 uint64_t *nnz_out)
{
    if (log2h < 0 || log2h > 15 || log2w < 0 || log2w > 15) return ORC_EINVAL;
    uint64_t *keys = NULL, nk = 0;
    int rc = orc_normalize(src, dst, m, n, flags, &keys, &nk, &n);
    if (rc) return rc;
    return build_2d_impl(keys, nk, NULL, 0, 0, n, log2h, log2w, image_out, image_bytes, offs_out, nbr_out,
                         n_out, nnz_out);
}

typedef struct { uint64_t key, idx; } kidx_t;

static int kidx_cmp(const void *a, const void *b)
{
    const kidx_t *x = (const kidx_t *)a, *y = (const kidx_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* Weighted image (attr_apply_operate::run_row, fm_utils.cpp:484-535):
 * per row, neighbours sorted ascending with their esz-byte values; self
 * edges dropped unless KEEP_SELF; with DEDUP one value per (row, column).
 * The reference sorts with std::sort, which leaves the order of equal
 * columns (and so the value kept) unspecified; this restatement keeps
 * input order among equal keys and keeps the first.  SYMMETRIZE adds
 * (dst, src) with the same value right after (src, dst). */
int orc_build_2d_w(const uint32_t *src, const uint32_t *dst, const void *vals, uint64_t esz, int prim,
                   uint64_t m, uint64_t n, int log2h, int log2w, unsigned flags,
                   uint8_t **image_out, uint64_t *image_bytes,
                   int64_t **offs_out, uint64_t *nbr_out, uint64_t *n_out, uint64_t *nnz_out)
{
    if (log2h < 0 || log2h > 15 || log2w < 0 || log2w > 15) return ORC_EINVAL;
    if (esz != 4 && esz != 8) return ORC_EINVAL;
    uint64_t maxid = 0;
    for (uint64_t i = 0; i < m; i++) {
        if (src[i] > maxid) maxid = src[i];
        if (dst[i] > maxid) maxid = dst[i];
    }
    if (n == 0) {
        if (m == 0) return ORC_EINVAL;
        n = maxid + 1;
    } else if (m > 0 && maxid >= n) {
        return ORC_EDIM;
    }
    const int sym = (flags & ORC_SYMMETRIZE) != 0, keep_self = (flags & ORC_KEEP_SELF) != 0;
    const uint64_t cap = sym ? 2 * m : m;
    kidx_t *kv = (kidx_t *)malloc((cap ? cap : 1) * sizeof(kidx_t));
    if (!kv) return ORC_ENOMEM;
    uint64_t c = 0;
    for (uint64_t i = 0; i < m; i++) {
        uint32_t s = src[i], d = dst[i];
        if (s == d && !keep_self) continue;
        kv[c].key = ((uint64_t)s << 32) | d;
        kv[c].idx = sym ? 2 * i : i;
        c++;
        if (sym && s != d) {
            kv[c].key = ((uint64_t)d << 32) | s;
            kv[c].idx = 2 * i + 1;
            c++;
        }
    }
    qsort(kv, c, sizeof(kidx_t), kidx_cmp);
    uint64_t *keys = (uint64_t *)malloc((c ? c : 1) * sizeof(uint64_t));
    uint8_t *sv = (uint8_t *)malloc((c ? c : 1) * esz);
    if (!keys || !sv) { free(kv); free(keys); free(sv); return ORC_ENOMEM; }
    uint64_t o = 0;
    for (uint64_t i = 0; i < c; i++) {
        if ((flags & ORC_DEDUP) && o > 0 && keys[o - 1] == kv[i].key) continue;
        keys[o] = kv[i].key;
        const uint64_t e = sym ? kv[i].idx / 2 : kv[i].idx;
        memcpy(sv + o * esz, (const uint8_t *)vals + e * esz, esz);
        o++;
    }
    free(kv);
    return build_2d_impl(keys, o, sv, esz, prim, n, log2h, log2w, image_out, image_bytes, offs_out, nbr_out,
                         n_out, nnz_out);
}

/* ------------------------------------------------------------------------ */
/* Image reader.                                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t nrows, ncols, H, W, esz;
    int log2h, log2w, prim;
} img_info_t;

static int log2_exact(uint64_t v)
{
    int b = 0;
    while (b < 16 && (1ULL << b) < v) b++;
    return ((1ULL << b) == v) ? b : -1;
}

/* matrix_header::verify (matrix_header.cpp:52-60) plus block_2d_size. */
static int read_info(const uint8_t *image, uint64_t bytes, const int64_t *offs,
                     uint64_t nbr, img_info_t *inf)
{
    if (bytes < HDR_BYTES) return ORC_EFORMAT;
    if ((int64_t)get_u64(image) != MAGIC) return ORC_EFORMAT;
    if (get_u32(image + 8) != 2) return ORC_EFORMAT;
    if (get_u32(image + 12) != 2 || get_u32(image + 40) != 2) return ORC_EFORMAT;
    inf->esz = get_u64(image + 16);
    inf->nrows = get_u64(image + 24);
    inf->ncols = get_u64(image + 32);
    inf->prim = (int)get_u32(image + 44);
    inf->H = get_u64(image + 48);
    inf->W = get_u64(image + 56);
    inf->log2h = log2_exact(inf->H);
    inf->log2w = log2_exact(inf->W);
    if (inf->log2h < 0 || inf->log2w < 0 || inf->H > 32768 || inf->W > 32768) return ORC_EFORMAT;
    if (nbr != (inf->nrows + inf->H - 1) / inf->H) return ORC_EFORMAT;
    if (offs[0] != HDR_BYTES || (uint64_t)offs[nbr] > bytes) return ORC_EFORMAT;
    for (uint64_t i = 1; i <= nbr; i++)
        if (offs[i] <= offs[i - 1]) return ORC_EFORMAT; /* SpM_2d_index::verify */
    return ORC_OK;
}

/* Walks one tile; calls fn(row_rel, col_rel, nz_index) for every non-zero
 * in storage order (sparse_matrix_format.cpp:62-88).  Returns the tile size,
 * or 0 on a malformed tile. */
typedef struct {
    uint32_t br, bc, nnz, nrow, ncoo;
    const uint8_t *scsr;   /* first SCSR word */
    const uint8_t *coo;    /* first COO pair */
    const uint8_t *vals;   /* values region (esz > 0) */
    uint64_t size;
} tile_t;

static int parse_tile(const uint8_t *t, const uint8_t *end, uint64_t esz, tile_t *tl)
{
    if (end - t < TILE_HDR) return ORC_EFORMAT;
    tl->br = get_u32(t);
    tl->bc = get_u32(t + 4);
    tl->nnz = get_u32(t + 8);
    tl->nrow = get_u16(t + 12);
    tl->ncoo = get_u16(t + 14);
    if (tl->nnz == 0) {           /* an empty tile is just its header */
        tl->size = TILE_HDR;
        tl->scsr = tl->coo = tl->vals = t + TILE_HDR;
        return ORC_OK;
    }
    if (tl->ncoo > tl->nrow || tl->ncoo > tl->nnz) return ORC_EFORMAT;
    uint64_t rheader = 2ULL * (tl->nrow - tl->ncoo) + 2ULL * (tl->nnz - tl->ncoo) + 2;
    tl->size = TILE_HDR + 2ULL * tl->nrow + 2ULL * tl->nnz + 2 + esz * tl->nnz;
    if ((uint64_t)(end - t) < tl->size) return ORC_EFORMAT;
    tl->scsr = t + TILE_HDR;
    tl->coo = t + TILE_HDR + rheader;
    tl->vals = t + TILE_HDR + 2ULL * tl->nrow + 2ULL * tl->nnz + 2;
    return ORC_OK;
}

/* sparse_block_2d::verify (sparse_matrix_format.cpp:36-60) plus the
 * structural facts the writer guarantees. */
int orc_verify_image(const uint8_t *image, uint64_t bytes,
                     const int64_t *offs, uint64_t nbr)
{
    img_info_t inf;
    int rc = read_info(image, bytes, offs, nbr, &inf);
    if (rc) return rc;
    for (uint64_t br = 0; br < nbr; br++) {
        const uint8_t *p = image + offs[br], *end = image + offs[br + 1];
        long prev_bc = -1;
        uint64_t rows_here = inf.nrows - br * inf.H < inf.H ? inf.nrows - br * inf.H : inf.H;
        uint64_t ntiles = 0, nempty = 0;
        while (p < end) {
            tile_t tl;
            if ((rc = parse_tile(p, end, inf.esz, &tl))) return rc;
            if (tl.br != br) return ORC_EFORMAT;
            ntiles++;
            if (!tl.nnz) nempty++;
            if (tl.nnz) {
                if ((long)tl.bc <= prev_bc) return ORC_EFORMAT;
                if ((uint64_t)tl.bc * inf.W >= inf.ncols) return ORC_EFORMAT;
                prev_bc = tl.bc;
                uint64_t cols_here = inf.ncols - (uint64_t)tl.bc * inf.W < inf.W ? inf.ncols - (uint64_t)tl.bc * inf.W : inf.W;
                const uint8_t *w = tl.scsr, *rend = tl.coo - 2;
                uint64_t nz = 0, nr = 0;
                long prev_row = -1;
                while (w < rend) {
                    uint16_t h = get_u16(w);
                    /* 0xFFFF is a valid head here: row 32767 of a 32768-row
                     * tile.  The region end is positional (is_rparts_end,
                     * sparse_matrix_format.h:339-347), not the marker value. */
                    if (!(h & 0x8000)) return ORC_EFORMAT;
                    uint64_t r = h & 0x7FFF;
                    if ((long)r <= prev_row || r >= rows_here) return ORC_EFORMAT;
                    prev_row = (long)r;
                    w += 2;
                    uint64_t len = 0;
                    while (w < rend && !(get_u16(w) & 0x8000)) {
                        if (get_u16(w) >= cols_here) return ORC_EFORMAT;
                        w += 2; len++;
                    }
                    if (len == 0) return ORC_EFORMAT;
                    nz += len; nr++;
                }
                if (get_u16(tl.coo - 2) != 0xFFFF) return ORC_EFORMAT;
                prev_row = -1;
                for (uint32_t i = 0; i < tl.ncoo; i++) {
                    uint16_t h = get_u16(tl.coo + 4ULL * i), c = get_u16(tl.coo + 4ULL * i + 2);
                    if (!(h & 0x8000) || (c & 0x8000)) return ORC_EFORMAT;
                    uint64_t r = h & 0x7FFF;
                    if ((long)r <= prev_row || r >= rows_here || c >= cols_here) return ORC_EFORMAT;
                    prev_row = (long)r;
                }
                if (nz + tl.ncoo != tl.nnz || nr + tl.ncoo != tl.nrow) return ORC_EFORMAT;
            }
            p += tl.size;
        }
        if (p != end) return ORC_EFORMAT;
        /* the writer stores no empty tiles, except one header for an empty
         * block row (fm_utils.cpp:317,340-350) */
        if (ntiles == 0 || (nempty && ntiles > 1)) return ORC_EFORMAT;
    }
    return ORC_OK;
}

int orc_image_nonzeros(const uint8_t *image, uint64_t bytes,
                       const int64_t *offs, uint64_t nbr,
                       uint64_t *rows, uint64_t *cols, uint64_t cap,
                       uint64_t *count)
{
    img_info_t inf;
    int rc = read_info(image, bytes, offs, nbr, &inf);
    if (rc) return rc;
    uint64_t c = 0;
    for (uint64_t br = 0; br < nbr; br++) {
        const uint8_t *p = image + offs[br], *end = image + offs[br + 1];
        while (p < end) {
            tile_t tl;
            if ((rc = parse_tile(p, end, inf.esz, &tl))) return rc;
            if (tl.nnz) {
                uint64_t r0 = (uint64_t)tl.br * inf.H, c0 = (uint64_t)tl.bc * inf.W;
                const uint8_t *w = tl.scsr, *rend = tl.coo - 2;
                uint64_t row = 0;
                while (w < rend) {
                    uint16_t v = get_u16(w);
                    if (v & 0x8000) row = v & 0x7FFF;
                    else {
                        if (c < cap) { rows[c] = r0 + row; cols[c] = c0 + v; }
                        c++;
                    }
                    w += 2;
                }
                for (uint32_t i = 0; i < tl.ncoo; i++) {
                    if (c < cap) {
                        rows[c] = r0 + (get_u16(tl.coo + 4ULL * i) & 0x7FFF);
                        cols[c] = c0 + get_u16(tl.coo + 4ULL * i + 2);
                    }
                    c++;
                }
            }
            p += tl.size;
        }
    }
    *count = c;
    return c <= cap ? ORC_OK : ORC_EINVAL;
}

/* ------------------------------------------------------------------------ */
/* SpMM.                                                                     */
/* ------------------------------------------------------------------------ */
static int elem_size(int dtype)
{
    switch (dtype) {
    case ORC_I32: case ORC_F32: return 4;
    case ORC_I64: case ORC_F64: return 8;
    default: return 0;
    }
}

/* Value of non-zero `idx` of a tile as the dense type (binary -> 1). */
#define NZ_VAL(T, tl, inf, idx) \
    ((inf)->esz == 0 ? (T)1 : nz_value_##T((tl)->vals, (inf)->prim, (idx)))

#define DEF_NZ_VALUE(T)                                                          \
    static T nz_value_##T(const uint8_t *vals, int prim, uint64_t i)            \
    {                                                                            \
        switch (prim) {                                                          \
        case 4: { int32_t v; memcpy(&v, vals + 4 * i, 4); return (T)v; }          \
        case 6: { int64_t v; memcpy(&v, vals + 8 * i, 8); return (T)v; }          \
        case 7: { uint64_t v; memcpy(&v, vals + 8 * i, 8); return (T)v; }         \
        case 8: { float v; memcpy(&v, vals + 4 * i, 4); return (T)v; }            \
        case 9: { double v; memcpy(&v, vals + 8 * i, 8); return (T)v; }           \
        default: return (T)1;                                                    \
        }                                                                        \
    }
typedef int32_t i32_t;
typedef int64_t i64_t;
typedef float f32_t;
typedef double f64_t;
DEF_NZ_VALUE(i32_t)
DEF_NZ_VALUE(i64_t)
DEF_NZ_VALUE(f32_t)
DEF_NZ_VALUE(f64_t)

/*
 * One tile, run_on_block (sparse_matrix.h:344-374): SCSR row parts first
 * (row_part_func, :219-276), then the COO region (coo_func, :278-331).
 * out points at the band row of the tile's first row; in at X row bc*W.
 * KC is the compile-time width (0 = runtime k), as spmm_creator::create
 * specialises k in {1,2,4,8,16} (sparse_matrix.h:425-435).
 */
#define DEF_RUN_TILE(T, KC)                                                     \
    static void run_tile_##T##_##KC(const tile_t *tl, const img_info_t *inf,     \
                                    int k, const T *in, int64_t ldx, T *out,    \
                                    int64_t ldo)                                \
    {                                                                           \
        const int K = KC ? KC : k;                                              \
        if (tl->nnz == 0) return;                                               \
        uint64_t nz = 0;                                                        \
        const uint8_t *w = tl->scsr, *rend = tl->coo - 2;                       \
        T *dest = out;                                                          \
        while (w < rend) {                                                      \
            uint16_t v = get_u16(w);                                            \
            w += 2;                                                             \
            if (v & 0x8000) { dest = out + (int64_t)(v & 0x7FFF) * ldo; continue; } \
            T d = NZ_VAL(T, tl, inf, nz);                                       \
            const T *s = in + (int64_t)v * ldx;                                 \
            for (int j = 0; j < K; j++) dest[j] += s[j] * d;                    \
            nz++;                                                               \
        }                                                                       \
        for (uint32_t i = 0; i < tl->ncoo; i++, nz++) {                         \
            uint16_t r = get_u16(tl->coo + 4ULL * i) & 0x7FFF;                  \
            uint16_t c = get_u16(tl->coo + 4ULL * i + 2);                       \
            T d = NZ_VAL(T, tl, inf, nz);                                       \
            const T *s = in + (int64_t)c * ldx;                                 \
            T *o = out + (int64_t)r * ldo;                                      \
            for (int j = 0; j < K; j++) o[j] += s[j] * d;                       \
        }                                                                       \
    }

#define DEF_RUN_TILES(T) \
    DEF_RUN_TILE(T, 0) DEF_RUN_TILE(T, 1) DEF_RUN_TILE(T, 2) DEF_RUN_TILE(T, 4) \
    DEF_RUN_TILE(T, 8) DEF_RUN_TILE(T, 16)                                     \
    static void run_tile_##T(const tile_t *tl, const img_info_t *inf, int k,    \
                             const T *in, int64_t ldx, T *out, int64_t ldo)    \
    {                                                                           \
        switch (k) {                                                            \
        case 1: run_tile_##T##_1(tl, inf, k, in, ldx, out, ldo); break;         \
        case 2: run_tile_##T##_2(tl, inf, k, in, ldx, out, ldo); break;         \
        case 4: run_tile_##T##_4(tl, inf, k, in, ldx, out, ldo); break;         \
        case 8: run_tile_##T##_8(tl, inf, k, in, ldx, out, ldo); break;         \
        case 16: run_tile_##T##_16(tl, inf, k, in, ldx, out, ldo); break;       \
        default: run_tile_##T##_0(tl, inf, k, in, ldx, out, ldo); break;        \
        }                                                                       \
    }
DEF_RUN_TILES(i32_t)
DEF_RUN_TILES(i64_t)
DEF_RUN_TILES(f32_t)
DEF_RUN_TILES(f64_t)

static void run_tile_any(int dtype, const tile_t *tl, const img_info_t *inf, int k,
                         const void *X, int64_t ldx, void *out, int64_t ldo)
{
    uint64_t c0 = (uint64_t)tl->bc * inf->W;
    switch (dtype) {
    case ORC_I32: run_tile_i32_t(tl, inf, k, (const int32_t *)X + c0 * ldx, ldx, (int32_t *)out, ldo); break;
    case ORC_I64: run_tile_i64_t(tl, inf, k, (const int64_t *)X + c0 * ldx, ldx, (int64_t *)out, ldo); break;
    case ORC_F32: run_tile_f32_t(tl, inf, k, (const float *)X + c0 * ldx, ldx, (float *)out, ldo); break;
    case ORC_F64: run_tile_f64_t(tl, inf, k, (const double *)X + c0 * ldx, ldx, (double *)out, ldo); break;
    }
}

/*
 * Y = A*X, sequential and in the reference's default order: block rows in
 * turn, each block-row band zeroed (sparse_matrix.cpp:339-345), its tiles in
 * ascending block column (seq_exec_order, sparse_matrix.h:161-176).  Every
 * output row therefore accumulates its neighbours in ascending column order.
 */
int orc_spmm(const uint8_t *image, uint64_t bytes, const int64_t *offs,
             uint64_t nbr, int dtype, int k, const void *X, int64_t ldx,
             void *Y, int64_t ldy)
{
    img_info_t inf;
    int rc = read_info(image, bytes, offs, nbr, &inf);
    if (rc) return rc;
    int es = elem_size(dtype);
    if (!es) return ORC_ETYPE;
    if (k <= 0 || ldx < k || ldy < k) return ORC_EDIM;
    for (uint64_t br = 0; br < nbr; br++) {
        uint64_t r0 = br * inf.H;
        uint64_t nr = inf.nrows - r0 < inf.H ? inf.nrows - r0 : inf.H;
        uint8_t *band = (uint8_t *)Y + r0 * ldy * es;
        for (uint64_t r = 0; r < nr; r++) memset(band + r * ldy * es, 0, (size_t)k * es);
        const uint8_t *p = image + offs[br], *end = image + offs[br + 1];
        while (p < end) {
            tile_t tl;
            if ((rc = parse_tile(p, end, inf.esz, &tl))) return rc;
            run_tile_any(dtype, &tl, &inf, k, X, ldx, band, ldy);
            p += tl.size;
        }
    }
    return ORC_OK;
}

/* ---- multi-threaded baseline -------------------------------------------- */
typedef struct {
    const uint8_t *image;
    const int64_t *offs;
    uint64_t nbr;
    img_info_t inf;
    int dtype, k, es, copy;
    const void *X;
    int64_t ldx;
    void *Y;
    int64_t ldy;
    uint64_t sblock, balance_point;
    atomic_ullong next;      /* next block row to hand out */
    pthread_mutex_t lock;    /* matrix_io.cpp:215 spinlock */
} mt_ctx_t;

/* b2d_io_generator::get_next_io (matrix_io.cpp:212-236). */
static int next_task(mt_ctx_t *c, uint64_t *b0, uint64_t *nb)
{
    pthread_mutex_lock(&c->lock);
    uint64_t cur = atomic_load(&c->next);
    int ok = cur < c->nbr;
    if (ok) {
        uint64_t take = cur >= c->balance_point ? 1 : c->sblock;
        if (take > c->nbr - cur) take = c->nbr - cur;
        *b0 = cur;
        *nb = take;
        atomic_store(&c->next, cur + take);
    }
    pthread_mutex_unlock(&c->lock);
    return ok;
}

/* block_compute_task::run (sparse_matrix.cpp:251-314) with the zeroed
 * per-task band (:329-350) and copy-out (:354-393). */
static void *mt_worker(void *arg)
{
    mt_ctx_t *c = (mt_ctx_t *)arg;
    uint8_t *buf = NULL;
    size_t bufcap = 0;
    uint8_t *band = NULL;
    size_t bandcap = 0;
    uint64_t b0, nb;
    while (next_task(c, &b0, &nb)) {
        uint64_t beg = (uint64_t)c->offs[b0], end = (uint64_t)c->offs[b0 + nb];
        const uint8_t *data = c->image + beg;
        if (c->copy) {  /* in_mem_io::access memcpy (in_mem_io.cpp:498-508) */
            if (bufcap < end - beg) { free(buf); bufcap = end - beg; buf = (uint8_t *)malloc(bufcap); }
            memcpy(buf, c->image + beg, end - beg);
            data = buf;
        }
        uint64_t r0 = b0 * c->inf.H;
        uint64_t nr = c->inf.nrows - r0 < nb * c->inf.H ? c->inf.nrows - r0 : nb * c->inf.H;
        size_t need = (size_t)nr * c->k * c->es;
        if (bandcap < need) { free(band); bandcap = need; band = (uint8_t *)malloc(bandcap); }
        memset(band, 0, need);
        /* Super-block walk: windows of nb block columns, block rows in turn. */
        const uint8_t *it[64];
        const uint8_t *itend[64];
        uint64_t nbl = nb < 64 ? nb : 64;
        for (uint64_t i = 0; i < nbl; i++) {
            it[i] = data + (c->offs[b0 + i] - (int64_t)beg);
            itend[i] = data + (c->offs[b0 + i + 1] - (int64_t)beg);
        }
        uint64_t sbc = 0;
        int more = 1;
        while (more) {
            more = 0;
            for (uint64_t i = 0; i < nbl; i++) {
                while (it[i] < itend[i]) {
                    tile_t tl;
                    if (parse_tile(it[i], itend[i], c->inf.esz, &tl)) { it[i] = itend[i]; break; }
                    if (tl.nnz && tl.bc >= sbc + nbl) break;
                    if (tl.nnz) {
                        uint8_t *o = band + (i * c->inf.H) * c->k * c->es;
                        run_tile_any(c->dtype, &tl, &c->inf, c->k, c->X, c->ldx, o, c->k);
                    }
                    it[i] += tl.size;
                }
                if (it[i] < itend[i]) more = 1;
            }
            sbc += nbl;
        }
        for (uint64_t r = 0; r < nr; r++)
            memcpy((uint8_t *)c->Y + (r0 + r) * c->ldy * c->es, band + r * c->k * c->es, (size_t)c->k * c->es);
    }
    free(buf);
    free(band);
    return NULL;
}

/* orc_spmm restricted to block rows [br0, br1): Y's row 0 is global row
 * br0*H.  Used to check a sample of a matrix too large for a full CPU pass. */
int orc_spmm_range(const uint8_t *image, uint64_t bytes, const int64_t *offs,
                   uint64_t nbr, uint64_t br0, uint64_t br1, int dtype, int k,
                   const void *X, int64_t ldx, void *Y, int64_t ldy)
{
    img_info_t inf;
    int rc = read_info(image, bytes, offs, nbr, &inf);
    if (rc) return rc;
    int es = elem_size(dtype);
    if (!es) return ORC_ETYPE;
    if (k <= 0 || ldx < k || ldy < k || br0 > br1 || br1 > nbr) return ORC_EDIM;
    for (uint64_t br = br0; br < br1; br++) {
        uint64_t r0 = br * inf.H;
        uint64_t nr = inf.nrows - r0 < inf.H ? inf.nrows - r0 : inf.H;
        uint8_t *band = (uint8_t *)Y + (r0 - br0 * inf.H) * ldy * es;
        for (uint64_t r = 0; r < nr; r++) memset(band + r * ldy * es, 0, (size_t)k * es);
        const uint8_t *p = image + offs[br], *end = image + offs[br + 1];
        while (p < end) {
            tile_t tl;
            if ((rc = parse_tile(p, end, inf.esz, &tl))) return rc;
            run_tile_any(dtype, &tl, &inf, k, X, ldx, band, ldy);
            p += tl.size;
        }
    }
    return ORC_OK;
}

/* ---- tile visiting order inside a task ------------------------------------
 * Position of cell (x, y) on the Hilbert curve filling an n x n grid (n a
 * power of two), the published recursive construction (the "xy2d" of the
 * Wikipedia article that matrix/hilbert_curve.cpp:1-34 names): at each
 * scale s the quadrant (rx, ry) contributes s*s*((3*rx) ^ ry) and the
 * coordinates below s are reflected/transposed into that quadrant's frame. */
uint64_t orc_hilbert_xy2d(uint64_t n, uint64_t x, uint64_t y)
{
    uint64_t d = 0;
    for (uint64_t s = n / 2; s > 0; s /= 2) {
        const uint64_t rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
        d += s * s * ((3 * rx) ^ ry);
        x &= s - 1;
        y &= s - 1;
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            const uint64_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

typedef struct { uint64_t d, i, j; } hcell_t;

static int hcell_cmp(const void *a, const void *b)
{
    const uint64_t x = ((const hcell_t *)a)->d, y = ((const hcell_t *)b)->d;
    return x < y ? -1 : x > y;
}

/*
 * Y = A*X sequentially, with the block rows taken in tasks of task_rows and
 * the tiles of each task visited as block_compute_task::run does
 * (matrix/sparse_matrix.cpp:246-314): super blocks of task_rows x task_rows
 * tiles (block rows of the task, the next task_rows block columns), each
 * visited in seq_exec_order (block row by block row, ascending column;
 * sparse_matrix.h:161-176) or, with hilbert != 0, in hilbert_exec_order
 * (cells sorted by their Hilbert position, sparse_matrix.cpp:144-186).  The
 * Hilbert order needs every task to be a full task_rows x task_rows square
 * (:170-173), so nbr must be a multiple of task_rows, a power of two.
 */
int orc_spmm_exec(const uint8_t *image, uint64_t bytes, const int64_t *offs,
                  uint64_t nbr, int dtype, int k, const void *X, int64_t ldx,
                  void *Y, int64_t ldy, uint64_t task_rows, int hilbert)
{
    img_info_t inf;
    int rc = read_info(image, bytes, offs, nbr, &inf);
    if (rc) return rc;
    const int es = elem_size(dtype);
    if (!es) return ORC_ETYPE;
    if (k <= 0 || ldx < k || ldy < k || task_rows == 0 || task_rows > 64) return ORC_EDIM;
    if (hilbert && ((task_rows & (task_rows - 1)) || nbr % task_rows)) return ORC_EDIM;
    const uint64_t n = task_rows, cells = n * n;
    hcell_t *order = (hcell_t *)malloc(cells * sizeof(hcell_t));
    tile_t *tiles = (tile_t *)malloc(cells * sizeof(tile_t));
    char *have = (char *)malloc(cells);
    if (!order || !tiles || !have) {
        free(order); free(tiles); free(have);
        return ORC_ENOMEM;
    }
    for (uint64_t i = 0; i < n; i++)
        for (uint64_t j = 0; j < n; j++) {
            hcell_t *c = &order[i * n + j];
            c->i = i;
            c->j = j;
            c->d = hilbert ? orc_hilbert_xy2d(n, i, j) : i * n + j;
        }
    qsort(order, cells, sizeof(hcell_t), hcell_cmp);
    for (uint64_t r = 0; r < inf.nrows; r++) memset((uint8_t *)Y + r * ldy * es, 0, (size_t)k * es);
    for (uint64_t b0 = 0; b0 < nbr && !rc; b0 += n) {
        const uint64_t nb = nbr - b0 < n ? nbr - b0 : n;
        const uint8_t *it[64], *itend[64];
        for (uint64_t i = 0; i < nb; i++) {
            it[i] = image + offs[b0 + i];
            itend[i] = image + offs[b0 + i + 1];
        }
        for (uint64_t sbc = 0;; sbc += n) {
            int more = 0;
            memset(have, 0, cells);
            for (uint64_t i = 0; i < nb && !rc; i++) {
                while (it[i] < itend[i]) {
                    tile_t tl;
                    if ((rc = parse_tile(it[i], itend[i], inf.esz, &tl))) break;
                    if (tl.nnz == 0) { it[i] += tl.size; continue; }
                    if (tl.bc >= sbc + n) break;
                    tiles[i * n + (tl.bc - sbc)] = tl;
                    have[i * n + (tl.bc - sbc)] = 1;
                    it[i] += tl.size;
                }
                if (it[i] < itend[i]) more = 1;
            }
            if (rc) break;
            for (uint64_t c = 0; c < cells; c++) {
                const uint64_t i = order[c].i, j = order[c].j;
                if (i >= nb || !have[i * n + j]) continue;
                uint8_t *band = (uint8_t *)Y + (b0 + i) * inf.H * ldy * es;
                run_tile_any(dtype, &tiles[i * n + j], &inf, k, X, ldx, band, ldy);
            }
            if (!more) break;
        }
    }
    free(order);
    free(tiles);
    free(have);
    return rc;
}

int orc_spmm_mt(const uint8_t *image, uint64_t bytes, const int64_t *offs,
                uint64_t nbr, int dtype, int k, const void *X, int64_t ldx,
                void *Y, int64_t ldy, int nthreads, int copy_block_rows)
{
    mt_ctx_t c;
    memset(&c, 0, sizeof(c));
    int rc = read_info(image, bytes, offs, nbr, &c.inf);
    if (rc) return rc;
    c.es = elem_size(dtype);
    if (!c.es) return ORC_ETYPE;
    if (k <= 0 || ldx < k || ldy < k) return ORC_EDIM;
    if (nthreads < 1) nthreads = 1;
    c.image = image; c.offs = offs; c.nbr = nbr; c.dtype = dtype; c.k = k;
    c.copy = copy_block_rows; c.X = X; c.ldx = ldx; c.Y = Y; c.ldy = ldy;
    /* cal_super_block_size (sparse_matrix.h:450-463): 1 MiB cpu cache,
     * rows rounded down to a power of two (:663-665), CHUNK_SIZE 16384. */
    uint64_t kp = 1;
    while (kp * 2 <= (uint64_t)k) kp *= 2;
    uint64_t sb = (1024 * 1024) / (kp * c.es) / c.inf.H / 2;
    if (sb < 1) sb = 1;
    uint64_t mx = 16384 / c.inf.H;
    if (mx < 1) mx = 1;
    c.sblock = sb < mx ? sb : mx;
    if (c.sblock > 64) c.sblock = 64;
    /* b2d_io_generator ctor (matrix_io.cpp:188-210). */
    if (c.sblock > 1)
        c.balance_point = (nbr > c.sblock * (uint64_t)nthreads) ? nbr - (uint64_t)nthreads : 0;
    else
        c.balance_point = nbr;
    atomic_init(&c.next, 0);
    pthread_mutex_init(&c.lock, NULL);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    if (!th) return ORC_ENOMEM;
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, mt_worker, &c);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&c.lock);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Tall-skinny dense ops (FP64, sequential sums).                            */
/* ------------------------------------------------------------------------ */
/* G[i][j] = sum_r A[r][i] * B[r][j]; MvTransMv (block_dense_matrix.cpp:1429-1513). */
void orc_gram_f64(const double *A, int64_t lda, const double *B, int64_t ldb,
                  uint64_t n, int ka, int kb, double *G)
{
    for (int i = 0; i < ka * kb; i++) G[i] = 0.0;
    for (uint64_t r = 0; r < n; r++)
        for (int i = 0; i < ka; i++) {
            double a = A[r * lda + i];
            for (int j = 0; j < kb; j++) G[i * kb + j] += a * B[r * ldb + j];
        }
}

/* Y = alpha*A*C + beta*Y, C is ka x kc row-major; gemm_op::run
 * (block_dense_matrix.cpp:822-914). */
void orc_gemm_tall_f64(const double *A, int64_t lda, uint64_t n, int ka,
                       const double *C, int kc, double alpha, double beta,
                       double *Y, int64_t ldy)
{
    for (uint64_t r = 0; r < n; r++)
        for (int j = 0; j < kc; j++) {
            double s = 0.0;
            for (int i = 0; i < ka; i++) s += A[r * lda + i] * C[i * kc + j];
            Y[r * ldy + j] = (beta == 0.0 ? 0.0 : beta * Y[r * ldy + j]) + alpha * s;
        }
}

/* MvNorm (FM_MultiVector.h:589-628): sqrt of the column sums of squares. */
void orc_colnorm_f64(const double *A, int64_t lda, uint64_t n, int k, double *norms)
{
    for (int j = 0; j < k; j++) norms[j] = 0.0;
    for (uint64_t r = 0; r < n; r++)
        for (int j = 0; j < k; j++) norms[j] += A[r * lda + j] * A[r * lda + j];
    for (int j = 0; j < k; j++) norms[j] = sqrt(norms[j]);
}

/* ------------------------------------------------------------------------ */
/* Learned warm start: the 3-layer message-passing network, FP64.  No        */
/* reference counterpart (the reference starts from a random block,          */
/* FM_MultiVector.h:164-168); this is the checker of the device kernels.     */
/*   H0 = X0 (n x 16); for l = 0..2:                                         */
/*     Agg_l = D^-1/2 A D^-1/2 H_l,  Z_l = H_l Ws_l + Agg_l Wn_l + b_l,      */
/*     H_{l+1} = relu(Z_l);                                                  */
/*   Y = H_3 Wo (n x 16).                                                    */
/* Parameters, flat: per layer Ws (din x 64), Wn (din x 64), b (64); then    */
/* Wo (64 x 16); din = 16, 64, 64.                                           */
/* Loss (rotation-invariant subspace loss against orthonormal U):            */
/*   L = k - tr(G^-1 B), G = Y^T Y, B = C^T C, C = U^T Y;                    */
/*   dL/dY = -2 (U C G^-1 - Y G^-1 B G^-1).                                  */
/* ------------------------------------------------------------------------ */
#define GNN_IN 16
#define GNN_H 64
#define GNN_OUT 16

static int gnn_din(int l) { return l == 0 ? GNN_IN : GNN_H; }

uint64_t orc_gnn_param_count(void)
{
    uint64_t c = 0;
    for (int l = 0; l < 3; l++) c += 2ull * gnn_din(l) * GNN_H + GNN_H;
    return c + (uint64_t)GNN_H * GNN_OUT;
}

static void gnn_agg(uint64_t n, const int64_t *rp, const int64_t *ci, const double *dinv,
                    const double *H, int d, double *out)
{
    for (uint64_t r = 0; r < n; r++) {
        for (int j = 0; j < d; j++) out[r * d + j] = 0.0;
        for (int64_t e = rp[r]; e < rp[r + 1]; e++) {
            const int64_t c = ci[e];
            const double w = dinv[r] * dinv[c];
            for (int j = 0; j < d; j++) out[r * d + j] += w * H[c * d + j];
        }
    }
}

/* small dense inverse (Gauss-Jordan with partial pivoting), k <= 64 */
static int inv_small(const double *A, int k, double *Ai)
{
    double M[64 * 128];
    for (int i = 0; i < k; i++)
        for (int j = 0; j < 2 * k; j++) M[i * 2 * k + j] = j < k ? A[i * k + j] : (j - k == i ? 1.0 : 0.0);
    for (int c = 0; c < k; c++) {
        int p = c;
        for (int r = c + 1; r < k; r++) if (fabs(M[r * 2 * k + c]) > fabs(M[p * 2 * k + c])) p = r;
        if (M[p * 2 * k + c] == 0.0) return ORC_EINVAL;
        for (int j = 0; j < 2 * k; j++) { double t = M[c * 2 * k + j]; M[c * 2 * k + j] = M[p * 2 * k + j]; M[p * 2 * k + j] = t; }
        const double d = M[c * 2 * k + c];
        for (int j = 0; j < 2 * k; j++) M[c * 2 * k + j] /= d;
        for (int r = 0; r < k; r++) {
            if (r == c) continue;
            const double f = M[r * 2 * k + c];
            if (f != 0.0) for (int j = 0; j < 2 * k; j++) M[r * 2 * k + j] -= f * M[c * 2 * k + j];
        }
    }
    for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) Ai[i * k + j] = M[i * 2 * k + k + j];
    return ORC_OK;
}

int orc_subspace_loss(uint64_t n, int k, const double *Y, const double *U, double *loss, double *dY)
{
    if (k > 64) return ORC_EINVAL;
    double G[64 * 64], C[64 * 64], B[64 * 64], Gi[64 * 64], M1[64 * 64], M2[64 * 64], T[64 * 64];
    orc_gram_f64(Y, k, Y, k, n, k, k, G);
    orc_gram_f64(U, k, Y, k, n, k, k, C);
    for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) {
        double s = 0; for (int t = 0; t < k; t++) s += C[t * k + i] * C[t * k + j];
        B[i * k + j] = s;
    }
    if (inv_small(G, k, Gi)) return ORC_EINVAL;
    double tr = 0;
    for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) tr += Gi[i * k + j] * B[j * k + i];
    *loss = k - tr;
    /* M1 = C Gi ; M2 = Gi B Gi */
    for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) {
        double s = 0, t2 = 0;
        for (int t = 0; t < k; t++) { s += C[i * k + t] * Gi[t * k + j]; t2 += Gi[i * k + t] * B[t * k + j]; }
        M1[i * k + j] = s; T[i * k + j] = t2;
    }
    for (int i = 0; i < k; i++) for (int j = 0; j < k; j++) {
        double s = 0; for (int t = 0; t < k; t++) s += T[i * k + t] * Gi[t * k + j];
        M2[i * k + j] = s;
    }
    if (dY)
        for (uint64_t r = 0; r < n; r++)
            for (int j = 0; j < k; j++) {
                double a = 0, b = 0;
                for (int t = 0; t < k; t++) { a += U[r * k + t] * M1[t * k + j]; b += Y[r * k + t] * M2[t * k + j]; }
                dY[r * k + j] = -2.0 * (a - b);
            }
    return ORC_OK;
}

/* Round-to-nearest-even to bfloat16 through FP32 (the GPU kernels round
 * their FP32 operands this way before a BF16 matrix-core product). */
static double bf16_rne(double x)
{
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) != 0x7F800000u) u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

void orc_bf16_round(const double *in, double *out, uint64_t n)
{
    for (uint64_t i = 0; i < n; i++) out[i] = bf16_rne(in[i]);
}

/* Copy of a[0..n) with every value rounded to BF16 when `on`, else a itself. */
static const double *maybe_bf16(const double *a, uint64_t n, int on, double **keep)
{
    *keep = NULL;
    if (!on) return a;
    double *b = (double *)malloc(n * sizeof(double));
    orc_bf16_round(a, b, n);
    *keep = b;
    return b;
}

int orc_gnn_loss_grad(uint64_t n, const int64_t *rp, const int64_t *ci, const double *dinv,
                      const double *X0, const double *U, const double *P, double *loss,
                      double *grad, double *Yout)
{
    return orc_gnn_loss_grad_ex(n, rp, ci, dinv, X0, U, P, loss, grad, Yout, 0);
}

/* bf16_operands = 1 rounds the operands of every dense product (activations,
 * aggregated features, weights, output and hidden gradients) to BF16 where
 * the BF16 matrix-core mode of the GPU kernels does; sums stay FP64. */
int orc_gnn_loss_grad_ex(uint64_t n, const int64_t *rp, const int64_t *ci, const double *dinv,
                         const double *X0, const double *U, const double *P, double *loss,
                         double *grad, double *Yout, int bf16_operands)
{
    const int rb = bf16_operands;
    double *tmp_a, *tmp_b, *tmp_w0, *tmp_w1;
    double *H[4], *Agg[3], *Z[3];
    H[0] = (double *)X0;
    for (int l = 0; l < 3; l++) {
        Agg[l] = (double *)malloc(n * gnn_din(l) * sizeof(double));
        Z[l] = (double *)malloc(n * GNN_H * sizeof(double));
        H[l + 1] = (double *)malloc(n * GNN_H * sizeof(double));
    }
    double *Y = (double *)malloc(n * GNN_OUT * sizeof(double));
    double *dY = (double *)malloc(n * GNN_OUT * sizeof(double));
    double *dH = (double *)malloc(n * GNN_H * sizeof(double));
    double *dZ = (double *)malloc(n * GNN_H * sizeof(double));
    double *T = (double *)malloc(n * GNN_H * sizeof(double));
    double *T2 = (double *)malloc(n * GNN_H * sizeof(double));
    const double *Wp[3][3];
    uint64_t off = 0, goff[3];
    for (int l = 0; l < 3; l++) {
        const int d = gnn_din(l);
        goff[l] = off;
        Wp[l][0] = P + off; off += (uint64_t)d * GNN_H;
        Wp[l][1] = P + off; off += (uint64_t)d * GNN_H;
        Wp[l][2] = P + off; off += GNN_H;
    }
    const double *Wo = P + off;
    const uint64_t wo_off = off;
    for (int l = 0; l < 3; l++) {
        const int d = gnn_din(l);
        gnn_agg(n, rp, ci, dinv, H[l], d, Agg[l]);
        const double *Hr = maybe_bf16(H[l], n * d, rb, &tmp_a);
        const double *Ar = maybe_bf16(Agg[l], n * d, rb, &tmp_b);
        const double *W0 = maybe_bf16(Wp[l][0], (uint64_t)d * GNN_H, rb, &tmp_w0);
        const double *W1 = maybe_bf16(Wp[l][1], (uint64_t)d * GNN_H, rb, &tmp_w1);
        for (uint64_t r = 0; r < n; r++)
            for (int j = 0; j < GNN_H; j++) {
                double s = 0;
                for (int t = 0; t < d; t++)
                    s += Hr[r * d + t] * W0[t * GNN_H + j] + Ar[r * d + t] * W1[t * GNN_H + j];
                s += Wp[l][2][j];
                Z[l][r * GNN_H + j] = s;
                H[l + 1][r * GNN_H + j] = s > 0 ? s : 0;
            }
        free(tmp_a); free(tmp_b); free(tmp_w0); free(tmp_w1);
    }
    {
        const double *Hr = maybe_bf16(H[3], n * GNN_H, rb, &tmp_a);
        const double *Wr = maybe_bf16(Wo, (uint64_t)GNN_H * GNN_OUT, rb, &tmp_w0);
        for (uint64_t r = 0; r < n; r++)
            for (int j = 0; j < GNN_OUT; j++) {
                double s = 0;
                for (int t = 0; t < GNN_H; t++) s += Hr[r * GNN_H + t] * Wr[t * GNN_OUT + j];
                Y[r * GNN_OUT + j] = s;
            }
        free(tmp_a); free(tmp_w0);
    }
    if (Yout) memcpy(Yout, Y, n * GNN_OUT * sizeof(double));
    int rc = orc_subspace_loss(n, GNN_OUT, Y, U, loss, dY);
    if (rc == ORC_OK && grad) {
        /* dWo = H3^T dY ; dH3 = dY Wo^T */
        {
            const double *Hr = maybe_bf16(H[3], n * GNN_H, rb, &tmp_a);
            const double *dYr = maybe_bf16(dY, n * GNN_OUT, rb, &tmp_b);
            const double *Wr = maybe_bf16(Wo, (uint64_t)GNN_H * GNN_OUT, rb, &tmp_w0);
            orc_gram_f64(Hr, GNN_H, dYr, GNN_OUT, n, GNN_H, GNN_OUT, grad + wo_off);
            for (uint64_t r = 0; r < n; r++)
                for (int t = 0; t < GNN_H; t++) {
                    double s = 0;
                    for (int j = 0; j < GNN_OUT; j++) s += dYr[r * GNN_OUT + j] * Wr[t * GNN_OUT + j];
                    dH[r * GNN_H + t] = s;
                }
            free(tmp_a); free(tmp_b); free(tmp_w0);
        }
        for (int l = 2; l >= 0; l--) {
            const int d = gnn_din(l);
            for (uint64_t e = 0; e < n * GNN_H; e++) dZ[e] = Z[l][e] > 0 ? dH[e] : 0.0;
            double *tmp_z;
            const double *dZr = maybe_bf16(dZ, n * GNN_H, rb, &tmp_z);
            const double *Hr = maybe_bf16(H[l], n * d, rb, &tmp_a);
            const double *Ar = maybe_bf16(Agg[l], n * d, rb, &tmp_b);
            orc_gram_f64(Hr, d, dZr, GNN_H, n, d, GNN_H, grad + goff[l]);
            orc_gram_f64(Ar, d, dZr, GNN_H, n, d, GNN_H, grad + goff[l] + (uint64_t)d * GNN_H);
            free(tmp_a); free(tmp_b);
            double *db = grad + goff[l] + 2ull * d * GNN_H;
            for (int j = 0; j < GNN_H; j++) db[j] = 0;
            for (uint64_t r = 0; r < n; r++) for (int j = 0; j < GNN_H; j++) db[j] += dZ[r * GNN_H + j];
            if (l == 0) { free(tmp_z); break; }
            /* dH_l = dZ Ws^T + Ahat (dZ Wn^T) */
            const double *W0 = maybe_bf16(Wp[l][0], (uint64_t)d * GNN_H, rb, &tmp_w0);
            const double *W1 = maybe_bf16(Wp[l][1], (uint64_t)d * GNN_H, rb, &tmp_w1);
            for (uint64_t r = 0; r < n; r++)
                for (int t = 0; t < d; t++) {
                    double a = 0, b = 0;
                    for (int j = 0; j < GNN_H; j++) {
                        a += dZr[r * GNN_H + j] * W0[t * GNN_H + j];
                        b += dZr[r * GNN_H + j] * W1[t * GNN_H + j];
                    }
                    T[r * d + t] = a;
                    T2[r * d + t] = b;
                }
            free(tmp_z); free(tmp_w0); free(tmp_w1);
            gnn_agg(n, rp, ci, dinv, T2, d, dH);
            for (uint64_t e = 0; e < n * (uint64_t)d; e++) dH[e] += T[e];
        }
    }
    for (int l = 0; l < 3; l++) { free(Agg[l]); free(Z[l]); free(H[l + 1]); }
    free(Y); free(dY); free(dH); free(dZ); free(T); free(T2);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* PageRank (matrix/libmatrix-algs/PageRank.cpp:49-137), FP32 throughout as  */
/* the reference computes it: out_deg = max(A*1, 1) (:49-60, :83); pr = 1/n  */
/* (:31-43); each iteration in = pr / out_deg, pr1 = in * out_deg (:101-105),*/
/* out = A^T * in (:110), pr2 = d*out + (1-d)/n (:118-119), converged =      */
/* #{|pr1 - pr2| < 0.001/n} (:97,120-124), pr = pr2; until every vertex has  */
/* converged or max_iters.  A and At are images of the matrix and of its     */
/* transpose (the same image for a symmetric matrix).                        */
/* ------------------------------------------------------------------------ */
int orc_pagerank(const uint8_t *a_img, uint64_t a_bytes, const int64_t *a_offs, uint64_t a_nbr,
                 const uint8_t *t_img, uint64_t t_bytes, const int64_t *t_offs, uint64_t t_nbr,
                 uint64_t n, int max_iters, float damping, float *pr, int *iters)
{
    float *deg = (float *)malloc(n * sizeof(float));
    float *ones = (float *)calloc(n ? n : 1, sizeof(float));
    float *in = (float *)malloc(n * sizeof(float));
    float *out = (float *)malloc(n * sizeof(float));
    float *pr1 = (float *)malloc(n * sizeof(float));
    int rc = ORC_OK;
    if (!deg || !ones || !in || !out || !pr1) { rc = ORC_ENOMEM; goto done; }
    for (uint64_t i = 0; i < n; i++) ones[i] = 1.0f;
    rc = orc_spmm(a_img, a_bytes, a_offs, a_nbr, ORC_F32, 1, ones, 1, deg, 1);
    if (rc) goto done;
    for (uint64_t i = 0; i < n; i++) {
        deg[i] = deg[i] > 1.0f ? deg[i] : 1.0f;
        pr[i] = (float)(1.0 / (double)n);
    }
    {
        const float eps = (float)(0.001 / (double)n);
        const float add = (float)((1.0 - (double)damping) / (double)n);
        uint64_t conv = 0;
        int it = 0;
        while (conv < n && it < max_iters) {
            for (uint64_t i = 0; i < n; i++) {
                in[i] = pr[i] / deg[i];
                pr1[i] = in[i] * deg[i];
            }
            rc = orc_spmm(t_img, t_bytes, t_offs, t_nbr, ORC_F32, 1, in, 1, out, 1);
            if (rc) goto done;
            conv = 0;
            for (uint64_t i = 0; i < n; i++) {
                const float p2 = out[i] * damping + add;
                conv += fabsf(pr1[i] - p2) < eps;
                pr[i] = p2;
            }
            it++;
        }
        *iters = it;
    }
done:
    free(deg); free(ones); free(in); free(out); free(pr1);
    return rc;
}
