// fmx_internal.h -- shared host/device declarations of libfmx.
//
// Layout of a matrix in HBM:
//   image    : the reference-format tile image, byte for byte (4096-byte
//              header + block rows; matrix/sparse_matrix_format.h:213-240).
//   dir      : one {start, bc} pair per stored tile, in image order; `start`
//              is the tile's first u16 word relative to its block row.
//   plans    : per sub-band height Hs, a list of work items (contiguous word
//              ranges of one block row) and the partial-band reduction list.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fmx.h"

namespace fmx {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint64_t kHeaderBytes = 4096;

// One unit of SpMM work: words [begin, begin+nwords) of a block row, whose
// non-zeros with tile-relative rows in [sb_base, sb_base + rows) accumulate
// into one LDS band.  `carry` is the tile-relative row in force at `begin`
// when the range starts inside a row part (kNone otherwise).
struct Item {
    uint64_t brow_word;   // absolute u16 index of the block row's first word
    uint32_t begin;       // first word, relative to the block row
    uint32_t nwords;
    uint32_t dir_begin;   // directory index of the tile holding `begin`
    uint32_t dir_end;     // one past the block row's last directory entry
    uint32_t row0_lo;     // first output row of the band (low / high 32 bits)
    uint32_t row0_hi;
    uint32_t sb_base;     // tile-relative row of band row 0
    uint32_t rows;        // rows in the band
    uint32_t carry;
    uint32_t slot;        // partial-band slot, kNone = write Y directly
    uint32_t rot_begin;   // the item is processed from this word (a tile start, or begin) ...
    uint32_t rot_dir;     // ... whose directory index this is, to its end, then from begin
};
static_assert(sizeof(Item) == 56, "Item layout");

// A band split over several items: Y rows [row0, row0+rows) are the ordered
// sum of partial slots [slot_begin, slot_begin+nslots).
struct Split {
    uint64_t row0;
    uint32_t rows;
    uint32_t slot_begin;
    uint32_t nslots;
    uint32_t pad;
};
static_assert(sizeof(Split) == 24, "Split layout");

struct Plan {
    uint32_t hs = 0;            // band height (rows)
    uint32_t nitems = 0;
    uint32_t nslots = 0;
    uint32_t nsplits = 0;
    Item *d_items = nullptr;
    Split *d_splits = nullptr;
    std::vector<Item> items;    // host copy (tests / diagnostics)
    // streamed matrices: the items of chunk c are [chunk_items[c], chunk_items[c+1])
    std::vector<uint32_t> chunk_items;
    ~Plan();
};

struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
    ~Scratch();
};

}  // namespace fmx

struct fmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    int num_cus = 0;
    size_t lds_per_block = 0;
    void *ws = nullptr;        // grow-only device workspace (dense-op partials)
    size_t ws_bytes = 0;
    fmx::Scratch xstage;       // grow-only: a dense operand restaged row-major / pre-scaled
    fmx::Scratch xhot;         // grow-only: the hot rows of X, densely (gather-order images)
    void *noalloc_ws = nullptr;  // this device's copy of the fixed scratch (fmx_gram_noalloc)
    size_t noalloc_bytes = 0;
};

struct fmx_spm {
    fmx_ctx *ctx = nullptr;
    uint64_t nrows = 0, ncols = 0, nnz = 0, H = 0, W = 0, nbr = 0, nbc = 0;
    uint64_t bytes = 0, esz = 0, ntiles = 0;
    int log2h = 0, log2w = 0, prim = 0;
    uint8_t *d_image = nullptr;         // image in HBM (padded to 4 KiB)
    uint2 *d_dir = nullptr;             // tile directory
    uint32_t *d_vdir = nullptr;         // weighted images: values-region word of each tile
    std::vector<uint8_t> h_image;       // host copy used to build plans (resident matrices)
    std::vector<int64_t> offs;          // block-row byte offsets
    std::vector<uint2> dir;             // host directory (empty for an image built on the device)
    std::vector<uint32_t> vdir;         // weighted images: values-region word (block-row relative)
    std::vector<uint32_t> brow_dir;     // first directory index per block row (+1)
    std::map<uint32_t, std::unique_ptr<fmx::Plan>> plans;
    fmx::Scratch partial;               // partial bands of split items
    // Out-of-core form (d_image == nullptr): the image stays in pinned host
    // memory and every SpMM streams it through two HBM staging buffers in
    // chunks of whole block rows; directory and plans stay in HBM.
    uint8_t *h_pinned = nullptr;        // image, padded with zeros to 4 KiB
    uint64_t h_pinned_len = 0;          // bytes given (file or buffer length)
    std::vector<uint64_t> chunk_br;     // first block row of each chunk, then nbr
    std::vector<uint64_t> chunk_lo, chunk_hi;  // 16-byte aligned byte span of each chunk
    uint8_t *d_stage[2] = {nullptr, nullptr};
    uint64_t stage_bytes = 0;
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_copied[2] = {nullptr, nullptr};
    hipEvent_t ev_free[2] = {nullptr, nullptr};   // last kernel reading each buffer
    bool free_recorded[2] = {false, false};
    hipEvent_t ev_start = nullptr;
    // Resident images with tiles taller than an SpMM band: the same matrix
    // re-cut into short tiles, which every SpMM runs on (retile.cpp).
    std::unique_ptr<fmx_spm> work;
    // Gather order (hotcols.cpp): the matrix with its nhot most used columns
    // relabelled 0..nhot-1 (hot_ids[i] is the original column of new column
    // i) and every other column c relabelled nhot + c.  Default-mode SpMMs
    // run on it with X's hot rows restaged densely first.
    std::unique_ptr<fmx_spm> fast;
    uint32_t *d_hot = nullptr;
    uint64_t nhot = 0;
    uint64_t build_ns = 0;              // device-built: construction kernels' time
    uint64_t edges_in = 0;              // device-built: edges of the input list
    bool streamed() const { return h_pinned != nullptr; }
    const uint8_t *host_image() const { return h_pinned ? h_pinned : h_image.data(); }
    uint64_t host_bytes() const { return h_pinned ? h_pinned_len : h_image.size(); }
    ~fmx_spm();
};

namespace fmx {

// Records `msg` on the context and returns `st`.
fmx_status fail(fmx_ctx *ctx, fmx_status st, const std::string &msg);
fmx_status hip_fail(fmx_ctx *ctx, hipError_t e, const char *what);
int elem_size(fmx_dtype dt);

// Validates a host image and fills the matrix fields + directory.
fmx_status analyse_image(fmx_ctx *ctx, fmx_spm *spm);
// Device construction (build_tiles.hip); ncols != 0: an n x ncols matrix;
// internal: the result gets no work images of its own.
fmx_status build_impl(fmx_ctx *ctx, const uint32_t *src, const uint32_t *dst, size_t m, uint64_t n,
                      uint32_t log2_h, uint32_t log2_w, uint32_t flags, const void *vals, uint32_t esz,
                      int32_t prim_type, fmx_spm **out, uint64_t ncols, bool internal);
// The gather-order image of a resident matrix (hotcols.cpp): columns
// relabelled so that the most used ones form a dense front block of X.
// stream_chunk != 0: the image is streamed from pinned host memory in
// chunks of that many bytes, as the matrix it belongs to.
fmx_status prepare_fast(fmx_ctx *ctx, fmx_spm *spm, uint64_t stream_chunk = 0);
// The non-zeros of the image in HBM of src (image_dev.hip) in stream order
// into device lists rows, cols (relabelled into gather order) and vals (esz
// bytes each, if weighted); *hot (device, allocated here) lists the nhot
// columns that cover `permille` of the non-zeros, in their new order.
// *nhot = 0 (and nothing allocated) when there are none.
fmx_status gather_order_lists(fmx_ctx *ctx, const fmx_spm *src, uint64_t ncols, uint64_t permille,
                              uint32_t *rows, uint32_t *cols, void *vals, uint32_t **hot, uint64_t *nhot);
// The non-zeros of the image in HBM of src in stream order into device lists
// rows, cols and vals (esz bytes each, if weighted); uses (ncols counters,
// zeroed by the caller, or NULL) counts each column's non-zeros
// (image_dev.hip).
fmx_status decode_lists(fmx_ctx *ctx, const fmx_spm *src, uint32_t *rows, uint32_t *cols, void *vals,
                        uint32_t *uses);
// carry[i]: the row in force at word p[i] of the block row whose first word
// is base[i] (an absolute word index into the image in HBM), kNone at a
// header word; p[i] lies past its tile's header (image_dev.hip).
fmx_status plan_carries(fmx_ctx *ctx, const fmx_spm *s, const std::vector<uint64_t> &base,
                        const std::vector<uint32_t> &p, std::vector<uint32_t> &carry);
fmx_status finish_streamed(fmx_ctx *ctx, fmx_spm *s, uint64_t chunk_bytes, bool whole_file);
// The chunks, staging buffers and directory of an analysed streamed image
// (or of its re-cut work image), without a gather-order image.
fmx_status finish_streamed_chunks(fmx_ctx *ctx, fmx_spm *s, uint64_t chunk_bytes);
fmx_status alloc_pinned(fmx_ctx *ctx, fmx_spm *s, uint64_t len);
// The work image of a resident matrix with tiles taller than the work
// tile height (retile.cpp): built, validated and uploaded into spm->work.
fmx_status prepare_work(fmx_ctx *ctx, fmx_spm *spm);
uint32_t work_rows_for(const fmx_spm *spm);
void retile_image(const fmx_spm *spm, uint32_t h2, std::vector<uint8_t> &img, std::vector<int64_t> &offs);
// Builds (or fetches) the work plan for band height hs.  An ordered plan
// neither splits nor rotates items, so every band is one item walked from
// its first word (the sequential mode needs this).
fmx_status get_plan(fmx_ctx *ctx, fmx_spm *spm, uint32_t hs, bool ordered, Plan **out);
// Band height for element size s and width k.
uint32_t band_rows(const fmx_spm *spm, int k, int s);

// Launches the SpMM kernel over plan items [item_begin, item_end) with the
// image at img (absolute u16 word w of the image is img[w]) and, if reduce,
// the in-order reduction of the plan's split bands.
// the kernels' part of the FMX_SPMM_* flags (FMX_SPMM_SEQ, FMX_SPMM_Y_COLMAJOR).
// Columns below nhot are read from Xhot (nhot x k, leading dimension k),
// column c >= nhot from row c - nhot of X.
fmx_status launch_spmm(fmx_ctx *ctx, const fmx_spm *spm, const Plan *plan,
                       fmx_dtype dt, int k, const void *X, int64_t ldx,
                       void *Y, int64_t ldy, void *partial,
                       hipStream_t stream, const uint16_t *img,
                       uint32_t item_begin, uint32_t item_end, bool reduce,
                       const double *rscale, uint32_t flags, const void *Xhot = nullptr,
                       uint64_t nhot = 0);
// out[i][0..k) = X[hot[i]][0..k) for i < nhot (out has leading dimension k).
fmx_status stage_hot_rows(fmx_ctx *ctx, fmx_dtype dt, uint64_t nhot, int k, const void *X, int64_t ldx,
                          const uint32_t *hot, void *out, hipStream_t stream);

// out (n x k, row-major, leading dimension k) = X, read row- or column-major
// (xcol), times pre[i] per row when pre is given (FP data only).
fmx_status stage_dense_rows(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int k, const void *X, int64_t ldx, bool xcol,
                            const double *pre, void *out, hipStream_t stream);

// Y = diag(post) A diag(pre) X with FMX_SPMM_* flags; pre / post may be null.
// The weighted-value type may differ from dtype only when mixed is set
// (FP64 dense data with FP32 values, spmm_creator<double, float>).
fmx_status spmm_exec(fmx_ctx *ctx, const fmx_spm *spm, fmx_dtype dtype, int k, const double *pre,
                     const double *post, const void *X, int64_t ldx, void *Y, int64_t ldy, uint32_t flags,
                     bool mixed, hipStream_t stream);

// fmx_spmm with each output row r scaled by rscale[r] (FP32/FP64 data; null:
// none), applied to the FP64 sum before its single rounding.
fmx_status spmm_scaled(fmx_ctx *ctx, const fmx_spm *spm, fmx_dtype dtype, int k, const void *X, int64_t ldx,
                       void *Y, int64_t ldy, const double *rscale, hipStream_t stream);

}  // namespace fmx

#define FMX_HIP(ctx, call)                                            \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) return fmx::hip_fail((ctx), e_, #call); \
    } while (0)
