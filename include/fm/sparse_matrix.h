/*
 * fm/sparse_matrix.h -- C++ drop-in for FlashMatrix's sparse-matrix SpMM API.
 *
 * Same class and method names as the reference (matrix/sparse_matrix.h:
 * 475-651, matrix/matrix_header.h:42-91, matrix/fm_utils.h:37-39), over the
 * C ABI of include/fmx.h.  Dense operands live in HBM (fm::dense_matrix is a
 * row-major device matrix here); multiply() is synchronous like the
 * reference's (matrix/sparse_matrix.cpp:465).
 *
 * Error behaviour follows the reference: recoverable problems return false
 * or a null pointer after logging (matrix/sparse_matrix.cpp:715-718,
 * 755-766); a malformed image throws fm::wrong_format
 * (matrix/matrix_header.cpp:52-60).
 *
 * Header-only; link with libfmx.so and the HIP runtime.
 */
#ifndef FM_SPARSE_MATRIX_H
#define FM_SPARSE_MATRIX_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../fmx.h"

/* safs::create_io_factory (libsafs/io_interface.h:518-525) names the file a
 * matrix is streamed from.  Here the file is a regular host file: it is read
 * once into pinned host memory and streamed through HBM in block-row chunks
 * of at most chunk_bytes on every multiply.  The access method is accepted
 * for the reference's signature and does not change the streaming. */
namespace safs {
enum { REMOTE_ACCESS = 3, GLOBAL_CACHE_ACCESS = 4, PART_GLOBAL_ACCESS = 5 };

class file_io_factory {
    std::string name_;
    size_t chunk_bytes_;

public:
    typedef std::shared_ptr<file_io_factory> shared_ptr;
    file_io_factory(const std::string &name, size_t chunk_bytes) : name_(name), chunk_bytes_(chunk_bytes) {}
    const std::string &get_name() const { return name_; }
    size_t get_chunk_bytes() const { return chunk_bytes_; }
};

inline file_io_factory::shared_ptr create_io_factory(const std::string &file_name, const int access_option,
                                                     size_t chunk_bytes = size_t(1) << 30)
{
    (void)access_option;
    return file_io_factory::shared_ptr(new file_io_factory(file_name, chunk_bytes));
}

inline bool exist_safs_file(const std::string &file_name)
{
    FILE *f = fopen(file_name.c_str(), "rb");
    if (f) fclose(f);
    return f != nullptr;
}
}  // namespace safs

namespace fm {

class wrong_format : public std::runtime_error {
public:
    explicit wrong_format(const std::string &msg) : std::runtime_error(msg) {}
};

namespace detail {
inline fmx_ctx *&global_ctx()
{
    static fmx_ctx *ctx = nullptr;
    return ctx;
}
inline void log_error(const std::string &what)
{
    fprintf(stderr, "[fm error] %s\n", what.c_str());
}
inline bool check(fmx_status st, const char *what)
{
    if (st == FMX_OK) return true;
    std::string msg = std::string(what) + ": " + fmx_strerror(st);
    if (global_ctx()) msg += std::string(" (") + fmx_ctx_last_error(global_ctx()) + ")";
    if (st == FMX_EFORMAT) throw wrong_format(msg);
    log_error(msg);
    return false;
}
}  // namespace detail

/* fm::init_flash_matrix / destroy_flash_matrix (matrix/sparse_matrix.cpp:835-872). */
inline bool init_flash_matrix(int device = 0)
{
    if (detail::global_ctx()) return true;
    return detail::check(fmx_ctx_create(device, &detail::global_ctx()), "init_flash_matrix");
}

inline void destroy_flash_matrix()
{
    if (detail::global_ctx()) fmx_ctx_destroy(detail::global_ctx());
    detail::global_ctx() = nullptr;
}

/* block_2d_size (matrix/matrix_header.h:45-91). */
class block_2d_size {
    uint16_t nrow_log = 0, ncol_log = 0;

public:
    block_2d_size() {}
    block_2d_size(size_t num_rows, size_t num_cols)
    {
        while ((1ull << nrow_log) < num_rows) nrow_log++;
        while ((1ull << ncol_log) < num_cols) ncol_log++;
        if ((1ull << nrow_log) != num_rows || (1ull << ncol_log) != num_cols || num_rows > 32768 ||
            num_cols > 32768)
            throw wrong_format("block sizes must be powers of two <= 32768");
    }
    size_t get_nrow_log() const { return nrow_log; }
    size_t get_ncol_log() const { return ncol_log; }
    size_t get_nrow_mask() const { return get_num_rows() - 1; }
    size_t get_ncol_mask() const { return get_num_cols() - 1; }
    size_t get_num_rows() const { return size_t(1) << nrow_log; }
    size_t get_num_cols() const { return size_t(1) << ncol_log; }
    size_t cal_num_block_rows(size_t tot) const { return (tot + get_num_rows() - 1) / get_num_rows(); }
};

/* matrix_layout_t (matrix/generic_type.h:36-43) for dense matrices. */
enum matrix_layout_t { L_COL = 0, L_ROW = 1 };

/* A dense matrix as the SpMM sees it (matrix/dense_matrix.h): row-major
 * (L_ROW, element (i, j) at data[i*ld + j]) or column-major (L_COL, at
 * data[j*ld + i]), in HBM, or -- for operands larger than the HBM given to
 * them, the reference's external-memory dense matrices -- in pinned host
 * memory (is_in_mem() false, row-major). */
class dense_matrix {
    size_t nrow_, ncol_;
    int64_t ld_;
    fmx_dtype type_;
    void *data_;
    bool owned_;
    matrix_layout_t layout_;
    bool host_;

    dense_matrix(size_t nrow, size_t ncol, int64_t ld, fmx_dtype t, void *d, bool owned, matrix_layout_t l,
                 bool host)
        : nrow_(nrow), ncol_(ncol), ld_(ld), type_(t), data_(d), owned_(owned), layout_(l), host_(host) {}

public:
    typedef std::shared_ptr<dense_matrix> ptr;

    static size_t elem_size(fmx_dtype t) { return (t == FMX_I32 || t == FMX_F32) ? 4 : 8; }

    static ptr create(size_t nrow, size_t ncol, fmx_dtype t, matrix_layout_t layout = L_ROW)
    {
        void *d = nullptr;
        size_t bytes = nrow * ncol * elem_size(t);
        if (hipMalloc(&d, bytes ? bytes : 1) != hipSuccess) {
            detail::log_error("dense_matrix::create: out of device memory");
            return ptr();
        }
        const int64_t ld = (int64_t)std::max<size_t>(1, layout == L_ROW ? ncol : nrow);
        return ptr(new dense_matrix(nrow, ncol, ld, t, d, true, layout, false));
    }
    /* A row-major matrix in pinned host memory: the SpMM streams it through
     * HBM in column panels (multiply(dense_matrix::ptr, mem_size)). */
    static ptr create_host(size_t nrow, size_t ncol, fmx_dtype t)
    {
        void *d = nullptr;
        size_t bytes = nrow * ncol * elem_size(t);
        if (hipHostMalloc(&d, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
            detail::log_error("dense_matrix::create_host: out of pinned host memory");
            return ptr();
        }
        return ptr(new dense_matrix(nrow, ncol, (int64_t)std::max<size_t>(1, ncol), t, d, true, L_ROW, true));
    }
    /* Wrap caller-owned device memory (not freed). */
    static ptr wrap(void *dev, size_t nrow, size_t ncol, int64_t ld, fmx_dtype t, matrix_layout_t layout = L_ROW)
    {
        return ptr(new dense_matrix(nrow, ncol, ld, t, dev, false, layout, false));
    }
    static ptr create_from_host(const void *host, size_t nrow, size_t ncol, fmx_dtype t)
    {
        ptr m = create(nrow, ncol, t);
        if (m && hipMemcpy(m->data_, host, nrow * ncol * elem_size(t), hipMemcpyHostToDevice) != hipSuccess)
            return ptr();
        return m;
    }
    ~dense_matrix()
    {
        if (owned_ && data_) (void)(host_ ? hipHostFree(data_) : hipFree(data_));
    }
    /* Copies the matrix to `host` as a row-major nrow x ncol array. */
    bool copy_to_host(void *host) const
    {
        const size_t es = elem_size(type_);
        if (host_) {
            for (size_t i = 0; i < nrow_; i++)
                memcpy((char *)host + i * ncol_ * es, (const char *)data_ + i * ld_ * es, ncol_ * es);
            return true;
        }
        if (layout_ == L_ROW)
            return hipMemcpy2D(host, ncol_ * es, data_, ld_ * es, ncol_ * es, nrow_, hipMemcpyDeviceToHost) ==
                   hipSuccess;
        std::vector<char> col(nrow_ * ncol_ * es);
        if (hipMemcpy2D(col.data(), nrow_ * es, data_, ld_ * es, nrow_ * es, ncol_, hipMemcpyDeviceToHost) !=
            hipSuccess)
            return false;
        for (size_t j = 0; j < ncol_; j++)
            for (size_t i = 0; i < nrow_; i++)
                memcpy((char *)host + (i * ncol_ + j) * es, col.data() + (j * nrow_ + i) * es, es);
        return true;
    }
    size_t get_num_rows() const { return nrow_; }
    size_t get_num_cols() const { return ncol_; }
    int64_t get_ld() const { return ld_; }
    fmx_dtype get_type() const { return type_; }
    matrix_layout_t store_layout() const { return layout_; }
    bool is_in_mem() const { return !host_; }
    void *get_raw_arr() { return data_; }
    const void *get_raw_arr() const { return data_; }
};

namespace detail {
/* A custom SpMM task creator (task_creator, matrix/sparse_matrix.h:52-62)
 * as the reference's operators build them (spmm_creator, :377-436,656-693,
 * with scale_rows around it, matrix/test/test-eigen.cpp:112-127): optional
 * FP64 row scales of the input (pre, ncols values) and of the output (post,
 * nrows values), both device arrays; seq sums in the reference's sequential
 * order (FMX_SPMM_SEQ).  An FP64 input over an FP32-weighted matrix is
 * spmm_creator<double, float>. */
class spmm_creator {
    const double *pre_ = nullptr, *post_ = nullptr;
    bool seq_ = false;

public:
    typedef std::shared_ptr<spmm_creator> ptr;
    static ptr create(const double *pre = nullptr, const double *post = nullptr, bool seq = false)
    {
        ptr p(new spmm_creator());
        p->pre_ = pre;
        p->post_ = post;
        p->seq_ = seq;
        return p;
    }
    const double *pre() const { return pre_; }
    const double *post() const { return post_; }
    bool seq() const { return seq_; }
};
}  // namespace detail

/* Global SpMM options (the role matrix_config plays for the reference,
 * matrix/matrix_config.h:81-98): seq_fp_sum makes every multiply sum
 * floating-point rows in the reference's sequential order, bit-identical
 * to its CPU SpMM (FMX_SPMM_SEQ); the default sums FP32 rows in FP64 and
 * rounds once. */
struct matrix_config {
    bool seq_fp_sum = false;
};
inline matrix_config &matrix_conf()
{
    static matrix_config conf;
    return conf;
}

/* I/O failure while loading (the reference throws safs::io_exception,
 * matrix/sparse_matrix_format.cpp:262-276). */
class io_exception : public std::runtime_error {
public:
    explicit io_exception(const std::string &msg) : std::runtime_error(msg) {}
};

namespace detail {
inline std::vector<char> read_all(const std::string &file)
{
    FILE *f = fopen(file.c_str(), "rb");
    if (!f) throw io_exception("can't open " + file);
    std::vector<char> buf;
    if (fseek(f, 0, SEEK_END) == 0) {
        long n = ftell(f);
        if (n > 0 && fseek(f, 0, SEEK_SET) == 0) {
            buf.resize((size_t)n);
            if (fread(buf.data(), (size_t)n, 1, f) != 1) buf.clear();
        }
    }
    fclose(f);
    if (buf.empty()) throw io_exception("can't read " + file);
    return buf;
}

inline bool write_all(const std::string &file, const void *a, size_t na, const void *b, size_t nb)
{
    FILE *f = fopen(file.c_str(), "wb");
    if (!f) {
        log_error("can't open " + file);
        return false;
    }
    bool ok = (na == 0 || fwrite(a, na, 1, f) == 1) && (nb == 0 || fwrite(b, nb, 1, f) == 1);
    ok = (fclose(f) == 0) && ok;
    if (!ok) log_error("can't write to " + file);
    return ok;
}
}  // namespace detail

/* SpM_2d_index (matrix/sparse_matrix_format.h:441-485): the matrix header
 * page and the nbr+1 block-row offsets, as stored in a .mat_idx file. */
class SpM_2d_index {
    std::vector<char> header_;
    std::vector<int64_t> offs_;

public:
    typedef std::shared_ptr<SpM_2d_index> ptr;

    static ptr create(const std::vector<char> &header_page, const std::vector<int64_t> &offs)
    {
        ptr p(new SpM_2d_index());
        p->header_.assign(header_page.begin(), header_page.begin() + std::min<size_t>(header_page.size(), 4096));
        p->offs_ = offs;
        return p;
    }

    /* SpM_2d_index::load (matrix/sparse_matrix_format.cpp:259-281). */
    static ptr load(const std::string &idx_file)
    {
        std::vector<char> raw = detail::read_all(idx_file);
        if (raw.size() < 4096 + 8 || (raw.size() - 4096) % 8) throw wrong_format("bad index file " + idx_file);
        ptr p(new SpM_2d_index());
        p->header_.assign(raw.begin(), raw.begin() + 4096);
        p->offs_.resize((raw.size() - 4096) / 8);
        memcpy(p->offs_.data(), raw.data() + 4096, p->offs_.size() * 8);
        return p;
    }

    /* SpM_2d_index::dump (matrix/sparse_matrix_format.cpp:164-181). */
    void dump(const std::string &file) const
    {
        detail::write_all(file, header_.data(), header_.size(), offs_.data(), offs_.size() * sizeof(int64_t));
    }

    size_t get_num_block_rows() const { return offs_.empty() ? 0 : offs_.size() - 1; }
    int64_t get_block_row_off(size_t i) const { return offs_[i]; }
    const std::vector<int64_t> &get_offs() const { return offs_; }
    const std::vector<char> &get_header_page() const { return header_; }
};

/* SpM_2d_storage (matrix/sparse_matrix_format.h:488-...): the image bytes
 * (header page, then the block rows) of a .mat file. */
class SpM_2d_storage {
    std::vector<char> data_;

public:
    typedef std::shared_ptr<SpM_2d_storage> ptr;

    static ptr create(std::vector<char> image)
    {
        ptr p(new SpM_2d_storage());
        p->data_ = std::move(image);
        return p;
    }

    /* SpM_2d_storage::load (matrix/sparse_matrix_format.cpp:296-306); the
     * index argument is kept for the reference's signature. */
    static ptr load(const std::string &mat_file, SpM_2d_index::ptr index = SpM_2d_index::ptr())
    {
        (void)index;
        return create(detail::read_all(mat_file));
    }

    /* SpM_2d_storage::dump (matrix/sparse_matrix_format.cpp:417-420). */
    void dump(const std::string &mat_file) const
    {
        detail::write_all(mat_file, data_.data(), data_.size(), nullptr, 0);
    }

    const std::vector<char> &get_data() const { return data_; }
};

/* fm::sparse_matrix (matrix/sparse_matrix.h:475-651) for a 2D-tiled image. */
class sparse_matrix {
    std::shared_ptr<fmx_spm> spm_;
    fmx_spm_info info_{};
    bool symmetric_ = true;
    std::shared_ptr<sparse_matrix> transpose_;  // asymmetric matrices carry A^T

    explicit sparse_matrix(fmx_spm *s, bool sym) : spm_(s, [](fmx_spm *p) { fmx_spm_destroy(p); }),
                                                   symmetric_(sym)
    {
        fmx_spm_info_get(s, &info_);
    }

public:
    typedef std::shared_ptr<sparse_matrix> ptr;

    /* sparse_matrix::create(SpM_2d_index::ptr, SpM_2d_storage::ptr)
     * (matrix/sparse_matrix.h:522): the image (header + block rows) and the
     * block-row offsets of its index. */
    static ptr create(const std::vector<char> &image, const std::vector<int64_t> &offs)
    {
        if (!init_flash_matrix()) return ptr();
        fmx_spm *s = nullptr;
        if (!detail::check(fmx_spm_upload(detail::global_ctx(), image.data(), image.size(), offs.data(),
                                          offs.size() - 1, &s),
                           "sparse_matrix::create"))
            return ptr();
        return ptr(new sparse_matrix(s, true));
    }

    /* sparse_matrix::create(SpM_2d_index::ptr, SpM_2d_storage::ptr)
     * (matrix/sparse_matrix.h:522, matrix/test/test-2d_multiply.cpp:147-161).
     * The index header page must agree with the image's. */
    static ptr create(SpM_2d_index::ptr index, SpM_2d_storage::ptr storage)
    {
        if (!index || !storage) return ptr();
        const std::vector<char> &h = index->get_header_page(), &img = storage->get_data();
        if (h.size() < 64 || img.size() < 64 || memcmp(h.data(), img.data(), 64) != 0)
            throw wrong_format("index and matrix headers differ");
        return create(img, index->get_offs());
    }

    static ptr create(SpM_2d_index::ptr index, SpM_2d_storage::ptr storage, SpM_2d_index::ptr t_index,
                      SpM_2d_storage::ptr t_storage)
    {
        ptr a = create(index, storage), t = create(t_index, t_storage);
        if (!a || !t) return ptr();
        a->symmetric_ = t->symmetric_ = false;
        a->transpose_ = t;
        return a;
    }

    /* sparse_matrix::create(SpM_2d_index::ptr, file_io_factory) (matrix/
     * sparse_matrix.h:534-535, .cpp:693-698; matrix/test/test-2d_multiply.cpp:
     * 156-158): out of core.  The image stays in host memory and every
     * multiply streams it through HBM (fmx_spm_open_index). */
    static ptr create(SpM_2d_index::ptr index, safs::file_io_factory::shared_ptr mat_io_fac)
    {
        if (!index || !mat_io_fac || !init_flash_matrix()) return ptr();
        const std::vector<char> &h = index->get_header_page();
        if (h.size() < 64) throw wrong_format("index without a header page");
        fmx_spm *s = nullptr;
        const fmx_status st = fmx_spm_open_index(detail::global_ctx(), mat_io_fac->get_name().c_str(), h.data(),
                                                 index->get_offs().data(), index->get_num_block_rows(),
                                                 mat_io_fac->get_chunk_bytes(), &s);
        if (st == FMX_EFORMAT) throw wrong_format(fmx_ctx_last_error(detail::global_ctx()));
        if (!detail::check(st, "sparse_matrix::create")) return ptr();
        return ptr(new sparse_matrix(s, true));
    }

    static ptr create(SpM_2d_index::ptr index, safs::file_io_factory::shared_ptr mat_io_fac,
                      SpM_2d_index::ptr t_index, safs::file_io_factory::shared_ptr t_mat_io_fac)
    {
        ptr a = create(index, mat_io_fac), t = create(t_index, t_mat_io_fac);
        if (!a || !t) return ptr();
        a->symmetric_ = t->symmetric_ = false;
        a->transpose_ = t;
        return a;
    }

    /* Whether the image is resident in HBM (false for the streamed form). */
    bool is_in_mem() const { return info_.stream_chunks == 0; }

    /* Writes this matrix as a .mat / .mat_idx pair (what the reference's
     * converters do with SpM_2d_storage::dump and SpM_2d_index::dump). */
    bool dump(const std::string &mat_file, const std::string &idx_file) const
    {
        return detail::check(fmx_spm_dump(spm_.get(), mat_file.c_str(), idx_file.c_str()), "sparse_matrix::dump");
    }

    /* The asymmetric form create(index, storage, t_index, t_storage)
     * (matrix/sparse_matrix.h:527-528). */
    static ptr create(const std::vector<char> &image, const std::vector<int64_t> &offs,
                      const std::vector<char> &t_image, const std::vector<int64_t> &t_offs)
    {
        ptr a = create(image, offs), t = create(t_image, t_offs);
        if (!a || !t) return ptr();
        a->symmetric_ = t->symmetric_ = false;
        a->transpose_ = t;
        return a;
    }

    /* fm::create_2d_matrix (matrix/fm_utils.h:37-39, fm_utils.cpp:726-841)
     * from a device edge list; builds A^T as well unless is_sym. */
    static ptr create_2d_matrix(const uint32_t *src, const uint32_t *dst, size_t m,
                                const block_2d_size &bs, bool is_sym, uint64_t n = 0,
                                uint32_t flags = FMX_BUILD_DEDUP)
    {
        if (!init_flash_matrix()) return ptr();
        fmx_spm *s = nullptr;
        if (!detail::check(fmx_spm_build(detail::global_ctx(), src, dst, m, n, (uint32_t)bs.get_nrow_log(),
                                         (uint32_t)bs.get_ncol_log(), flags, &s),
                           "create_2d_matrix"))
            return ptr();
        ptr a(new sparse_matrix(s, is_sym));
        if (!is_sym) {
            fmx_spm *t = nullptr;
            if (!detail::check(fmx_spm_build(detail::global_ctx(), dst, src, m, a->info_.nrows,
                                             (uint32_t)bs.get_nrow_log(), (uint32_t)bs.get_ncol_log(),
                                             flags, &t),
                               "create_2d_matrix (transpose)"))
                return ptr();
            a->transpose_ = ptr(new sparse_matrix(t, false));
        }
        return a;
    }

    size_t get_num_rows() const { return info_.nrows; }
    size_t get_num_cols() const { return info_.ncols; }
    size_t get_nnz() const { return info_.nnz; }
    bool is_symmetric() const { return symmetric_; }
    size_t get_entry_size() const { return info_.entry_size; }
    block_2d_size get_block_size() const { return block_2d_size(info_.block_h, info_.block_w); }
    const fmx_spm *handle() const { return spm_.get(); }

    /* sparse_matrix::transpose (matrix/sparse_matrix.h:630): a symmetric
     * matrix is its own transpose. */
    ptr transpose() const
    {
        if (symmetric_) return ptr(new sparse_matrix(*this));  // shares the image
        return transpose_;
    }

    /* multiply(in, out, creator) (matrix/sparse_matrix.cpp:709-741): in and
     * out in HBM, each row- or column-major (a column-major input is
     * converted first, as the reference does, :720-736); false on a type or
     * dimension mismatch. */
    bool multiply(const dense_matrix &in, dense_matrix &out, detail::spmm_creator::ptr create) const
    {
        if (in.get_num_rows() != get_num_cols() || in.get_num_cols() != out.get_num_cols() ||
            out.get_num_rows() != get_num_rows()) {
            detail::log_error("the input and output matrix have incompatible dimensions");
            return false;
        }
        if (in.get_type() != out.get_type()) {
            detail::log_error("input and output matrices need to have the same type");
            return false;
        }
        if (!in.is_in_mem() || !out.is_in_mem()) {
            detail::log_error("multiply(in, out) needs operands in HBM");
            return false;
        }
        uint32_t flags = (in.store_layout() == L_COL ? FMX_SPMM_X_COLMAJOR : 0u) |
                         (out.store_layout() == L_COL ? FMX_SPMM_Y_COLMAJOR : 0u);
        fmx_status st;
        if (create) {
            if (create->seq() || matrix_conf().seq_fp_sum) flags |= FMX_SPMM_SEQ;
            st = fmx_spmm_scaled(detail::global_ctx(), spm_.get(), in.get_type(), (int)in.get_num_cols(),
                                 create->pre(), create->post(), in.get_raw_arr(), in.get_ld(), out.get_raw_arr(),
                                 out.get_ld(), flags, nullptr);
        } else {
            if (matrix_conf().seq_fp_sum) flags |= FMX_SPMM_SEQ;
            st = fmx_spmm_ex(detail::global_ctx(), spm_.get(), in.get_type(), (int)in.get_num_cols(),
                             in.get_raw_arr(), in.get_ld(), out.get_raw_arr(), out.get_ld(), flags, nullptr);
        }
        if (!detail::check(st, "sparse_matrix::multiply")) return false;
        return hipDeviceSynchronize() == hipSuccess;
    }

    /* multiply(in, out) (matrix/sparse_matrix.h:636-637, .cpp:751-771): out
     * must be allocated; false on a type or dimension mismatch. */
    bool multiply(const dense_matrix &in, dense_matrix &out) const
    {
        return multiply(in, out, detail::spmm_creator::ptr());
    }

    /* multiply(dense_matrix::ptr, mem_size) (matrix/sparse_matrix.h:643-644,
     * .cpp:773-831): returns a new row-major Y = A * X, or null.  An X in
     * HBM gives a Y in HBM.  An X in host memory (create_host) gives a Y in
     * host memory and goes through HBM in column panels that fit mem_size
     * bytes (the reference's panel branch, :796-830; "The memory is too
     * small for the SpMM" when not even one column fits). */
    dense_matrix::ptr multiply(dense_matrix::ptr right,
                               size_t mem_size = std::numeric_limits<size_t>::max()) const
    {
        if (!right) return dense_matrix::ptr();
        if (!right->is_in_mem()) {
            if (right->get_num_rows() != get_num_cols()) {
                detail::log_error("the matrix size doesn't match");
                return dense_matrix::ptr();
            }
            dense_matrix::ptr out = dense_matrix::create_host(get_num_rows(), right->get_num_cols(),
                                                              right->get_type());
            if (!out ||
                !detail::check(fmx_spmm_host(detail::global_ctx(), spm_.get(), right->get_type(),
                                             (int)right->get_num_cols(), right->get_raw_arr(), right->get_ld(),
                                             out->get_raw_arr(), out->get_ld(), (uint64_t)mem_size, nullptr),
                               "sparse_matrix::multiply"))
                return dense_matrix::ptr();
            return out;
        }
        dense_matrix::ptr out = dense_matrix::create(get_num_rows(), right->get_num_cols(), right->get_type());
        if (!out || !multiply(*right, *out)) return dense_matrix::ptr();
        return out;
    }

};

/* fm::PageRank (matrix/libmatrix-algs/PageRank.cpp:78-137): returns the
 * n x 1 FP32 rank vector, or null.  num_in_mem is kept for the reference's
 * signature; everything is resident in HBM. */
inline dense_matrix::ptr PageRank(sparse_matrix::ptr mat, size_t max_niters, float damping_factor,
                                  size_t num_in_mem = 1)
{
    (void)num_in_mem;
    if (!mat || mat->get_num_rows() != mat->get_num_cols()) return dense_matrix::ptr();
    dense_matrix::ptr pr = dense_matrix::create(mat->get_num_rows(), 1, FMX_F32);
    sparse_matrix::ptr t = mat->transpose();
    int iters = 0;
    if (!pr || !t ||
        !detail::check(fmx_pagerank(detail::global_ctx(), mat->handle(), t->handle(), (int)max_niters,
                                    damping_factor, (float *)pr->get_raw_arr(), &iters, nullptr),
                       "PageRank"))
        return dense_matrix::ptr();
    return pr;
}

namespace alg {
/* fm::alg::NMF (matrix/libmatrix-algs/NMF.cpp:259-316): Lee-Seung updates of
 * a binary matrix from a U[0,1) start (a SplitMix64 stream from `seed`; the
 * reference's own random stream is not reproducible here).  Returns W
 * (n x k) and H transposed (m x k row-major, which is the reference's
 * column-major k x m H byte for byte), or null pointers.  num_in_mem is kept
 * for the reference's signature; everything is resident in HBM (or streamed,
 * for a matrix created from a file factory). */
inline std::pair<dense_matrix::ptr, dense_matrix::ptr> NMF(sparse_matrix::ptr mat, size_t k, size_t max_niters,
                                                           size_t num_in_mem = 0, uint64_t seed = 1)
{
    (void)num_in_mem;
    std::pair<dense_matrix::ptr, dense_matrix::ptr> none;
    if (!mat) return none;
    const size_t n = mat->get_num_rows(), m = mat->get_num_cols();
    std::vector<float> w(n * k), h(m * k);
    uint64_t x = seed;
    auto next = [&x]() {  // SplitMix64 -> U[0,1) with 24 random bits
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (float)((z ^ (z >> 31)) >> 40) * (1.0f / 16777216.0f);
    };
    for (float &v : w) v = next();
    for (float &v : h) v = next();
    dense_matrix::ptr W = dense_matrix::create_from_host(w.data(), n, k, FMX_F32);
    dense_matrix::ptr tH = dense_matrix::create_from_host(h.data(), m, k, FMX_F32);
    sparse_matrix::ptr t = mat->transpose();
    if (!W || !tH || !t ||
        !detail::check(fmx_nmf(detail::global_ctx(), mat->handle(), t->handle(), (int)k, (int)max_niters,
                               (float *)W->get_raw_arr(), (float *)tH->get_raw_arr(), nullptr, nullptr),
                       "NMF"))
        return none;
    return std::make_pair(W, tH);
}
}  // namespace alg

}  // namespace fm

#endif
