/* fm::eigen -- the reference's eigensolver interface
 * (matrix/eigensolver/eigensolver.h:29-88) over the native block Krylov
 * solver of libfmx (fmx_eigen).
 *
 * compute_eigen(func, sym, opts) takes the operators the reference's driver
 * defines (matrix/test/test-eigen.cpp): eigen_Operator (A), NA_eigen_Operator
 * (D^-1/2 A D^-1/2) and SVD_Operator (A^T A).  They are declared here as
 * spm_function subclasses that carry their matrix; the solver recognises
 * them and runs the whole solve on the GPU.  run() applies the operator to a
 * device FP64 block, as the reference's run() does, so an operator can also
 * be used on its own.  A user-defined spm_function is not accepted by the
 * native solver (compute_eigen returns an empty result and logs why).
 *
 * Differences from the reference, by design:
 *   - which is "LA" (largest algebraic); the reference defaults to "LM";
 *   - the solver is the block Krylov-Schur form for symmetric operators
 *     (thick-restart block Lanczos); sym must be true;
 *   - eigenvectors come back as a device FP64 n x nev row-major matrix;
 *   - opts.precision selects FP32 (default), mixed (BF16 first-pass Grams)
 *     or FP64 arithmetic; the residual check is always FP64. */
#ifndef FM_EIGENSOLVER_H
#define FM_EIGENSOLVER_H

#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "sparse_matrix.h"

namespace fm {
namespace eigen {

/* spm_function (eigensolver.h:36-47). */
class spm_function {
public:
    virtual ~spm_function() {}
    virtual dense_matrix::ptr run(dense_matrix::ptr &x) const = 0;
    virtual size_t get_num_cols() const = 0;
    virtual size_t get_num_rows() const = 0;
};

namespace detail {
/* The operators the native solver runs itself. */
class native_operator : public spm_function {
public:
    native_operator(int op, sparse_matrix::ptr mat) : op_(op), mat_(mat) {}
    int op() const { return op_; }
    const sparse_matrix::ptr &mat() const { return mat_; }
    size_t get_num_cols() const override { return mat_ ? mat_->get_num_cols() : 0; }
    size_t get_num_rows() const override { return mat_ ? mat_->get_num_rows() : 0; }

protected:
    /* y = A x on an FP64 block (spmm_creator<double, float>) */
    static dense_matrix::ptr spmm64(const sparse_matrix &a, const dense_matrix &x, const double *pre,
                                    const double *post)
    {
        dense_matrix::ptr y = dense_matrix::create(a.get_num_rows(), x.get_num_cols(), FMX_F64);
        if (!y) return y;
        if (!a.multiply(x, *y, fm::detail::spmm_creator::create(pre, post))) return dense_matrix::ptr();
        return y;
    }
    int op_;
    sparse_matrix::ptr mat_;
};
}  // namespace detail

/* eigen_Operator (matrix/test/test-eigen.cpp:32-58): y = A x, A symmetric. */
class eigen_Operator : public detail::native_operator {
public:
    explicit eigen_Operator(sparse_matrix::ptr mat) : native_operator(FMX_EIG_OP_PLAIN, mat) {}
    dense_matrix::ptr run(dense_matrix::ptr &x) const override
    {
        if (!x || !mat_) return dense_matrix::ptr();
        dense_matrix::ptr y = spmm64(*mat_, *x, nullptr, nullptr);
        x.reset();
        return y;
    }
};

/* NA_eigen_Operator (test-eigen.cpp:62-136): y = D^-1/2 A D^-1/2 x with the
 * degrees of A (D^-1/2 := 0 for isolated vertices). */
class NA_eigen_Operator : public detail::native_operator {
public:
    explicit NA_eigen_Operator(sparse_matrix::ptr mat) : native_operator(FMX_EIG_OP_NORMALIZED, mat)
    {
        if (!mat_) return;
        const size_t n = mat_->get_num_rows();
        dense_matrix::ptr ones = dense_matrix::create(mat_->get_num_cols(), 1, FMX_F64);
        dinv_ = dense_matrix::create(n, 1, FMX_F64);
        if (!ones || !dinv_) return;
        std::vector<double> h1(mat_->get_num_cols(), 1.0);
        if (hipMemcpy(ones->get_raw_arr(), h1.data(), h1.size() * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess)
            return;
        dense_matrix::ptr deg = spmm64(*mat_, *ones, nullptr, nullptr);
        if (!deg) return;
        std::vector<double> d(n);
        if (!deg->copy_to_host(d.data())) return;
        for (double &v : d) v = v > 0 ? 1.0 / std::sqrt(v) : 0.0;
        (void)hipMemcpy(dinv_->get_raw_arr(), d.data(), n * sizeof(double), hipMemcpyHostToDevice);
    }
    dense_matrix::ptr run(dense_matrix::ptr &x) const override
    {
        if (!x || !mat_ || !dinv_) return dense_matrix::ptr();
        const double *d = (const double *)dinv_->get_raw_arr();
        dense_matrix::ptr y = spmm64(*mat_, *x, d, d);
        x.reset();
        return y;
    }

private:
    dense_matrix::ptr dinv_;
};

/* SVD_Operator (test-eigen.cpp:138-177): y = A^T (A x), through the
 * matrix's transpose (sparse_matrix::transpose). */
class SVD_Operator : public detail::native_operator {
public:
    explicit SVD_Operator(sparse_matrix::ptr mat)
        : native_operator(FMX_EIG_OP_SVD, mat), t_mat_(mat ? mat->transpose() : sparse_matrix::ptr())
    {
    }
    const sparse_matrix::ptr &t_mat() const { return t_mat_; }
    dense_matrix::ptr run(dense_matrix::ptr &x) const override
    {
        if (!x || !mat_ || !t_mat_) return dense_matrix::ptr();
        dense_matrix::ptr tmp = spmm64(*mat_, *x, nullptr, nullptr);
        x.reset();
        if (!tmp) return tmp;
        return spmm64(*t_mat_, *tmp, nullptr, nullptr);
    }
    size_t get_num_rows() const override { return mat_ ? mat_->get_num_cols() : 0; }

private:
    sparse_matrix::ptr t_mat_;
};

/* eigen_res (eigensolver.h:49-56). */
struct eigen_res {
    std::vector<double> vals;
    dense_matrix::ptr vecs;
    struct eigen_status {
        size_t num_iters;  /* restarts */
        size_t num_ops;    /* block SpMMs */
    } status = {0, 0};
    std::vector<double> resid;  /* FP64 ||N x - theta x|| / |theta| per pair */
    bool converged = false;
};

/* eigen_options (eigensolver.h:58-73; defaults as eigensolver.cpp:78-96
 * where they apply). */
struct eigen_options {
    double tol = 1e-8;
    int num_blocks = 6;
    int max_restarts = 100;
    int max_iters = 0;  /* kept for the reference's signature */
    int block_size = 16;
    int nev = 1;
    std::string solver = "KrylovSchur";
    std::string which = "LA";
    bool in_mem = true;
    int precision = FMX_EIG_FP32;
    uint64_t seed = 1;

    /* eigen_options::init (eigensolver.cpp:92-116): only the block
     * Krylov-Schur solver exists here, and it starts from blocks of 16
     * (the reference's KrylovSchur uses single vectors, 2 nev of them):
     * enough blocks to keep nev Ritz vectors and one more block at a
     * restart, and to expand by one; false if that exceeds 128 columns. */
    bool init(int nev_, std::string solver_ = "KrylovSchur")
    {
        if (nev_ <= 0) return false;
        if (!solver_.empty() && solver_ != "KrylovSchur") return false;
        nev = nev_;
        solver = "KrylovSchur";
        block_size = 16;
        num_blocks = 6;
        const int kept = (nev + block_size - 1) / block_size * block_size;
        while (nev > block_size * (num_blocks - 2) || kept + 2 * block_size > block_size * num_blocks) num_blocks++;
        return block_size * num_blocks <= 128;
    }
};

/* compute_eigen (eigensolver.h:78-82): `func' is destroyed by this function,
 * as in the reference. */
inline eigen_res compute_eigen(spm_function *func, bool sym, eigen_options &opts, bool verbose = false)
{
    eigen_res res;
    std::unique_ptr<spm_function> own(func);
    const detail::native_operator *op = dynamic_cast<const detail::native_operator *>(func);
    if (!op || !op->mat()) {
        fm::detail::log_error("compute_eigen: only eigen_Operator, NA_eigen_Operator and SVD_Operator run natively");
        return res;
    }
    if (!sym) {
        fm::detail::log_error("compute_eigen: the native solver needs a symmetric operator");
        return res;
    }
    if (opts.which != "LA") {
        fm::detail::log_error("compute_eigen: only which = \"LA\" is supported");
        return res;
    }
    fmx_eigen_opts o;
    if (!fm::detail::check(fmx_eigen_opts_init(&o, opts.nev), "compute_eigen")) return res;
    o.block_size = opts.block_size;
    o.num_blocks = opts.num_blocks;
    o.max_restarts = opts.max_restarts;
    o.tol = opts.tol;
    o.precision = opts.precision;
    o.seed = opts.seed;
    o.verbose = verbose ? 1 : 0;
    const fmx_spm *at = nullptr;
    if (op->op() == FMX_EIG_OP_SVD) {
        const SVD_Operator *svd = static_cast<const SVD_Operator *>(op);
        if (!svd->t_mat()) {
            fm::detail::log_error("compute_eigen: the SVD operator's matrix has no transpose");
            return res;
        }
        if (!op->mat()->is_symmetric()) at = svd->t_mat()->handle();
    }
    const size_t n = op->get_num_cols();
    res.vecs = dense_matrix::create(n, opts.nev, FMX_F64);
    if (!res.vecs) return res;
    res.vals.assign(opts.nev, 0.0);
    res.resid.assign(opts.nev, 0.0);
    fmx_eigen_stats st;
    if (!fm::detail::check(fmx_eigen(fm::detail::global_ctx(), op->op(), op->mat()->handle(), at, &o, nullptr, 0,
                                     FMX_F64, res.vals.data(), (double *)res.vecs->get_raw_arr(), opts.nev,
                                     res.resid.data(), &st),
                           "compute_eigen")) {
        res.vals.clear();
        res.resid.clear();
        res.vecs.reset();
        return res;
    }
    res.status.num_iters = (size_t)st.restarts;
    res.status.num_ops = (size_t)st.spmm_blocks;
    res.converged = st.converged != 0;
    return res;
}

}  // namespace eigen
}  // namespace fm

#endif
