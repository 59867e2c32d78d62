// eigen.cpp -- block Krylov eigensolver for the three operators of the
// reference's eigensolver driver (matrix/test/test-eigen.cpp), replacing
// fm::eigen::compute_eigen with Anasazi's BlockKrylovSchur
// (matrix/eigensolver/eigensolver.cpp:118-308):
//   NORMALIZED  N = D^-1/2 A D^-1/2 of a symmetric binary matrix
//               (NA_eigen_Operator, test-eigen.cpp:62-136)
//   PLAIN       A itself, symmetric (eigen_Operator, test-eigen.cpp:32-58)
//   SVD         A^T A, x -> A^T (A x) through the matrix and its transpose
//               (SVD_Operator, test-eigen.cpp:138-177); the eigenvalues are
//               the squared singular values of A, the vectors its right
//               singular vectors
//
// Method: thick-restart block Lanczos with full reorthogonalisation, the
// symmetric form of block Krylov-Schur.  The basis V (n x m, m = blocks *
// block_size) lives in HBM; the projected matrix H = V^T N V (m x m, FP64) is
// assembled from the projection coefficients, so after a restart it simply
// becomes diag(kept Ritz values) and fills again as the next blocks are
// expanded.  Each expansion (N the operator):
//   W = N V_j                              (NORMALIZED: row scaling, then the
//                                           SpMM with the outer scaling
//                                           applied to its sums; SVD: two
//                                           SpMMs through an m_rows x b block)
//   C1 = V^T W, W -= V C1                  (first pass; in the mixed mode a
//                                           BF16 MFMA Gram over a BF16 copy of
//                                           V kept beside it, FP64 otherwise)
//   C2 = V^T W, W -= V C2                  (FP64 reorthogonalisation)
//   H[:, j] = C1 + C2
//   V_{j+1} = W R^-1, R from Cholesky-QR twice (FP64 Gram + Cholesky)
// When the basis is full: Rayleigh-Ritz by FP64 Jacobi on the device, and
// residual estimates ||W s_last|| from the Lanczos relation N V = V H + W E^T.
// Once every wanted estimate is within twice the tolerance: Ritz vectors
// X = V S in FP64 and residuals ||N X - X Theta|| with FP64 SpMMs.
// Converged when every wanted pair has ||N x - theta x|| / |theta| <= tol in
// that FP64 check.
//
// D^-1/2 := 0 for isolated vertices: the reference's 1/sqrt(deg)
// (test-eigen.cpp:69-71) would be inf there and turn their rows into NaN.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

#include "dense.h"
#include "small.h"

namespace fmx {
namespace {

template <typename T> constexpr fmx_dtype dt_of();
template <> constexpr fmx_dtype dt_of<float>() { return FMX_F32; }
template <> constexpr fmx_dtype dt_of<double>() { return FMX_F64; }

// Cholesky-QR of an FP32 block stops after one pass when ||R||_F ||R^-1||_F
// is at most this
constexpr double kOnePassCond = 1e3;

#define TRY(x)                             \
    do {                                   \
        fmx_status st_ = (x);              \
        if (st_ != FMX_OK) return st_;     \
    } while (0)

template <typename T>
struct Dev {
    T *p = nullptr;
    ~Dev() { if (p) (void)hipFree(p); }
};

template <typename T>
class Solver {
public:
    Solver(fmx_ctx *ctx, int op, fmx_spm *spm, fmx_spm *spm_t, const fmx_eigen_opts &o, hipStream_t s)
        : ctx_(ctx), op_(op), spm_(spm), spm_t_(spm_t), o_(o), s_(s),
          n_(op == FMX_EIG_OP_SVD ? spm->ncols : spm->nrows), nmid_(spm->nrows), b_(o.block_size),
          m_(o.block_size * o.num_blocks)
    {
    }

    fmx_status run(const void *X0, int64_t ldx0, fmx_dtype x0_dt, double *evals, double *evecs,
                   int64_t ldv, double *resid, fmx_eigen_stats *stats);

private:
    fmx_status alloc();
    fmx_status apply(const T *X, int64_t ldx, T *Y, int64_t ldy);
    fmx_status apply64(const double *X, double *Y, int k);
    fmx_status project(T *W, int cols, double *Cacc);
    fmx_status cholqr(T *W, int64_t ldw, T *dst, int64_t ldd, bool &ok, unsigned short *dstb = nullptr,
                      bool scaled = false);
    fmx_status orth_block(T *W, int cols, int dst_col);
    fmx_status sync() { FMX_HIP(ctx_, hipStreamSynchronize(s_)); return FMX_OK; }

    // one SpMM of the operator's matrix (or its transpose) on FP32 or FP64
    // data; FP64 data over FP32 values is spmm_creator<double, float>
    fmx_status spmm(const fmx_spm *a, fmx_dtype dt, int k, const void *X, int64_t ldx, void *Y, int64_t ldy,
                    const double *post)
    {
        return spmm_exec(ctx_, a, dt, k, nullptr, post, X, ldx, Y, ldy, 0, dt == FMX_F64, s_);
    }

    fmx_ctx *ctx_;
    int op_;
    fmx_spm *spm_, *spm_t_;
    fmx_eigen_opts o_;
    hipStream_t s_;
    uint64_t n_, nmid_;  // basis rows; rows of A (the SVD's middle product)
    int b_, m_;
    Dev<T> V_, W_, T1_, T2_;
    // mixed mode: a BF16 copy of the basis, written with it, which the
    // first-pass Grams read (half the bytes of the FP32 basis)
    Dev<unsigned short> Vb_;
    bool shadow_ = false;
    // the block whose D^-1/2-scaled copy T1_ holds (the normalised operator's
    // input, written with the block by Cholesky-QR), or null
    const T *scaled_of_ = nullptr;
    Dev<double> dinv_, H_, S_, ev_, C1_, C2_, G_, Rinv_, info_, X64_, Y64_, Z64_, M64_, Ssel_, Scols_, theta_, rn_,
        part_;
    Dev<float> deg_;
    Dev<unsigned char> rot_;  // the Rayleigh-Ritz rotation log
    int spmm_blocks_ = 0, spmm64_blocks_ = 0;
    int history_len_ = 0;
    void record(double r)
    {
        if (o_.history && history_len_ < o_.history_cap) o_.history[history_len_++] = r;
    }
};

template <typename T>
fmx_status Solver<T>::alloc()
{
    const uint64_t n = n_ ? n_ : 1;
    FMX_HIP(ctx_, hipMalloc(&V_.p, n * m_ * sizeof(T)));
    FMX_HIP(ctx_, hipMalloc(&W_.p, n * b_ * sizeof(T)));
    FMX_HIP(ctx_, hipMalloc(&T1_.p, n * b_ * sizeof(T)));
    FMX_HIP(ctx_, hipMalloc(&dinv_.p, n * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&deg_.p, n * sizeof(float)));
    FMX_HIP(ctx_, hipMalloc(&H_.p, (size_t)m_ * m_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&S_.p, (size_t)m_ * m_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&Ssel_.p, (size_t)m_ * m_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&Scols_.p, (size_t)m_ * m_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&ev_.p, (size_t)m_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&C1_.p, (size_t)m_ * b_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&C2_.p, (size_t)m_ * b_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&G_.p, (size_t)b_ * b_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&Rinv_.p, (size_t)b_ * b_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&info_.p, 4 * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&rot_.p, small_sym_eig_ws(m_)));
    FMX_HIP(ctx_, hipMalloc(&X64_.p, n * o_.nev * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&Y64_.p, n * o_.nev * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&Z64_.p, n * o_.nev * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&theta_.p, (size_t)m_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&rn_.p, (size_t)m_ * sizeof(double)));
    FMX_HIP(ctx_, hipMalloc(&part_.p, (size_t)1024 * 256 * sizeof(double)));
    const int keep = std::min(m_ - b_, ((o_.nev + b_ - 1) / b_) * b_ + b_);
    shadow_ = std::is_same<T, float>::value && o_.precision == FMX_EIG_MIXED && b_ % 8 == 0 && m_ % 8 == 0 &&
              m_ <= 128 && keep <= 32;
    if (shadow_) FMX_HIP(ctx_, hipMalloc(&Vb_.p, n * m_ * sizeof(unsigned short)));
    if (op_ == FMX_EIG_OP_SVD) {
        const uint64_t nm = nmid_ ? nmid_ : 1;
        FMX_HIP(ctx_, hipMalloc(&T2_.p, nm * b_ * sizeof(T)));
        FMX_HIP(ctx_, hipMalloc(&M64_.p, nm * o_.nev * sizeof(double)));
    }
    return FMX_OK;
}

// Y = N X for one block: NA_eigen_Operator::run, eigen_Operator::run or
// SVD_Operator::run
template <typename T>
fmx_status Solver<T>::apply(const T *X, int64_t ldx, T *Y, int64_t ldy)
{
    if (op_ == FMX_EIG_OP_NORMALIZED) {
        // D^-1/2 X, then A (D^-1/2 X) with the outer D^-1/2 applied to each
        // row's FP64 sum before its rounding
        if (X != scaled_of_) TRY(row_scale(ctx_, dt_of<T>(), n_, b_, dinv_.p, X, ldx, T1_.p, b_, s_));
        scaled_of_ = X;
        TRY(spmm(spm_, dt_of<T>(), b_, T1_.p, b_, Y, ldy, dinv_.p));
    } else if (op_ == FMX_EIG_OP_PLAIN) {
        TRY(spmm(spm_, dt_of<T>(), b_, X, ldx, Y, ldy, nullptr));
    } else {
        // A X (rows of A), then A^T (A X)
        TRY(spmm(spm_, dt_of<T>(), b_, X, ldx, T2_.p, b_, nullptr));
        TRY(spmm(spm_t_, dt_of<T>(), b_, T2_.p, b_, Y, ldy, nullptr));
        spmm_blocks_++;
    }
    spmm_blocks_++;
    return FMX_OK;
}

// FP64 operator on k <= nev columns (residual check) into Y, through the
// persistent Z64_ (the scaled input) and M64_ (the SVD's A X)
template <typename T>
fmx_status Solver<T>::apply64(const double *X, double *Y, int k)
{
    if (k > o_.nev) return fail(ctx_, FMX_EDIM, "residual block wider than nev");
    if (op_ == FMX_EIG_OP_NORMALIZED) {
        TRY(row_scale(ctx_, FMX_F64, n_, k, dinv_.p, X, k, Z64_.p, k, s_));
        TRY(spmm(spm_, FMX_F64, k, Z64_.p, k, Y, k, dinv_.p));
    } else if (op_ == FMX_EIG_OP_PLAIN) {
        TRY(spmm(spm_, FMX_F64, k, X, k, Y, k, nullptr));
    } else {
        TRY(spmm(spm_, FMX_F64, k, X, k, M64_.p, k, nullptr));
        TRY(spmm(spm_t_, FMX_F64, k, M64_.p, k, Y, k, nullptr));
        spmm64_blocks_++;
    }
    spmm64_blocks_++;
    return FMX_OK;
}

// W -= V[:, :cols] (V^T W) twice; Cacc = total coefficients (cols x b)
template <typename T>
fmx_status Solver<T>::project(T *W, int cols, double *Cacc)
{
    const int first_mode = (std::is_same<T, float>::value && o_.precision == FMX_EIG_MIXED)
                               ? FMX_GRAM_BF16 : FMX_GRAM_F64;
    if (shadow_)
        TRY(gram_bf16_shadow(ctx_, n_, cols, b_, Vb_.p, m_, (const float *)W, b_, Cacc, s_));
    else
        TRY(gram(ctx_, dt_of<T>(), first_mode, n_, cols, b_, V_.p, m_, W, b_, Cacc, s_));
    TRY(gemm_tall(ctx_, dt_of<T>(), n_, cols, b_, -1.0, V_.p, m_, Cacc, 1.0, W, b_, s_));
    TRY(gram(ctx_, dt_of<T>(), FMX_GRAM_F64, n_, cols, b_, V_.p, m_, W, b_, C2_.p, s_));
    TRY(gemm_tall(ctx_, dt_of<T>(), n_, cols, b_, -1.0, V_.p, m_, C2_.p, 1.0, W, b_, s_));
    return FMX_OK;
}

// dst = orthonormal basis of W (Cholesky-QR applied twice, FP64)
template <typename T>
fmx_status Solver<T>::cholqr(T *W, int64_t ldw, T *dst, int64_t ldd, bool &ok, unsigned short *dstb, bool scaled)
{
    ok = true;
    double info[2];
    // the first pass works in place in W (contiguous rows), the second
    // writes the basis block (every b-th stretch of V's rows) once
    for (int pass = 0; pass < 2; pass++) {
        T *src = W;
        const int64_t lds = ldw;
        TRY(gram(ctx_, dt_of<T>(), FMX_GRAM_F64, n_, b_, b_, src, lds, src, lds, G_.p, s_));
        TRY(small_chol_inv(ctx_, G_.p, b_, Rinv_.p, info_.p, s_));
        FMX_HIP(ctx_, hipMemcpyAsync(info, info_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s_));
        TRY(sync());
        if (info[0] != 0.0) { ok = false; return FMX_OK; }
        // One pass suffices for a well-conditioned FP32 block: with the Gram
        // summed in FP64, W R^-1 is orthonormal to about kappa^2 * 1e-16,
        // below FP32 rounding while kappa <= info[1] <= 1e3 (the second pass
        // only repairs what the first lost to conditioning).  An FP64 basis
        // always takes both.
        const bool last = pass == 1 || (sizeof(T) == 4 && info[1] <= kOnePassCond);
        if (last && (dstb || scaled))
            TRY(gemm_tall_copies(ctx_, dt_of<T>(), n_, b_, b_, 1.0, src, lds, Rinv_.p, 0.0, dst, ldd, dstb, m_,
                                 scaled ? dinv_.p : nullptr, scaled ? T1_.p : nullptr, b_, s_));
        else if (last)
            TRY(gemm_tall(ctx_, dt_of<T>(), n_, b_, b_, 1.0, src, lds, Rinv_.p, 0.0, dst, ldd, s_));
        else
            TRY(gemm_tall(ctx_, dt_of<T>(), n_, b_, b_, 1.0, W, ldw, Rinv_.p, 0.0, W, ldw, s_));
        if (last) break;
    }
    return FMX_OK;
}

// Orthonormalise W (already projected against V[:, :cols]) into V[:, dst_col:]
// replacing a rank-deficient block by a fresh random one.
template <typename T>
fmx_status Solver<T>::orth_block(T *W, int cols, int dst_col)
{
    bool ok = false;
    for (int attempt = 0; attempt < 4 && !ok; attempt++) {
        if (attempt > 0) {
            TRY(rand_uniform(ctx_, dt_of<T>(), W, b_, n_, b_, o_.seed + 7919ull * (attempt + dst_col), s_));
            if (cols) TRY(project(W, cols, C1_.p));
        }
        const bool scaled = op_ == FMX_EIG_OP_NORMALIZED;
        TRY(cholqr(W, b_, V_.p + dst_col, m_, ok, shadow_ ? Vb_.p + dst_col : nullptr, scaled));
        scaled_of_ = (ok && scaled) ? V_.p + dst_col : nullptr;
    }
    if (!ok) return fail(ctx_, FMX_EINVAL, "Krylov basis breakdown: operator rank too small");
    return FMX_OK;
}

template <typename T>
fmx_status Solver<T>::run(const void *X0, int64_t ldx0, fmx_dtype x0_dt, double *evals,
                          double *evecs, int64_t ldv, double *resid, fmx_eigen_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    TRY(alloc());
    const int nev = o_.nev, b = b_, m = m_;
    // degrees and D^-1/2 (NA_eigen_Operator ctor, test-eigen.cpp:93-110)
    if (op_ == FMX_EIG_OP_NORMALIZED) {
        Dev<float> ones;
        FMX_HIP(ctx_, hipMalloc(&ones.p, (n_ ? n_ : 1) * sizeof(float)));
        std::vector<float> h1(n_, 1.0f);
        FMX_HIP(ctx_, hipMemcpyAsync(ones.p, h1.data(), n_ * sizeof(float), hipMemcpyHostToDevice, s_));
        TRY(fmx_spmm(ctx_, spm_, FMX_F32, 1, ones.p, 1, deg_.p, 1, s_));
        TRY(degree_inv_sqrt(ctx_, deg_.p, n_, dinv_.p, s_));
        TRY(sync());
    }
    // starting block: caller's warm start or uniform [-1, 1)
    if (X0) {
        if (x0_dt == dt_of<T>()) {
            FMX_HIP(ctx_, hipMemcpy2DAsync(W_.p, b * sizeof(T), X0, ldx0 * sizeof(T), b * sizeof(T), n_,
                                           hipMemcpyDeviceToDevice, s_));
        } else if (x0_dt == FMX_F32 && std::is_same<T, double>::value) {
            TRY(cast_f32_f64(ctx_, (const float *)X0, ldx0, n_, b, (double *)W_.p, b, s_));
        } else if (x0_dt == FMX_F64 && std::is_same<T, float>::value) {
            TRY(cast_f64_f32(ctx_, (const double *)X0, ldx0, n_, b, (float *)W_.p, b, s_));
        } else {
            return fail(ctx_, FMX_ETYPE, "warm start must be FP32 or FP64");
        }
    } else {
        TRY(rand_uniform(ctx_, dt_of<T>(), W_.p, b, n_, b, o_.seed, s_));
    }
    TRY(orth_block(W_.p, 0, 0));
    FMX_HIP(ctx_, hipMemsetAsync(H_.p, 0, (size_t)m * m * sizeof(double), s_));

    int jstart = 0;
    int restarts = 0;
    bool converged = false;
    std::vector<double> hev(m), hS((size_t)m * m), hsel((size_t)m * m), hres(m), hG((size_t)b * b);
    std::vector<int> order(m);
    const int keep = std::min(m - b, ((nev + b - 1) / b) * b + b);
    double worst = INFINITY;
    for (;;) {
        for (int j = jstart; j < o_.num_blocks; j++) {
            const int cols = (j + 1) * b;
            TRY(apply(V_.p + (size_t)j * b, m, W_.p, b));
            TRY(project(W_.p, cols, C1_.p));
            TRY(h_update(ctx_, H_.p, m, cols, j * b, b, C1_.p, C2_.p, s_));
            if (j + 1 < o_.num_blocks) TRY(orth_block(W_.p, cols, (j + 1) * b));
        }
        // Rayleigh-Ritz; with it the Gram of the last residual block W, for
        // the residual estimates of the Lanczos relation N V = V H + W E^T
        TRY(small_sym_eig(ctx_, H_.p, m, ev_.p, S_.p, info_.p, rot_.p, s_));
        TRY(gram(ctx_, dt_of<T>(), FMX_GRAM_F64, n_, b, b, W_.p, b, W_.p, b, G_.p, s_));
        FMX_HIP(ctx_, hipMemcpyAsync(hev.data(), ev_.p, m * sizeof(double), hipMemcpyDeviceToHost, s_));
        FMX_HIP(ctx_, hipMemcpyAsync(hS.data(), S_.p, (size_t)m * m * sizeof(double), hipMemcpyDeviceToHost, s_));
        FMX_HIP(ctx_, hipMemcpyAsync(hG.data(), G_.p, (size_t)b * b * sizeof(double), hipMemcpyDeviceToHost, s_));
        TRY(sync());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return hev[a] > hev[c]; });
        for (int i = 0; i < m; i++)
            for (int c = 0; c < m; c++) hsel[(size_t)i * m + c] = hS[(size_t)i * m + order[c]];
        std::vector<double> th(m);
        for (int c = 0; c < m; c++) th[c] = hev[order[c]];
        FMX_HIP(ctx_, hipMemcpyAsync(Ssel_.p, hsel.data(), (size_t)m * m * sizeof(double),
                                     hipMemcpyHostToDevice, s_));
        FMX_HIP(ctx_, hipMemcpyAsync(theta_.p, th.data(), m * sizeof(double), hipMemcpyHostToDevice, s_));
        // Estimated residual of Ritz pair i (y = V s_i): ||N y - theta y|| =
        // ||W s_i,last|| with s_i,last the coordinates of s_i on the last
        // block.  The FP64 check below (Ritz vectors, an FP64 SpMM) runs only
        // once every wanted estimate is within twice the tolerance, and at
        // the restart limit; it alone decides convergence.
        double est_worst = 0.0;
        for (int i = 0; i < nev; i++) {
            double q = 0.0;
            for (int r = 0; r < b; r++)
                for (int c = 0; c < b; c++)
                    q += hsel[(size_t)(m - b + r) * m + i] * hG[(size_t)r * b + c] * hsel[(size_t)(m - b + c) * m + i];
            est_worst = std::max(est_worst, std::sqrt(std::max(q, 0.0)) / std::max(std::fabs(th[i]), 1e-300));
        }
        if (est_worst > 2.0 * o_.tol && restarts < o_.max_restarts) {
            worst = est_worst;
            record(est_worst);
            if (o_.verbose)
                fprintf(stderr, "[fmx eigen] restart %d: spmm blocks %d, estimated max rel resid %.3e\n",
                        restarts, spmm_blocks_, est_worst);
        } else {
            // Ritz vectors (FP64) of the wanted pairs and their FP64 residuals;
            // Ssel is m x m, the first nev columns are wanted: copy them to m x nev
            {
                double *Sw = Scols_.p;  // m x nev, stream-ordered reuse
                FMX_HIP(ctx_, hipMemcpy2DAsync(Sw, nev * sizeof(double), Ssel_.p, m * sizeof(double),
                                               nev * sizeof(double), m, hipMemcpyDeviceToDevice, s_));
                if (std::is_same<T, float>::value)
                    TRY(gemm_tall_f32_f64(ctx_, n_, m, nev, 1.0, (const float *)V_.p, m, Sw, 0.0, X64_.p, nev, s_));
                else
                    TRY(gemm_tall(ctx_, FMX_F64, n_, m, nev, 1.0, V_.p, m, Sw, 0.0, X64_.p, nev, s_));
                TRY(apply64(X64_.p, Y64_.p, nev));
            }
            TRY(resid_norms(ctx_, Y64_.p, nev, X64_.p, nev, theta_.p, n_, nev, rn_.p, part_.p, 1024, s_));
            FMX_HIP(ctx_, hipMemcpyAsync(hres.data(), rn_.p, nev * sizeof(double), hipMemcpyDeviceToHost, s_));
            TRY(sync());
            worst = 0.0;
            for (int i = 0; i < nev; i++) {
                hres[i] /= std::max(std::fabs(th[i]), 1e-300);
                worst = std::max(worst, hres[i]);
            }
            record(worst);
            if (o_.verbose)
                fprintf(stderr, "[fmx eigen] restart %d: spmm blocks %d, theta[0]=%.12f theta[nev-1]=%.12f "
                                "max rel resid %.3e\n", restarts, spmm_blocks_, th[0], th[nev - 1], worst);
            if (worst <= o_.tol) {
                converged = true;
                for (int i = 0; i < nev; i++) {
                    evals[i] = th[i];
                    if (resid) resid[i] = hres[i];
                }
                break;
            }
            if (restarts >= o_.max_restarts) {
                for (int i = 0; i < nev; i++) {
                    evals[i] = th[i];
                    if (resid) resid[i] = hres[i];
                }
                break;
            }
        }
        // thick restart: V[:, :keep] = V S[:, :keep], H = diag(theta), then the
        // last residual block W (orthogonal to V) becomes block keep/b
        {
            double *Sk = Scols_.p;  // m x keep, stream-ordered after the Ritz-vector product
            FMX_HIP(ctx_, hipMemcpy2DAsync(Sk, keep * sizeof(double), Ssel_.p, m * sizeof(double),
                                           keep * sizeof(double), m, hipMemcpyDeviceToDevice, s_));
            scaled_of_ = nullptr;  // the kept blocks are rotated
            if (shadow_)
                TRY(gemm_tall_copies(ctx_, FMX_F32, n_, m, keep, 1.0, V_.p, m, Sk, 0.0, V_.p, m, Vb_.p, m, nullptr,
                                     nullptr, 0, s_));
            else
                TRY(gemm_tall(ctx_, dt_of<T>(), n_, m, keep, 1.0, V_.p, m, Sk, 0.0, V_.p, m, s_));
            std::vector<double> hH((size_t)m * m, 0.0);
            for (int i = 0; i < keep; i++) hH[(size_t)i * m + i] = th[i];
            FMX_HIP(ctx_, hipMemcpyAsync(H_.p, hH.data(), (size_t)m * m * sizeof(double),
                                         hipMemcpyHostToDevice, s_));
            TRY(sync());
        }
        TRY(orth_block(W_.p, keep, keep));
        jstart = keep / b;
        restarts++;
    }
    if (evecs) {
        FMX_HIP(ctx_, hipMemcpy2DAsync(evecs, ldv * sizeof(double), X64_.p, nev * sizeof(double),
                                       nev * sizeof(double), n_, hipMemcpyDeviceToDevice, s_));
        TRY(sync());
    }
    if (stats) {
        stats->spmm_blocks = spmm_blocks_;
        stats->fp64_spmm_blocks = spmm64_blocks_;
        stats->restarts = restarts;
        stats->converged = converged ? 1 : 0;
        stats->max_resid = worst;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        stats->history_len = history_len_;
    }
    return FMX_OK;
}

}  // namespace
}  // namespace fmx

using namespace fmx;

extern "C" fmx_status fmx_eigen_opts_init(fmx_eigen_opts *o, int nev)
{
    if (!o || nev <= 0) return FMX_EINVAL;
    o->nev = nev;
    o->block_size = 16;
    o->num_blocks = 6;
    o->max_restarts = 100;
    o->tol = 1e-6;
    o->precision = FMX_EIG_FP32;
    o->seed = 1;
    o->verbose = 0;
    o->history = nullptr;
    o->history_cap = 0;
    return FMX_OK;
}

namespace {

fmx_status eigen_run(fmx_ctx *ctx, int op, const fmx_spm *ca, const fmx_spm *cat, const fmx_eigen_opts *o,
                     const void *X0, int64_t ldx0, fmx_dtype x0_dtype, double *evals, double *evecs, int64_t ldv,
                     double *resid, fmx_eigen_stats *stats)
{
    if (!ctx || !ca || !o || !evals) return fail(ctx, FMX_EINVAL, "null argument");
    fmx_spm *a = const_cast<fmx_spm *>(ca), *at = const_cast<fmx_spm *>(cat);
    if (a->ctx != ctx || (at && at->ctx != ctx)) return fail(ctx, FMX_EINVAL, "matrix belongs to another context");
    if (op == FMX_EIG_OP_SVD) {
        // a symmetric matrix is its own transpose (sparse_matrix::transpose)
        if (!at && a->nrows != a->ncols)
            return fail(ctx, FMX_EINVAL, "the SVD operator of a non-square matrix needs its transpose");
        if (!at) at = a;
        if (at->nrows != a->ncols || at->ncols != a->nrows)
            return fail(ctx, FMX_EDIM, "the transpose's shape does not match the matrix");
    } else if (op == FMX_EIG_OP_NORMALIZED || op == FMX_EIG_OP_PLAIN) {
        if (at) return fail(ctx, FMX_EINVAL, "only the SVD operator takes a transpose");
        if (a->nrows != a->ncols) return fail(ctx, FMX_EDIM, "the operator needs a square matrix");
    } else {
        return fail(ctx, FMX_EINVAL, "unknown eigen operator");
    }
    if (o->nev <= 0 || o->block_size <= 0 || o->block_size > 32 || o->num_blocks < 2)
        return fail(ctx, FMX_EINVAL, "bad eigen options");
    const int m = o->block_size * o->num_blocks;
    if (m > 128 || (m & 1)) return fail(ctx, FMX_EINVAL, "block_size * num_blocks must be even and <= 128");
    const int keep = ((o->nev + o->block_size - 1) / o->block_size) * o->block_size + o->block_size;
    if (o->nev > o->block_size * (o->num_blocks - 2) || keep > m - o->block_size)
        return fail(ctx, FMX_EINVAL, "nev too large for the basis");
    if (o->nev > 256) return fail(ctx, FMX_EINVAL, "nev <= 256");
    if (evecs && ldv < o->nev) return fail(ctx, FMX_EDIM, "ldv < nev");
    if (X0 && ldx0 < o->block_size) return fail(ctx, FMX_EDIM, "ldx0 < block_size");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    FMX_HIP(ctx, hipDeviceSynchronize());
    hipStream_t s = ctx->stream;
    if (o->precision == FMX_EIG_FP64) {
        Solver<double> sv(ctx, op, a, at, *o, s);
        return sv.run(X0, ldx0, x0_dtype, evals, evecs, ldv, resid, stats);
    }
    if (o->precision != FMX_EIG_MIXED && o->precision != FMX_EIG_FP32)
        return fail(ctx, FMX_EINVAL, "unknown precision");
    Solver<float> sv(ctx, op, a, at, *o, s);
    return sv.run(X0, ldx0, x0_dtype, evals, evecs, ldv, resid, stats);
}

}  // namespace

extern "C" fmx_status fmx_eigen(fmx_ctx *ctx, int op, const fmx_spm *A, const fmx_spm *At,
                                const fmx_eigen_opts *o, const void *X0, int64_t ldx0, fmx_dtype x0_dtype,
                                double *evals, double *evecs, int64_t ldv, double *resid, fmx_eigen_stats *stats)
{
    return eigen_run(ctx, op, A, At, o, X0, ldx0, x0_dtype, evals, evecs, ldv, resid, stats);
}

extern "C" fmx_status fmx_eigen_normalized(fmx_ctx *ctx, const fmx_spm *cspm, const fmx_eigen_opts *o,
                                           const void *X0, int64_t ldx0, fmx_dtype x0_dtype,
                                           double *evals, double *evecs, int64_t ldv, double *resid,
                                           fmx_eigen_stats *stats)
{
    return eigen_run(ctx, FMX_EIG_OP_NORMALIZED, cspm, nullptr, o, X0, ldx0, x0_dtype, evals, evecs, ldv, resid,
                     stats);
}
