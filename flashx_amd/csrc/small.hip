// small.hip -- small dense FP64 kernels of the block eigensolver, each run by
// one workgroup: the symmetric Rayleigh-Ritz eigenproblem (parallel cyclic
// Jacobi), Cholesky with triangular inverse for Cholesky-QR, the projected
// matrix update, residual norms and precision casts.
//
// In the reference these steps run inside Trilinos/Anasazi on the CPU
// (LAPACK through Teuchos, matrix/eigensolver/eigensolver.cpp:215-235).
#include "small.h"

namespace fmx {
namespace {

constexpr int kJT = 1024;  // threads of the Jacobi workgroup (16 waves; 8 were 25% slower)
constexpr int kJB = (64 * 65 / 2 + kJT - 1) / kJT;  // (pair, pair) blocks per thread at m = 128

// Cyclic Jacobi for a symmetric m x m matrix (m even, m <= 128), the upper
// triangle of A in LDS (entry (i, j), i <= j, at i * (m + 1) + j).  Each
// round rotates m/2 disjoint pairs (round-robin ordering: index 0 stays, the
// others move one place per round), rows first then columns, so
// A' = P^T A P exactly as the sequential method would.  Round t's rotations
// go to rot[t * m/2 + pair] as (c, s): the eigenvector matrix is their
// product, which jacobi_vectors_kernel forms row by row on other CUs
// afterwards (rows of V are independent), so this single workgroup, the
// serial part, carries only A.  info[0] = sweeps that rotated (the rounds
// logged are info[0] * (m - 1)), info[1] the final relative off-diagonal
// norm.
__global__ void __launch_bounds__(kJT) jacobi_kernel(const double *Ain, int m, double *evals, double2 *rot,
                                                     int max_sweeps, double *info)
{
    extern __shared__ double A[];        // m x (m + 1), upper triangle used
    __shared__ double cs[64], sn[64];
    __shared__ int pp[64], qq[64];
    __shared__ double red[2][kJT / 64];
    const int tid = threadIdx.x;
    const int S = m + 1;                 // LDS row stride
    for (int e = tid; e < m * m; e += kJT) {
        const int r = e / m, c = e - r * m;
        if (c >= r) A[r * S + c] = 0.5 * (Ain[e] + Ain[c * m + r]);  // symmetrise
    }
    __syncthreads();
    const int half = m / 2;
    // this thread's (pair k, pair l) blocks, k <= l, the same every round:
    // element tid + j * kJT of the half (half + 1) / 2 triangle (at most kJB
    // with m <= 128); block (l, k) holds the mirrored entries of (k, l)
    int bk[kJB], bl[kJB];
#pragma unroll
    for (int j = 0; j < kJB; j++) {
        int e = tid + j * kJT, k = 0;
        while (k < half && e >= half - k) {
            e -= half - k;
            k++;
        }
        bk[j] = k < half ? k : -1;
        bl[j] = k + e;
    }
    int sweep = 0;
    double off = 0.0, tot = 0.0;
    for (; sweep < max_sweeps; sweep++) {
        // convergence test: off-diagonal Frobenius norm vs whole norm
        double so = 0.0, st = 0.0;
        for (int e = tid; e < m * m; e += kJT) {
            const int r = e / m, c = e - r * m;
            if (c < r) continue;
            const double v = A[r * S + c] * A[r * S + c];
            st += r == c ? v : 2.0 * v;
            if (r != c) so += 2.0 * v;
        }
        for (int o = 32; o; o >>= 1) {
            so += __shfl_xor(so, o);
            st += __shfl_xor(st, o);
        }
        if ((tid & 63) == 0) {
            red[0][tid >> 6] = so;
            red[1][tid >> 6] = st;
        }
        __syncthreads();
        off = 0.0;
        tot = 0.0;
        for (int w = 0; w < kJT / 64; w++) {  // every thread sums in the same order
            off += red[0][w];
            tot += red[1][w];
        }
        __syncthreads();
        // off-diagonal norm below 1e-12 of the whole: a Ritz pair's residual
        // then moves by at most about that fraction of ||H|| (the product of
        // rotations stays orthogonal to rounding whatever the threshold)
        if (off <= 1e-24 * tot || off == 0.0) break;
        // pair k of round r: (0 or 1 + (r + k - 1) mod (m - 1), 1 + (r + m - 2 - k) mod (m - 1)),
        // both residues stepped by one per round
        int ia = tid - 1, ib = m - 2 - tid;
        for (int r = 0; r < m - 1; r++) {
            if (tid < half) {
                const int k = tid;
                int a = (k == 0) ? 0 : ia + 1;
                int b = ib + 1;
                if (a > b) { int t = a; a = b; b = t; }
                pp[k] = a;
                qq[k] = b;
                const double apq = A[a * S + b];
                double c = 1.0, s = 0.0;
                if (apq != 0.0) {
                    const double theta = (A[b * S + b] - A[a * S + a]) / (2.0 * apq);
                    const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    c = 1.0 / sqrt(t * t + 1.0);
                    s = t * c;
                }
                cs[k] = c;
                sn[k] = s;
                rot[((size_t)sweep * (m - 1) + r) * half + k] = make_double2(c, s);
                ia = ia + 1 == m - 1 ? 0 : ia + 1;
                ib = ib + 1 == m - 1 ? 0 : ib + 1;
            }
            __syncthreads();
            // A <- G A G^T, one 2x2 block (rows of pair k, columns of pair l)
            // per element of the triangle: the blocks are disjoint, so the
            // row and the column rotation of a block need no barrier between
            // them (the same operations as rotating all rows, then all
            // columns), and every load of the round may precede every store.
            // In a diagonal block (k = l) entries (p, q) and (q, p) share one
            // slot, which takes the (q, p) result.
            int o00[kJB], o01[kJB], o10[kJB], o11[kJB];
            double x00[kJB], x01[kJB], x10[kJB], x11[kJB];
#pragma unroll
            for (int j = 0; j < kJB; j++) {
                if (bk[j] < 0) continue;
                const int p = pp[bk[j]], q = qq[bk[j]], u = pp[bl[j]], v = qq[bl[j]];
                o00[j] = p < u ? p * S + u : u * S + p;
                o01[j] = p < v ? p * S + v : v * S + p;
                o10[j] = q < u ? q * S + u : u * S + q;
                o11[j] = q < v ? q * S + v : v * S + q;
                x00[j] = A[o00[j]];
                x01[j] = A[o01[j]];
                x10[j] = A[o10[j]];
                x11[j] = A[o11[j]];
            }
#pragma unroll
            for (int j = 0; j < kJB; j++) {
                if (bk[j] < 0) continue;
                const double ck = cs[bk[j]], sk = sn[bk[j]], cl = cs[bl[j]], sl = sn[bl[j]];
                const double a = x00[j], b = x01[j], c = x10[j], d = x11[j];
                const double a1 = ck * a - sk * c, b1 = ck * b - sk * d;
                const double c1 = sk * a + ck * c, d1 = sk * b + ck * d;
                A[o00[j]] = cl * a1 - sl * b1;
                A[o01[j]] = sl * a1 + cl * b1;
                A[o10[j]] = cl * c1 - sl * d1;
                A[o11[j]] = sl * c1 + cl * d1;
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < m; i += kJT) evals[i] = A[i * S + i];
    if (tid == 0) {
        info[0] = sweep;
        info[1] = sqrt(off / (tot > 0 ? tot : 1.0));
    }
}

// Row blockIdx.x of V = G_1^T G_2^T ... (the logged rotations in order, from
// the identity): V[row][p] <- c V[row][p] - s V[row][q], V[row][q] <-
// s V[row][p] + c V[row][q] for each pair (p, q) of each round, the
// sequential method's eigenvector update.  One wave per row; the next
// round's (c, s) is loaded while this one is applied.
__global__ void __launch_bounds__(64) jacobi_vectors_kernel(const double2 *rot, const double *info, int m,
                                                            double *V)
{
    __shared__ double row[128];
    const int r = blockIdx.x, lane = threadIdx.x;
    const int half = m / 2;
    for (int c = lane; c < m; c += 64) row[c] = (c == r) ? 1.0 : 0.0;
    __syncthreads();
    const long rounds = (long)info[0] * (m - 1);
    int ia = lane - 1, ib = m - 2 - lane, rr = 0;
    double2 nxt = (lane < half && rounds > 0) ? rot[lane] : make_double2(1.0, 0.0);
    for (long t = 0; t < rounds; t++) {
        const double2 g = nxt;
        if (lane < half && t + 1 < rounds) nxt = rot[(t + 1) * half + lane];
        if (lane < half) {
            int a = (lane == 0) ? 0 : ia + 1;
            int b = ib + 1;
            if (a > b) { int x = a; a = b; b = x; }
            const double vp = row[a], vq = row[b];
            row[a] = g.x * vp - g.y * vq;
            row[b] = g.y * vp + g.x * vq;
            ia = ia + 1 == m - 1 ? 0 : ia + 1;
            ib = ib + 1 == m - 1 ? 0 : ib + 1;
        }
        if (++rr == m - 1) {  // a new sweep starts the pairing again
            rr = 0;
            ia = lane - 1;
            ib = m - 2 - lane;
        }
        __syncthreads();
    }
    for (int c = lane; c < m; c += 64) V[(size_t)r * m + c] = row[c];
}

// G = R^T R (R upper), Rinv = R^-1; info[0] = 0 on success, else 1-based
// index of the first non-positive pivot (relative threshold eps); info[1] =
// ||R||_F ||R^-1||_F = sqrt(trace G) ||R^-1||_F, an upper bound on the
// 2-norm condition number of R (and of any W with W^T W = G).
__global__ void __launch_bounds__(64) chol_inv_kernel(const double *G, int b, double *Rinv, double eps,
                                                      double *info)
{
    __shared__ double R[32 * 32];
    __shared__ double I[32 * 32];
    __shared__ int bad;
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    for (int e = tid; e < b * b; e += 64) {
        R[e] = 0.5 * (G[e] + G[(e % b) * b + e / b]);
        I[e] = 0.0;
    }
    __syncthreads();
    double dmax = 0.0, tr = 0.0;
    for (int i = 0; i < b; i++) {
        dmax = fmax(dmax, R[i * b + i]);
        tr += R[i * b + i];
    }
    // right-looking Cholesky on the upper triangle
    for (int k = 0; k < b; k++) {
        if (tid == 0) {
            double d = R[k * b + k];
            if (!(d > eps * dmax) && !bad) bad = k + 1;
            R[k * b + k] = sqrt(d > eps * dmax ? d : eps * dmax + 1e-300);
        }
        __syncthreads();
        const double rkk = R[k * b + k];
        for (int j = k + 1 + tid; j < b; j += 64) R[k * b + j] /= rkk;
        __syncthreads();
        for (int e = tid; e < b * b; e += 64) {
            const int i = e / b, j = e - i * b;
            if (i > k && j >= i) R[i * b + j] -= R[k * b + i] * R[k * b + j];
        }
        __syncthreads();
    }
    // inverse of the upper-triangular R, column by column (back substitution)
    for (int j = tid; j < b; j += 64) {
        for (int i = j; i >= 0; i--) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int l = i + 1; l <= j; l++) s -= R[i * b + l] * I[l * b + j];
            I[i * b + j] = s / R[i * b + i];
        }
    }
    __syncthreads();
    double f2 = 0.0;
    for (int e = tid; e < b * b; e += 64) {
        const int i = e / b, j = e - i * b;
        const double v = (j >= i) ? I[e] : 0.0;
        Rinv[e] = v;
        f2 += v * v;
    }
    for (int o = 32; o; o >>= 1) f2 += __shfl_xor(f2, o);
    if (tid == 0) {
        info[0] = bad;
        info[1] = bad ? INFINITY : sqrt(tr * f2);
    }
}

// H[0:rows, c0:c0+b] = C1 + C2 and the mirrored block, H is ld x ld.  In
// the diagonal block (rows c0 .. c0+b-1) entry (i, j) and its mirror are
// both set to the mean of C(i, j) and C(j, i), so that every H entry gets
// one value whichever thread writes it last.
__global__ void __launch_bounds__(256) h_update_kernel(double *H, int ld, int rows, int c0, int b,
                                                       const double *C1, const double *C2)
{
    for (int e = threadIdx.x + blockIdx.x * 256; e < rows * b; e += gridDim.x * 256) {
        const int i = e / b, j = e - i * b;
        double v = C1[e] + (C2 ? C2[e] : 0.0);
        if (i >= c0 && i < c0 + b) {
            const int t = (c0 + j) * b + (i - c0);  // C(j, i) of the diagonal block
            v = 0.5 * (v + (C1[t] + (C2 ? C2[t] : 0.0)));
        }
        H[i * ld + c0 + j] = v;
        H[(c0 + j) * ld + i] = v;
    }
}

__global__ void __launch_bounds__(256) cast_kernel(const float *X, int64_t ldx, uint64_t n, int k,
                                                   double *Y, int64_t ldy)
{
    const uint64_t tot = n * (uint64_t)k;
    for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * 256) {
        const uint64_t r = e / k, c = e - r * k;
        Y[r * ldy + c] = (double)X[r * ldx + c];
    }
}

__global__ void __launch_bounds__(256) cast_down_kernel(const double *X, int64_t ldx, uint64_t n, int k,
                                                        float *Y, int64_t ldy)
{
    const uint64_t tot = n * (uint64_t)k;
    for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * 256) {
        const uint64_t r = e / k, c = e - r * k;
        Y[r * ldy + c] = (float)X[r * ldx + c];
    }
}

// partial[b][j] = sum over the block's rows of (Y[r][j] - theta[j] X[r][j])^2
__global__ void __launch_bounds__(256) resid_kernel(const double *Y, int64_t ldy, const double *X,
                                                    int64_t ldx, const double *theta, uint64_t n, int k,
                                                    uint64_t rows_per_blk, double *partial)
{
    const int per = 256 / k;
    const int c = threadIdx.x % k, rr = threadIdx.x / k;
    double s = 0.0;
    if (rr < per) {
        const double th = theta[c];
        const uint64_t r0 = blockIdx.x * rows_per_blk;
        const uint64_t r1 = min(n, r0 + rows_per_blk);
        for (uint64_t r = r0 + rr; r < r1; r += per) {
            const double v = Y[r * ldy + c] - th * X[r * ldx + c];
            s = fma(v, v, s);
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < k) {
        double t = 0.0;
        for (int q = 0; q < per; q++) t += red[q * k + threadIdx.x];
        partial[(uint64_t)blockIdx.x * k + threadIdx.x] = t;
    }
}

__global__ void __launch_bounds__(256) sum_cols_kernel(const double *partial, int nblk, int k, double *out)
{
    for (int j = threadIdx.x; j < k; j += 256) {
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += partial[(uint64_t)b * k + j];
        out[j] = sqrt(s);
    }
}

// dinv[i] = deg[i] > 0 ? 1/sqrt(deg[i]) : 0  (deg as float from A*1)
__global__ void __launch_bounds__(256) dinv_kernel(const float *deg, uint64_t n, double *dinv)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const double d = (double)deg[i];
        dinv[i] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
    }
}

// X[r][c] = uniform [-1, 1) from SplitMix64(seed, r * k + c) (the random
// start of FM_MultiVector::Random, FM_MultiVector.h:164-168)
template <typename T>
__global__ void __launch_bounds__(256) rand_kernel(T *X, int64_t ldx, uint64_t n, int k, uint64_t seed)
{
    const uint64_t tot = n * (uint64_t)k;
    for (uint64_t e = blockIdx.x * 256ull + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * 256) {
        uint64_t z = seed + (e + 1) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
        const uint64_t r = e / k, c = e - r * k;
        X[r * ldx + c] = (T)(2.0 * u - 1.0);
    }
}

}  // namespace

fmx_status rand_uniform(fmx_ctx *ctx, fmx_dtype dt, void *X, int64_t ldx, uint64_t n, int k, uint64_t seed,
                        hipStream_t s)
{
    uint64_t blocks = std::min<uint64_t>((n * k + 255) / 256, (uint64_t)ctx->num_cus * 16);
    if (!blocks) return FMX_OK;
    if (dt == FMX_F32)
        hipLaunchKernelGGL(rand_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, s, (float *)X, ldx, n, k,
                           seed);
    else
        hipLaunchKernelGGL(rand_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, s, (double *)X, ldx, n,
                           k, seed);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

size_t small_sym_eig_ws(int m)
{
    return (size_t)kJacobiSweeps * std::max(m - 1, 1) * std::max(m / 2, 1) * sizeof(double2);
}

fmx_status small_sym_eig(fmx_ctx *ctx, const double *H, int m, double *evals, double *V, double *info,
                         void *ws, hipStream_t s)
{
    if (m <= 0 || m > 128 || (m & 1)) return fail(ctx, FMX_EDIM, "Rayleigh-Ritz size must be even and <= 128");
    if (!info || !ws) return fail(ctx, FMX_EINVAL, "Rayleigh-Ritz needs its info and rotation buffers");
    static bool attr = false;
    if (!attr) {
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)jacobi_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         150 * 1024));
        attr = true;
    }
    const size_t mm = (size_t)m * (m + 1) * sizeof(double);  // padded rows (at most 132,096 bytes)
    hipLaunchKernelGGL(jacobi_kernel, dim3(1), dim3(kJT), mm, s, H, m, evals, (double2 *)ws, kJacobiSweeps, info);
    FMX_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(jacobi_vectors_kernel, dim3(m), dim3(64), 0, s, (const double2 *)ws, (const double *)info,
                       m, V);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status small_chol_inv(fmx_ctx *ctx, const double *G, int b, double *Rinv, double *info, hipStream_t s)
{
    if (b <= 0 || b > 32) return fail(ctx, FMX_EDIM, "Cholesky block must be <= 32");
    hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(64), 0, s, G, b, Rinv, 1e-13, info);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status h_update(fmx_ctx *ctx, double *H, int ld, int rows, int c0, int b, const double *C1,
                    const double *C2, hipStream_t s)
{
    hipLaunchKernelGGL(h_update_kernel, dim3(8), dim3(256), 0, s, H, ld, rows, c0, b, C1, C2);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status cast_f32_f64(fmx_ctx *ctx, const float *X, int64_t ldx, uint64_t n, int k, double *Y,
                        int64_t ldy, hipStream_t s)
{
    uint64_t blocks = std::min<uint64_t>((n * k + 255) / 256, (uint64_t)ctx->num_cus * 16);
    if (!blocks) return FMX_OK;
    hipLaunchKernelGGL(cast_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, X, ldx, n, k, Y, ldy);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status cast_f64_f32(fmx_ctx *ctx, const double *X, int64_t ldx, uint64_t n, int k, float *Y,
                        int64_t ldy, hipStream_t s)
{
    uint64_t blocks = std::min<uint64_t>((n * k + 255) / 256, (uint64_t)ctx->num_cus * 16);
    if (!blocks) return FMX_OK;
    hipLaunchKernelGGL(cast_down_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, X, ldx, n, k, Y, ldy);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status resid_norms(fmx_ctx *ctx, const double *Y, int64_t ldy, const double *X, int64_t ldx,
                       const double *theta, uint64_t n, int k, double *out, double *partial,
                       int nblk, hipStream_t s)
{
    if (k <= 0 || k > 256) return fail(ctx, FMX_EDIM, "residual width must be <= 256");
    const uint64_t rpb = (n + nblk - 1) / nblk;
    hipLaunchKernelGGL(resid_kernel, dim3(nblk), dim3(256), 0, s, Y, ldy, X, ldx, theta, n, k, rpb, partial);
    FMX_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(sum_cols_kernel, dim3(1), dim3(256), 0, s, partial, nblk, k, out);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status degree_inv_sqrt(fmx_ctx *ctx, const float *deg, uint64_t n, double *dinv, hipStream_t s)
{
    uint64_t blocks = std::min<uint64_t>((n + 255) / 256, (uint64_t)ctx->num_cus * 16);
    if (!blocks) return FMX_OK;
    hipLaunchKernelGGL(dinv_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, deg, n, dinv);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

}  // namespace fmx
