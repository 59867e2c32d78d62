// dense.hip -- tall-skinny dense operations of the block eigensolver.
//
// Reference operations (all HBM-bound on an n x k operand, k <= 256):
//   Gram        B = A^T X      MvTransMv (FM_MultiVector.h:499-558,
//                              block_dense_matrix.cpp:1429-1513)
//   gemm_tall   C = a*A*B+b*C  MvTimesMatAddMv (FM_MultiVector.h:389-433,
//                              block_dense_matrix.cpp:822-914)
//   colnorm     ||a_j||_2      MvNorm (FM_MultiVector.h:589-628)
//   row_scale   diag(d)*X      NA_eigen_Operator's scale_rows
//                              (matrix/test/test-eigen.cpp:112-127)
// The Gram comes in two precisions: FP64 accumulation (the reorthogonalising
// pass) and BF16 MFMA with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
// Partial results per row block are reduced in a fixed order, so every
// result is deterministic.
#include "dense.h"

namespace fmx {
namespace {

constexpr int kT = 256;

template <typename T>
__device__ __forceinline__ double ld64(const T *p) { return (double)*p; }

template <typename T>
struct alignas(16) Vec16 {
    T v[16 / sizeof(T)];
};

// ---- FP64 Gram ----------------------------------------------------------------
// Block b accumulates G_b = sum over its rows of A[r][:]^T B[r][:] for the
// output slice [o0, o0 + 4096) of the ka*kb outputs, rows staged in LDS.
constexpr int kGRows = 32;

template <typename T>
__global__ void __launch_bounds__(kT) gram64_kernel(const T *A, int64_t lda, const T *B, int64_t ldb,
                                                    uint64_t n, int ka, int kb, uint64_t rows_per_blk,
                                                    double *partial)
{
    extern __shared__ double sm[];
    double *As = sm;                  // kGRows x ka
    double *Bs = sm + kGRows * ka;    // kGRows x kb
    const int nout = ka * kb;
    const int o0 = blockIdx.y * 4096;
    double acc[16];
#pragma unroll
    for (int t = 0; t < 16; t++) acc[t] = 0.0;
    const uint64_t r0 = blockIdx.x * rows_per_blk;
    const uint64_t r1 = min(n, r0 + rows_per_blk);
    for (uint64_t rb = r0; rb < r1; rb += kGRows) {
        const int nr = (int)min((uint64_t)kGRows, r1 - rb);
        for (int e = threadIdx.x; e < kGRows * ka; e += kT) {
            int r = e / ka, c = e - r * ka;
            As[e] = r < nr ? ld64(A + (rb + r) * lda + c) : 0.0;
        }
        for (int e = threadIdx.x; e < kGRows * kb; e += kT) {
            int r = e / kb, c = e - r * kb;
            Bs[e] = r < nr ? ld64(B + (rb + r) * ldb + c) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int o = o0 + t * kT + threadIdx.x;
            if (o < nout) {
                const int i = o / kb, j = o - i * kb;
                double s = acc[t];
                for (int r = 0; r < nr; r++) s = fma(As[r * ka + i], Bs[r * kb + j], s);
                acc[t] = s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const int o = o0 + t * kT + threadIdx.x;
        if (o < nout) partial[(uint64_t)blockIdx.x * nout + o] = acc[t];
    }
}

// out[o] = sum_b partial[b][o] in block order (deterministic)
template <typename TP, typename TO>
__global__ void __launch_bounds__(kT) reduce_partials(const TP *partial, int nblk, int nout,
                                                      TO *out, int sqrt_out)
{
    for (int o = blockIdx.x * kT + threadIdx.x; o < nout; o += gridDim.x * kT) {
        TP s = 0;
        for (int b = 0; b < nblk; b++) s += partial[(uint64_t)b * nout + o];
        double v = (double)s;
        out[o] = (TO)(sqrt_out ? sqrt(v) : v);
    }
}

// out[o] = the sum over 16 interleaved segments of the partials (segment q
// holds blocks q, q + 16, ...), each summed in block order, then the
// segments in order; FP64 throughout (deterministic).  16 outputs x 16
// segments per block: the many-partial reduce of the BF16 Gram.
template <typename TP>
__global__ void __launch_bounds__(kT) reduce_partials_seg(const TP *partial, int nblk, int nout, double *out)
{
    __shared__ double red[kT];
    const int ol = threadIdx.x & 15, sg = threadIdx.x >> 4;
    const int o = blockIdx.x * 16 + ol;
    double acc = 0.0;
    if (o < nout)
        for (int b = sg; b < nblk; b += 16) acc += (double)partial[(uint64_t)b * nout + o];
    red[threadIdx.x] = acc;
    __syncthreads();
    if (sg == 0 && o < nout) {
        double t = 0.0;
        for (int q = 0; q < 16; q++) t += red[q * 16 + ol];
        out[o] = t;
    }
}

// ---- BF16 operands ---------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f)
{
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (unsigned short)(u >> 16);  // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// ---- FP64 MFMA Gram and tall GEMM ------------------------------------------------
// v_mfma_f64_16x16x4f64 with operands A: 16 x 4 (lane l holds A[l%16][l/16]),
// B: 4 x 16 (lane l holds B[l/16][l%16]); products of FP32 inputs are exact
// in FP64.  For the Gram G = A^T B the rows are the K dimension, so every
// output tile accumulates over a wave's rows in registers; per-wave tiles go
// to partial[wave] and are reduced in wave order.
typedef double doublex4 __attribute__((ext_vector_type(4)));
constexpr int kGramTiles = 16;  // output tiles per wave (64 FP64 accumulators per lane)

// Staged forms: each wave copies 16 rows of its operands into its own LDS
// slice with 16-byte loads (the next 16 rows are in flight while the current
// ones feed the MFMAs), and the MFMA operands are read from LDS.  Rows are
// padded in LDS so that the lanes of one operand read spread over the banks.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

constexpr int kStageRows = 16;

// Copies rows [r, r + 16) x columns [0, k) of M (row stride ld) as 16-byte
// pieces into registers; pieces past the matrix are zero.  The slice is
// 16 x KP elements; lane l takes pieces l, l + 64, ...
template <typename T, int KP>
struct Stage {
    static constexpr int V = 16 / sizeof(T);               // elements per piece
    static constexpr int NP = kStageRows * KP / V / 64;    // pieces per lane
    Vec16<T> reg[NP > 0 ? NP : 1];
    __device__ __forceinline__ void load(const T *M, int64_t ld, int k, uint64_t r, uint64_t r1, int lane, int vec)
    {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int e = (p * 64 + lane) * V, row = e / KP, col = e - row * KP;
            const uint64_t gr = r + row;
            Vec16<T> v;
#pragma unroll
            for (int u = 0; u < V; u++) v.v[u] = T(0);
            if (gr < r1 && col < k) {  // pieces wholly past the width stay zero
                const T *src = M + gr * ld + col;
                if (vec && col + V <= k) {
                    v = *(const Vec16<T> *)src;
                } else {
#pragma unroll
                    for (int u = 0; u < V; u++)
                        if (col + u < k) v.v[u] = src[u];
                }
            }
            reg[p] = v;
        }
    }
    // Every piece loaded with no branch, so that the count of loads in
    // flight is the same on every path (a wait for a later load then need
    // not wait for these): a piece past row r1 holds row r's, a piece past
    // the width the row's first.  Only for a product whose rows past r1 are
    // not stored and whose factor is zero past the width (the finite values
    // there then add nothing).  M 16-byte aligned, ld and k multiples of V,
    // r < r1.
    __device__ __forceinline__ void load_all(const T *M, int64_t ld, int k, uint64_t r, uint64_t r1, int lane)
    {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int e = (p * 64 + lane) * V, row = e / KP, col = e - row * KP;
            const uint64_t gr = r + row;
            const Vec16<T> t = *(const Vec16<T> *)(M + (gr < r1 ? gr : r) * ld + (col < k ? col : 0));
#pragma unroll
            for (int u = 0; u < V; u++) reg[p].v[u] = t.v[u];  // (a whole-struct copy stays in scratch)
        }
    }
    // into LDS rows of stride LS elements (16-byte stores when the rows are
    // 16-byte aligned; the caller's slice is)
    template <int LS>
    __device__ __forceinline__ void store(T *lds, int lane) const
    {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int e = (p * 64 + lane) * V, row = e / KP, col = e - row * KP;
            if constexpr ((LS * sizeof(T)) % 16 == 0) {
                *(Vec16<T> *)(lds + row * LS + col) = reg[p];
            } else {
#pragma unroll
                for (int u = 0; u < V; u++) lds[row * LS + col + u] = reg[p].v[u];
            }
        }
    }
};

// G = A^T B, ka <= 16*TI, kb <= 16*TJ: per 16 staged rows, four K = 4 steps.
// SYM (B is A: same pointer, stride and width, TI == TJ) stages the rows once
// and reads both MFMA operands from them.
template <typename T, int TI, int TJ, bool SYM>
__device__ __forceinline__ void gram_stage_body(const T *A, int64_t lda, const T *B, int64_t ldb,
                                                        uint64_t n, int ka, int kb, uint64_t rows_per_wave,
                                                        uint64_t nwaves, int veca, int vecb, double *partial)
{
    static_assert(!SYM || TI == TJ, "a symmetric Gram has square tiles");
    constexpr int KAP = TI * 16, KBP = TJ * 16;
    constexpr int LA = KAP + 16, LB = KBP + 16;  // padded LDS row strides (elements)
    extern __shared__ __align__(16) unsigned char smg[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T *As = (T *)smg + w * kStageRows * (LA + LB);
    T *Bs = As + kStageRows * LA;
    const uint64_t wave = (blockIdx.x * (uint64_t)kT + threadIdx.x) >> 6;  // past nwaves: no rows
    (void)nwaves;
    const int li = lane & 15, lk = lane >> 4;
    doublex4 acc[TI * TJ];
#pragma unroll
    for (int t = 0; t < TI * TJ; t++) acc[t] = doublex4{0.0, 0.0, 0.0, 0.0};
    const uint64_t r0 = wave * rows_per_wave;
    const uint64_t r1 = min(n, r0 + rows_per_wave);
    Stage<T, KAP> sa;
    Stage<T, KBP> sb;
    if (r0 < r1) {
        sa.load(A, lda, ka, r0, r1, lane, veca);
        if constexpr (!SYM) sb.load(B, ldb, kb, r0, r1, lane, vecb);
    }
    for (uint64_t r = r0; r < r1; r += kStageRows) {
        sa.template store<LA>(As, lane);
        if constexpr (!SYM) sb.template store<LB>(Bs, lane);
        wave_lds_sync();
        if (r + kStageRows < r1) {
            sa.load(A, lda, ka, r + kStageRows, r1, lane, veca);
            if constexpr (!SYM) sb.load(B, ldb, kb, r + kStageRows, r1, lane, vecb);
        }
#pragma unroll
        for (int ks = 0; ks < kStageRows / 4; ks++) {
            const int row = ks * 4 + lk;
            double a[TI], b[TJ];
#pragma unroll
            for (int t = 0; t < TI; t++) a[t] = (double)As[row * LA + t * 16 + li];
#pragma unroll
            for (int t = 0; t < TJ; t++) {
                if constexpr (SYM) b[t] = a[t];
                else b[t] = (double)Bs[row * LB + t * 16 + li];
            }
#pragma unroll
            for (int ti = 0; ti < TI; ti++)
#pragma unroll
                for (int tj = 0; tj < TJ; tj++)
                    acc[ti * TJ + tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], b[tj], acc[ti * TJ + tj], 0, 0, 0);
        }
        wave_lds_sync();
    }
    // the block's four waves combine in LDS as (w0 + w2) + (w1 + w3); the
    // block's sum goes to partial[block]
    const int nout = ka * kb;
    double *red = (double *)smg;  // 2 x nout
    auto exchange = [&](int writer_lo, int writers, bool add) {
        __syncthreads();
        if (w >= writer_lo && w < writer_lo + writers) {
            double *slot = red + (w - writer_lo) * nout;
#pragma unroll
            for (int ti = 0; ti < TI; ti++)
#pragma unroll
                for (int tj = 0; tj < TJ; tj++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        // FP64 16x16x4 output: value q of lane l is row l/16 + 4q, column l%16
                        const int i = ti * 16 + lk + 4 * q, j = tj * 16 + li;
                        if (i < ka && j < kb) slot[i * kb + j] = acc[ti * TJ + tj][q];
                    }
        }
        __syncthreads();
        if (add && w < writers) {
            const double *slot = red + w * nout;
#pragma unroll
            for (int ti = 0; ti < TI; ti++)
#pragma unroll
                for (int tj = 0; tj < TJ; tj++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int i = ti * 16 + lk + 4 * q, j = tj * 16 + li;
                        if (i < ka && j < kb) acc[ti * TJ + tj][q] += slot[i * kb + j];
                    }
        }
    };
    exchange(2, 2, true);   // w0 += w2, w1 += w3
    exchange(1, 1, true);   // w0 += w1
    if (w == 0) {
#pragma unroll
        for (int ti = 0; ti < TI; ti++)
#pragma unroll
            for (int tj = 0; tj < TJ; tj++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = ti * 16 + lk + 4 * q, j = tj * 16 + li;
                    if (i < ka && j < kb) partial[(uint64_t)blockIdx.x * nout + i * kb + j] = acc[ti * TJ + tj][q];
                }
    }
}

// The Gram kernel, and the same kernel asked to fit three waves per SIMD (168
// registers) for the FP32 shapes whose body fits that budget without
// spilling (at most 8 output tiles); the launcher runs the latter with up to
// 12 waves per CU.
template <typename T, int TI, int TJ, bool SYM = false>
__global__ void __launch_bounds__(kT) gram_stage_kernel(const T *A, int64_t lda, const T *B, int64_t ldb,
                                                        uint64_t n, int ka, int kb, uint64_t rows_per_wave,
                                                        uint64_t nwaves, int veca, int vecb, double *partial)
{
    gram_stage_body<T, TI, TJ, SYM>(A, lda, B, ldb, n, ka, kb, rows_per_wave, nwaves, veca, vecb, partial);
}
template <typename T, int TI, int TJ, bool SYM = false>
__global__ void __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(3))) gram_stage_kernel3(const T *A, int64_t lda, const T *B, int64_t ldb,
                                                        uint64_t n, int ka, int kb, uint64_t rows_per_wave,
                                                        uint64_t nwaves, int veca, int vecb, double *partial)
{
    gram_stage_body<T, TI, TJ, SYM>(A, lda, B, ldb, n, ka, kb, rows_per_wave, nwaves, veca, vecb, partial);
}

// Y = alpha A C + beta Y for kc <= 16*TJ, ka <= KAP: per 16 staged rows of A,
// ka/4 MFMA steps against C held in registers (lane l: C[4s + l/16][l%16]).
// A row is staged whole before any output of it is written, so Y may be A.
template <typename T, typename TY, int KAP, int TJ, bool VEC, int SC = -1, int OLD = -1>
__device__ __forceinline__ void gemm_stage_body(const T *A, int64_t lda, uint64_t n, int ka, int kc,
                                                        const double *C, double alpha, double beta, TY *Y,
                                                        int64_t ldy, uint64_t rows_per_wave, uint64_t nwaves,
                                                        int veca, unsigned short *Yb, int64_t ldyb,
                                                        const double *ds, TY *Ys, int64_t ldys)
{
    // LDS row stride: FP32 rows KAP + 4 floats put the 16 rows of a k column
    // on 16 distinct 4-bank groups (and keep rows 16-byte aligned for vector
    // stores); FP64 rows an odd count of doubles
    constexpr int LA = sizeof(T) == 4 ? KAP + 4 : KAP + 1;
    constexpr int KS = KAP / 4;       // MFMA K steps
    extern __shared__ __align__(16) unsigned char smg[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T *As = (T *)smg + w * kStageRows * LA;
    const uint64_t wave = (blockIdx.x * (uint64_t)kT + threadIdx.x) >> 6;
    if (wave >= nwaves) return;
    const int li = lane & 15, lk = lane >> 4;
    double c[TJ][KS];
#pragma unroll
    for (int tj = 0; tj < TJ; tj++)
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int k = 4 * s + lk, j = tj * 16 + li;
            c[tj][s] = (k < ka && j < kc) ? C[k * kc + j] : 0.0;
        }
    const uint64_t r0 = wave * rows_per_wave;
    const uint64_t r1 = min(n, r0 + rows_per_wave);
    Stage<T, KAP> sa;
    if (r0 < r1) {
        if constexpr (VEC) sa.load_all(A, lda, ka, r0, r1, lane);
        else sa.load(A, lda, ka, r0, r1, lane, veca);
    }
    for (uint64_t r = r0; r < r1; r += kStageRows) {
        sa.template store<LA>(As, lane);
        wave_lds_sync();
        // This stage's old Y values (read when beta != 0) and row scales,
        // loaded from in-bounds addresses whatever the rows, and requested
        // before the next stage's rows: the epilogue's wait for them then
        // leaves those rows in flight through the MFMAs.
        // (SC: whether the row-scaled copy Ys is written, 1 or 0 when known
        // at compile time, else -1; a run-time branch here would wait for
        // these loads at once)
        const bool sc = SC < 0 ? Ys != nullptr : SC > 0;
        double dsc[4] = {0.0, 0.0, 0.0, 0.0};
        if (sc) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint64_t gr = r + lk + 4 * q;
                dsc[q] = ds[gr < r1 ? gr : r];
            }
        }
        // (OLD: whether Y's old values are read, that is beta != 0, known
        // at compile time as SC is, or -1)
        const bool old = OLD < 0 ? beta != 0.0 : OLD > 0;
        TY yold[TJ][4];
#pragma unroll
        for (int tj = 0; tj < TJ; tj++) {
            const int j = tj * 16 + li;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint64_t gr = r + lk + 4 * q;
                yold[tj][q] = old ? Y[(gr < r1 ? gr : r) * ldy + (j < kc ? j : 0)] : TY(0);
            }
        }
        if constexpr (VEC) {
            // the last stage reloads its own rows (unused), so that every
            // stage has the same loads in flight; the scheduling barriers
            // keep the loads in this order, ahead of the MFMAs
            __builtin_amdgcn_sched_barrier(0);
            sa.load_all(A, lda, ka, r + kStageRows < r1 ? r + kStageRows : r, r1, lane);
            __builtin_amdgcn_sched_barrier(0);
        } else {
            if (r + kStageRows < r1) sa.load(A, lda, ka, r + kStageRows, r1, lane, veca);
        }
        doublex4 acc[TJ];
#pragma unroll
        for (int tj = 0; tj < TJ; tj++) acc[tj] = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const double a = (double)As[li * LA + 4 * s + lk];
#pragma unroll
            for (int tj = 0; tj < TJ; tj++) acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, c[tj][s], acc[tj], 0, 0, 0);
        }
        if constexpr (VEC) __builtin_amdgcn_sched_barrier(0);  // the old values' uses stay after the MFMAs
        wave_lds_sync();
#pragma unroll
        for (int tj = 0; tj < TJ; tj++) {
            const int j = tj * 16 + li;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint64_t gr = r + lk + 4 * q;
                if (gr < r1 && j < kc) {
                    const double prev = old ? beta * (double)yold[tj][q] : 0.0;
                    const TY y = (TY)(prev + alpha * acc[tj][q]);
                    Y[gr * ldy + j] = y;
                    if (Yb) Yb[gr * ldyb + j] = f2bf((float)y);  // the BF16 copy of what was stored
                    if (sc) Ys[gr * ldys + j] = (TY)((double)y * dsc[q]);  // and its row-scaled copy
                }
            }
        }
    }
}

// The kernel, and the same kernel asked to fit three waves per SIMD (168
// registers): more 16-row stages in flight.  The launcher takes the second
// for FP32 rows with one column tile and KAP <= 96, where the body fits that
// budget without spilling (the 96-column product 8 percent faster, the
// 64-column 11 percent); wider variants would spill.
template <typename T, typename TY, int KAP, int TJ, bool VEC = false, int SC = -1, int OLD = -1>
__global__ void __launch_bounds__(kT) gemm_stage_kernel(const T *A, int64_t lda, uint64_t n, int ka, int kc,
                                                        const double *C, double alpha, double beta, TY *Y,
                                                        int64_t ldy, uint64_t rows_per_wave, uint64_t nwaves,
                                                        int veca, unsigned short *Yb, int64_t ldyb,
                                                        const double *ds, TY *Ys, int64_t ldys)
{
    gemm_stage_body<T, TY, KAP, TJ, VEC, SC, OLD>(A, lda, n, ka, kc, C, alpha, beta, Y, ldy, rows_per_wave, nwaves, veca, Yb, ldyb, ds, Ys, ldys);
}
template <typename T, typename TY, int KAP, int TJ, bool VEC, int SC, int OLD>
__global__ void __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(3))) gemm_stage_kernel3(const T *A, int64_t lda, uint64_t n, int ka, int kc,
                                                        const double *C, double alpha, double beta, TY *Y,
                                                        int64_t ldy, uint64_t rows_per_wave, uint64_t nwaves,
                                                        int veca, unsigned short *Yb, int64_t ldyb,
                                                        const double *ds, TY *Ys, int64_t ldys)
{
    gemm_stage_body<T, TY, KAP, TJ, VEC, SC, OLD>(A, lda, n, ka, kc, C, alpha, beta, Y, ldy, rows_per_wave, nwaves, veca, Yb, ldyb, ds, Ys, ldys);
}

// out[o] = sum over p of partial[p][o], p in order 0..np-1 within four
// interleaved quarters that are then added in order (deterministic).
__global__ void __launch_bounds__(kT) reduce_partials4(const double *partial, int np, int nout, double *out)
{
    __shared__ double red[kT];
    const int ol = threadIdx.x & 63, qt = threadIdx.x >> 6;
    const int o = blockIdx.x * 64 + ol;
    double s = 0.0;
    if (o < nout)
        for (int p = qt; p < np; p += 4) s += partial[(uint64_t)p * nout + o];
    red[threadIdx.x] = s;
    __syncthreads();
    if (qt == 0 && o < nout) out[o] = ((red[ol] + red[64 + ol]) + red[128 + ol]) + red[192 + ol];
}

// ---- BF16 MFMA Gram ------------------------------------------------------------
constexpr int kBK = 64;          // rows per LDS stage (two MFMA k-steps)
constexpr int kBPad = kBK + 8;   // padded row length (bf16 elements)

// ka, kb multiples of 16 (caller pads); tiles of 16x16 spread over 4 waves.
__global__ void __launch_bounds__(kT) gram_bf16_kernel(const float *A, int64_t lda, const float *B,
                                                       int64_t ldb, uint64_t n, int ka, int kb,
                                                       int ka_real, int kb_real,
                                                       uint64_t rows_per_blk, float *partial)
{
    extern __shared__ unsigned short smb[];
    unsigned short *At = smb;                    // ka x kBPad  (A transposed)
    unsigned short *Bt = smb + ka * kBPad;       // kb x kBPad
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tiles_i = ka / 16, tiles_j = kb / 16, ntiles = tiles_i * tiles_j;
    constexpr int kMaxTiles = 16;                // per wave (ka*kb <= 64*256*16)
    floatx4 acc[kMaxTiles];
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    const uint64_t r0 = blockIdx.x * rows_per_blk;
    const uint64_t r1 = min(n, r0 + rows_per_blk);
    for (uint64_t rb = r0; rb < r1; rb += kBK) {
        const int nr = (int)min((uint64_t)kBK, r1 - rb);
        for (int e = threadIdx.x; e < kBK * ka; e += kT) {
            int r = e / ka, c = e - r * ka;
            float v = (r < nr && c < ka_real) ? A[(rb + r) * lda + c] : 0.f;
            At[c * kBPad + r] = f2bf(v);
        }
        for (int e = threadIdx.x; e < kBK * kb; e += kT) {
            int r = e / kb, c = e - r * kb;
            float v = (r < nr && c < kb_real) ? B[(rb + r) * ldb + c] : 0.f;
            Bt[c * kBPad + r] = f2bf(v);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kMaxTiles; t++) {
            const int tile = w + 4 * t;
            if (tile < ntiles) {
                const int ti = tile / tiles_j, tj = tile - ti * tiles_j;
#pragma unroll
                for (int ks = 0; ks < kBK / 32; ks++) {
                    const int k0 = ks * 32 + (lane >> 4) * 8;
                    bf16x8 a = *(const bf16x8 *)(At + (ti * 16 + (lane & 15)) * kBPad + k0);
                    bf16x8 b = *(const bf16x8 *)(Bt + (tj * 16 + (lane & 15)) * kBPad + k0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    const int nout = ka_real * kb_real;
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) {
        const int tile = w + 4 * t;
        if (tile < ntiles) {
            const int ti = tile / tiles_j, tj = tile - ti * tiles_j;
            const int j = tj * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = ti * 16 + (lane >> 4) * 4 + q;
                if (i < ka_real && j < kb_real)
                    partial[(uint64_t)blockIdx.x * nout + i * kb_real + j] = acc[t][q];
            }
        }
    }
}

// Fast path of gram_bf16_kernel for 16-byte aligned rows whose real widths
// are multiples of 4 (ka <= 128, kb <= 32): each stage's rows are read as
// 16-byte pieces, and the next stage's pieces are requested before the
// current stage's MFMAs, so the loads overlap the products.  AB16: A is
// already BF16 (widths a multiple of 8), so its pieces carry 8 values and
// the stage moves half the bytes.
constexpr int kBVA = kBK * 128 / 4 / kT;  // 16-byte pieces of FP32 A per thread and stage (8)
constexpr int kBVB = kBK * 32 / 4 / kT;   // and of B (2)
struct alignas(16) Bf8 {
    unsigned short v[8];
};
template <bool AB16>
__global__ void __launch_bounds__(kT) gram_bf16_vec_kernel(const void *Av, int64_t lda, const float *B,
                                                           int64_t ldb, uint64_t n, int ka, int kb,
                                                           int ka_real, int kb_real,
                                                           uint64_t rows_per_blk, float *partial)
{
    extern __shared__ unsigned short smb[];
    unsigned short *At = smb;                    // ka x kBPad  (A transposed)
    unsigned short *Bt = smb + ka * kBPad;       // kb x kBPad
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tiles_i = ka / 16, tiles_j = kb / 16, ntiles = tiles_i * tiles_j;
    constexpr int kMaxTiles = 4;                 // ka <= 128, kb <= 32: at most 16 tiles
    constexpr int VA = AB16 ? 8 : 4;             // A values per piece
    constexpr int NA = AB16 ? kBVA / 2 : kBVA;   // A pieces per thread and stage
    floatx4 acc[kMaxTiles];
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    const uint64_t r0 = blockIdx.x * rows_per_blk;
    const uint64_t r1 = min(n, r0 + rows_per_blk);
    const int qa = ka / VA, qb = kb / 4;         // pieces per (padded) row
    const int pa = kBK * qa, pb = kBK * qb;      // pieces per stage
    float4 ra[AB16 ? 1 : NA], rbv[kBVB];
    Bf8 rh[AB16 ? NA : 1];
    auto load = [&](uint64_t rb) {
#pragma unroll
        for (int t = 0; t < NA; t++) {
            const int e = threadIdx.x + t * kT;
            const int r = e / qa, ca = e - r * qa;
            const bool in = e < pa && rb + r < r1 && ca * VA < ka_real;
            if constexpr (AB16) {
                Bf8 v = {};
                if (in) v = *(const Bf8 *)((const unsigned short *)Av + (rb + r) * lda + ca * 8);
                rh[t] = v;
            } else {
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (in) v = *(const float4 *)((const float *)Av + (rb + r) * lda + ca * 4);
                ra[t] = v;
            }
        }
#pragma unroll
        for (int t = 0; t < kBVB; t++) {
            const int e = threadIdx.x + t * kT;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (e < pb) {
                const int r = e / qb, c4 = e - r * qb;
                if (rb + r < r1 && c4 * 4 < kb_real) v = *(const float4 *)(B + (rb + r) * ldb + c4 * 4);
            }
            rbv[t] = v;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int t = 0; t < NA; t++) {
            const int e = threadIdx.x + t * kT;
            if (e < pa) {
                const int r = e / qa, c = (e - r * qa) * VA;
                if constexpr (AB16) {
#pragma unroll
                    for (int u = 0; u < 8; u++) At[(c + u) * kBPad + r] = rh[t].v[u];
                } else {
                    At[(c + 0) * kBPad + r] = f2bf(ra[t].x);
                    At[(c + 1) * kBPad + r] = f2bf(ra[t].y);
                    At[(c + 2) * kBPad + r] = f2bf(ra[t].z);
                    At[(c + 3) * kBPad + r] = f2bf(ra[t].w);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < kBVB; t++) {
            const int e = threadIdx.x + t * kT;
            if (e < pb) {
                const int r = e / qb, c = (e - r * qb) * 4;
                Bt[(c + 0) * kBPad + r] = f2bf(rbv[t].x);
                Bt[(c + 1) * kBPad + r] = f2bf(rbv[t].y);
                Bt[(c + 2) * kBPad + r] = f2bf(rbv[t].z);
                Bt[(c + 3) * kBPad + r] = f2bf(rbv[t].w);
            }
        }
    };
    if (r0 < r1) load(r0);
    for (uint64_t rb = r0; rb < r1; rb += kBK) {
        store();
        __syncthreads();
        if (rb + kBK < r1) load(rb + kBK);
#pragma unroll
        for (int t = 0; t < kMaxTiles; t++) {
            const int tile = w + 4 * t;
            if (tile < ntiles) {
                const int ti = tile / tiles_j, tj = tile - ti * tiles_j;
#pragma unroll
                for (int ks = 0; ks < kBK / 32; ks++) {
                    const int k0 = ks * 32 + (lane >> 4) * 8;
                    bf16x8 a = *(const bf16x8 *)(At + (ti * 16 + (lane & 15)) * kBPad + k0);
                    bf16x8 b = *(const bf16x8 *)(Bt + (tj * 16 + (lane & 15)) * kBPad + k0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    const int nout = ka_real * kb_real;
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) {
        const int tile = w + 4 * t;
        if (tile < ntiles) {
            const int ti = tile / tiles_j, tj = tile - ti * tiles_j;
            const int j = tj * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = ti * 16 + (lane >> 4) * 4 + q;
                if (i < ka_real && j < kb_real)
                    partial[(uint64_t)blockIdx.x * nout + i * kb_real + j] = acc[t][q];
            }
        }
    }
}

// ---- gemm_tall -------------------------------------------------------------------
// Y[r][j] = alpha * sum_i A[r][i] C[i][j] + beta * Y[r][j], FP64 accumulation.
constexpr int kMRows = 32;

template <typename T, typename TY = T>
__global__ void __launch_bounds__(kT) gemm_tall_kernel(const T *A, int64_t lda, uint64_t n, int ka,
                                                       int kc, const double *C, double alpha,
                                                       double beta, TY *Y, int64_t ldy)
{
    extern __shared__ double sg[];
    double *Cs = sg;                 // ka x kc
    double *As = sg + ka * kc;       // kMRows x ka
    for (int e = threadIdx.x; e < ka * kc; e += kT) Cs[e] = C[e];
    for (uint64_t rb = blockIdx.x * (uint64_t)kMRows; rb < n; rb += (uint64_t)gridDim.x * kMRows) {
        const int nr = (int)min((uint64_t)kMRows, n - rb);
        __syncthreads();
        for (int e = threadIdx.x; e < kMRows * ka; e += kT) {
            int r = e / ka, c = e - r * ka;
            As[e] = r < nr ? ld64(A + (rb + r) * lda + c) : 0.0;
        }
        __syncthreads();
        for (int o = threadIdx.x; o < nr * kc; o += kT) {
            const int r = o / kc, j = o - r * kc;
            double s = 0.0;
            for (int i = 0; i < ka; i++) s = fma(As[r * ka + i], Cs[i * kc + j], s);
            TY *y = Y + (rb + r) * ldy + j;
            const double prev = (beta == 0.0) ? 0.0 : beta * (double)*y;
            *y = (TY)(prev + alpha * s);
        }
    }
}

// One thread per row: the row of A is read once (16-byte loads when rows are
// aligned), C sits in LDS zero-padded to whole JT-column tiles (every lane
// reads the same C words: broadcasts), and each tile's JT sums stay in
// registers in FP64.  With Y == A a row is overwritten only after its last
// tile, so the in-place form needs kc <= JT.
template <typename T, typename TY, int JT>
__global__ void __launch_bounds__(kT) gemm_rows_kernel(const T *A, int64_t lda, uint64_t n, int ka, int kc,
                                                       int kcp, const double *C, double alpha, double beta,
                                                       TY *Y, int64_t ldy, int vec)
{
    extern __shared__ double sg[];
    double *Cs = sg;  // ka x kcp
    for (int e = threadIdx.x; e < ka * kcp; e += kT) {
        const int i = e / kcp, j = e - i * kcp;
        Cs[e] = j < kc ? C[i * kc + j] : 0.0;
    }
    __syncthreads();
    constexpr int V = 16 / sizeof(T);  // elements per 16-byte load
    for (uint64_t r = blockIdx.x * (uint64_t)kT + threadIdx.x; r < n; r += (uint64_t)gridDim.x * kT) {
        const T *a = A + r * lda;
        for (int j0 = 0; j0 < kc; j0 += JT) {
            double acc[JT];
#pragma unroll
            for (int j = 0; j < JT; j++) acc[j] = 0.0;
            int i = 0;
            if (vec) {
                for (; i + V <= ka; i += V) {
                    T av[V];
                    *(Vec16<T> *)av = *(const Vec16<T> *)(a + i);
#pragma unroll
                    for (int u = 0; u < V; u++) {
                        const double x = (double)av[u];
                        const double *c = Cs + (i + u) * kcp + j0;
#pragma unroll
                        for (int j = 0; j < JT; j++) acc[j] = fma(x, c[j], acc[j]);
                    }
                }
            }
            for (; i < ka; i++) {
                const double x = (double)a[i];
                const double *c = Cs + i * kcp + j0;
#pragma unroll
                for (int j = 0; j < JT; j++) acc[j] = fma(x, c[j], acc[j]);
            }
            TY *y = Y + r * ldy + j0;
            const int w = kc - j0 < JT ? kc - j0 : JT;
#pragma unroll
            for (int j = 0; j < JT; j++)
                if (j < w) {
                    const double prev = (beta == 0.0) ? 0.0 : beta * (double)y[j];
                    y[j] = (TY)(prev + alpha * acc[j]);
                }
        }
    }
}

// ---- column norms / sums of squares ----------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kT) colsq_kernel(const T *A, int64_t lda, uint64_t n, int k,
                                                   uint64_t rows_per_blk, double *partial)
{
    // thread t handles column t % k for rows t / k, t / k + kT / k, ...
    const int per = kT / k;
    const int c = threadIdx.x % k, rr = threadIdx.x / k;
    double s = 0.0;
    if (rr < per) {
        const uint64_t r0 = blockIdx.x * rows_per_blk;
        const uint64_t r1 = min(n, r0 + rows_per_blk);
        for (uint64_t r = r0 + rr; r < r1; r += per) {
            double v = ld64(A + r * lda + c);
            s = fma(v, v, s);
        }
    }
    __shared__ double red[kT];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < k) {
        double t = 0.0;
        for (int q = 0; q < per; q++) t += red[q * k + threadIdx.x];
        partial[(uint64_t)blockIdx.x * k + threadIdx.x] = t;
    }
}

template <typename T, typename TD>
__global__ void __launch_bounds__(kT) row_scale_kernel(uint64_t n, int k, const TD *d, const T *X,
                                                       int64_t ldx, T *Y, int64_t ldy)
{
    const uint64_t tot = n * (uint64_t)k;
    for (uint64_t e = blockIdx.x * (uint64_t)kT + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * kT) {
        const uint64_t r = e / k, c = e - r * k;
        Y[r * ldy + c] = (T)((TD)X[r * ldx + c] * d[r]);
    }
}

// Y = diag(d) X for FP32 rows of k = 4 * k4 values (k4 dividing 256): one
// float4 per thread, kT / k4 rows per block and step.
__global__ void __launch_bounds__(kT) row_scale_vec_kernel(uint64_t n, int k4, const double *d, const float *X,
                                                           int64_t ldx, float *Y, int64_t ldy)
{
    const int rpi = kT / k4, rr = threadIdx.x / k4, c = (threadIdx.x - rr * k4) * 4;
    for (uint64_t r = blockIdx.x * (uint64_t)rpi + rr; r < n; r += (uint64_t)gridDim.x * rpi) {
        const float4 x = *(const float4 *)(X + r * ldx + c);
        const double dr = d[r];
        float4 y;
        y.x = (float)((double)x.x * dr);
        y.y = (float)((double)x.y * dr);
        y.z = (float)((double)x.z * dr);
        y.w = (float)((double)x.w * dr);
        *(float4 *)(Y + r * ldy + c) = y;
    }
}

uint32_t blocks_for(fmx_ctx *ctx, uint64_t n, uint64_t rows_min)
{
    uint64_t b = (n + rows_min - 1) / rows_min;
    uint64_t cap = (uint64_t)ctx->num_cus * 4;
    if (b > cap) b = cap;
    return (uint32_t)(b ? b : 1);
}

}  // namespace

namespace {
// The fixed scratch of fmx_gram_noalloc: per-block partial sums of the Gram
// kernels.  Every GPU the library runs on holds its own copy until the
// process ends.  In host code the name is only a stand-in for the device
// array: kernels get the address that noalloc_scratch_lookup finds.
constexpr size_t kNoAllocBytes = size_t(4) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];
}  // namespace

fmx_status noalloc_scratch_lookup(fmx_ctx *ctx)
{
    void *p = nullptr;
    FMX_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ctx->noalloc_ws = p;
    ctx->noalloc_bytes = kNoAllocBytes;
    return FMX_OK;
}

fmx_status ctx_scratch(fmx_ctx *ctx, size_t bytes, void **out)
{
    if (bytes > ctx->ws_bytes) {
        if (ctx->ws) FMX_HIP(ctx, hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
        FMX_HIP(ctx, hipMalloc(&ctx->ws, bytes));
        ctx->ws_bytes = bytes;
    }
    *out = ctx->ws;
    return FMX_OK;
}

// Per-block partial sums of the Gram kernels: the context's grow-only
// workspace, or -- fixed != null -- a caller-given buffer of fixed_bytes,
// in which case the block count is capped so that the partials fit.
static fmx_status gram_ws(fmx_ctx *ctx, void *fixed, size_t fixed_bytes, size_t per_block, uint32_t &nblk,
                          void **ws)
{
    if (!fixed) return ctx_scratch(ctx, (size_t)nblk * per_block, ws);
    const size_t fit = fixed_bytes / per_block;
    if (fit == 0) return fail(ctx, FMX_EDIM, "Gram output larger than the fixed scratch");
    if (nblk > fit) nblk = (uint32_t)fit;
    *ws = fixed;
    return FMX_OK;
}

fmx_status gram(fmx_ctx *ctx, fmx_dtype dt, int mode, uint64_t n, int ka, int kb, const void *A,
                int64_t lda, const void *B, int64_t ldb, double *G, hipStream_t s, void *fixed_ws,
                size_t fixed_bytes)
{
    if (ka <= 0 || kb <= 0 || ka > 256 || kb > 256) return fail(ctx, FMX_EDIM, "gram widths must be in [1, 256]");
    if (lda < ka || ldb < kb) return fail(ctx, FMX_EDIM, "leading dimension too small");
    const int nout = ka * kb;
    if (mode == FMX_GRAM_BF16) {
        if (dt != FMX_F32) return fail(ctx, FMX_ETYPE, "BF16 Gram takes FP32 operands");
        const int kap = (ka + 15) & ~15, kbp = (kb + 15) & ~15;
        if ((kap / 16) * (kbp / 16) > 64) return fail(ctx, FMX_EDIM, "BF16 Gram output too large");
        uint32_t nblk = blocks_for(ctx, n, 1024);
        void *ws;
        fmx_status st = gram_ws(ctx, fixed_ws, fixed_bytes, (size_t)nout * sizeof(float), nblk, &ws);
        if (st) return st;
        const uint64_t rpb = (n + nblk - 1) / nblk;
        const size_t lds = (size_t)(kap + kbp) * kBPad * sizeof(unsigned short);
        static bool attr = false;
        if (!attr) {
            FMX_HIP(ctx, hipFuncSetAttribute((const void *)gram_bf16_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            FMX_HIP(ctx, hipFuncSetAttribute((const void *)gram_bf16_vec_kernel<false>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr = true;
        }
        const bool vec = ka % 4 == 0 && kb % 4 == 0 && kap <= 128 && kbp <= 32 && (uintptr_t)A % 16 == 0 &&
                         (uintptr_t)B % 16 == 0 && lda % 4 == 0 && ldb % 4 == 0;
        if (vec)
            hipLaunchKernelGGL(gram_bf16_vec_kernel<false>, dim3(nblk), dim3(kT), lds, s, A, lda,
                               (const float *)B, ldb, n, kap, kbp, ka, kb, rpb, (float *)ws);
        else
            hipLaunchKernelGGL(gram_bf16_kernel, dim3(nblk), dim3(kT), lds, s, (const float *)A, lda,
                               (const float *)B, ldb, n, kap, kbp, ka, kb, rpb, (float *)ws);
        FMX_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(reduce_partials_seg<float>, dim3((nout + 15) / 16), dim3(kT), 0, s, (const float *)ws,
                           (int)nblk, nout, G);
        FMX_HIP(ctx, hipGetLastError());
        return FMX_OK;
    }
    if (mode != FMX_GRAM_F64) return fail(ctx, FMX_EINVAL, "unknown Gram mode");
    const int tjn = (kb + 15) / 16, tin = (ka + 15) / 16;
    const int es = dt == FMX_F64 ? 8 : 4;
    const int tjp = tjn <= 1 ? 1 : tjn <= 2 ? 2 : 4;
    // row tiles of the launched kernel (3 and 6 exist for one column tile: a
    // 96-column Gram padded to 128 took 0.90 ms against 0.41 ms unpadded)
    const int tip = (tjp == 1 && (tin == 3 || tin == 5 || tin == 6)) ? (tin == 3 ? 3 : 6)
                    : tin <= 1 ? 1 : tin <= 2 ? 2 : tin <= 4 ? 4 : tin <= 8 ? 8 : 16;
    const size_t stage_lds = (size_t)4 * kStageRows * (tip * 16 + 16 + tjp * 16 + 16) * es;
    if (n == 0) {  // an empty sum: every entry 0
        hipLaunchKernelGGL(reduce_partials4, dim3((nout + 63) / 64), dim3(kT), 0, s, (const double *)nullptr, 0,
                           nout, G);
        FMX_HIP(ctx, hipGetLastError());
        return FMX_OK;
    }
    if (tjn <= 4 && tip * tjp <= kGramTiles && std::max(stage_lds, (size_t)2 * nout * sizeof(double)) <= 96 * 1024 &&
        (dt == FMX_F32 || dt == FMX_F64)) {
        // A^T A with one staged operand (the Cholesky-QR Grams)
        const bool sym = A == B && lda == ldb && ka == kb && tip == tjp && tip <= 2;
        // FP32 A^T A: the three-waves-per-SIMD kernel, up to 12 waves per CU
        // (1.2x on the eigensolver's 16- and 32-column Grams); otherwise at
        // most 8 per CU (the 48- to 96-row products ran 5-8% slower with three
        // waves, and more waves only add partials to reduce)
        const bool occ3 = dt == FMX_F32 && sym;
        // waves of at least 1024 rows (a multiple of 16)
        uint64_t waves = (n + 1023) / 1024;
        uint64_t cap = (uint64_t)ctx->num_cus * (occ3 ? 12 : 8);
        if (fixed_ws) {
            const uint64_t fit = fixed_bytes / ((size_t)nout * sizeof(double));  // blocks whose partials fit
            if (fit == 0) return fail(ctx, FMX_EDIM, "Gram output larger than the fixed scratch");
            cap = std::min<uint64_t>(cap, fit * (kT / 64));
        }
        if (waves > cap) waves = cap;
        if (waves < 1) waves = 1;
        const uint64_t rpw = ((n + waves - 1) / waves + 15) & ~15ull;
        waves = (n + rpw - 1) / rpw;
        if (waves < 1) waves = 1;
        uint32_t nblk = (uint32_t)((waves * 64 + kT - 1) / kT);  // waves past the rows add zeros
        void *ws;
        fmx_status st = gram_ws(ctx, fixed_ws, fixed_bytes, (size_t)nout * sizeof(double), nblk, &ws);
        if (st) return st;
        const size_t glds = std::max(stage_lds, (size_t)2 * nout * sizeof(double));
        const int veca = ((uintptr_t)A % 16 == 0) && ((lda * es) % 16 == 0);
        const int vecb = ((uintptr_t)B % 16 == 0) && ((ldb * es) % 16 == 0);
        static bool attr_g = false;
        if (!attr_g) {
            const void *ks[] = {
                (const void *)gram_stage_kernel<float, 1, 1, true>, (const void *)gram_stage_kernel<float, 2, 2, true>,
                (const void *)gram_stage_kernel<double, 1, 1, true>, (const void *)gram_stage_kernel<double, 2, 2, true>,
                (const void *)gram_stage_kernel<float, 3, 1>, (const void *)gram_stage_kernel<float, 6, 1>,
                (const void *)gram_stage_kernel<double, 3, 1>, (const void *)gram_stage_kernel<double, 6, 1>,
                (const void *)gram_stage_kernel<float, 1, 1>, (const void *)gram_stage_kernel<float, 2, 1>,
                (const void *)gram_stage_kernel<float, 4, 1>, (const void *)gram_stage_kernel<float, 8, 1>,
                (const void *)gram_stage_kernel<float, 16, 1>, (const void *)gram_stage_kernel<float, 1, 2>,
                (const void *)gram_stage_kernel<float, 2, 2>, (const void *)gram_stage_kernel<float, 4, 2>,
                (const void *)gram_stage_kernel<float, 8, 2>, (const void *)gram_stage_kernel<float, 1, 4>,
                (const void *)gram_stage_kernel<float, 2, 4>, (const void *)gram_stage_kernel<float, 4, 4>,
                (const void *)gram_stage_kernel<double, 1, 1>, (const void *)gram_stage_kernel<double, 2, 1>,
                (const void *)gram_stage_kernel<double, 4, 1>, (const void *)gram_stage_kernel<double, 8, 1>,
                (const void *)gram_stage_kernel<double, 16, 1>, (const void *)gram_stage_kernel<double, 1, 2>,
                (const void *)gram_stage_kernel<double, 2, 2>, (const void *)gram_stage_kernel<double, 4, 2>,
                (const void *)gram_stage_kernel<double, 8, 2>, (const void *)gram_stage_kernel<double, 1, 4>,
                (const void *)gram_stage_kernel<double, 2, 4>, (const void *)gram_stage_kernel<double, 4, 4>,
                (const void *)gram_stage_kernel3<float, 1, 1, true>, (const void *)gram_stage_kernel3<float, 2, 2, true>};
            for (const void *k : ks)
                FMX_HIP(ctx, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
            attr_g = true;
        }
#define FMX_GRAMS(TT, I, J)                                                                                   \
    hipLaunchKernelGGL((gram_stage_kernel<TT, I, J>), dim3(nblk), dim3(kT), glds, s, (const TT *)A, lda,        \
                       (const TT *)B, ldb, n, ka, kb, rpw, waves, veca, vecb, (double *)ws)
#define FMX_GRAMS_I(TT, J)                                                                                    \
    do {                                                                                                      \
        if (tip == 1) FMX_GRAMS(TT, 1, J);                                                                    \
        else if (tip == 2) FMX_GRAMS(TT, 2, J);                                                               \
        else if (tip == 3 && J == 1) FMX_GRAMS(TT, 3, 1);                                                     \
        else if (tip == 4) FMX_GRAMS(TT, 4, J);                                                               \
        else if (tip == 6 && J == 1) FMX_GRAMS(TT, 6, 1);                                                     \
        else if (J == 4) FMX_GRAMS(TT, 4, 4);                                                                 \
        else if (tip == 8 || J == 2) FMX_GRAMS(TT, 8, (J == 2 ? 2 : 1));                                      \
        else FMX_GRAMS(TT, 16, 1);                                                                            \
    } while (0)
        if (sym && dt == FMX_F32) {
            if (tip == 1) hipLaunchKernelGGL((gram_stage_kernel3<float, 1, 1, true>), dim3(nblk), dim3(kT), glds, s,
                                             (const float *)A, lda, (const float *)B, ldb, n, ka, kb, rpw, waves,
                                             veca, vecb, (double *)ws);
            else hipLaunchKernelGGL((gram_stage_kernel3<float, 2, 2, true>), dim3(nblk), dim3(kT), glds, s,
                                    (const float *)A, lda, (const float *)B, ldb, n, ka, kb, rpw, waves, veca,
                                    vecb, (double *)ws);
        } else if (sym) {
            if (tip == 1) hipLaunchKernelGGL((gram_stage_kernel<double, 1, 1, true>), dim3(nblk), dim3(kT), glds, s,
                                             (const double *)A, lda, (const double *)B, ldb, n, ka, kb, rpw, waves,
                                             veca, vecb, (double *)ws);
            else hipLaunchKernelGGL((gram_stage_kernel<double, 2, 2, true>), dim3(nblk), dim3(kT), glds, s,
                                    (const double *)A, lda, (const double *)B, ldb, n, ka, kb, rpw, waves, veca,
                                    vecb, (double *)ws);
        } else if (dt == FMX_F32) {
            if (tjp == 1) FMX_GRAMS_I(float, 1);
            else if (tjp == 2) FMX_GRAMS_I(float, 2);
            else FMX_GRAMS_I(float, 4);
        } else {
            if (tjp == 1) FMX_GRAMS_I(double, 1);
            else if (tjp == 2) FMX_GRAMS_I(double, 2);
            else FMX_GRAMS_I(double, 4);
        }
#undef FMX_GRAMS_I
#undef FMX_GRAMS
        FMX_HIP(ctx, hipGetLastError());
        if (nblk > 256)  // many partials: 16 segments per output instead of 4
            hipLaunchKernelGGL(reduce_partials_seg<double>, dim3((nout + 15) / 16), dim3(kT), 0, s,
                               (const double *)ws, (int)nblk, nout, G);
        else
            hipLaunchKernelGGL(reduce_partials4, dim3((nout + 63) / 64), dim3(kT), 0, s, (const double *)ws,
                               (int)nblk, nout, G);
        FMX_HIP(ctx, hipGetLastError());
        return FMX_OK;
    }
    uint32_t nblk = blocks_for(ctx, n, 512);
    void *ws;
    fmx_status st = gram_ws(ctx, fixed_ws, fixed_bytes, (size_t)nout * sizeof(double), nblk, &ws);
    if (st) return st;
    const uint64_t rpb = (n + nblk - 1) / nblk;
    const size_t lds = (size_t)kGRows * (ka + kb) * sizeof(double);
    dim3 grid(nblk, (nout + 4095) / 4096);
    if (dt == FMX_F32)
        hipLaunchKernelGGL(gram64_kernel<float>, grid, dim3(kT), lds, s, (const float *)A, lda,
                           (const float *)B, ldb, n, ka, kb, rpb, (double *)ws);
    else if (dt == FMX_F64)
        hipLaunchKernelGGL(gram64_kernel<double>, grid, dim3(kT), lds, s, (const double *)A, lda,
                           (const double *)B, ldb, n, ka, kb, rpb, (double *)ws);
    else
        return fail(ctx, FMX_ETYPE, "Gram takes FP32 or FP64 operands");
    FMX_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL((reduce_partials<double, double>), dim3((nout + kT - 1) / kT), dim3(kT), 0, s,
                       (const double *)ws, (int)nblk, nout, G, 0);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

// Staged MFMA tall GEMM (kc <= 32, ka <= 128): 16-row slices of A in LDS,
// C in registers; A may be Y.
template <typename T, typename TY>
fmx_status gemm_staged(fmx_ctx *ctx, uint64_t n, int ka, int kc, double alpha, const T *A, int64_t lda,
                       const double *C, double beta, TY *Y, int64_t ldy, hipStream_t s,
                       unsigned short *Yb = nullptr, int64_t ldyb = 0, const double *ds = nullptr,
                       TY *Ys = nullptr, int64_t ldys = 0)
{
    if (n == 0) return FMX_OK;
    const int kap = ka <= 16 ? 16 : ka <= 32 ? 32 : ka <= 64 ? 64 : ka <= 96 ? 96 : 128;
    const size_t lds = (size_t)4 * kStageRows * (kap + (sizeof(T) == 4 ? 4 : 1)) * sizeof(T);
    uint64_t waves = (n + 255) / 256;
    // up to 32 waves per CU (LDS permitting): each wave has one 16-row stage
    // in flight, so the grid must be wide to keep enough rows in flight
    const uint64_t cap = (uint64_t)ctx->num_cus * 32;
    if (waves > cap) waves = cap;
    if (waves < 1) waves = 1;
    const uint64_t rpw = ((n + waves - 1) / waves + 15) & ~15ull;
    waves = (n + rpw - 1) / rpw;
    if (waves < 1) waves = 1;
    const uint32_t nblk = (uint32_t)((waves * 64 + kT - 1) / kT);
    const int veca = ((uintptr_t)A % 16 == 0) && ((lda * (int64_t)sizeof(T)) % 16 == 0);
    static bool attr = false;
    if (!attr) {
        const void *ks[] = {(const void *)gemm_stage_kernel<T, TY, 16, 1>, (const void *)gemm_stage_kernel<T, TY, 32, 1>,
                            (const void *)gemm_stage_kernel<T, TY, 64, 1>, (const void *)gemm_stage_kernel<T, TY, 96, 1>,
                            (const void *)gemm_stage_kernel<T, TY, 128, 1>, (const void *)gemm_stage_kernel<T, TY, 16, 2>,
                            (const void *)gemm_stage_kernel<T, TY, 32, 2>, (const void *)gemm_stage_kernel<T, TY, 64, 2>,
                            (const void *)gemm_stage_kernel<T, TY, 96, 2>, (const void *)gemm_stage_kernel<T, TY, 128, 2>};
        for (const void *k : ks)
            FMX_HIP(ctx, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
        if constexpr (sizeof(T) == 4) {
            const void *k3[] = {(const void *)gemm_stage_kernel3<T, TY, 16, 1, false, -1, -1>,
                                (const void *)gemm_stage_kernel3<T, TY, 32, 1, false, -1, -1>,
                                (const void *)gemm_stage_kernel3<T, TY, 64, 1, false, -1, -1>,
                                (const void *)gemm_stage_kernel3<T, TY, 16, 1, true, 0, 1>,
                                (const void *)gemm_stage_kernel3<T, TY, 32, 1, true, 0, 1>,
                                (const void *)gemm_stage_kernel3<T, TY, 64, 1, true, 0, 1>,
                                (const void *)gemm_stage_kernel3<T, TY, 96, 1, true, 0, 1>,
                                (const void *)gemm_stage_kernel3<T, TY, 16, 1, true, 0, 0>,
                                (const void *)gemm_stage_kernel3<T, TY, 32, 1, true, 0, 0>,
                                (const void *)gemm_stage_kernel3<T, TY, 64, 1, true, 0, 0>,
                                (const void *)gemm_stage_kernel3<T, TY, 96, 1, true, 0, 0>,
                                (const void *)gemm_stage_kernel3<T, TY, 16, 1, true, 1, 0>,
                                (const void *)gemm_stage_kernel3<T, TY, 32, 1, true, 1, 0>,
                                (const void *)gemm_stage_kernel3<T, TY, 64, 1, true, 1, 0>,
                                (const void *)gemm_stage_kernel3<T, TY, 96, 1, true, 1, 0>};
            for (const void *k : k3)
                FMX_HIP(ctx, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
            const void *k2v[] = {(const void *)gemm_stage_kernel<T, TY, 16, 2, true, 0, 0>,
                                 (const void *)gemm_stage_kernel<T, TY, 32, 2, true, 0, 0>,
                                 (const void *)gemm_stage_kernel<T, TY, 64, 2, true, 0, 0>,
                                 (const void *)gemm_stage_kernel<T, TY, 96, 2, true, 0, 0>,
                                 (const void *)gemm_stage_kernel<T, TY, 128, 2, true, 0, 0>};
            for (const void *k : k2v)
                FMX_HIP(ctx, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
        }
        attr = true;
    }
#define FMX_GEMMS(K, P, J)                                                                                    \
    hipLaunchKernelGGL((K<T, TY, P, J>), dim3(nblk), dim3(kT), lds, s, A, lda, n, ka, kc, C, alpha, beta, Y, ldy, \
                       rpw, waves, veca, Yb, ldyb, ds, Ys, ldys)
#define FMX_GEMMS3(P, VEC, SC, OLD)                                                                           \
    hipLaunchKernelGGL((gemm_stage_kernel3<T, TY, P, 1, VEC, SC, OLD>), dim3(nblk), dim3(kT), lds, s, A, lda, n, \
                       ka, kc, C, alpha, beta, Y, ldy, rpw, waves, veca, Yb, ldyb, ds, Ys, ldys)
#define FMX_GEMMS3W(SC, OLD)                                                                                  \
    do {                                                                                                      \
        if (kap == 16) FMX_GEMMS3(16, true, SC, OLD);                                                         \
        else if (kap == 32) FMX_GEMMS3(32, true, SC, OLD);                                                    \
        else if (kap == 64) FMX_GEMMS3(64, true, SC, OLD);                                                    \
        else FMX_GEMMS3(96, true, SC, OLD);                                                                   \
    } while (0)
    // FP32 rows and one column tile: the three-waves-per-SIMD form up to 96
    // columns, with whole-row vector loads when the rows allow them
    bool done = false;
    if constexpr (sizeof(T) == 4) {
        if (kc <= 16 && kap <= 96) {
            // whole-row vector loads, with the row-scaled copy and the old
            // values' loads fixed at compile time: Y = A C + beta Y, Y = A C,
            // and Y = A C with its scaled copy
            const bool vec = veca && ka % 4 == 0;
            if (vec && !Ys && beta != 0.0) {
                FMX_GEMMS3W(0, 1);
                done = true;
            } else if (vec && !Ys) {
                FMX_GEMMS3W(0, 0);
                done = true;
            } else if (vec && beta == 0.0) {
                FMX_GEMMS3W(1, 0);
                done = true;
            } else if (kap <= 64) {  // (the 96-column form of these loads would spill)
                if (kap == 16) FMX_GEMMS3(16, false, -1, -1);
                else if (kap == 32) FMX_GEMMS3(32, false, -1, -1);
                else FMX_GEMMS3(64, false, -1, -1);
                done = true;
            }
        }
    }
    // FP32 rows, two column tiles, beta = 0, no scaled copy (the thick
    // restart's V = V S): whole-row vector loads as above
    if constexpr (sizeof(T) == 4) {
        if (!done && kc > 16 && veca && ka % 4 == 0 && beta == 0.0 && !Ys) {
#define FMX_GEMMS2V(P)                                                                                        \
    hipLaunchKernelGGL((gemm_stage_kernel<T, TY, P, 2, true, 0, 0>), dim3(nblk), dim3(kT), lds, s, A, lda, n, ka, \
                       kc, C, alpha, beta, Y, ldy, rpw, waves, veca, Yb, ldyb, ds, Ys, ldys)
            if (kap == 16) FMX_GEMMS2V(16);
            else if (kap == 32) FMX_GEMMS2V(32);
            else if (kap == 64) FMX_GEMMS2V(64);
            else if (kap == 96) FMX_GEMMS2V(96);
            else FMX_GEMMS2V(128);
#undef FMX_GEMMS2V
            done = true;
        }
    }
    if (!done) {
#define FMX_GEMMS_P(J)                                                                                        \
    do {                                                                                                      \
        if (kap == 16) FMX_GEMMS(gemm_stage_kernel, 16, J);                                                   \
        else if (kap == 32) FMX_GEMMS(gemm_stage_kernel, 32, J);                                              \
        else if (kap == 64) FMX_GEMMS(gemm_stage_kernel, 64, J);                                              \
        else if (kap == 96) FMX_GEMMS(gemm_stage_kernel, 96, J);                                              \
        else FMX_GEMMS(gemm_stage_kernel, 128, J);                                                            \
    } while (0)
        if (kc <= 16) FMX_GEMMS_P(1);
        else FMX_GEMMS_P(2);
#undef FMX_GEMMS_P
    }
#undef FMX_GEMMS
#undef FMX_GEMMS3
#undef FMX_GEMMS3W
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status gram_bf16_shadow(fmx_ctx *ctx, uint64_t n, int ka, int kb, const unsigned short *Ab, int64_t lda,
                            const float *B, int64_t ldb, double *G, hipStream_t s)
{
    const int kap = (ka + 15) & ~15, kbp = (kb + 15) & ~15;
    if (ka <= 0 || kb <= 0 || ka % 8 || kb % 4 || kap > 128 || kbp > 32 || lda < ka || ldb < kb || lda % 8 ||
        ldb % 4 || (uintptr_t)Ab % 16 || (uintptr_t)B % 16)
        return fail(ctx, FMX_EDIM, "BF16-operand Gram: widths, strides or alignment unsupported");
    const int nout = ka * kb;
    uint32_t nblk = blocks_for(ctx, n, 1024);
    void *ws;
    fmx_status st = ctx_scratch(ctx, (size_t)nblk * nout * sizeof(float), &ws);
    if (st) return st;
    const uint64_t rpb = (n + nblk - 1) / nblk;
    const size_t lds = (size_t)(kap + kbp) * kBPad * sizeof(unsigned short);
    static bool attr = false;
    if (!attr) {
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)gram_bf16_vec_kernel<true>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    hipLaunchKernelGGL(gram_bf16_vec_kernel<true>, dim3(nblk), dim3(kT), lds, s, (const void *)Ab, lda, B, ldb, n,
                       kap, kbp, ka, kb, rpb, (float *)ws);
    FMX_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(reduce_partials_seg<float>, dim3((nout + 15) / 16), dim3(kT), 0, s, (const float *)ws,
                       (int)nblk, nout, G);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status gemm_tall_copies(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int ka, int kc, double alpha, const void *A,
                            int64_t lda, const double *C, double beta, void *Y, int64_t ldy, unsigned short *Yb,
                            int64_t ldyb, const double *ds, void *Ys, int64_t ldys, hipStream_t s)
{
    if (ka <= 0 || kc <= 0 || ka > 128 || kc > 32 || lda < ka || ldy < kc || (Yb && ldyb < kc) ||
        (Ys && (ldys < kc || !ds)))
        return fail(ctx, FMX_EDIM, "gemm_tall with copies needs kc <= 32, ka <= 128");
    if (dt == FMX_F32)
        return gemm_staged<float, float>(ctx, n, ka, kc, alpha, (const float *)A, lda, C, beta, (float *)Y, ldy, s,
                                         Yb, ldyb, ds, (float *)Ys, ldys);
    if (dt == FMX_F64 && !Yb)
        return gemm_staged<double, double>(ctx, n, ka, kc, alpha, (const double *)A, lda, C, beta, (double *)Y,
                                           ldy, s, nullptr, 0, ds, (double *)Ys, ldys);
    return fail(ctx, FMX_ETYPE, "gemm_tall with copies: FP32, or FP64 without a BF16 copy");
}

fmx_status gemm_tall_f32_f64(fmx_ctx *ctx, uint64_t n, int ka, int kc, double alpha, const float *A,
                             int64_t lda, const double *C, double beta, double *Y, int64_t ldy,
                             hipStream_t s)
{
    if (ka <= 0 || kc <= 0 || (size_t)ka * kc > 8192 || ka > 512)
        return fail(ctx, FMX_EDIM, "gemm_tall needs ka*kc <= 8192");
    if (kc <= 32 && ka <= 128)
        return gemm_staged<float, double>(ctx, n, ka, kc, alpha, A, lda, C, beta, Y, ldy, s);
    const size_t lds = ((size_t)ka * kc + (size_t)kMRows * ka) * sizeof(double);
    const uint32_t nblk = blocks_for(ctx, n, kMRows);
    static bool attr = false;
    if (!attr) {
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)gemm_tall_kernel<float, double>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    hipLaunchKernelGGL((gemm_tall_kernel<float, double>), dim3(nblk), dim3(kT), lds, s, A, lda, n, ka, kc,
                       C, alpha, beta, Y, ldy);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status gemm_tall(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int ka, int kc, double alpha,
                     const void *A, int64_t lda, const double *C, double beta, void *Y, int64_t ldy,
                     hipStream_t s)
{
    if (ka <= 0 || kc <= 0 || (size_t)ka * kc > 8192 || ka > 512)
        return fail(ctx, FMX_EDIM, "gemm_tall needs ka*kc <= 8192");
    if (lda < ka || ldy < kc) return fail(ctx, FMX_EDIM, "leading dimension too small");
    if (dt != FMX_F32 && dt != FMX_F64) return fail(ctx, FMX_ETYPE, "gemm_tall takes FP32 or FP64 operands");
    if (kc <= 32 && ka <= 128) {
        if (dt == FMX_F32)
            return gemm_staged<float, float>(ctx, n, ka, kc, alpha, (const float *)A, lda, C, beta, (float *)Y, ldy, s);
        return gemm_staged<double, double>(ctx, n, ka, kc, alpha, (const double *)A, lda, C, beta, (double *)Y,
                                           ldy, s);
    }
    const bool in_place = (const void *)A == (const void *)Y;
    const int jt = kc <= 8 ? 8 : (kc <= 16 ? 16 : 32);
    if (!in_place || kc <= jt) {
        const int kcp = (kc + jt - 1) / jt * jt;
        const size_t lds_r = (size_t)ka * kcp * sizeof(double);
        if (lds_r <= 64 * 1024) {
            const int es = dt == FMX_F32 ? 4 : 8;
            const int vec = ((uintptr_t)A % 16 == 0) && ((lda * es) % 16 == 0);
            const uint32_t nblk = blocks_for(ctx, n, kT);
            static bool attr_r = false;
            if (!attr_r) {
                const void *ks[] = {(const void *)gemm_rows_kernel<float, float, 8>,
                                    (const void *)gemm_rows_kernel<float, float, 16>,
                                    (const void *)gemm_rows_kernel<float, float, 32>,
                                    (const void *)gemm_rows_kernel<double, double, 8>,
                                    (const void *)gemm_rows_kernel<double, double, 16>,
                                    (const void *)gemm_rows_kernel<double, double, 32>};
                for (const void *k : ks)
                    FMX_HIP(ctx, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024));
                attr_r = true;
            }
#define FMX_GEMM_ROWS(TT, J)                                                                                  \
    hipLaunchKernelGGL((gemm_rows_kernel<TT, TT, J>), dim3(nblk), dim3(kT), lds_r, s, (const TT *)A, lda, n, ka, kc, \
                       kcp, C, alpha, beta, (TT *)Y, ldy, vec)
            if (dt == FMX_F32) {
                if (jt == 8) FMX_GEMM_ROWS(float, 8);
                else if (jt == 16) FMX_GEMM_ROWS(float, 16);
                else FMX_GEMM_ROWS(float, 32);
            } else {
                if (jt == 8) FMX_GEMM_ROWS(double, 8);
                else if (jt == 16) FMX_GEMM_ROWS(double, 16);
                else FMX_GEMM_ROWS(double, 32);
            }
#undef FMX_GEMM_ROWS
            FMX_HIP(ctx, hipGetLastError());
            return FMX_OK;
        }
    }
    const size_t lds = ((size_t)ka * kc + (size_t)kMRows * ka) * sizeof(double);
    const uint32_t nblk = blocks_for(ctx, n, kMRows);
    static bool attr = false;
    if (!attr) {
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)gemm_tall_kernel<float>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        FMX_HIP(ctx, hipFuncSetAttribute((const void *)gemm_tall_kernel<double>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    if (dt == FMX_F32)
        hipLaunchKernelGGL(gemm_tall_kernel<float>, dim3(nblk), dim3(kT), lds, s, (const float *)A, lda, n,
                           ka, kc, C, alpha, beta, (float *)Y, ldy);
    else if (dt == FMX_F64)
        hipLaunchKernelGGL(gemm_tall_kernel<double>, dim3(nblk), dim3(kT), lds, s, (const double *)A, lda,
                           n, ka, kc, C, alpha, beta, (double *)Y, ldy);
    else
        return fail(ctx, FMX_ETYPE, "gemm_tall takes FP32 or FP64 operands");
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status colnorm(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int k, const void *A, int64_t lda,
                   double *norms, int squared, hipStream_t s)
{
    if (k <= 0 || k > kT) return fail(ctx, FMX_EDIM, "colnorm width must be in [1, 256]");
    if (lda < k) return fail(ctx, FMX_EDIM, "leading dimension too small");
    const uint32_t nblk = blocks_for(ctx, n, 2048);
    const uint64_t rpb = (n + nblk - 1) / nblk;
    void *ws;
    fmx_status st = ctx_scratch(ctx, (size_t)nblk * k * sizeof(double), &ws);
    if (st) return st;
    if (dt == FMX_F32)
        hipLaunchKernelGGL(colsq_kernel<float>, dim3(nblk), dim3(kT), 0, s, (const float *)A, lda, n, k,
                           rpb, (double *)ws);
    else if (dt == FMX_F64)
        hipLaunchKernelGGL(colsq_kernel<double>, dim3(nblk), dim3(kT), 0, s, (const double *)A, lda, n, k,
                           rpb, (double *)ws);
    else
        return fail(ctx, FMX_ETYPE, "colnorm takes FP32 or FP64 operands");
    FMX_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL((reduce_partials<double, double>), dim3(1), dim3(kT), 0, s, (const double *)ws,
                       (int)nblk, k, norms, squared ? 0 : 1);
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

fmx_status row_scale(fmx_ctx *ctx, fmx_dtype dt, uint64_t n, int k, const double *d, const void *X,
                     int64_t ldx, void *Y, int64_t ldy, hipStream_t s)
{
    if (k <= 0 || ldx < k || ldy < k) return fail(ctx, FMX_EDIM, "bad width or leading dimension");
    uint64_t tot = n * (uint64_t)k;
    uint64_t blocks = (tot + kT - 1) / kT;
    if (blocks > (uint64_t)ctx->num_cus * 16) blocks = (uint64_t)ctx->num_cus * 16;
    if (!blocks) return FMX_OK;
    if (dt == FMX_F32 && k % 4 == 0 && kT % (k / 4) == 0 && ldx % 4 == 0 && ldy % 4 == 0 &&
        (uintptr_t)X % 16 == 0 && (uintptr_t)Y % 16 == 0) {
        const uint64_t rpi = kT / (k / 4);
        const uint64_t vb = std::min<uint64_t>((n + rpi - 1) / rpi, (uint64_t)ctx->num_cus * 16);
        hipLaunchKernelGGL(row_scale_vec_kernel, dim3((uint32_t)vb), dim3(kT), 0, s, n, k / 4, d, (const float *)X,
                           ldx, (float *)Y, ldy);
    } else if (dt == FMX_F32)
        hipLaunchKernelGGL((row_scale_kernel<float, double>), dim3((uint32_t)blocks), dim3(kT), 0, s, n, k,
                           d, (const float *)X, ldx, (float *)Y, ldy);
    else if (dt == FMX_F64)
        hipLaunchKernelGGL((row_scale_kernel<double, double>), dim3((uint32_t)blocks), dim3(kT), 0, s, n,
                           k, d, (const double *)X, ldx, (double *)Y, ldy);
    else
        return fail(ctx, FMX_ETYPE, "row_scale takes FP32 or FP64 operands");
    FMX_HIP(ctx, hipGetLastError());
    return FMX_OK;
}

}  // namespace fmx

using namespace fmx;

extern "C" {

fmx_status fmx_gram(fmx_ctx *ctx, fmx_dtype dtype, int mode, uint64_t n, int ka, int kb,
                    const void *A, int64_t lda, const void *B, int64_t ldb, double *G, void *stream)
{
    if (!ctx || (n && (!A || !B)) || !G) return fail(ctx, FMX_EINVAL, "null argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    return gram(ctx, dtype, mode, n, ka, kb, A, lda, B, ldb, G,
                (hipStream_t)stream);
}

fmx_status fmx_gram_noalloc(fmx_ctx *ctx, fmx_dtype dtype, int mode, uint64_t n, int ka, int kb, const void *A,
                            int64_t lda, const void *B, int64_t ldb, double *G, void *stream)
{
    if (!ctx || (n && (!A || !B)) || !G) return fail(ctx, FMX_EINVAL, "null argument");
    if (!ctx->noalloc_ws) return fail(ctx, FMX_EINVAL, "the context has no fixed scratch");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    return gram(ctx, dtype, mode, n, ka, kb, A, lda, B, ldb, G, (hipStream_t)stream, ctx->noalloc_ws,
                ctx->noalloc_bytes);
}

fmx_status fmx_gemm_tall(fmx_ctx *ctx, fmx_dtype dtype, uint64_t n, int ka, int kc, double alpha,
                         const void *A, int64_t lda, const double *C, double beta, void *Y,
                         int64_t ldy, void *stream)
{
    if (!ctx || (n && (!A || !Y)) || !C) return fail(ctx, FMX_EINVAL, "null argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    return gemm_tall(ctx, dtype, n, ka, kc, alpha, A, lda, C, beta, Y, ldy,
                     (hipStream_t)stream);
}

fmx_status fmx_colnorm(fmx_ctx *ctx, fmx_dtype dtype, uint64_t n, int k, const void *A,
                       int64_t lda, double *norms, void *stream)
{
    if (!ctx || (n && !A) || !norms) return fail(ctx, FMX_EINVAL, "null argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    return colnorm(ctx, dtype, n, k, A, lda, norms, 0, (hipStream_t)stream);
}

fmx_status fmx_row_scale(fmx_ctx *ctx, fmx_dtype dtype, uint64_t n, int k, const double *d,
                         const void *X, int64_t ldx, void *Y, int64_t ldy, void *stream)
{
    if (!ctx || (n && (!X || !Y || !d))) return fail(ctx, FMX_EINVAL, "null argument");
    FMX_HIP(ctx, hipSetDevice(ctx->device));
    return row_scale(ctx, dtype, n, k, d, X, ldx, Y, ldy, (hipStream_t)stream);
}

}  // extern "C"
