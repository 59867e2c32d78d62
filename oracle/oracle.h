/*
 * oracle.h -- CPU restatement of FlashMatrix's 2D-tiled SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker and the CPU
 * baseline: it is loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, and by nothing else.  The product library
 * (flashx_amd/lib/libfmx.so) never links, loads or calls it.
 *
 * Each function cites the reference file:line whose behaviour it restates.
 * Reference paths are relative to the FlashX source tree.
 */
#ifndef FMX_ORACLE_H
#define FMX_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (mirror the product's fmx_status numbering) ---------- */
#define ORC_OK       0
#define ORC_EINVAL   1
#define ORC_ETYPE    2
#define ORC_EDIM     3
#define ORC_EFORMAT  4
#define ORC_ENOMEM   5

/* ---- dtypes ------------------------------------------------------------ */
#define ORC_I32  0
#define ORC_I64  1
#define ORC_F32  2
#define ORC_F64  3

/* ---- build flags ------------------------------------------------------- */
#define ORC_DEDUP      1u  /* fm_utils.cpp:37 set_deduplicate(true)          */
#define ORC_KEEP_SELF  2u  /* fm_utils.cpp:39 set_remove_self_edge(false)    */
#define ORC_SYMMETRIZE 4u  /* add (dst,src) for every (src,dst)              */

/* ---- counter-based RNG (SplitMix64 output function) -------------------- */
uint64_t orc_splitmix64(uint64_t key, uint64_t ctr);

/* ---- seeded generators -------------------------------------------------- */
/* R-MAT edge list; see the generator spec in the .c file. */
void orc_rmat_edges(int scale, uint64_t m, uint64_t seed,
                    uint32_t thr_a, uint32_t thr_ab, uint32_t thr_abc,
                    int permute, uint32_t *src, uint32_t *dst);
uint32_t orc_rmat_perm(uint32_t v, int scale, uint64_t seed);

/* Degree-corrected SBM edge list (project spec, see the .c file). */
int orc_dcsbm_edges(uint64_t n, int scale, int nblocks, const uint32_t *bstart,
                    const uint32_t *wtab, uint32_t ratio, uint64_t m, uint64_t seed,
                    uint32_t *src, uint32_t *dst);

/* ---- tile construction (fm_utils.cpp:617-841) ---------------------------- */
int orc_build_2d(const uint32_t *src, const uint32_t *dst, uint64_t m,
                 uint64_t n, int log2h, int log2w, unsigned flags,
                 uint8_t **image, uint64_t *image_bytes,
                 int64_t **offs, uint64_t *nbr, uint64_t *n_out,
                 uint64_t *nnz_out);
/* Weighted form: esz = 4 or 8 value bytes per edge, prim the header's
 * prim_type (4 P_INTEGER, 6 P_LONG, 8 P_FLOAT, 9 P_DOUBLE). */
int orc_build_2d_w(const uint32_t *src, const uint32_t *dst, const void *vals, uint64_t esz, int prim,
                   uint64_t m, uint64_t n, int log2h, int log2w, unsigned flags,
                   uint8_t **image_out, uint64_t *image_bytes,
                   int64_t **offs_out, uint64_t *nbr_out, uint64_t *n_out, uint64_t *nnz_out);
void orc_free(void *p);

/* Sorted, normalised edge list (the rows fed to the tile writer). */
int orc_normalize(const uint32_t *src, const uint32_t *dst, uint64_t m,
                  uint64_t n, unsigned flags, uint64_t **keys,
                  uint64_t *nkeys, uint64_t *n_out);

/* ---- tile decoding / checks (sparse_matrix_format.cpp:36-88,128-136) ---- */
int orc_verify_image(const uint8_t *image, uint64_t bytes,
                     const int64_t *offs, uint64_t nbr);
/* Every non-zero in image order (per tile: SCSR runs, then COO). */
int orc_image_nonzeros(const uint8_t *image, uint64_t bytes,
                       const int64_t *offs, uint64_t nbr,
                       uint64_t *rows, uint64_t *cols, uint64_t cap,
                       uint64_t *count);

/* ---- SpMM (sparse_matrix.h:219-375, sparse_matrix.cpp:197-393) ---------- */
/* Y = A*X in the reference's default order (seq_exec_order). */
int orc_spmm(const uint8_t *image, uint64_t bytes, const int64_t *offs,
             uint64_t nbr, int dtype, int k, const void *X, int64_t ldx,
             void *Y, int64_t ldy);
/* Multi-threaded restatement of the reference engine, timed as baseline. */
/* Hilbert position of (x, y) in an n x n grid (hilbert_curve.cpp:24-34). */
uint64_t orc_hilbert_xy2d(uint64_t n, uint64_t x, uint64_t y);
/* Sequential SpMM with tasks of task_rows block rows whose super blocks are
 * visited in seq (hilbert = 0) or Hilbert order (sparse_matrix.cpp:120-314). */
int orc_spmm_exec(const uint8_t *image, uint64_t bytes, const int64_t *offs,
                  uint64_t nbr, int dtype, int k, const void *X, int64_t ldx,
                  void *Y, int64_t ldy, uint64_t task_rows, int hilbert);
int orc_spmm_range(const uint8_t *image, uint64_t bytes, const int64_t *offs,
                   uint64_t nbr, uint64_t br0, uint64_t br1, int dtype, int k,
                   const void *X, int64_t ldx, void *Y, int64_t ldy);
int orc_spmm_mt(const uint8_t *image, uint64_t bytes, const int64_t *offs,
                uint64_t nbr, int dtype, int k, const void *X, int64_t ldx,
                void *Y, int64_t ldy, int nthreads, int copy_block_rows);

/* ---- tall-skinny dense ops (block_dense_matrix.cpp, FM_MultiVector.h) ---- */
void orc_gram_f64(const double *A, int64_t lda, const double *B, int64_t ldb,
                  uint64_t n, int ka, int kb, double *G);
void orc_gemm_tall_f64(const double *A, int64_t lda, uint64_t n, int ka,
                       const double *C, int kc, double alpha, double beta,
                       double *Y, int64_t ldy);
void orc_colnorm_f64(const double *A, int64_t lda, uint64_t n, int k,
                     double *norms);

/* ---- PageRank (matrix/libmatrix-algs/PageRank.cpp:49-137) ------------- */
int orc_pagerank(const uint8_t *a_img, uint64_t a_bytes, const int64_t *a_offs, uint64_t a_nbr,
                 const uint8_t *t_img, uint64_t t_bytes, const int64_t *t_offs, uint64_t t_nbr,
                 uint64_t n, int max_iters, float damping, float *pr, int *iters);

/* ---- learned warm start: network and loss (FP64 checker) ----------------- */
uint64_t orc_gnn_param_count(void);
int orc_subspace_loss(uint64_t n, int k, const double *Y, const double *U, double *loss,
                      double *dY);
int orc_gnn_loss_grad(uint64_t n, const int64_t *rowptr, const int64_t *col, const double *dinv,
                      const double *X0, const double *U, const double *params, double *loss,
                      double *grad, double *Y);
int orc_gnn_loss_grad_ex(uint64_t n, const int64_t *rowptr, const int64_t *col, const double *dinv,
                         const double *X0, const double *U, const double *params, double *loss,
                         double *grad, double *Y, int bf16_operands);
void orc_bf16_round(const double *in, double *out, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
