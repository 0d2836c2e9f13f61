/*
 * oracle.h -- CPU restatement of the reference's tall-skinny products.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library
 * libskinny.so, the C++ drop-in headers) may include, link or call this code.
 * It is used by tests/, by __graft_entry__.smoke() as the checker and by
 * bench.py's cpu_baseline leg, and by nothing else.
 *
 * Semantics restated (reference file:line):
 *   TSMTTSM  C(m,n) = alpha * sum_k A(k,m) B(k,n) + beta * C(m,n)
 *            A row-major K x M (lda), B row-major K x N (ldb), C column-major
 *            M x N (ldc); the cuBLAS mapping in tsmttsm/gen_cublas.cuh:22-24,
 *            stage-2 epilogue in tsmttsm/genv3.cuh:25-26.  Complex is a plain
 *            transpose (CUBLAS_OP_T, gen_cublas.cuh:30-33), never conjugated.
 *   TSMM     C(k,n) = alpha * sum_m A(k,m) B(m,n) + beta * C(k,n)
 *            A row-major K x M, B column-major M x N (ldb), C row-major K x N;
 *            the cuBLAS mapping in tsmm/cublas.cuh:37-39, epilogue
 *            tsmm/fix3.cuh:31-35 (axpby: beta*C + alpha*sum).  In place
 *            (C aliases A) reads the old A, as tsmm/cublas.cuh:27-33 does by
 *            copying A first.
 *   beta == 0 follows BLAS: C is not read (the reference's stage 2 would read
 *            it; this is a documented deviation).
 *   PseudoQuad (double-double) ops restated from PseudoQuad.cuh:10-157.
 */
#ifndef SKINNY_ORACLE_H
#define SKINNY_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter RNG shared with the device generator ---------------------- */
uint64_t orc_stream_key(uint64_t seed, uint64_t stream);
uint64_t orc_rng_bits(uint64_t key, uint64_t index);
void orc_fill_uniform(double* out, int64_t n, uint64_t seed, uint64_t stream,
                      int64_t offset);
void orc_fill_uniform_f(float* out, int64_t n, uint64_t seed, uint64_t stream,
                        int64_t offset);
void orc_fill_int(double* out, int64_t n, uint64_t seed, uint64_t stream,
                  int64_t offset, int lo, int hi);

/* ---- FP64 / FP32 / complex loops (sequential k order) ------------------- */
void orc_dtsmttsm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                  double alpha, const double* B, int64_t ldb, double beta,
                  double* C, int64_t ldc);
void orc_stsmttsm(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                  float alpha, const float* B, int64_t ldb, float beta,
                  float* C, int64_t ldc);
/* complex: interleaved (re, im); alpha/beta as (re, im) pairs */
void orc_ztsmttsm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                  const double* alpha, const double* B, int64_t ldb,
                  const double* beta, double* C, int64_t ldc);
void orc_ctsmttsm(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                  const float* alpha, const float* B, int64_t ldb,
                  const float* beta, float* C, int64_t ldc);

void orc_dtsmm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
               double alpha, const double* B, int64_t ldb, double beta,
               double* C, int64_t ldc);
void orc_stsmm(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
               float alpha, const float* B, int64_t ldb, float beta, float* C,
               int64_t ldc);
void orc_ztsmm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
               const double* alpha, const double* B, int64_t ldb,
               const double* beta, double* C, int64_t ldc);
void orc_ctsmm(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
               const float* alpha, const float* B, int64_t ldb,
               const float* beta, float* C, int64_t ldc);

/* ---- OpenMP CPU baseline (FP64): K-chunks with private accumulators ---- */
void orc_dtsmttsm_par(int nthreads, int64_t M, int64_t N, int64_t K,
                      const double* A, int64_t lda, double alpha,
                      const double* B, int64_t ldb, double beta, double* C,
                      int64_t ldc);
void orc_dtsmm_par(int nthreads, int64_t M, int64_t N, int64_t K,
                   const double* A, int64_t lda, double alpha, const double* B,
                   int64_t ldb, double beta, double* C, int64_t ldc);
int orc_max_threads(void);

/* ---- PseudoQuad (double-double) primitives, PseudoQuad.cuh ------------ */
void orc_dd_fast2sum(double a, double b, double* s, double* t);
void orc_dd_gen2sum(double a, double b, double* s, double* t);
void orc_dd_fma2mult(double a, double b, double* s, double* t);
void orc_dd_accu(const double* x, const double* y, double* r);
void orc_dd_axpy2(const double* c, double a, double b, double* r);
void orc_dd_slow_axpy2(const double* c, double a, double b, double* r);
void orc_dd_scale2(const double* x, double b, double* r);
void orc_dd_renormalize(const double* x, double* r);
void orc_dda_add(const double* x3, const double* y2, double* r3);
void orc_dda_add3(const double* x3, const double* y3, double* r3);
void orc_dda_value(const double* x3, double* r2);

/* HR: double inputs, PseudoQuad accumulation (slow_axpy2, the accurate
 * TwoProd + dd add of PseudoQuad.cuh:115-118) in sequential k order, then
 * C = beta*C + convert(scale2(renormalize(sum), alpha))  (genv3.cuh:19-26). */
void orc_htsmttsm(int64_t M, int64_t N, int64_t K, const double* A,
                  int64_t lda, double alpha, const double* B, int64_t ldb,
                  double beta, double* C, int64_t ldc);
/* same accumulation, dd result scale2(renormalize(sum), alpha) kept as the
 * pair (C_hi, C_lo) without the final conversion to double */
void orc_htsmttsm_dd(int64_t M, int64_t N, int64_t K, const double* A,
                     int64_t lda, double alpha, const double* B, int64_t ldb,
                     double* C_hi, double* C_lo, int64_t ldc);

/* ---- exact products (superaccumulator) --------------------------------- */
/* exact sum_k a[k*inca] * b[k*incb]; returned as hi = RN(v), lo = RN(v - hi)
 * (accurate to ~2^-110 relative, computed through binary128). */
void orc_xdot(int64_t n, const double* a, int64_t inca, const double* b,
              int64_t incb, double* hi, double* lo);
/* exact A^T B (no alpha/beta), column-major M x N outputs hi and lo */
void orc_xtsmttsm(int64_t M, int64_t N, int64_t K, const double* A,
                  int64_t lda, const double* B, int64_t ldb, double* C_hi,
                  double* C_lo, int64_t ldc);
/* exact A^T B for a list of (m, n) entries only */
void orc_xtsmttsm_entries(int64_t K, const double* A, int64_t lda,
                          const double* B, int64_t ldb, int64_t nentries,
                          const int64_t* ms, const int64_t* ns, double* hi,
                          double* lo);

/* ---- Ozaki split (bit-exact twin of the device path at blockCount 1) ----- */
void orc_ozaki_tsmttsm(int64_t M, int64_t N, int64_t K, const double* A,
                       int64_t lda, double alpha, const double* B, int64_t ldb,
                       double beta, double* C, int64_t ldc, int slices);

/* ---- SVQB restatement (svqb/svqb.cu:72-130) ----------------------------- */
int orc_jacobi_eig(int64_t M, double* S, int64_t lds, double* w, double* V,
                   int64_t ldv, int max_sweeps);
void orc_svqb_transform(int64_t M, const double* S, int64_t lds, double* T,
                        int64_t ldt);
void orc_svqb_sweep(int64_t M, int64_t K, double* X, int64_t ldx);
void orc_gram_error(int64_t M, int64_t K, const double* X, int64_t ldx,
                    double* fro, double* maxabs);

#ifdef __cplusplus
}
#endif
#endif
