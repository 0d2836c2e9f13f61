/*
 * oracle.c -- CPU restatement of the reference's tall-skinny products.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg.  The product (libskinny.so) never links or
 * calls it.  Built with -ffp-contract=off so every a*b+c below is two
 * roundings unless fma() is written out, exactly as PseudoQuad.cuh spells its
 * __dadd_rn/__dmul_rn/__fma_rn sequence.
 *
 * Reference semantics followed (file:line in te42kyfo/skinny_matmul):
 *   TSMTTSM: tsmttsm/gen_cublas.cuh:22-24 (cublasXgemm(N, T, M, N, K, ...)),
 *            epilogue tsmttsm/genv3.cuh:25-26 (C = beta*C + alpha*sum).
 *   TSMM:    tsmm/cublas.cuh:37-39 (cublasXgemm(T, N, N, K, M, ...)),
 *            in-place copy tsmm/cublas.cuh:27-33, epilogue tsmm/fix3.cuh:31-35.
 *   complex: cu_complex.h axpy/axpby/scale (plain products, no conjugate).
 *   PseudoQuad: PseudoQuad.cuh:10-157.
 *   beta == 0: BLAS semantics, C is not read.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* counter RNG: SplitMix64 finaliser over (stream key + index * golden)      */
/* ------------------------------------------------------------------------ */
static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t orc_stream_key(uint64_t seed, uint64_t stream) {
  return mix64(seed ^ mix64(stream + 0x632BE59BD9B4E019ull));
}

uint64_t orc_rng_bits(uint64_t key, uint64_t index) {
  return mix64(key + index * 0xD1B54A32D192ED03ull);
}

/* U[-1, 1) on a 2^-52 grid: exactly representable, identical on the device */
static inline double bits_to_uniform(uint64_t b) {
  return (double)(b >> 11) * 0x1p-52 - 1.0;
}

void orc_fill_uniform(double* out, int64_t n, uint64_t seed, uint64_t stream,
                      int64_t offset) {
  uint64_t key = orc_stream_key(seed, stream);
  for (int64_t i = 0; i < n; i++)
    out[i] = bits_to_uniform(orc_rng_bits(key, (uint64_t)(offset + i)));
}

void orc_fill_uniform_f(float* out, int64_t n, uint64_t seed, uint64_t stream,
                        int64_t offset) {
  uint64_t key = orc_stream_key(seed, stream);
  for (int64_t i = 0; i < n; i++) {
    uint64_t b = orc_rng_bits(key, (uint64_t)(offset + i));
    out[i] = (float)((double)(b >> 40) * 0x1p-23 - 1.0);
  }
}

void orc_fill_int(double* out, int64_t n, uint64_t seed, uint64_t stream,
                  int64_t offset, int lo, int hi) {
  uint64_t key = orc_stream_key(seed, stream);
  uint64_t span = (uint64_t)(hi - lo + 1);
  for (int64_t i = 0; i < n; i++) {
    uint64_t b = orc_rng_bits(key, (uint64_t)(offset + i));
    out[i] = (double)(lo + (int64_t)((b >> 32) % span));
  }
}

/* ------------------------------------------------------------------------ */
/* TSMTTSM, real                                                            */
/* ------------------------------------------------------------------------ */
void orc_dtsmttsm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                  double alpha, const double* B, int64_t ldb, double beta,
                  double* C, int64_t ldc) {
  for (int64_t n = 0; n < N; n++) {
    for (int64_t m = 0; m < M; m++) {
      double s = 0.0;
      for (int64_t k = 0; k < K; k++) s += A[k * lda + m] * B[k * ldb + n];
      double* c = &C[n * ldc + m];
      *c = (beta == 0.0) ? alpha * s : beta * *c + alpha * s;
    }
  }
}

void orc_stsmttsm(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                  float alpha, const float* B, int64_t ldb, float beta,
                  float* C, int64_t ldc) {
  /* accumulated in double: a more accurate oracle than the fp32 kernel */
  for (int64_t n = 0; n < N; n++) {
    for (int64_t m = 0; m < M; m++) {
      double s = 0.0;
      for (int64_t k = 0; k < K; k++)
        s += (double)A[k * lda + m] * (double)B[k * ldb + n];
      float* c = &C[n * ldc + m];
      double r = (double)alpha * s;
      if (beta != 0.0f) r += (double)beta * (double)*c;
      *c = (float)r;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* complex helpers (interleaved re, im)                                      */
/* ------------------------------------------------------------------------ */
static inline void cmul(double ar, double ai, double br, double bi, double* r,
                        double* i) {
  *r = ar * br - ai * bi;
  *i = ar * bi + ai * br;
}

void orc_ztsmttsm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                  const double* alpha, const double* B, int64_t ldb,
                  const double* beta, double* C, int64_t ldc) {
  int bzero = (beta[0] == 0.0 && beta[1] == 0.0);
  for (int64_t n = 0; n < N; n++) {
    for (int64_t m = 0; m < M; m++) {
      double sr = 0.0, si = 0.0;
      for (int64_t k = 0; k < K; k++) {
        const double* a = &A[2 * (k * lda + m)];
        const double* b = &B[2 * (k * ldb + n)];
        double pr, pi;
        cmul(a[0], a[1], b[0], b[1], &pr, &pi);
        sr += pr;
        si += pi;
      }
      double* c = &C[2 * (n * ldc + m)];
      double rr, ri;
      cmul(alpha[0], alpha[1], sr, si, &rr, &ri);
      if (!bzero) {
        double br, bi;
        cmul(beta[0], beta[1], c[0], c[1], &br, &bi);
        rr = br + rr;
        ri = bi + ri;
      }
      c[0] = rr;
      c[1] = ri;
    }
  }
}

void orc_ctsmttsm(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                  const float* alpha, const float* B, int64_t ldb,
                  const float* beta, float* C, int64_t ldc) {
  int bzero = (beta[0] == 0.0f && beta[1] == 0.0f);
  for (int64_t n = 0; n < N; n++) {
    for (int64_t m = 0; m < M; m++) {
      double sr = 0.0, si = 0.0;
      for (int64_t k = 0; k < K; k++) {
        const float* a = &A[2 * (k * lda + m)];
        const float* b = &B[2 * (k * ldb + n)];
        double pr, pi;
        cmul(a[0], a[1], b[0], b[1], &pr, &pi);
        sr += pr;
        si += pi;
      }
      float* c = &C[2 * (n * ldc + m)];
      double rr, ri;
      cmul(alpha[0], alpha[1], sr, si, &rr, &ri);
      if (!bzero) {
        double br, bi;
        cmul(beta[0], beta[1], c[0], c[1], &br, &bi);
        rr += br;
        ri += bi;
      }
      c[0] = (float)rr;
      c[1] = (float)ri;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* TSMM                                                                     */
/* ------------------------------------------------------------------------ */
void orc_dtsmm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
               double alpha, const double* B, int64_t ldb, double beta,
               double* C, int64_t ldc) {
  /* row k of C depends only on row k of A: copying the row first gives the
   * in-place (A == C) semantics of tsmm/cublas.cuh:27-33 */
  double* arow = (double*)malloc(sizeof(double) * (M > 0 ? M : 1));
  double* crow = (double*)malloc(sizeof(double) * (N > 0 ? N : 1));
  for (int64_t k = 0; k < K; k++) {
    memcpy(arow, &A[k * lda], sizeof(double) * M);
    for (int64_t n = 0; n < N; n++) {
      double s = 0.0;
      for (int64_t m = 0; m < M; m++) s += arow[m] * B[n * ldb + m];
      crow[n] = (beta == 0.0) ? alpha * s : beta * C[k * ldc + n] + alpha * s;
    }
    memcpy(&C[k * ldc], crow, sizeof(double) * N);
  }
  free(arow);
  free(crow);
}

void orc_stsmm(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
               float alpha, const float* B, int64_t ldb, float beta, float* C,
               int64_t ldc) {
  float* arow = (float*)malloc(sizeof(float) * (M > 0 ? M : 1));
  float* crow = (float*)malloc(sizeof(float) * (N > 0 ? N : 1));
  for (int64_t k = 0; k < K; k++) {
    memcpy(arow, &A[k * lda], sizeof(float) * M);
    for (int64_t n = 0; n < N; n++) {
      double s = 0.0;
      for (int64_t m = 0; m < M; m++)
        s += (double)arow[m] * (double)B[n * ldb + m];
      double r = (double)alpha * s;
      if (beta != 0.0f) r += (double)beta * (double)C[k * ldc + n];
      crow[n] = (float)r;
    }
    memcpy(&C[k * ldc], crow, sizeof(float) * N);
  }
  free(arow);
  free(crow);
}

void orc_ztsmm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
               const double* alpha, const double* B, int64_t ldb,
               const double* beta, double* C, int64_t ldc) {
  int bzero = (beta[0] == 0.0 && beta[1] == 0.0);
  double* arow = (double*)malloc(sizeof(double) * 2 * (M > 0 ? M : 1));
  double* crow = (double*)malloc(sizeof(double) * 2 * (N > 0 ? N : 1));
  for (int64_t k = 0; k < K; k++) {
    memcpy(arow, &A[2 * k * lda], sizeof(double) * 2 * M);
    for (int64_t n = 0; n < N; n++) {
      double sr = 0.0, si = 0.0;
      for (int64_t m = 0; m < M; m++) {
        const double* b = &B[2 * (n * ldb + m)];
        double pr, pi;
        cmul(arow[2 * m], arow[2 * m + 1], b[0], b[1], &pr, &pi);
        sr += pr;
        si += pi;
      }
      double rr, ri;
      cmul(alpha[0], alpha[1], sr, si, &rr, &ri);
      if (!bzero) {
        const double* c = &C[2 * (k * ldc + n)];
        double br, bi;
        cmul(beta[0], beta[1], c[0], c[1], &br, &bi);
        rr = br + rr;
        ri = bi + ri;
      }
      crow[2 * n] = rr;
      crow[2 * n + 1] = ri;
    }
    memcpy(&C[2 * k * ldc], crow, sizeof(double) * 2 * N);
  }
  free(arow);
  free(crow);
}

void orc_ctsmm(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
               const float* alpha, const float* B, int64_t ldb,
               const float* beta, float* C, int64_t ldc) {
  int bzero = (beta[0] == 0.0f && beta[1] == 0.0f);
  float* arow = (float*)malloc(sizeof(float) * 2 * (M > 0 ? M : 1));
  float* crow = (float*)malloc(sizeof(float) * 2 * (N > 0 ? N : 1));
  for (int64_t k = 0; k < K; k++) {
    memcpy(arow, &A[2 * k * lda], sizeof(float) * 2 * M);
    for (int64_t n = 0; n < N; n++) {
      double sr = 0.0, si = 0.0;
      for (int64_t m = 0; m < M; m++) {
        const float* b = &B[2 * (n * ldb + m)];
        double pr, pi;
        cmul(arow[2 * m], arow[2 * m + 1], b[0], b[1], &pr, &pi);
        sr += pr;
        si += pi;
      }
      double rr, ri;
      cmul(alpha[0], alpha[1], sr, si, &rr, &ri);
      if (!bzero) {
        const float* c = &C[2 * (k * ldc + n)];
        double br, bi;
        cmul(beta[0], beta[1], c[0], c[1], &br, &bi);
        rr += br;
        ri += bi;
      }
      crow[2 * n] = (float)rr;
      crow[2 * n + 1] = (float)ri;
    }
    memcpy(&C[2 * k * ldc], crow, sizeof(float) * 2 * N);
  }
  free(arow);
  free(crow);
}

/* ------------------------------------------------------------------------ */
/* OpenMP CPU baseline                                                      */
/* ------------------------------------------------------------------------ */
int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_dtsmttsm_par(int nthreads, int64_t M, int64_t N, int64_t K,
                      const double* A, int64_t lda, double alpha,
                      const double* B, int64_t ldb, double beta, double* C,
                      int64_t ldc) {
  if (nthreads < 1) nthreads = 1;
  double* part = (double*)calloc((size_t)nthreads * M * N, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
    int t = 0, nt = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    int64_t k0 = K * t / nt, k1 = K * (t + 1) / nt;
    double* acc = part + (size_t)t * M * N;
    for (int64_t k = k0; k < k1; k++) {
      const double* a = &A[k * lda];
      const double* b = &B[k * ldb];
      for (int64_t n = 0; n < N; n++) {
        double bn = b[n];
        double* col = acc + n * M;
        for (int64_t m = 0; m < M; m++) col[m] += a[m] * bn;
      }
    }
  }
  for (int64_t n = 0; n < N; n++)
    for (int64_t m = 0; m < M; m++) {
      double s = 0.0;
      for (int t = 0; t < nthreads; t++) s += part[(size_t)t * M * N + n * M + m];
      double* c = &C[n * ldc + m];
      *c = (beta == 0.0) ? alpha * s : beta * *c + alpha * s;
    }
  free(part);
}

void orc_dtsmm_par(int nthreads, int64_t M, int64_t N, int64_t K,
                   const double* A, int64_t lda, double alpha, const double* B,
                   int64_t ldb, double beta, double* C, int64_t ldc) {
  if (nthreads < 1) nthreads = 1;
  if (M > 512 || N > 512) return; /* baseline widths only */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
  for (int64_t k = 0; k < K; k++) {
    double arow[512];
    double crow[512];
    int64_t mm = M < 512 ? M : 512;
    for (int64_t m = 0; m < mm; m++) arow[m] = A[k * lda + m];
    for (int64_t n = 0; n < N && n < 512; n++) {
      double s = 0.0;
      for (int64_t m = 0; m < mm; m++) s += arow[m] * B[n * ldb + m];
      crow[n] = (beta == 0.0) ? alpha * s : beta * C[k * ldc + n] + alpha * s;
    }
    for (int64_t n = 0; n < N && n < 512; n++) C[k * ldc + n] = crow[n];
  }
}

/* ------------------------------------------------------------------------ */
/* PseudoQuad (double-double), PseudoQuad.cuh:10-157                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  double s, t;
} dd_t;

/* Fast2Sum (Dekker), PseudoQuad.cuh:10-15 */
static inline dd_t fast2sum(double a, double b) {
  double s = a + b;
  double z = s - a;
  double t = b - z;
  dd_t r = {s, t};
  return r;
}
/* Gen2Sum = TwoSum (Knuth), PseudoQuad.cuh:17-25 */
static inline dd_t gen2sum(double a, double b) {
  double s = a + b;
  double ap = s - b;
  double bp = s - ap;
  double da = a - ap;
  double db = b - bp;
  dd_t r = {s, da + db};
  return r;
}
/* FMA2Mult = TwoProd, PseudoQuad.cuh:27-31 */
static inline dd_t fma2mult(double a, double b) {
  double s = a * b;
  dd_t r = {s, fma(a, b, -s)};
  return r;
}
/* accu, the "book version" dd + dd, PseudoQuad.cuh:66-75 */
static inline dd_t dd_accu(dd_t x, dd_t y) {
  dd_t S = gen2sum(x.s, y.s);
  dd_t T = gen2sum(x.t, y.t);
  double c = S.t + T.s;
  dd_t V = fast2sum(S.s, c);
  double w = T.t + V.t;
  return fast2sum(V.s, w);
}
/* scale2, dd * d, PseudoQuad.cuh:80-86 */
static inline dd_t dd_scale2(dd_t a, double b) {
  dd_t p = fma2mult(a.s, b);
  p.t = fma(a.s, 0.0, p.t);
  p.t = fma(a.t, b, p.t);
  return fast2sum(p.s, p.t);
}
/* axpy2, dd += d * d, PseudoQuad.cuh:96-108 */
static inline dd_t dd_axpy2(dd_t c, double a, double b) {
  dd_t r;
  r.s = fma(a, b, c.s);
  double abp = r.s - c.s;
  double cp = r.s - abp;
  double dab = fma(a, b, -abp);
  double dc = c.s - cp;
  r.t = dc + dab;
  r.t = r.t + c.t;
  return r;
}
/* slow_axpy2, dd += d * d through TwoProd and the accurate dd + dd,
 * PseudoQuad.cuh:115-118.  Unlike axpy2 above it stays accurate to ~2^-104
 * when the sum cancels, so the HR accumulation below uses it. */
static inline dd_t dd_slow_axpy2(dd_t c, double a, double b) {
  return dd_accu(c, fma2mult(a, b));
}
/* renormalize, PseudoQuad.cuh:137-140 */
static inline dd_t dd_renorm(dd_t a) { return fast2sum(a.s, a.t); }

void orc_dd_fast2sum(double a, double b, double* s, double* t) {
  dd_t r = fast2sum(a, b);
  *s = r.s;
  *t = r.t;
}
void orc_dd_gen2sum(double a, double b, double* s, double* t) {
  dd_t r = gen2sum(a, b);
  *s = r.s;
  *t = r.t;
}
void orc_dd_fma2mult(double a, double b, double* s, double* t) {
  dd_t r = fma2mult(a, b);
  *s = r.s;
  *t = r.t;
}
void orc_dd_accu(const double* x, const double* y, double* r) {
  dd_t a = {x[0], x[1]}, b = {y[0], y[1]};
  dd_t o = dd_accu(a, b);
  r[0] = o.s;
  r[1] = o.t;
}
void orc_dd_axpy2(const double* c, double a, double b, double* r) {
  dd_t x = {c[0], c[1]};
  dd_t o = dd_axpy2(x, a, b);
  r[0] = o.s;
  r[1] = o.t;
}
void orc_dd_slow_axpy2(const double* c, double a, double b, double* r) {
  dd_t x = {c[0], c[1]};
  dd_t o = dd_slow_axpy2(x, a, b);
  r[0] = o.s;
  r[1] = o.t;
}
void orc_dd_scale2(const double* x, double b, double* r) {
  dd_t a = {x[0], x[1]};
  dd_t o = dd_scale2(a, b);
  r[0] = o.s;
  r[1] = o.t;
}
void orc_dd_renormalize(const double* x, double* r) {
  dd_t a = {x[0], x[1]};
  dd_t o = dd_renorm(a);
  r[0] = o.s;
  r[1] = o.t;
}

/* dd plus the rounding remainders of dd_accu's two inexact additions (c and
 * w) in a third word u; (s, t) is bitwise dd_accu's result.  The HR kernel
 * combines its partial sums with it (no reference counterpart). */
typedef struct {
  dd_t v;
  double u;
} dda_t;

static inline dda_t dda_add(dda_t x, dd_t y) {
  dd_t S = gen2sum(x.v.s, y.s);
  dd_t T = gen2sum(x.v.t, y.t);
  dd_t Cc = gen2sum(S.t, T.s);
  dd_t V = fast2sum(S.s, Cc.s);
  dd_t W = gen2sum(T.t, V.t);
  dda_t r;
  r.v = fast2sum(V.s, W.s);
  r.u = x.u + (Cc.t + W.t);
  return r;
}

void orc_dda_add(const double* x, const double* y, double* r) {
  dda_t a = {{x[0], x[1]}, x[2]};
  dd_t b = {y[0], y[1]};
  dda_t o = dda_add(a, b);
  r[0] = o.v.s;
  r[1] = o.v.t;
  r[2] = o.u;
}
/* a triple plus a triple: the dd part as above, then the third words */
void orc_dda_add3(const double* x, const double* y, double* r) {
  dda_t a = {{x[0], x[1]}, x[2]};
  dd_t b = {y[0], y[1]};
  dda_t o = dda_add(a, b);
  o.u = o.u + y[2];
  r[0] = o.v.s;
  r[1] = o.v.t;
  r[2] = o.u;
}
/* the triple's value as dd: dd_accu((s, t), (u, 0)) */
void orc_dda_value(const double* x, double* r) {
  dd_t a = {x[0], x[1]}, b = {x[2], 0.0};
  dd_t o = dd_accu(a, b);
  r[0] = o.s;
  r[1] = o.t;
}

static dd_t hr_sum(int64_t K, const double* A, int64_t lda, const double* B,
                   int64_t ldb, int64_t m, int64_t n, double alpha) {
  dd_t s = {0.0, 0.0};
  for (int64_t k = 0; k < K; k++)
    s = dd_slow_axpy2(s, A[k * lda + m], B[k * ldb + n]);
  return dd_scale2(dd_renorm(s), alpha);
}

void orc_htsmttsm(int64_t M, int64_t N, int64_t K, const double* A,
                  int64_t lda, double alpha, const double* B, int64_t ldb,
                  double beta, double* C, int64_t ldc) {
  for (int64_t n = 0; n < N; n++)
    for (int64_t m = 0; m < M; m++) {
      dd_t r = hr_sum(K, A, lda, B, ldb, m, n, alpha);
      double v = r.s + r.t; /* convert, PseudoQuad.cuh:149-152 */
      double* c = &C[n * ldc + m];
      *c = (beta == 0.0) ? v : beta * *c + v;
    }
}

void orc_htsmttsm_dd(int64_t M, int64_t N, int64_t K, const double* A,
                     int64_t lda, double alpha, const double* B, int64_t ldb,
                     double* C_hi, double* C_lo, int64_t ldc) {
  for (int64_t n = 0; n < N; n++)
    for (int64_t m = 0; m < M; m++) {
      dd_t r = hr_sum(K, A, lda, B, ldb, m, n, alpha);
      C_hi[n * ldc + m] = r.s;
      C_lo[n * ldc + m] = r.t;
    }
}

/* ------------------------------------------------------------------------ */
/* exact superaccumulator (fixed point over the whole double range)         */
/* ------------------------------------------------------------------------ */
/* limb i holds a signed multiple of 2^(32 i - XBIAS) */
#define XBIAS 1088
#define XLIMBS 70

typedef struct {
  int64_t l[XLIMBS];
  int64_t adds;
} xacc_t;

static void xacc_norm(xacc_t* x) {
  for (int i = 0; i < XLIMBS - 1; i++) {
    int64_t c = x->l[i] >> 32; /* floor division by 2^32 */
    x->l[i] -= c * ((int64_t)1 << 32);
    x->l[i + 1] += c;
  }
  x->adds = 0;
}

static void xacc_add(xacc_t* x, double v) {
  if (v == 0.0) return;
  int e;
  double f = frexp(v, &e);                 /* v = f * 2^e, 0.5 <= |f| < 1 */
  int64_t mant = (int64_t)ldexp(f, 53);    /* exact, |mant| < 2^53 */
  int p = e - 53 + XBIAS;                  /* bit position of mant's lsb */
  int i = p >> 5, s = p & 31;
  __int128 w = (__int128)mant * ((__int128)1 << s);
  for (int j = 0; j < 3; j++) {
    int64_t chunk = (int64_t)(w & 0xFFFFFFFF);
    w >>= 32; /* arithmetic */
    if (j == 2) chunk += (int64_t)w * ((int64_t)1 << 32);
    x->l[i + j] += chunk;
  }
  if (++x->adds >= ((int64_t)1 << 28)) xacc_norm(x);
}

static void xacc_add_prod(xacc_t* x, double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  xacc_add(x, p);
  xacc_add(x, e);
}

static void xacc_to_dd(xacc_t* x, double* hi, double* lo) {
  xacc_norm(x);
  int neg = x->l[XLIMBS - 1] < 0;
  if (neg) {
    for (int i = 0; i < XLIMBS; i++) x->l[i] = -x->l[i];
    xacc_norm(x);
  }
  /* all limbs now in [0, 2^32): sum top-down in binary128 (no cancellation) */
  __float128 q = 0;
  for (int i = XLIMBS - 1; i >= 0; i--) {
    if (x->l[i] == 0) continue;
    q += (__float128)x->l[i] * (__float128)ldexp(1.0, 32 * i - XBIAS);
  }
  if (neg) q = -q;
  double h = (double)q;
  double r = (double)(q - (__float128)h);
  *hi = h;
  *lo = r;
}

void orc_xdot(int64_t n, const double* a, int64_t inca, const double* b,
              int64_t incb, double* hi, double* lo) {
  xacc_t x;
  memset(&x, 0, sizeof(x));
  for (int64_t k = 0; k < n; k++) xacc_add_prod(&x, a[k * inca], b[k * incb]);
  xacc_to_dd(&x, hi, lo);
}

void orc_xtsmttsm(int64_t M, int64_t N, int64_t K, const double* A,
                  int64_t lda, const double* B, int64_t ldb, double* C_hi,
                  double* C_lo, int64_t ldc) {
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(dynamic)
#endif
  for (int64_t n = 0; n < N; n++)
    for (int64_t m = 0; m < M; m++)
      orc_xdot(K, A + m, lda, B + n, ldb, &C_hi[n * ldc + m],
               &C_lo[n * ldc + m]);
}

void orc_xtsmttsm_entries(int64_t K, const double* A, int64_t lda,
                          const double* B, int64_t ldb, int64_t nentries,
                          const int64_t* ms, const int64_t* ns, double* hi,
                          double* lo) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic)
#endif
  for (int64_t e = 0; e < nentries; e++)
    orc_xdot(K, A + ms[e], lda, B + ns[e], ldb, &hi[e], &lo[e]);
}

/* ------------------------------------------------------------------------ */
/* SVQB, svqb/svqb.cu:72-130, FP64 throughout                               */
/* ------------------------------------------------------------------------ */
/* Cyclic Jacobi eigensolver for a symmetric M x M column-major S (destroyed;
 * its diagonal ends as the eigenvalues).  Textbook row-cyclic ordering
 * (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.5.3) with the
 * stable rotation t = sign(tau) / (|tau| + sqrt(1 + tau^2)).  V receives the
 * eigenvectors as columns.  Returns the sweeps used. */
int orc_jacobi_eig(int64_t M, double* S, int64_t lds, double* w, double* V,
                   int64_t ldv, int max_sweeps) {
  for (int64_t j = 0; j < M; j++)
    for (int64_t i = 0; i < M; i++) V[j * ldv + i] = (i == j) ? 1.0 : 0.0;
  int sweep = 0;
  for (; sweep < max_sweeps; sweep++) {
    double off = 0.0, dia = 0.0;
    for (int64_t j = 0; j < M; j++)
      for (int64_t i = 0; i < M; i++) {
        double v = S[j * lds + i];
        if (i == j) dia += v * v; else off += v * v;
      }
    if (off <= 1e-30 * dia) break;
    for (int64_t p = 0; p < M - 1; p++)
      for (int64_t q = p + 1; q < M; q++) {
        double apq = S[q * lds + p];
        if (apq == 0.0) continue;
        double app = S[p * lds + p], aqq = S[q * lds + q];
        double tau = (aqq - app) / (2.0 * apq);
        double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int64_t j = 0; j < M; j++) { /* rows p, q */
          double a = S[j * lds + p], b = S[j * lds + q];
          S[j * lds + p] = c * a - s * b;
          S[j * lds + q] = s * a + c * b;
        }
        for (int64_t i = 0; i < M; i++) { /* columns p, q */
          double a = S[p * lds + i], b = S[q * lds + i];
          S[p * lds + i] = c * a - s * b;
          S[q * lds + i] = s * a + c * b;
          double va = V[p * ldv + i], vb = V[q * ldv + i];
          V[p * ldv + i] = c * va - s * vb;
          V[q * ldv + i] = s * va + c * vb;
        }
        S[q * lds + p] = 0.0;
        S[p * lds + q] = 0.0;
      }
  }
  for (int64_t i = 0; i < M; i++) w[i] = S[i * lds + i];
  return sweep;
}

/* T = D U L^-1/2 from the Gram S (svqb.cu:99-119); T column-major M x M */
void orc_svqb_transform(int64_t M, const double* S, int64_t lds, double* T,
                        int64_t ldt) {
  double* A = (double*)malloc(sizeof(double) * M * M);
  double* V = (double*)malloc(sizeof(double) * M * M);
  double* w = (double*)malloc(sizeof(double) * M);
  double* d = (double*)malloc(sizeof(double) * M);
  for (int64_t i = 0; i < M; i++) {
    double sii = S[i * lds + i];
    d[i] = (fabs(sii) > 1.0e-15) ? 1.0 / sqrt(sii) : 0.0;
  }
  for (int64_t j = 0; j < M; j++)
    for (int64_t i = 0; i < M; i++) A[j * M + i] = S[j * lds + i] * (d[i] * d[j]);
  orc_jacobi_eig(M, A, M, w, V, M, 40);
  for (int64_t n = 0; n < M; n++)
    for (int64_t m = 0; m < M; m++)
      T[n * ldt + m] = V[n * M + m] * (d[m] * (1.0 / sqrt(fmax(1.0e-15, w[n]))));
  free(A); free(V); free(w); free(d);
}

/* one sweep: X (K x M row-major) <- X T with T from X^T X */
void orc_svqb_sweep(int64_t M, int64_t K, double* X, int64_t ldx) {
  double* S = (double*)calloc(M * M, sizeof(double));
  double* T = (double*)malloc(sizeof(double) * M * M);
  orc_dtsmttsm(M, M, K, X, ldx, 1.0, X, ldx, 0.0, S, M);
  orc_svqb_transform(M, S, M, T, M);
  orc_dtsmm(M, M, K, X, ldx, 1.0, T, M, 0.0, X, ldx);
  free(S); free(T);
}

/* ||X^T X - I||_F (svqb.cu:184-211) and max |X^T X - I| (svqb.cu:213-239),
 * the Gram accumulated exactly */
void orc_gram_error(int64_t M, int64_t K, const double* X, int64_t ldx,
                    double* fro, double* maxabs) {
  double l2 = 0.0, mx = 0.0;
  for (int64_t n = 0; n < M; n++)
    for (int64_t m = 0; m < M; m++) {
      double hi, lo;
      orc_xdot(K, X + m, ldx, X + n, ldx, &hi, &lo);
      double e = (m == n ? 1.0 : 0.0) - hi - lo;
      l2 += e * e;
      if (fabs(e) > mx) mx = fabs(e);
    }
  *fro = sqrt(l2);
  *maxabs = mx;
}

/* ------------------------------------------------------------------------ */
/* Ozaki split TSMTTSM, restated for blockCount = 1 (one workgroup per tile */
/* walking all row chunks in order): the bit-exact twin of the device path.  */
/* Tiles are 128 columns of A by 128 columns of B for 3 limbs and 128 by 64  */
/* for 4 and 5 (a flush is decided over the tile's columns, so the width    */
/* shapes the result).  Chunks are 32 rows, or 64 rows for the diagonal     */
/* tiles of a Gram product with 3 limbs (A == B, lda == ldb, M == N,        */
/* lda <= 2^21), where both operands share one scale per column, and for    */
/* every other product with M, N <= 64 and lda, ldb <= 2^21 (one 128 x 64   */
/* tile).  Gram products with M <= 64 take 128-row chunks whose two 64-row  */
/* blocks share each column's scale, tested for shrinking block by block.   */
/* Algorithm: Ozaki, Ogita, Oishi, Rump, "Error-free                         */
/* transformations of matrix multiplication by using fast routines of matrix */
/* multiplication and its applications", Numer. Algorithms 59 (2012), with   */
/* the rounded integer q = rint(x 2^(8S - 2 - E)) written in S signed bytes  */
/* (balanced base-256 digits), levels s + t < L (L = 3 for 3 limbs, S + 1   */
/* for 4 and 5) and a running per-column scale.                              */
/* ------------------------------------------------------------------------ */
#define OZ_TILE 128
#define OZ_ROWS 32      /* rows per chunk */
#define OZ_ROWS_SELF 64 /* rows per chunk, diagonal tiles of a Gram product */
#define OZ_FLUSH 16384
#define OZ_EMIN (-4000)

/* Level accumulators kept per limb count (levels g = s + t < L). */
static int oz_levels(int S) { return S == 3 ? 3 : S + 1; }

/* L level accumulators per entry, level g = s + t at index g; TN columns of
 * B in the tile.  x y = 2^(Ea + Eb - 2 (8S - 2)) sum d_s d'_t
 * 2^(8 (2S - 2 - s - t)) = 2^(Ea + Eb - 12) sum_g acc_g 2^-8g. */
static void oz_flush_tile(int L, int TN, int64_t* acc, double* cacc,
                          const int* ea, const int* eb) {
  for (int n = 0; n < TN; n++)
    for (int m = 0; m < OZ_TILE; m++) {
      int64_t* a = acc + ((int64_t)n * OZ_TILE + m) * L;
      if (ea[m] != OZ_EMIN && eb[n] != OZ_EMIN) {
        double v = ldexp((double)a[L - 1], -8 * (L - 1));
        for (int g = L - 2; g >= 0; g--) v = v + ldexp((double)a[g], -8 * g);
        cacc[n * OZ_TILE + m] = cacc[n * OZ_TILE + m] + ldexp(v, ea[m] + eb[n] - 12);
      }
      for (int g = 0; g < L; g++) a[g] = 0;
    }
}

/* The limb integer of x against 2^E (|x| < 2^E): q = rint(x 2^(8S - 2 - E)),
 * round to nearest even; |q| <= 2^(8S - 2).  The device computes the same
 * with one FMA. */
static int64_t oz_q(double x, int E, int S) {
  return (int64_t)nearbyint(ldexp(x, 8 * S - 2 - E));
}

/* Balanced base-256 digits of q, most significant first: the least
 * significant digit is the low byte of q read as a signed byte, and the rest
 * are the digits of (q - d) / 256.  For |q| <= 2^(8S - 2) the top digit lies
 * in [-64, 64]. */
static void oz_digits(int64_t q, int S, int* d) {
  for (int s = S - 1; s >= 0; s--) {
    const int dig = (int)(int8_t)(uint8_t)(q & 0xFF);
    d[s] = dig;
    q = (q - dig) / 256;
  }
  if (q != 0) abort(); /* |q| beyond S digits: the scale was not respected */
}

void orc_ozaki_tsmttsm(int64_t M, int64_t N, int64_t K, const double* A,
                       int64_t lda, double alpha, const double* B, int64_t ldb,
                       double beta, double* C, int64_t ldc, int S) {
  const int L = oz_levels(S);
  int64_t* acc = (int64_t*)malloc(sizeof(int64_t) * OZ_TILE * OZ_TILE * L);
  double* cacc = (double*)malloc(sizeof(double) * OZ_TILE * OZ_TILE);
  signed char* da = (signed char*)malloc((size_t)S * 2 * OZ_ROWS_SELF * OZ_TILE);
  signed char* db = (signed char*)malloc((size_t)S * 2 * OZ_ROWS_SELF * OZ_TILE);
  const int gram = A == B && lda == ldb && M == N && lda <= ((int64_t)1 << 21);
  /* Gram products with M <= 64: one tile of 128-row chunks whose two 64-row
   * blocks share each column's scale and are tested for shrinking block by
   * block (the device lays the blocks side by side) */
  const int gcat = gram && M <= 64;
  /* other products with M, N <= 64: one tile of 64-row chunks (the device
   * lays A and B side by side as the columns of one Gram tile) */
  const int cat = gcat || (!(gram && S == 3) && M <= 64 && N <= 64 &&
                           lda <= ((int64_t)1 << 21) && ldb <= ((int64_t)1 << 21));
  const int TN = S == 3 && !cat ? OZ_TILE : 64; /* the tile's columns of B */
  for (int64_t m0 = 0; m0 < M; m0 += OZ_TILE)
    for (int64_t n0 = 0; n0 < N; n0 += TN) {
      const int R = gcat ? 2 * OZ_ROWS_SELF
                    : (gram && m0 == n0 && S == 3) || cat ? OZ_ROWS_SELF : OZ_ROWS;
      const int RB = gcat ? OZ_ROWS_SELF : R; /* rows per scale test */
      int erun[2][OZ_TILE], ecur[2][OZ_TILE];
      for (int c = 0; c < OZ_TILE; c++) erun[0][c] = erun[1][c] = OZ_EMIN;
      memset(acc, 0, sizeof(int64_t) * OZ_TILE * OZ_TILE * L);
      memset(cacc, 0, sizeof(double) * OZ_TILE * OZ_TILE);
      int since = 0;
      for (int64_t r0 = 0; r0 < K; r0 += R) {
        int flush = (since + R > OZ_FLUSH);
        for (int mat = 0; mat < 2; mat++)
          for (int c = 0; c < (mat == 0 ? OZ_TILE : TN); c++) {
            int64_t gc = (mat == 0 ? m0 : n0) + c;
            int valid = gc < (mat == 0 ? M : N);
            ecur[mat][c] = OZ_EMIN;
            for (int b0 = 0; b0 < R; b0 += RB) {
              double mx = 0.0;
              for (int i = b0; i < b0 + RB && valid; i++) {
                int64_t r = r0 + i;
                if (r >= K) break;
                double x = mat == 0 ? A[r * lda + gc] : B[r * ldb + gc];
                if (fabs(x) > mx) mx = fabs(x);
              }
              int e = OZ_EMIN;
              if (mx > 0.0) frexp(mx, &e);
              if (e > ecur[mat][c]) ecur[mat][c] = e;
              if (e > erun[mat][c]) flush = 1; /* growth */
              /* shrink: the block's largest limb integer q (oz_q of its
               * largest |x|) is nonzero but 9 or more bits below the scale,
               * so the scale restarts from this chunk (an outlier coarsens
               * only its own chunk) */
              if (e <= erun[mat][c] && erun[mat][c] != OZ_EMIN && mx > 0.0) {
                int64_t q = oz_q(mx, erun[mat][c], S);
                if (q != 0 && q < ((int64_t)1 << (8 * S - 11))) flush = 1;
              }
            }
          }
        if (flush) {
          /* the accumulators are empty after a flush, so every column's
           * scale restarts from this chunk: an outlier only coarsens the
           * window it falls in */
          oz_flush_tile(L, TN, acc, cacc, erun[0], erun[1]);
          since = 0;
          for (int mat = 0; mat < 2; mat++)
            for (int c = 0; c < (mat == 0 ? OZ_TILE : TN); c++) erun[mat][c] = ecur[mat][c];
        }
        /* limbs */
        for (int mat = 0; mat < 2; mat++) {
          signed char* d = mat == 0 ? da : db;
          for (int c = 0; c < (mat == 0 ? OZ_TILE : TN); c++) {
            int64_t gc = (mat == 0 ? m0 : n0) + c;
            int valid = gc < (mat == 0 ? M : N);
            int E = erun[mat][c];
            for (int i = 0; i < R; i++) {
              int64_t r = r0 + i;
              double x = (valid && r < K) ? (mat == 0 ? A[r * lda + gc] : B[r * ldb + gc]) : 0.0;
              int digit[5] = {0, 0, 0, 0, 0};
              if (E != OZ_EMIN && x != 0.0) oz_digits(oz_q(x, E, S), S, digit);
              for (int s2 = 0; s2 < S; s2++)
                d[((size_t)s2 * OZ_TILE + c) * R + i] = (signed char)digit[s2];
            }
          }
        }
        /* exact integer products per level g = s + t (limb 0 the most
         * significant) */
        for (int g = 0; g < L; g++)
          for (int s1 = 0; s1 < S; s1++) {
            int t1 = g - s1;
            if (t1 < 0 || t1 >= S) continue;
            const signed char* pa = da + (size_t)s1 * OZ_TILE * R;
            const signed char* pb = db + (size_t)t1 * OZ_TILE * R;
            for (int n = 0; n < TN; n++)
              for (int m = 0; m < OZ_TILE; m++) {
                int64_t sum = 0;
                const signed char* x = pa + (size_t)m * R;
                const signed char* y = pb + (size_t)n * R;
                for (int i = 0; i < R; i++) sum += (int64_t)x[i] * y[i];
                acc[((int64_t)n * OZ_TILE + m) * L + g] += sum;
              }
          }
        since += R;
      }
      oz_flush_tile(L, TN, acc, cacc, erun[0], erun[1]);
      for (int64_t n = 0; n < TN && n0 + n < N; n++)
        for (int64_t m = 0; m < OZ_TILE && m0 + m < M; m++) {
          double sres = cacc[n * OZ_TILE + m];
          double* c = &C[(n0 + n) * ldc + (m0 + m)];
          double r = alpha * sres;
          if (beta != 0.0) r = beta * *c + r;
          *c = r;
        }
    }
  free(acc);
  free(cacc);
  free(da);
  free(db);
}
