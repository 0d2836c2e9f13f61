// extern "C" product entry points: argument validation, planning, dispatch.
#include <algorithm>

#include "internal.hpp"

using namespace skinny;

namespace {

skinny_status_t check_common(const char* fn, int64_t M, int64_t N, int64_t K,
                             const void* A, const void* B, const void* C) {
  if (M < 0 || N < 0 || K < 0)
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": negative size");
  if (M > 0 && N > 0 && !C)
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": null C");
  if (M > 0 && N > 0 && K > 0 && (!A || !B))
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": null A or B");
  return SKINNY_OK;
}

template <typename T>
skinny_status_t tsmttsm_real(const char* fn, skinny_handle_t h,
                             size_t blockCount, int64_t M, int64_t N, int64_t K,
                             const T* A, int64_t lda, T alpha, const T* B,
                             int64_t ldb, T beta, T* C, int64_t ldc,
                             skinny_stream_t stream) {
  skinny_status_t s = check_common(fn, M, N, K, A, B, C);
  if (s != SKINNY_OK) return s;
  if (lda < std::max<int64_t>(1, M) || ldb < std::max<int64_t>(1, N) ||
      ldc < std::max<int64_t>(1, M))
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": leading dimension too small");
  if (M == 0 || N == 0) return SKINNY_OK;
  if (M > 1024 || N > 1024)
    return fail(SKINNY_UNSUPPORTED, std::string(fn) + ": M, N <= 1024 only");
  h = resolve(h);
  if (!h) return no_handle(fn);
  TsmttsmPlan p = plan_tsmttsm(h, blockCount, M, N, K, (int)sizeof(T), (int)sizeof(T));
  return launch_tsmttsm_real<T>(h, p, M, N, K, A, lda, alpha, B, ldb, beta, C,
                                ldc, (hipStream_t)stream);
}

template <typename T>
skinny_status_t tsmm_real(const char* fn, skinny_handle_t h, size_t blockCount,
                          int64_t M, int64_t N, int64_t K, const T* A,
                          int64_t lda, T alpha, const T* B, int64_t ldb, T beta,
                          T* C, int64_t ldc, skinny_stream_t stream) {
  skinny_status_t s = check_common(fn, M, N, K, A, B, C);
  if (s != SKINNY_OK) return s;
  if (lda < std::max<int64_t>(1, M) || ldb < std::max<int64_t>(1, M) ||
      ldc < std::max<int64_t>(1, N))
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": leading dimension too small");
  if (K == 0 || N == 0) return SKINNY_OK;
  // in place: C aliases A row for row (test.cu:130-132, perf.cu:182)
  const bool inplace = (const void*)A == (const void*)C;
  if (inplace && lda != ldc)
    return fail(SKINNY_UNSUPPORTED, std::string(fn) + ": in place needs lda == ldc");
  if (!inplace && K > 0 && M > 0) {
    // any other overlap of A and C is not supported
    const char* a0 = (const char*)A;
    const char* a1 = a0 + sizeof(T) * ((K - 1) * lda + M);
    const char* c0 = (const char*)C;
    const char* c1 = c0 + sizeof(T) * ((K - 1) * ldc + N);
    if (a0 < c1 && c0 < a1)
      return fail(SKINNY_UNSUPPORTED, std::string(fn) + ": partially overlapping A and C");
  }
  h = resolve(h);
  if (!h) return no_handle(fn);
  if (M == 0) {
    // C = beta * C: the M-loop is empty, the kernel handles it
  }
  return launch_tsmm_real<T>(h, blockCount, M, N, K, A, lda, alpha, B, ldb,
                             beta, C, ldc, (hipStream_t)stream);
}

template <typename T, typename CT>
skinny_status_t tsmttsm_cplx(const char* fn, skinny_handle_t h,
                             size_t blockCount, int64_t M, int64_t N, int64_t K,
                             const CT* A, int64_t lda, CT alpha, const CT* B,
                             int64_t ldb, CT beta, CT* C, int64_t ldc,
                             skinny_stream_t stream) {
  skinny_status_t s = check_common(fn, M, N, K, A, B, C);
  if (s != SKINNY_OK) return s;
  if (lda < std::max<int64_t>(1, M) || ldb < std::max<int64_t>(1, N) ||
      ldc < std::max<int64_t>(1, M))
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": leading dimension too small");
  if (M == 0 || N == 0) return SKINNY_OK;
  if (M > 512 || N > 512)
    return fail(SKINNY_UNSUPPORTED, std::string(fn) + ": M, N <= 512 only");
  h = resolve(h);
  if (!h) return no_handle(fn);
  TsmttsmPlan p = plan_tsmttsm(h, blockCount, 2 * M, 2 * N, K, (int)sizeof(T),
                               (int)sizeof(T), sizeof(T) == 8 ? TT_DC : TT_FC, M, N);
  return launch_tsmttsm<T>(h, p, 2 * M, 2 * N, K, (const T*)A, 2 * lda,
                           alpha.x, alpha.y, (const T*)B, 2 * ldb, beta.x,
                           beta.y, (T*)C, ldc, true, (hipStream_t)stream);
}

template <typename T, typename CT>
skinny_status_t tsmm_cplx(const char* fn, skinny_handle_t h, size_t blockCount,
                          int64_t M, int64_t N, int64_t K, const CT* A,
                          int64_t lda, CT alpha, const CT* B, int64_t ldb,
                          CT beta, CT* C, int64_t ldc, skinny_stream_t stream) {
  skinny_status_t s = check_common(fn, M, N, K, A, B, C);
  if (s != SKINNY_OK) return s;
  if (lda < std::max<int64_t>(1, M) || ldb < std::max<int64_t>(1, M) ||
      ldc < std::max<int64_t>(1, N))
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": leading dimension too small");
  if (K == 0 || N == 0) return SKINNY_OK;
  if (M > kTsmmWideMax / 2 || N > kTsmmWideMax / 2)
    return fail(SKINNY_UNSUPPORTED, std::string(fn) + ": M, N <= 512 only");
  const bool inplace = (const void*)A == (const void*)C;
  if (inplace && lda != ldc)
    return fail(SKINNY_UNSUPPORTED, std::string(fn) + ": in place needs lda == ldc");
  if (!inplace && M > 0) {
    const char* a0 = (const char*)A;
    const char* a1 = a0 + sizeof(CT) * ((K - 1) * lda + M);
    const char* c0 = (const char*)C;
    const char* c1 = c0 + sizeof(CT) * ((K - 1) * ldc + N);
    if (a0 < c1 && c0 < a1)
      return fail(SKINNY_UNSUPPORTED, std::string(fn) + ": partially overlapping A and C");
  }
  h = resolve(h);
  if (!h) return no_handle(fn);
  // real beta, 2..32 complex columns: the real matrix-core kernel on the
  // interleaved view (alpha folded into the operand); M = N = 1: the real
  // whole-row kernel on the same view
  if (beta.y == 0 && ((M >= 2 && M <= 32 && N <= 32) || (M == 1 && N == 1)))
    return launch_tsmm_cview<T>(h, blockCount, M, N, K, (const T*)A, lda, alpha.x,
                                alpha.y, (const T*)B, ldb, beta.x, (T*)C, ldc,
                                (hipStream_t)stream);
  return launch_tsmm_lds<T>(h, blockCount, 2 * M, 2 * N, K, (const T*)A,
                            2 * lda, alpha.x, alpha.y, (const T*)B, ldb,
                            beta.x, beta.y, (T*)C, 2 * ldc, true,
                            (hipStream_t)stream);
}

skinny_status_t htsmttsm_any(const char* fn, skinny_handle_t h,
                             size_t blockCount, int64_t M, int64_t N, int64_t K,
                             const double* A, int64_t lda, double alpha,
                             const double* B, int64_t ldb, double beta,
                             double* C, bool dd_out, int64_t ldc,
                             skinny_stream_t stream) {
  skinny_status_t s = check_common(fn, M, N, K, A, B, C);
  if (s != SKINNY_OK) return s;
  if (lda < std::max<int64_t>(1, M) || ldb < std::max<int64_t>(1, N) ||
      ldc < std::max<int64_t>(1, M))
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": leading dimension too small");
  if (M == 0 || N == 0) return SKINNY_OK;
  if (M > 1024 || N > 1024)
    return fail(SKINNY_UNSUPPORTED, std::string(fn) + ": M, N <= 1024 only");
  h = resolve(h);
  if (!h) return no_handle(fn);
  return launch_htsmttsm(h, blockCount, M, N, K, A, lda, alpha, B, ldb, beta, C,
                         dd_out, ldc, (hipStream_t)stream);
}

}  // namespace

extern "C" {

size_t skinny_tsmttsm_workspace(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                int elem_bytes) {
  h = resolve(h);
  if (!h || M <= 0 || N <= 0) return 0;
  // complex double runs on the interleaved real view (2M x 2N, 8-byte
  // partials), as the call itself plans it
  TsmttsmPlan p = elem_bytes == 16
                      ? plan_tsmttsm(h, blockCount, 2 * M, 2 * N, K, 8, 8, TT_DC, M, N)
                      : plan_tsmttsm(h, blockCount, M, N, K, elem_bytes, elem_bytes);
  return p.scratch_bytes;
}

int skinny_tuned_blockcount_tsmttsm(skinny_handle_t h, int64_t M, int64_t N,
                                    int64_t K, int elem_bytes) {
  h = resolve(h);
  if (!h || M <= 0 || N <= 0) return 0;
  const TsmttsmPlan p = plan_tsmttsm(h, 0, M, N, K, elem_bytes, elem_bytes);
  if (!p.fixedG && elem_bytes == 8) {
    const int g = tsmttsm_pair_default_blocks(h, M, N, K);
    if (g > 0) return g;
  }
  return p.G;
}

int skinny_tuned_blockcount_tsmm(skinny_handle_t h, int64_t M, int64_t N,
                                 int64_t K, int elem_bytes) {
  h = resolve(h);
  if (!h) return 0;
  return tuned_blocks_tsmm(h, M, N, K, elem_bytes);
}

int skinny_tuned_blockcount_tsmm_ex(skinny_handle_t h, int64_t M, int64_t N,
                                    int64_t K, int elem_bytes, int beta_zero,
                                    int in_place) {
  h = resolve(h);
  if (!h || M < 0 || N < 0 || K < 0) return 0;
  return tuned_blocks_tsmm(h, M, N, K, elem_bytes, true, beta_zero != 0,
                           in_place != 0);
}

skinny_status_t skinny_dtsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const double* A, int64_t lda, double alpha,
                                const double* B, int64_t ldb, double beta,
                                double* C, int64_t ldc,
                                skinny_stream_t stream) {
  return tsmttsm_real<double>("skinny_dtsmttsm", h, blockCount, M, N, K, A,
                              lda, alpha, B, ldb, beta, C, ldc, stream);
}

skinny_status_t skinny_stsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const float* A, int64_t lda, float alpha,
                                const float* B, int64_t ldb, float beta,
                                float* C, int64_t ldc, skinny_stream_t stream) {
  return tsmttsm_real<float>("skinny_stsmttsm", h, blockCount, M, N, K, A, lda,
                             alpha, B, ldb, beta, C, ldc, stream);
}

skinny_status_t skinny_dtsmm(skinny_handle_t h, size_t blockCount, int64_t M,
                             int64_t N, int64_t K, const double* A,
                             int64_t lda, double alpha, const double* B,
                             int64_t ldb, double beta, double* C, int64_t ldc,
                             skinny_stream_t stream) {
  return tsmm_real<double>("skinny_dtsmm", h, blockCount, M, N, K, A, lda,
                           alpha, B, ldb, beta, C, ldc, stream);
}

skinny_status_t skinny_stsmm(skinny_handle_t h, size_t blockCount, int64_t M,
                             int64_t N, int64_t K, const float* A, int64_t lda,
                             float alpha, const float* B, int64_t ldb,
                             float beta, float* C, int64_t ldc,
                             skinny_stream_t stream) {
  return tsmm_real<float>("skinny_stsmm", h, blockCount, M, N, K, A, lda,
                          alpha, B, ldb, beta, C, ldc, stream);
}

skinny_status_t skinny_ztsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const skinny_double_complex* A, int64_t lda,
                                skinny_double_complex alpha,
                                const skinny_double_complex* B, int64_t ldb,
                                skinny_double_complex beta,
                                skinny_double_complex* C, int64_t ldc,
                                skinny_stream_t stream) {
  return tsmttsm_cplx<double>("skinny_ztsmttsm", h, blockCount, M, N, K, A,
                              lda, alpha, B, ldb, beta, C, ldc, stream);
}

skinny_status_t skinny_ctsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const skinny_float_complex* A, int64_t lda,
                                skinny_float_complex alpha,
                                const skinny_float_complex* B, int64_t ldb,
                                skinny_float_complex beta,
                                skinny_float_complex* C, int64_t ldc,
                                skinny_stream_t stream) {
  return tsmttsm_cplx<float>("skinny_ctsmttsm", h, blockCount, M, N, K, A, lda,
                             alpha, B, ldb, beta, C, ldc, stream);
}

skinny_status_t skinny_ztsmm(skinny_handle_t h, size_t blockCount, int64_t M,
                             int64_t N, int64_t K,
                             const skinny_double_complex* A, int64_t lda,
                             skinny_double_complex alpha,
                             const skinny_double_complex* B, int64_t ldb,
                             skinny_double_complex beta,
                             skinny_double_complex* C, int64_t ldc,
                             skinny_stream_t stream) {
  return tsmm_cplx<double>("skinny_ztsmm", h, blockCount, M, N, K, A, lda,
                           alpha, B, ldb, beta, C, ldc, stream);
}

skinny_status_t skinny_ctsmm(skinny_handle_t h, size_t blockCount, int64_t M,
                             int64_t N, int64_t K,
                             const skinny_float_complex* A, int64_t lda,
                             skinny_float_complex alpha,
                             const skinny_float_complex* B, int64_t ldb,
                             skinny_float_complex beta,
                             skinny_float_complex* C, int64_t ldc,
                             skinny_stream_t stream) {
  return tsmm_cplx<float>("skinny_ctsmm", h, blockCount, M, N, K, A, lda,
                          alpha, B, ldb, beta, C, ldc, stream);
}

// In-place block-vector update in the argument order of the GHOST kernels
// (tsmm/fix_ip_ghost.cuh:4-21, var_ip_ghost.cuh:7-27): x is nrows x ncols
// row-major with row stride stridex (ncols = max(ncolsin, ncolsout)), w is
// ncolsin x ncolsout with column stride stridew; x <- alpha x w + beta x.
// The same call as the TSMM entry points with A = C = x.
skinny_status_t skinny_dtsmm_inplace(skinny_handle_t h, double* x,
                                     const double* w, double alpha,
                                     double beta, int64_t nrows,
                                     int64_t ncolsin, int64_t ncolsout,
                                     int64_t stridex, int64_t stridew,
                                     skinny_stream_t stream) {
  return tsmm_real<double>("skinny_dtsmm_inplace", h, 0, ncolsin, ncolsout,
                           nrows, x, stridex, alpha, w, stridew, beta, x,
                           stridex, stream);
}

skinny_status_t skinny_stsmm_inplace(skinny_handle_t h, float* x,
                                     const float* w, float alpha, float beta,
                                     int64_t nrows, int64_t ncolsin,
                                     int64_t ncolsout, int64_t stridex,
                                     int64_t stridew, skinny_stream_t stream) {
  return tsmm_real<float>("skinny_stsmm_inplace", h, 0, ncolsin, ncolsout,
                          nrows, x, stridex, alpha, w, stridew, beta, x,
                          stridex, stream);
}

skinny_status_t skinny_ztsmm_inplace(skinny_handle_t h,
                                     skinny_double_complex* x,
                                     const skinny_double_complex* w,
                                     skinny_double_complex alpha,
                                     skinny_double_complex beta, int64_t nrows,
                                     int64_t ncolsin, int64_t ncolsout,
                                     int64_t stridex, int64_t stridew,
                                     skinny_stream_t stream) {
  return tsmm_cplx<double>("skinny_ztsmm_inplace", h, 0, ncolsin, ncolsout,
                           nrows, x, stridex, alpha, w, stridew, beta, x,
                           stridex, stream);
}

skinny_status_t skinny_ctsmm_inplace(skinny_handle_t h,
                                     skinny_float_complex* x,
                                     const skinny_float_complex* w,
                                     skinny_float_complex alpha,
                                     skinny_float_complex beta, int64_t nrows,
                                     int64_t ncolsin, int64_t ncolsout,
                                     int64_t stridex, int64_t stridew,
                                     skinny_stream_t stream) {
  return tsmm_cplx<float>("skinny_ctsmm_inplace", h, 0, ncolsin, ncolsout,
                          nrows, x, stridex, alpha, w, stridew, beta, x,
                          stridex, stream);
}

skinny_status_t skinny_htsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const double* A, int64_t lda, double alpha,
                                const double* B, int64_t ldb, double beta,
                                double* C, int64_t ldc,
                                skinny_stream_t stream) {
  return htsmttsm_any("skinny_htsmttsm", h, blockCount, M, N, K, A, lda, alpha,
                      B, ldb, beta, C, false, ldc, stream);
}

skinny_status_t skinny_htsmttsm_dd(skinny_handle_t h, size_t blockCount,
                                   int64_t M, int64_t N, int64_t K,
                                   const double* A, int64_t lda, double alpha,
                                   const double* B, int64_t ldb, double beta,
                                   skinny_dd* C, int64_t ldc,
                                   skinny_stream_t stream) {
  return htsmttsm_any("skinny_htsmttsm_dd", h, blockCount, M, N, K, A, lda,
                      alpha, B, ldb, beta, (double*)C, true, ldc, stream);
}

}  // extern "C"
