// rocBLAS versions of TSMTTSM and TSMM with the operator signature, the
// drivers' library reference (the role tsmttsm_cublas / tsmm_cublas play in
// the reference: tsmttsm/gen_cublas.cuh:11-47, tsmm/cublas.cuh:18-62).
//
// Column-major view of the operands:
//   TSMTTSM  A (K x M row-major, lda) is the column-major M x K matrix A', so
//            C = alpha A'' B'^T + beta C = gemm(N, T, M, N, K).
//   TSMM     C (K x N row-major, ldc) is the column-major N x K matrix C', and
//            C' = alpha B^T A' + beta C' = gemm(T, N, N, K, M).
// Complex types use the plain transpose, as the library does.  In-place TSMM
// (C == A) copies A first: gemm does not allow aliasing.
#pragma once

#include <rocblas/rocblas.h>

#include "common.hpp"

namespace drv {

inline rocblas_handle rocblas() {
  static rocblas_handle h = [] {
    rocblas_handle x;
    if (rocblas_create_handle(&x) != rocblas_status_success) {
      std::fprintf(stderr, "rocblas_create_handle failed\n");
      std::exit(3);
    }
    return x;
  }();
  return h;
}

inline bool gemm(rocblas_operation ta, rocblas_operation tb, int m, int n,
                 int k, double alpha, const double* A, int lda,
                 const double* B, int ldb, double beta, double* C, int ldc) {
  return rocblas_dgemm(rocblas(), ta, tb, m, n, k, &alpha, A, lda, B, ldb,
                       &beta, C, ldc) == rocblas_status_success;
}
inline bool gemm(rocblas_operation ta, rocblas_operation tb, int m, int n,
                 int k, float alpha, const float* A, int lda, const float* B,
                 int ldb, float beta, float* C, int ldc) {
  return rocblas_sgemm(rocblas(), ta, tb, m, n, k, &alpha, A, lda, B, ldb,
                       &beta, C, ldc) == rocblas_status_success;
}
inline bool gemm(rocblas_operation ta, rocblas_operation tb, int m, int n,
                 int k, skinny_double_complex alpha,
                 const skinny_double_complex* A, int lda,
                 const skinny_double_complex* B, int ldb,
                 skinny_double_complex beta, skinny_double_complex* C,
                 int ldc) {
  typedef rocblas_double_complex z;
  return rocblas_zgemm(rocblas(), ta, tb, m, n, k, (const z*)&alpha,
                       (const z*)A, lda, (const z*)B, ldb, (const z*)&beta,
                       (z*)C, ldc) == rocblas_status_success;
}
inline bool gemm(rocblas_operation ta, rocblas_operation tb, int m, int n,
                 int k, skinny_float_complex alpha,
                 const skinny_float_complex* A, int lda,
                 const skinny_float_complex* B, int ldb,
                 skinny_float_complex beta, skinny_float_complex* C, int ldc) {
  typedef rocblas_float_complex c;
  return rocblas_cgemm(rocblas(), ta, tb, m, n, k, (const c*)&alpha,
                       (const c*)A, lda, (const c*)B, ldb, (const c*)&beta,
                       (c*)C, ldc) == rocblas_status_success;
}

template <typename T>
bool tsmttsm_rocblas(size_t, int M, int N, int K, const T* A, int lda,
                     T alpha, const T* B, int ldb, T beta, T* C, int ldc) {
  return gemm(rocblas_operation_none, rocblas_operation_transpose, M, N, K,
              alpha, A, lda, B, ldb, beta, C, ldc);
}

template <typename T>
bool tsmm_rocblas(size_t, int M, int N, int K, const T* A, int lda, T alpha,
                  const T* B, int ldb, T beta, T* C, int ldc) {
  if (A == C) {
    DevBuf<T> copy((size_t)K * lda);
    HIP_CHECK(hipMemcpy(copy.p, A, sizeof(T) * (size_t)K * lda,
                        hipMemcpyDeviceToDevice));
    const bool ok = gemm(rocblas_operation_transpose, rocblas_operation_none,
                         N, K, M, alpha, B, ldb, copy.p, lda, beta, C, ldc);
    HIP_CHECK(hipDeviceSynchronize());
    return ok;
  }
  return gemm(rocblas_operation_transpose, rocblas_operation_none, N, K, M,
              alpha, B, ldb, A, lda, beta, C, ldc);
}

}  // namespace drv
