// C++ drop-in surface of libskinny.
//
// Mirrors the reference operator boundary (versions.hpp:43-177, types.hpp:9-59):
//
//   * the type set FR / DR / FC / DC / HR with htype, dtype, dItype and
//     flopsPerCell (types.hpp:9-57), as structs skinny::TypeFR .. TypeHR
//     (the bare names are the selecting macros);
//   * MatmulFunctionType, the std::function every variant is called through
//     (versions.hpp:43-45);
//   * templates skinny::tsmttsm<T, iT> and skinny::tsmm<T> with the reference
//     argument order (blockCount, M, N, K, A, lda, alpha, B, ldb, beta, C, ldc),
//     plus skinny::fixed::tsmttsm<T, iT, M, N> / tsmm<T, M, N> that accept only
//     their compile-time widths, as the reference's templates do
//     (e.g. genv3.cuh:86-89: `if (varM != M || varN != N) return false;`);
//   * getEnabledTSMTTSMVersions() / getEnabledTSMMVersions() registries when
//     one of the FR/DR/FC/DC/HR macros selects the type (versions.hpp:47-177).
//
// Every call forwards to the C ABI in skinny.h on the default handle and the
// null stream.  `true` means success, `false` means the arguments are not
// supported (the reference's SKIP); any other failure prints the library's
// message and exits, as GPU_ERROR does (gpu_error.cuh:4-13), unless
// skinny::set_abort_on_error(false) is called.
#pragma once

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../skinny.h"
#include "pseudoquad.hpp"

namespace skinny {

inline bool& abort_on_error_flag() {
  static bool v = true;
  return v;
}
inline void set_abort_on_error(bool v) { abort_on_error_flag() = v; }

inline bool status_to_bool(skinny_status_t s) {
  if (s == SKINNY_OK) return true;
  if (s == SKINNY_UNSUPPORTED) return false;
  if (abort_on_error_flag()) {
    std::fprintf(stderr, "skinny error %d (%s): %s\n", (int)s,
                 skinny_status_string(s), skinny_last_error());
    std::exit(int(s));
  }
  return false;
}

// ---- type set (types.hpp:9-57) --------------------------------------------
struct TypeFR {
  typedef float htype;
  typedef float dtype;
  typedef float dItype;
  static constexpr int flopsPerCell = 2;
  static const char* name() { return "FR"; }
};
struct TypeDR {
  typedef double htype;
  typedef double dtype;
  typedef double dItype;
  static constexpr int flopsPerCell = 2;
  static const char* name() { return "DR"; }
};
struct TypeFC {
  typedef std::complex<float> htype;
  typedef skinny_float_complex dtype;
  typedef skinny_float_complex dItype;
  static constexpr int flopsPerCell = 8;
  static const char* name() { return "FC"; }
};
struct TypeDC {
  typedef std::complex<double> htype;
  typedef skinny_double_complex dtype;
  typedef skinny_double_complex dItype;
  static constexpr int flopsPerCell = 8;
  static const char* name() { return "DC"; }
};
struct TypeHR {
  typedef double htype;
  typedef double dtype;
  typedef PseudoQuad dItype;
  static constexpr int flopsPerCell = 9;
  static const char* name() { return "HR"; }
};

inline skinny_float_complex makeDtype(std::complex<float> v) {
  return {v.real(), v.imag()};
}
inline skinny_double_complex makeDtype(std::complex<double> v) {
  return {v.real(), v.imag()};
}
inline float makeDtype(float v) { return v; }
inline double makeDtype(double v) { return v; }

// ---- runtime-width entry points -------------------------------------------
template <typename T, typename iT = T>
bool tsmttsm(size_t blockCount, int M, int N, int K, const T* A, int lda,
             T alpha, const T* B, int ldb, T beta, T* C, int ldc);

template <>
inline bool tsmttsm<double, double>(size_t bc, int M, int N, int K,
                                    const double* A, int lda, double alpha,
                                    const double* B, int ldb, double beta,
                                    double* C, int ldc) {
  return status_to_bool(skinny_dtsmttsm(nullptr, bc, M, N, K, A, lda, alpha, B,
                                        ldb, beta, C, ldc, nullptr));
}
template <>
inline bool tsmttsm<float, float>(size_t bc, int M, int N, int K,
                                  const float* A, int lda, float alpha,
                                  const float* B, int ldb, float beta,
                                  float* C, int ldc) {
  return status_to_bool(skinny_stsmttsm(nullptr, bc, M, N, K, A, lda, alpha, B,
                                        ldb, beta, C, ldc, nullptr));
}

template <>
inline bool tsmttsm<skinny_double_complex, skinny_double_complex>(
    size_t bc, int M, int N, int K, const skinny_double_complex* A, int lda,
    skinny_double_complex alpha, const skinny_double_complex* B, int ldb,
    skinny_double_complex beta, skinny_double_complex* C, int ldc) {
  return status_to_bool(skinny_ztsmttsm(nullptr, bc, M, N, K, A, lda, alpha, B,
                                        ldb, beta, C, ldc, nullptr));
}
template <>
inline bool tsmttsm<skinny_float_complex, skinny_float_complex>(
    size_t bc, int M, int N, int K, const skinny_float_complex* A, int lda,
    skinny_float_complex alpha, const skinny_float_complex* B, int ldb,
    skinny_float_complex beta, skinny_float_complex* C, int ldc) {
  return status_to_bool(skinny_ctsmttsm(nullptr, bc, M, N, K, A, lda, alpha, B,
                                        ldb, beta, C, ldc, nullptr));
}
// HR: double data, PseudoQuad accumulation (types.hpp:49-57)
template <>
inline bool tsmttsm<double, PseudoQuad>(size_t bc, int M, int N, int K,
                                        const double* A, int lda, double alpha,
                                        const double* B, int ldb, double beta,
                                        double* C, int ldc) {
  return status_to_bool(skinny_htsmttsm(nullptr, bc, M, N, K, A, lda, alpha, B,
                                        ldb, beta, C, ldc, nullptr));
}

// FP64 TSMTTSM through S INT8 limbs on the matrix cores (no reference
// counterpart); same signature, so it registers as one more DR variant
template <int S>
inline bool tsmttsm_ozaki(size_t bc, int M, int N, int K, const double* A,
                          int lda, double alpha, const double* B, int ldb,
                          double beta, double* C, int ldc) {
  return status_to_bool(skinny_dtsmttsm_ozaki(nullptr, bc, M, N, K, A, lda,
                                              alpha, B, ldb, beta, C, ldc, S,
                                              nullptr));
}

template <typename T>
bool tsmm(size_t blockCount, int M, int N, int K, const T* A, int lda, T alpha,
          const T* B, int ldb, T beta, T* C, int ldc);

template <>
inline bool tsmm<double>(size_t bc, int M, int N, int K, const double* A,
                         int lda, double alpha, const double* B, int ldb,
                         double beta, double* C, int ldc) {
  return status_to_bool(skinny_dtsmm(nullptr, bc, M, N, K, A, lda, alpha, B,
                                     ldb, beta, C, ldc, nullptr));
}
template <>
inline bool tsmm<float>(size_t bc, int M, int N, int K, const float* A,
                        int lda, float alpha, const float* B, int ldb,
                        float beta, float* C, int ldc) {
  return status_to_bool(skinny_stsmm(nullptr, bc, M, N, K, A, lda, alpha, B,
                                     ldb, beta, C, ldc, nullptr));
}

template <>
inline bool tsmm<skinny_double_complex>(
    size_t bc, int M, int N, int K, const skinny_double_complex* A, int lda,
    skinny_double_complex alpha, const skinny_double_complex* B, int ldb,
    skinny_double_complex beta, skinny_double_complex* C, int ldc) {
  return status_to_bool(skinny_ztsmm(nullptr, bc, M, N, K, A, lda, alpha, B,
                                     ldb, beta, C, ldc, nullptr));
}
template <>
inline bool tsmm<skinny_float_complex>(
    size_t bc, int M, int N, int K, const skinny_float_complex* A, int lda,
    skinny_float_complex alpha, const skinny_float_complex* B, int ldb,
    skinny_float_complex beta, skinny_float_complex* C, int ldc) {
  return status_to_bool(skinny_ctsmm(nullptr, bc, M, N, K, A, lda, alpha, B,
                                     ldb, beta, C, ldc, nullptr));
}

// In-place variant (the reference's tsmm_var_ip_ghost, var_ip_ghost.cuh:29-50,
// registered as VARIPG): false unless C aliases A.
template <typename T>
bool tsmm_inplace(size_t bc, int M, int N, int K, const T* A, int lda, T alpha,
                  const T* B, int ldb, T beta, T* C, int ldc) {
  if ((const void*)A != (const void*)C || lda != ldc) return false;
  return tsmm<T>(bc, M, N, K, A, lda, alpha, B, ldb, beta, C, ldc);
}

// ---- compile-time-width shims (the reference's template shape) ------------
namespace fixed {
template <typename T, typename iT, int M, int N>
bool tsmttsm(size_t bc, int varM, int varN, int K, const T* A, int lda,
             T alpha, const T* B, int ldb, T beta, T* C, int ldc) {
  if (varM != M || varN != N) return false;
  return skinny::tsmttsm<T, iT>(bc, M, N, K, A, lda, alpha, B, ldb, beta, C,
                                ldc);
}
template <typename T, int M, int N>
bool tsmm(size_t bc, int varM, int varN, int K, const T* A, int lda, T alpha,
          const T* B, int ldb, T beta, T* C, int ldc) {
  if (varM != M || varN != N) return false;
  return skinny::tsmm<T>(bc, M, N, K, A, lda, alpha, B, ldb, beta, C, ldc);
}
}  // namespace fixed

template <typename dtype>
using MatmulFunction = std::function<bool(
    const size_t, const int, const int, const int, const dtype*, const int,
    const dtype, const dtype*, const int, const dtype, dtype*, const int)>;

template <typename dtype>
using VersionList = std::vector<std::pair<MatmulFunction<dtype>, std::string>>;

// Registries per type set (versions.hpp:47-177 selects them by macro; the
// macro-selected globals below forward here).  "SKINNY" is the library's
// variant for the type; DR TSMTTSM also lists the Ozaki INT8 split.
template <class Types>
VersionList<typename Types::dtype> tsmttsm_versions() {
  typedef typename Types::dtype dt;
  VersionList<dt> v;
  v.push_back({tsmttsm<dt, typename Types::dItype>, "SKINNY"});
  if constexpr (std::is_same<Types, TypeDR>::value) {
    v.push_back({tsmttsm_ozaki<3>, "OZAKI3"});
  }
  return v;
}
template <class Types>
VersionList<typename Types::dtype> tsmm_versions() {
  typedef typename Types::dtype dt;
  VersionList<dt> v;
  v.push_back({tsmm<dt>, "SKINNY"});
  v.push_back({tsmm_inplace<dt>, "SKINNY_IP"});
  return v;
}

}  // namespace skinny

// ---- reference-style globals, selected by one type macro -------------------
#if defined(FR) || defined(DR) || defined(FC) || defined(DC) || defined(HR)
#if defined(FR)
typedef skinny::TypeFR skinny_types_t;
#elif defined(DR)
typedef skinny::TypeDR skinny_types_t;
#elif defined(FC)
typedef skinny::TypeFC skinny_types_t;
#elif defined(DC)
typedef skinny::TypeDC skinny_types_t;
#else
typedef skinny::TypeHR skinny_types_t;
#endif
typedef skinny_types_t::htype htype;
typedef skinny_types_t::dtype dtype;
typedef skinny_types_t::dItype dItype;
static const std::string types = skinny_types_t::name();
static const int flopsPerCell = skinny_types_t::flopsPerCell;

using MatmulFunctionType = skinny::MatmulFunction<dtype>;

inline std::vector<std::pair<MatmulFunctionType, std::string>>
getEnabledTSMTTSMVersions() {
  return skinny::tsmttsm_versions<skinny_types_t>();
}
inline std::vector<std::pair<MatmulFunctionType, std::string>>
getEnabledTSMMVersions() {
  return skinny::tsmm_versions<skinny_types_t>();
}
#endif
