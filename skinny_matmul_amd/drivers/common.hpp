// Shared plumbing of the C++ drivers (test, perf, numeric_tests, svqb):
// HIP error checks, device buffers, event timing, the type-set dispatch and
// device data generation through the library's counter RNG.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "skinny/skinny.hpp"

#define HIP_CHECK(expr)                                                     \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      std::fprintf(stderr, "%s:%d: %s: %s\n", __FILE__, __LINE__, #expr,    \
                   hipGetErrorString(e_));                                  \
      std::exit(3);                                                         \
    }                                                                       \
  } while (0)

#define SKINNY_CHECK(expr)                                                  \
  do {                                                                      \
    skinny_status_t s_ = (expr);                                            \
    if (s_ != SKINNY_OK) {                                                  \
      std::fprintf(stderr, "%s:%d: %s: %s (%s)\n", __FILE__, __LINE__,      \
                   #expr, skinny_status_string(s_), skinny_last_error());   \
      std::exit(3);                                                         \
    }                                                                       \
  } while (0)

namespace drv {

// Device buffer of n elements of T (no initialisation).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) : n(count) {
    HIP_CHECK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// Elapsed milliseconds of f() on the null stream, measured with HIP events.
template <typename F>
double time_ms(F&& f) {
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, nullptr));
  f();
  HIP_CHECK(hipEventRecord(e1, nullptr));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  return ms;
}

inline double median(std::vector<double> v) {
  if (v.empty()) return -1.0;
  std::sort(v.begin(), v.end());
  return v[v.size() / 2];
}

// ---- per-type helpers -------------------------------------------------------
// Number of doubles (or floats) per element and the real scalar type.
template <typename T>
struct Elem;
template <>
struct Elem<double> {
  typedef double real;
  static constexpr int reals = 1;
};
template <>
struct Elem<float> {
  typedef float real;
  static constexpr int reals = 1;
};
template <>
struct Elem<skinny_double_complex> {
  typedef double real;
  static constexpr int reals = 2;
};
template <>
struct Elem<skinny_float_complex> {
  typedef float real;
  static constexpr int reals = 2;
};

// makeDtype over a real value for every dtype
inline double dval(double v, double*) { return v; }
inline float dval(double v, float*) { return (float)v; }
inline skinny_double_complex dval(double v, skinny_double_complex*) {
  return {v, 0.0};
}
inline skinny_float_complex dval(double v, skinny_float_complex*) {
  return {(float)v, 0.f};
}
template <typename T>
inline T make(double v) {
  return dval(v, (T*)nullptr);
}

// exact element comparison on the host (integer data: any correct order of
// summation gives the same value)
inline bool same(double a, double b) { return a == b; }
inline bool same(float a, float b) { return a == b; }
inline bool same(const skinny_double_complex& a, const skinny_double_complex& b) {
  return a.x == b.x && a.y == b.y;
}
inline bool same(const skinny_float_complex& a, const skinny_float_complex& b) {
  return a.x == b.x && a.y == b.y;
}
inline std::string show(double a) { return std::to_string(a); }
inline std::string show(float a) { return std::to_string(a); }
inline std::string show(const skinny_double_complex& a) {
  return "(" + std::to_string(a.x) + "," + std::to_string(a.y) + ")";
}
inline std::string show(const skinny_float_complex& a) {
  return "(" + std::to_string(a.x) + "," + std::to_string(a.y) + ")";
}

// Fill n elements with integers in [lo, hi] (both parts for complex types)
// or with U[-1,1), from the device counter RNG.
template <typename T>
void fill_int(T* p, size_t n, uint64_t seed, uint64_t stream, int lo, int hi) {
  const int64_t nr = (int64_t)n * Elem<T>::reals;
  if constexpr (std::is_same<typename Elem<T>::real, double>::value) {
    SKINNY_CHECK(skinny_fill_int((double*)p, nr, seed, stream, 0, lo, hi, nullptr));
  } else {
    // float data: the same integers (exact in float) narrowed on the host
    DevBuf<double> tmp(nr);
    SKINNY_CHECK(skinny_fill_int(tmp.p, nr, seed, stream, 0, lo, hi, nullptr));
    std::vector<double> h(nr);
    HIP_CHECK(hipMemcpy(h.data(), tmp.p, sizeof(double) * nr, hipMemcpyDeviceToHost));
    std::vector<float> f(h.begin(), h.end());
    HIP_CHECK(hipMemcpy(p, f.data(), sizeof(float) * nr, hipMemcpyHostToDevice));
  }
}

template <typename T>
void fill_uniform(T* p, size_t n, uint64_t seed, uint64_t stream) {
  const int64_t nr = (int64_t)n * Elem<T>::reals;
  if constexpr (std::is_same<typename Elem<T>::real, double>::value) {
    SKINNY_CHECK(skinny_fill_uniform((double*)p, nr, seed, stream, 0, nullptr));
  } else {
    SKINNY_CHECK(skinny_fill_uniform_f((float*)p, nr, seed, stream, 0, nullptr));
  }
}

inline std::string device_name() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  // the marketing name is empty where libdrm's device table is not installed
  if (prop.name[0] == '\0') return prop.gcnArchName;
  return std::string(prop.name) + " (" + prop.gcnArchName + ")";
}

inline int cu_count() {
  int dev = 0, cus = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus;
}

// "tsmttsm" / "tsmm" and FR / DR / FC / DC / HR from the command line; calls
// f.template run<Types>(is_tsmm).
template <typename F>
int dispatch_type(const std::string& type, bool is_tsmm, F&& f) {
  if (type == "FR") return f.template run<skinny::TypeFR>(is_tsmm);
  if (type == "DR") return f.template run<skinny::TypeDR>(is_tsmm);
  if (type == "FC") return f.template run<skinny::TypeFC>(is_tsmm);
  if (type == "DC") return f.template run<skinny::TypeDC>(is_tsmm);
  if (type == "HR") {
    if (is_tsmm) {
      std::fprintf(stderr, "HR is a TSMTTSM accumulation type\n");
      return 2;
    }
    return f.template run<skinny::TypeHR>(is_tsmm);
  }
  std::fprintf(stderr, "unknown type %s (FR, DR, FC, DC, HR)\n", type.c_str());
  return 2;
}

// M and N ranges as the reference drivers take them: none -> defaults,
// "m2" -> 1..m2 (N = M), "m1 m2" -> m1..m2 (N = M), "m1 m2 n1 n2".
struct Range {
  int m1 = 1, m2 = 8, n1 = 0, n2 = 0;  // n1 == n2 == 0: N follows M
};
inline bool parse_range(const std::vector<std::string>& pos, Range& r) {
  if (pos.size() == 1) {
    r.m1 = 1;
    r.m2 = std::stoi(pos[0]);
  } else if (pos.size() == 2 || pos.size() == 4) {
    r.m1 = std::stoi(pos[0]);
    r.m2 = std::stoi(pos[1]);
    if (pos.size() == 4) {
      r.n1 = std::stoi(pos[2]);
      r.n2 = std::stoi(pos[3]);
    }
  } else if (!pos.empty()) {
    return false;
  }
  return r.m1 >= 1 && r.m2 >= r.m1 && r.n1 >= 0 && r.n2 >= r.n1;
}

}  // namespace drv
