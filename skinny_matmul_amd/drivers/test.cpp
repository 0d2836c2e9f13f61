// skinny_test: exact-parity fuzzer of every registered variant against
// rocBLAS, the role test.cu plays for the reference (test.cu:82-95,
// 128-293).
//
//   skinny_test <tsmttsm|tsmm> <FR|DR|FC|DC|HR> [m2 | m1 m2 | m1 m2 n1 n2]
//               [--samples S] [--mib MB] [--seed X]
//
// For every (M, N) in the range and every variant the registry lists
// (skinny::tsmttsm_versions / tsmm_versions), `samples` random cases per
// (A*B or self, beta in {0, 1}) draw blockCount in 1..200, leading
// dimensions padded by 0..4 and K up to the buffer budget.  Self is the Gram
// A^T A for TSMTTSM and the in-place C = A B for TSMM.  The data are
// integers in [-2, 2] from the device counter RNG, so every correct
// summation order gives the same result and the comparison is exact.  A
// variant returning false for a case is a SKIP ("-"), a match "#", a
// mismatch "X".  The exit status is 1 if any case failed.
#include <iostream>
#include <random>

#include "common.hpp"
#include "rocblas_ref.hpp"

namespace {

struct Opts {
  int samples = 20;
  size_t mib = 128;  // budget per matrix
  uint64_t seed = 0;
  drv::Range range;
};

struct Tester {
  Opts o;

  template <class Types>
  int run(bool is_tsmm) {
    typedef typename Types::dtype T;
    const size_t total = o.mib * (1u << 20) / sizeof(T);
    // the small operand (TSMTTSM C, TSMM B): N columns of ld <= M + 4
    const int mmax = o.range.m2, nmax = o.range.n2 ? o.range.n2 : o.range.m2;
    const size_t smallC = (size_t)(nmax + 4) * (size_t)(mmax + 4);
    const size_t totalA = total;
    const size_t totalB = is_tsmm ? smallC : total;
    const size_t totalC = is_tsmm ? total : smallC;
    drv::DevBuf<T> A0(totalA), B0(totalB), C0(totalC);
    drv::DevBuf<T> A(totalA), B(totalB), C(totalC);
    drv::fill_int(A0.p, totalA, 1000 + o.seed, 0, -2, 2);
    drv::fill_int(B0.p, totalB, 1000 + o.seed, 1, -2, 2);
    drv::fill_int(C0.p, totalC, 1000 + o.seed, 2, -2, 2);
    std::vector<T> got(totalC), want(totalC);

    auto versions = is_tsmm ? skinny::tsmm_versions<Types>()
                            : skinny::tsmttsm_versions<Types>();
    skinny::MatmulFunction<T> reference =
        is_tsmm ? skinny::MatmulFunction<T>(drv::tsmm_rocblas<T>)
                : skinny::MatmulFunction<T>(drv::tsmttsm_rocblas<T>);
    skinny::set_abort_on_error(true);

    std::mt19937_64 gen(o.seed * 7919 + 17);
    auto uni = [&](int64_t lo, int64_t hi) {
      return std::uniform_int_distribution<int64_t>(lo, hi)(gen);
    };
    int failures = 0;
    const drv::Range& r = o.range;
    for (int M = r.m1; M <= r.m2; M++) {
      const int n1 = r.n1 ? r.n1 : M, n2 = r.n2 ? r.n2 : M;
      for (int N = n1; N <= n2; N++) {
        for (const auto& ver : versions) {
          std::cout << M << "xKx" << N << "  " << ver.second << " "
                    << Types::name() << " " << std::flush;
          bool passed = true;
          for (int self = 0; self <= 1; self++) {
            for (int b = 0; b <= 1; b++) {
              for (int t = 0; t < o.samples; t++) {
                const size_t bc = (size_t)uni(1, 200);
                int64_t lda = M + uni(0, 4), ldb, ldc, K;
                if (is_tsmm) {
                  ldb = M + uni(0, 4);
                  ldc = self ? std::max<int64_t>(N + uni(0, 4), M) : N + uni(0, 4);
                  if (self) lda = ldc;
                  K = uni(1, (int64_t)(total / std::max(lda, ldc)));
                } else {
                  ldb = self ? lda : N + uni(0, 4);
                  ldc = M + uni(0, 4);
                  K = uni(1, (int64_t)(total / std::max(lda, ldb)));
                }
                const char res = one_case<T>(ver.first, reference, is_tsmm, M, N,
                                             K, lda, ldb, ldc, bc, self, b,
                                             A0, B0, C0, A, B, C, got, want);
                if (res == 'X') passed = false;
                std::cout << res << std::flush;
              }
            }
          }
          if (!passed) failures++;
          std::cout << (passed ? " Passed\n" : " Failed\n");
        }
      }
    }
    return failures ? 1 : 0;
  }

  template <typename T>
  char one_case(const skinny::MatmulFunction<T>& fn,
                const skinny::MatmulFunction<T>& ref, bool is_tsmm, int M,
                int N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc,
                size_t bc, int self, int beta, drv::DevBuf<T>& A0,
                drv::DevBuf<T>& B0, drv::DevBuf<T>& C0, drv::DevBuf<T>& A,
                drv::DevBuf<T>& B, drv::DevBuf<T>& C, std::vector<T>& got,
                std::vector<T>& want) {
    if (self && !is_tsmm && M != N) return '-';
    const T alpha = drv::make<T>(1.0), bt = drv::make<T>(beta);
    // the region the call writes: TSMTTSM N columns of ldc, TSMM K rows of ldc
    const size_t outN = is_tsmm ? (size_t)K * ldc : (size_t)N * ldc;
    auto call = [&](const skinny::MatmulFunction<T>& f, std::vector<T>& dst) {
      HIP_CHECK(hipMemcpy(A.p, A0.p, sizeof(T) * A.n, hipMemcpyDeviceToDevice));
      HIP_CHECK(hipMemcpy(B.p, B0.p, sizeof(T) * B.n, hipMemcpyDeviceToDevice));
      HIP_CHECK(hipMemcpy(C.p, C0.p, sizeof(T) * C.n, hipMemcpyDeviceToDevice));
      bool ok;
      if (!self)
        ok = f(bc, M, N, (int)K, A.p, (int)lda, alpha, B.p, (int)ldb, bt, C.p, (int)ldc);
      else if (is_tsmm)  // in place: C = C * B
        ok = f(bc, M, N, (int)K, C.p, (int)ldc, alpha, B.p, (int)ldb, bt, C.p, (int)ldc);
      else  // Gram
        ok = f(bc, M, N, (int)K, A.p, (int)lda, alpha, A.p, (int)lda, bt, C.p, (int)ldc);
      HIP_CHECK(hipDeviceSynchronize());
      HIP_CHECK(hipMemcpy(dst.data(), C.p, sizeof(T) * outN, hipMemcpyDeviceToHost));
      return ok;
    };
    if (!call(fn, got)) return '-';
    if (!call(ref, want)) {
      std::fprintf(stderr, "rocBLAS reference failed\n");
      std::exit(3);
    }
    const size_t rows = is_tsmm ? (size_t)K : (size_t)N;
    const size_t cols = is_tsmm ? (size_t)N : (size_t)M;
    for (size_t i = 0; i < rows; i++)
      for (size_t j = 0; j < cols; j++) {
        const size_t e = i * ldc + j;
        if (!drv::same(got[e], want[e])) {
          std::cout << "\n(" << bc << " blocks, " << (self ? "self" : "A*B")
                    << ", beta=" << beta << ", lda=" << lda << ", ldb=" << ldb
                    << ", ldc=" << ldc << ", K=" << K << ") mismatch at " << i
                    << ", " << j << ": " << drv::show(got[e])
                    << " != " << drv::show(want[e]) << " ";
          return 'X';
        }
      }
    return '#';
  }
};

int usage() {
  std::fprintf(stderr,
               "usage: skinny_test <tsmttsm|tsmm> <FR|DR|FC|DC|HR> "
               "[m2 | m1 m2 | m1 m2 n1 n2] [--samples S] [--mib MB] [--seed X]\n");
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return usage();
  const std::string op = argv[1], type = argv[2];
  if (op != "tsmttsm" && op != "tsmm") return usage();
  Tester t;
  std::vector<std::string> pos;
  for (int i = 3; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--samples" && i + 1 < argc) t.o.samples = std::stoi(argv[++i]);
    else if (a == "--mib" && i + 1 < argc) t.o.mib = std::stoul(argv[++i]);
    else if (a == "--seed" && i + 1 < argc) t.o.seed = std::stoull(argv[++i]);
    else pos.push_back(a);
  }
  if (!drv::parse_range(pos, t.o.range)) return usage();
  return drv::dispatch_type(type, op == "tsmm", t);
}
