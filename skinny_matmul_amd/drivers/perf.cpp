// skinny_perf: throughput of every registered variant over a range of block
// widths, recorded in BenchDB (the role perf.cu plays for the reference,
// perf.cu:96-315).
//
//   skinny_perf <tsmttsm|tsmm> <FR|DR|FC|DC|HR> [m2 | m1 m2 | m1 m2 n1 n2]
//               [--db benchmarks.db] [--mib MB] [--tune] [--usr1 NAME=VALUE]
//
// * Buffers of --mib MiB per tall matrix (default 4096) hold U[-1,1) data
//   from the device counter RNG.
// * Per (M, N, variant): K doubles from maxK / 100 until a call takes 5 ms
//   (or K reaches the buffer), then for A*B and self (the Gram / in place)
//   and the beta values the reference measures (TSMTTSM beta = 1, TSMM
//   beta = 0 and 1), blockCount 0 (the library's tuned choice) and
//   1..8 x CUs are timed, median of 3 HIP-event timings each; the best is
//   printed and stored.
// * Bytes and flops follow perf.cu:209-231:
//     TSMTTSM  bw = (M + N [A != B]) K s / t,  flops = M N K fpc / t;
//     TSMM     beta = 0: (M K + N K + M N) s / t,  M N K fpc / t;
//              beta = 1: (M K + 2 N K + M N) s / t (in place: one N K),
//                        (M N K + N K) fpc / t.
//   The reference's CUPTI columns (l2bw, sharedbw, occupancy, drambw) are
//   stored as 0 by the sweep; --counters fills drambw and occupancy later.
// * Variants: the library's registry (SKINNY, OZAKI3, SKINNY_IP) and, for
//   every type but HR, ROCBLAS: the vendor library's gemm with the same
//   operand mapping, timed like the others (the CUBLAS entries of the
//   reference registries, versions.hpp:70-71, 105-106); it takes no grid.
// * --tune runs the library autotuner for each shape of the type
//   before timing and keeps the tuned table in BenchDB's `tuning` table; the
//   stored table is loaded at start.
// * --counters MARKS replays the rows already stored for this operation,
//   device and type set, once each, for a profiler collecting hardware
//   counters per dispatch (rocprofv3 --pmc): first one HBM copy
//   microbenchmark over 1 GiB arrays (the byte-counter calibration), then per
//   row a one-element fill_uniform (start marker), the row's call with its
//   stored K, blockCount, beta and aliasing, and a one-element fill_int (end
//   marker).  MARKS gets one tab-separated line per replayed row (the row's
//   key columns and stored time).  skinny_matmul_amd.benchdb_counters turns
//   the counters into the drambw and occupancy columns.
#include <iomanip>
#include <iostream>

#include <fstream>

#include "benchdb.hpp"
#include "common.hpp"
#include "rocblas_ref.hpp"

#ifndef SKINNY_BRANCH
#define SKINNY_BRANCH "main"
#endif

namespace {

struct Opts {
  std::string db = "benchmarks.db";
  size_t mib = 4096;
  bool tune = false;
  std::string counters;  // marks file of a --counters replay
  std::string usr1_name, usr1_val;
  drv::Range range;
};

constexpr int kTuneDtype(const char* name) {
  return name[0] == 'D' && name[1] == 'R' ? 0
         : name[0] == 'F' && name[1] == 'R' ? 1
         : name[0] == 'D' && name[1] == 'C' ? 2
         : name[0] == 'F' && name[1] == 'C' ? 3
                                            : 4;
}

struct Perf {
  Opts o;

  template <class Types>
  int run(bool is_tsmm) {
    typedef typename Types::dtype T;
    const std::string multype = is_tsmm ? "TSMM" : "TSMTTSM";
    const std::string device = drv::device_name();
    const int cus = drv::cu_count();
    drv::BenchDB db(o.db);
    const std::string stored = db.load_tuning(device);
    if (!stored.empty()) SKINNY_CHECK(skinny_tune_import(nullptr, stored.c_str()));

    const size_t total = o.mib * (1u << 20) / sizeof(T);
    const int mmax = o.range.m2, nmax = o.range.n2 ? o.range.n2 : o.range.m2;
    const size_t small = (size_t)std::max(mmax, nmax) * std::max(mmax, nmax);
    drv::DevBuf<T> A(total), B(is_tsmm ? small : total), C(is_tsmm ? total : small);
    drv::fill_uniform(A.p, A.n, 1, 0);
    drv::fill_uniform(B.p, B.n, 1, 1);
    drv::fill_uniform(C.p, C.n, 1, 2);
    HIP_CHECK(hipDeviceSynchronize());

    auto versions = is_tsmm ? skinny::tsmm_versions<Types>()
                            : skinny::tsmttsm_versions<Types>();
    const std::string tname = Types::name();
    if (tname != "HR") {
      if (is_tsmm)
        versions.push_back({drv::tsmm_rocblas<T>, "ROCBLAS"});
      else
        versions.push_back({drv::tsmttsm_rocblas<T>, "ROCBLAS"});
    }
    skinny::set_abort_on_error(true);
    const double s = sizeof(T), fpc = Types::flopsPerCell;
    if (!o.counters.empty())
      return replay(db, versions, is_tsmm, multype, device, tname, A, B, C, total);

    for (int M = o.range.m1; M <= o.range.m2; M++) {
      const int n1 = o.range.n1 ? o.range.n1 : M, n2 = o.range.n2 ? o.range.n2 : M;
      for (int N = n1; N <= n2; N++) {
        for (const auto& ver : versions) {
          int64_t lda = M, ldb = is_tsmm ? M : N, ldc = is_tsmm ? N : M;
          const int64_t maxK = (int64_t)(total / std::max(lda, is_tsmm ? ldc : ldb));
          const T two = drv::make<T>(2.0);
          // one call; negative time when the variant does not support it
          auto measure = [&](int64_t K, size_t bc, bool self, double beta) {
            const T bt = drv::make<T>(beta);
            bool ok = true;
            const double ms = drv::time_ms([&] {
              if (!self)
                ok = ver.first(bc, M, N, (int)K, A.p, (int)lda, two, B.p, (int)ldb, bt,
                               C.p, (int)ldc);
              else if (is_tsmm)
                ok = ver.first(bc, M, N, (int)K, C.p, (int)ldc, two, B.p, (int)ldb, bt,
                               C.p, (int)ldc);
              else
                ok = ver.first(bc, M, N, (int)K, A.p, (int)lda, two, A.p, (int)lda, bt,
                               C.p, (int)ldc);
            });
            return ok ? ms * 1e-3 : -1.0;
          };
          int64_t K = maxK / 100 + 1;
          (void)measure(K, 0, false, 1.0);
          double t = measure(K, 0, false, 1.0);
          while (t > 0 && t < 0.005 && K < maxK) {
            K = std::min(maxK, 2 * K);
            t = measure(K, 0, false, 1.0);
          }
          if (t < 0) continue;  // variant does not support this width
          if (o.tune && ver.second == "SKINNY") {
            double best = 0;
            SKINNY_CHECK(skinny_autotune(nullptr, is_tsmm ? 1 : 0, kTuneDtype(Types::name()),
                                         M, N, K, 3, &best));
          }
          const int selfs = (M == N || is_tsmm) ? 1 : 0;
          for (int self = 0; self <= selfs; self++) {
            if (self && is_tsmm) ldc = lda = std::max(M, N);
            for (double beta = is_tsmm ? 0.0 : 1.0; beta <= 1.0; beta += 1.0) {
              std::vector<size_t> counts = {0};
              if (ver.second != "ROCBLAS")  // the vendor library takes no grid
                for (int k = 1; k <= 8; k++) counts.push_back((size_t)k * cus);
              double bestT = -1;
              size_t bestBC = 0;
              for (size_t bc : counts) {
                std::vector<double> ts;
                for (int rep = 0; rep < 3; rep++) {
                  const double x = measure(K, bc, self, beta);
                  if (x > 0) ts.push_back(x);
                }
                const double m = drv::median(ts);
                if (m > 0 && (bestT < 0 || m < bestT)) {
                  bestT = m;
                  bestBC = bc;
                }
              }
              if (bestT <= 0) continue;
              double bw, flops;
              const double Md = M, Nd = N, Kd = (double)K;
              if (is_tsmm) {
                if (beta == 0.0) {
                  flops = Md * Nd * Kd * fpc / bestT * 1e-9;
                  bw = (Md * Kd + Nd * Kd + Md * Nd) * s / bestT * 1e-9;
                } else {
                  flops = (Md * Nd * Kd + Nd * Kd) * fpc / bestT * 1e-9;
                  bw = (self ? (Md * Kd + Nd * Kd + Md * Nd)
                             : (Md * Kd + 2 * Nd * Kd + Md * Nd)) * s / bestT * 1e-9;
                }
              } else {
                bw = (Md + (self ? 0.0 : Nd)) * Kd * s / bestT * 1e-9;
                flops = Md * Nd * Kd * fpc / bestT * 1e-9;
              }
              std::cout << multype << " " << device << " " << std::setw(4) << M << " "
                        << std::setw(4) << N << " " << std::setw(2) << beta << "  "
                        << (self ? "self" : "A*B ") << " " << std::setw(8) << ver.second
                        << " " << std::setw(10) << K << " " << std::setw(5) << bestBC
                        << " " << std::setw(9) << std::setprecision(4) << bestT * 1e3
                        << " ms " << std::setw(8) << std::setprecision(4) << flops
                        << " GF/s " << std::setw(8) << bw << " GB/s\n"
                        << std::flush;
              db.insert({{"multype", multype}, {"device", device}, {"types", Types::name()},
                         {"M", M}, {"N", N}, {"name", ver.second}, {"inplace", self},
                         {"zerobeta", beta == 0.0 ? 1 : 0}, {"branch", SKINNY_BRANCH},
                         {"usr1_name", o.usr1_name}, {"usr1_val", o.usr1_val}},
                        {{"blockCount", (long long)bestBC}, {"K", (long long)K},
                         {"time", bestT}, {"flops", flops}, {"bw", bw}, {"l2bw", 0.0},
                         {"sharedbw", 0.0}, {"occupancy", 0.0}, {"drambw", 0.0}});
            }
          }
        }
      }
    }
    if (o.tune) {
      const size_t n = skinny_tune_export(nullptr, nullptr, 0);
      std::string text(n, '\0');
      skinny_tune_export(nullptr, &text[0], n);
      text.resize(n ? n - 1 : 0);
      db.save_tuning(device, text);
    }
    db.flush();
    return 0;
  }

  // --counters: replay the stored rows between marker kernels (see above).
  template <class VL, class T>
  int replay(drv::BenchDB& db, const VL& versions, bool is_tsmm, const std::string& multype,
             const std::string& device, const std::string& tname, drv::DevBuf<T>& A,
             drv::DevBuf<T>& B, drv::DevBuf<T>& C, size_t total) {
    std::ofstream marks(o.counters);
    if (!marks) {
      std::fprintf(stderr, "cannot write %s\n", o.counters.c_str());
      return 3;
    }
    double gbs = 0.0;
    SKINNY_CHECK(skinny_mb_stream(1, (size_t)1 << 30, 1, &gbs, nullptr));
    HIP_CHECK(hipDeviceSynchronize());
    drv::DevBuf<double> mark(1);
    const T two = drv::make<T>(2.0);
    int n = 0;
    for (const auto& r : db.rows(multype, device, tname)) {
      if (r.M < o.range.m1 || r.M > o.range.m2) continue;
      const auto* ver = &versions[0];
      bool found = false;
      for (const auto& v : versions)
        if (v.second == r.name) {
          ver = &v;
          found = true;
        }
      if (!found) continue;
      const int M = r.M, N = r.N;
      int64_t lda = M, ldb = is_tsmm ? M : N, ldc = is_tsmm ? N : M;
      if (r.inplace && is_tsmm) ldc = lda = std::max(M, N);
      if ((size_t)r.K * std::max(lda, is_tsmm ? ldc : ldb) > total) continue;
      const T bt = drv::make<T>(r.zerobeta ? 0.0 : 1.0);
      SKINNY_CHECK(skinny_fill_uniform(mark.p, 1, 1, 0, 0, nullptr));
      bool ok;
      if (!r.inplace)
        ok = ver->first((size_t)r.blockCount, M, N, (int)r.K, A.p, (int)lda, two, B.p,
                        (int)ldb, bt, C.p, (int)ldc);
      else if (is_tsmm)
        ok = ver->first((size_t)r.blockCount, M, N, (int)r.K, C.p, (int)ldc, two, B.p,
                        (int)ldb, bt, C.p, (int)ldc);
      else
        ok = ver->first((size_t)r.blockCount, M, N, (int)r.K, A.p, (int)lda, two, A.p,
                        (int)lda, bt, C.p, (int)ldc);
      SKINNY_CHECK(skinny_fill_int(mark.p, 1, 1, 0, 0, 0, 1, nullptr));
      HIP_CHECK(hipDeviceSynchronize());
      if (!ok) continue;
      marks << multype << '\t' << device << '\t' << tname << '\t' << M << '\t' << N << '\t'
            << r.name << '\t' << r.inplace << '\t' << r.zerobeta << '\t' << r.branch << '\t'
            << r.usr1_name << '\t' << r.usr1_val << '\t' << r.time << '\n';
      n++;
    }
    std::cout << "replayed " << n << " rows; marks in " << o.counters << "\n";
    return 0;
  }
};

int usage() {
  std::fprintf(stderr,
               "usage: skinny_perf <tsmttsm|tsmm> <FR|DR|FC|DC|HR> "
               "[m2 | m1 m2 | m1 m2 n1 n2] [--db FILE] [--mib MB] [--tune] "
               "[--usr1 NAME=VALUE] [--counters MARKS]\n");
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return usage();
  const std::string op = argv[1], type = argv[2];
  if (op != "tsmttsm" && op != "tsmm") return usage();
  Perf p;
  std::vector<std::string> pos;
  for (int i = 3; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--db" && i + 1 < argc) p.o.db = argv[++i];
    else if (a == "--mib" && i + 1 < argc) p.o.mib = std::stoul(argv[++i]);
    else if (a == "--tune") p.o.tune = true;
    else if (a == "--counters" && i + 1 < argc) p.o.counters = argv[++i];
    else if (a == "--usr1" && i + 1 < argc) {
      const std::string kv = argv[++i];
      const size_t eq = kv.find('=');
      p.o.usr1_name = kv.substr(0, eq);
      p.o.usr1_val = eq == std::string::npos ? "" : kv.substr(eq + 1);
    } else pos.push_back(a);
  }
  if (!drv::parse_range(pos, p.o.range)) return usage();
  return drv::dispatch_type(type, op == "tsmm", p);
}
