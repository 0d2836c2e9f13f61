// Autotuner: pick the unroll variant and the split-K / grid size per
// (op, type, M, N, K class) by timing the candidates on the device.
//
// Replaces the reference's hand-fitted blends (tsmm/fix_blend.cuh:13-28,
// tsmm/varip_blend.cuh:13-26) and perf.cu's grid sweep (perf.cu:186-204,
// blockCount = 1..8 x SMs, median of 3): the same sweep is run once per
// shape class and remembered in the handle, and calls with blockCount == 0
// use the remembered choice.  The table exports to and imports from text, so
// a benchmark database can keep it between runs.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <vector>

#include "internal.hpp"

namespace skinny {

static int log2_class(int64_t K) {
  int c = 0;
  while ((1LL << (c + 1)) <= K && c < 62) c++;
  return c;
}

bool tune_lookup(skinny_handle_t h, int op, int dtype, int64_t M, int64_t N,
                 int64_t K, TuneEntry* out) {
  std::lock_guard<std::mutex> lock(h->tune_mu);
  auto it = h->tune.find(TuneKey{op, dtype, (int)M, (int)N, log2_class(K)});
  if (it == h->tune.end()) return false;
  *out = it->second;
  return true;
}

static void tune_store(skinny_handle_t h, int op, int dtype, int64_t M,
                       int64_t N, int64_t K, TuneEntry e) {
  std::lock_guard<std::mutex> lock(h->tune_mu);
  h->tune[TuneKey{op, dtype, (int)M, (int)N, log2_class(K)}] = e;
}

template <typename F>
static float median_ms(F&& launch, hipStream_t st, int reps) {
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  launch();
  std::vector<float> t;
  for (int r = 0; r < reps; r++) {
    (void)hipEventRecord(e0, st);
    launch();
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    t.push_back(ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

}  // namespace skinny

using namespace skinny;

extern "C" {

// op 0: TSMTTSM, 1: TSMM.  dtype 0: double (the only tuned type so far).
skinny_status_t skinny_autotune(skinny_handle_t h, int op, int dtype,
                                int64_t M, int64_t N, int64_t K, int reps,
                                double* best_ms) {
  if (op < 0 || op > 1 || dtype != 0 || M < 1 || N < 1 || K < 1 || M > 64 ||
      N > 64)
    return fail(SKINNY_UNSUPPORTED, "autotune: fp64, 1 <= M, N <= 64");
  h = resolve(h);
  if (!h) return no_handle("autotune");
  if (reps < 1) reps = 3;
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "autotune stream");
  const int64_t W = std::max(M, N);
  double *A = nullptr, *B = nullptr, *C = nullptr;
  e = hipMalloc(&A, sizeof(double) * K * W);
  if (e == hipSuccess) e = hipMalloc(&B, sizeof(double) * K * W);
  if (e == hipSuccess) e = hipMalloc(&C, sizeof(double) * std::max<int64_t>(W * W, 64));
  skinny_status_t s = SKINNY_OK;
  if (e != hipSuccess) {
    s = hip_fail(e, "autotune alloc");
  } else {
    s = skinny_fill_uniform(A, K * W, 17, 0, 0, st);
    if (s == SKINNY_OK) s = skinny_fill_uniform(B, K * W, 17, 1, 0, st);
    if (s == SKINNY_OK) s = skinny_fill_uniform(C, W * W, 17, 2, 0, st);
  }
  TuneEntry best{0, 0};
  float best_t = 1e30f;
  if (s == SKINNY_OK) {
    const double mults[] = {0.5, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0, 8.0};
    for (int uvar = 0; uvar < (op == 0 ? 2 : 1); uvar++) {
      for (double mu : mults) {
        const int G = std::max(1, (int)(mu * h->cu_count));
        skinny_status_t ls = SKINNY_OK;
        auto launch = [&]() {
          if (op == 0) {
            TsmttsmPlan p = plan_tsmttsm(h, G, M, N, K, 8, 8);
            p.uvar = uvar;
            ls = launch_tsmttsm_real<double>(h, p, M, N, K, A, M, 1.0, B, N,
                                             0.0, C, M, st);
          } else {
            ls = launch_tsmm_real<double>(h, G, M, N, K, A, M, 1.0, C, M, 0.0,
                                          B, N, st);
          }
        };
        float t = median_ms(launch, st, reps);
        if (ls != SKINNY_OK) {
          s = ls;
          break;
        }
        if (t < best_t) {
          best_t = t;
          best = TuneEntry{uvar, G};
        }
      }
      if (s != SKINNY_OK) break;
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(A);
  (void)hipFree(B);
  (void)hipFree(C);
  (void)hipStreamDestroy(st);
  if (s != SKINNY_OK) return s;
  tune_store(h, op, dtype, M, N, K, best);
  if (best_ms) *best_ms = best_t;
  return SKINNY_OK;
}

// Lines "op dtype M N log2K uvar G"; returns the bytes needed (incl. NUL).
size_t skinny_tune_export(skinny_handle_t h, char* buf, size_t len) {
  h = resolve(h);
  if (!h) return 0;
  std::ostringstream os;
  {
    std::lock_guard<std::mutex> lock(h->tune_mu);
    for (const auto& kv : h->tune)
      os << kv.first.op << ' ' << kv.first.dtype << ' ' << kv.first.M << ' '
         << kv.first.N << ' ' << kv.first.kclass << ' ' << kv.second.uvar
         << ' ' << kv.second.G << '\n';
  }
  const std::string str = os.str();
  if (buf && len > 0) {
    const size_t n = std::min(len - 1, str.size());
    std::memcpy(buf, str.data(), n);
    buf[n] = '\0';
  }
  return str.size() + 1;
}

skinny_status_t skinny_tune_import(skinny_handle_t h, const char* text) {
  h = resolve(h);
  if (!h || !text) return fail(SKINNY_INVALID_ARG, "tune import: bad arguments");
  std::istringstream is(text);
  int op, dtype, M, N, kc, uvar, G;
  std::lock_guard<std::mutex> lock(h->tune_mu);
  while (is >> op >> dtype >> M >> N >> kc >> uvar >> G) {
    if (G < 1 || uvar < 0 || uvar > 1)
      return fail(SKINNY_INVALID_ARG, "tune import: bad entry");
    h->tune[TuneKey{op, dtype, M, N, kc}] = TuneEntry{uvar, G};
  }
  return SKINNY_OK;
}

void skinny_tune_clear(skinny_handle_t h) {
  h = resolve(h);
  if (!h) return;
  std::lock_guard<std::mutex> lock(h->tune_mu);
  h->tune.clear();
}

}  // extern "C"
