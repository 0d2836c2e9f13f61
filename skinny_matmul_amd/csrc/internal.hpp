// Internal declarations shared by the HIP translation units of libskinny.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "skinny.h"

namespace skinny {
struct TuneKey {
  int op, dtype, M, N, kclass;  // kclass = floor(log2 K)
  bool operator<(const TuneKey& o) const {
    return std::tie(op, dtype, M, N, kclass) <
           std::tie(o.op, o.dtype, o.M, o.N, o.kclass);
  }
};
struct TuneEntry {
  int uvar;  // TSMTTSM: bit 0 doubled unroll, bit 1 one launch; real TSMM: bit 0 other form
  int G;     // workgroups along K
};
// Device workspace of one stream: [counter region (zeroed at allocation,
// reset to zero by the kernels that use it)] [scratch region].
struct Workspace {
  char* base = nullptr;
  size_t bytes = 0;
};
}  // namespace skinny

struct skinny_handle_s {
  int device = 0;
  int cu_count = 256;
  // One workspace per stream: calls on one stream are ordered by it, calls
  // on different streams may overlap.  Outgrown buffers are retired, not
  // freed, until the handle is destroyed, so a call enqueued earlier never
  // sees its buffer disappear.
  std::map<hipStream_t, skinny::Workspace> ws;
  std::vector<void*> retired;
  std::mutex mu;
  // autotuned choices for blockCount == 0
  std::map<skinny::TuneKey, skinny::TuneEntry> tune;
  std::mutex tune_mu;
};

namespace skinny {

// bytes at the start of the workspace reserved for arrival counters
constexpr size_t kCounterBytes = 64 * 1024;
constexpr int kMaxCounters = kCounterBytes / sizeof(unsigned int);

void set_error(const std::string& msg);
skinny_status_t fail(skinny_status_t s, const std::string& msg);
skinny_status_t hip_fail(hipError_t e, const char* what);
// failure of a call that found no handle (the default handle's creation
// error is included)
skinny_status_t no_handle(const std::string& fn);

// The workspace of `stream` with at least `scratch_bytes` of scratch: made on
// first use and grown geometrically when too small (growth allocates, zeroes
// the new counters and synchronises the device; it never happens once the
// stream's workspace is large enough, e.g. after skinny_reserve_stream).
skinny_status_t acquire_workspace(skinny_handle_t h, hipStream_t stream,
                                  size_t scratch_bytes, Workspace* out);

// Drop the workspace of `stream` (a stream about to be destroyed): waits for
// the stream, then frees its buffer.
void release_workspace(skinny_handle_t h, hipStream_t stream);

inline unsigned int* ws_counters(const Workspace& w) {
  return reinterpret_cast<unsigned int*>(w.base);
}
inline void* ws_scratch(const Workspace& w) { return w.base + kCounterBytes; }

inline skinny_handle_t resolve(skinny_handle_t h) {
  return h ? h : skinny_default_handle();
}

// autotuner table lookup (op 0 TSMTTSM, 1 TSMM; dtype a TuneType; M, N the
// widths the caller passed, complex ones for complex types)
enum TuneType { TT_DR = 0, TT_FR = 1, TT_DC = 2, TT_FC = 3, TT_HR = 4 };
bool tune_lookup(skinny_handle_t h, int op, int dtype, int64_t M, int64_t N,
                 int64_t K, TuneEntry* out);
void tune_store(skinny_handle_t h, int op, int dtype, int64_t M, int64_t N,
                int64_t K, TuneEntry e);

// ---- TSMTTSM ------------------------------------------------------------
struct TsmttsmPlan {
  int P = 1;          // packed blocks per 16-lane MFMA tile (small widths)
  int MT = 1, NT = 1; // 16-wide subtiles per wave tile
  int tilesM = 1, tilesN = 1;
  int tileM = 16, tileN = 16;  // C rows / cols per tile
  int G = 1;          // workgroups along K per tile
  int GS = 1;         // first-level reduction group size
  int ngroups = 1;
  int outM = 1, outN = 1;  // entries of one partial (outM x outN)
  int uvar = 0;            // unroll variant: 0 base, 1 doubled
  bool one_pass = false;   // G > 1 partials summed in the same launch
  bool tuned = false;      // G and uvar come from the autotuner table
  bool fixedG = false;     // G given by the caller or the autotuner table
  size_t scratch_bytes = 0;
};

// tdtype, tM, tN: the autotuner key of the call (TuneType and its widths);
// tdtype < 0 derives it from a real type of elem_bytes and M, N.
TsmttsmPlan plan_tsmttsm(skinny_handle_t h, size_t blockCount, int64_t M,
                         int64_t N, int64_t K, int elem_bytes, int partial_bytes,
                         int tdtype = -1, int64_t tM = 0, int64_t tN = 0);

// Default grid of a real FP64 TSMTTSM that runs on the paired-row kernel
// (one operand 5..16 wide, the other narrow), 0 for other shapes.
int tsmttsm_pair_default_blocks(skinny_handle_t h, int64_t M, int64_t N, int64_t K);

// complex (cplx = true): A, B, C point at interleaved (re, im) data and M, N,
// lda, ldb are the real-view sizes (2x the complex ones); ldc stays complex.
// `fixed` (optional) is a workspace of the same layout that the call uses
// instead of the stream's; it must hold kCounterBytes + p.scratch_bytes, with
// its counters at zero.
template <typename T>
skinny_status_t launch_tsmttsm(skinny_handle_t h, const TsmttsmPlan& p,
                               int64_t M, int64_t N, int64_t K, const T* A,
                               int64_t lda, T alpha, T alpha_i, const T* B,
                               int64_t ldb, T beta, T beta_i, T* C,
                               int64_t ldc, bool cplx, hipStream_t stream,
                               const Workspace* fixed = nullptr);

template <typename T>
skinny_status_t launch_tsmttsm_real(skinny_handle_t h, const TsmttsmPlan& p,
                                    int64_t M, int64_t N, int64_t K,
                                    const T* A, int64_t lda, T alpha,
                                    const T* B, int64_t ldb, T beta, T* C,
                                    int64_t ldc, hipStream_t stream,
                                    const Workspace* fixed = nullptr);

// ---- TSMM ---------------------------------------------------------------
// Default grid of a TSMM call; real = false for the complex kernels, whose
// whole-row variant covers only M, N <= 2 of the real view.  beta0 and
// inplace select between the real kernels as the call does.
// tdtype, tM, tN as for plan_tsmttsm (complex kernels pass their complex
// widths; tdtype < 0 with real = false looks nothing up).
int tuned_blocks_tsmm(skinny_handle_t h, int64_t M, int64_t N, int64_t K,
                      int elem_bytes, bool real = true, bool beta0 = true,
                      bool inplace = false, int tdtype = -1, int64_t tM = 0,
                      int64_t tN = 0);

template <typename T>
skinny_status_t launch_tsmm_real(skinny_handle_t h, size_t blockCount,
                                 int64_t M, int64_t N, int64_t K, const T* A,
                                 int64_t lda, T alpha, const T* B, int64_t ldb,
                                 T beta, T* C, int64_t ldc, hipStream_t stream);

// Complex TSMM with a real beta on the real kernels over the interleaved view
// (Mc, Nc complex widths: 2 <= Mc <= 32 with Nc <= 32 on the matrix cores, or
// Mc = Nc = 1 on whole rows; lda, ldb, ldc in complex units).
template <typename T>
skinny_status_t launch_tsmm_cview(skinny_handle_t h, size_t blockCount,
                                  int64_t Mc, int64_t Nc, int64_t K, const T* A,
                                  int64_t lda, T alpha_r, T alpha_i, const T* B,
                                  int64_t ldb, T beta, T* C, int64_t ldc,
                                  hipStream_t stream);

// HR: double in, PseudoQuad accumulation; dd_out writes (s, t) pairs to C
skinny_status_t launch_htsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const double* A, int64_t lda, double alpha,
                                const double* B, int64_t ldb, double beta,
                                double* C, bool dd_out, int64_t ldc,
                                hipStream_t stream);

// B staged in LDS: complex (cplx, real-view M, N, lda, ldc; ldb in complex
// units) and real widths 65..128.
template <typename T>
skinny_status_t launch_tsmm_lds(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K, const T* A,
                                int64_t lda, T alpha, T alpha_i, const T* B,
                                int64_t ldb, T beta, T beta_i, T* C,
                                int64_t ldc, bool cplx, hipStream_t stream);

// FP64 on the 4x4x4 matrix-core instruction with B staged in LDS (real-view
// M in 29..128, N <= 128; same arguments as launch_tsmm_lds).  Returns
// SKINNY_UNSUPPORTED, having launched nothing, outside that range.
skinny_status_t launch_tsmm_q4(skinny_handle_t h, size_t blockCount, int64_t M, int64_t N,
                               int64_t K, const double* A, int64_t lda, double alpha,
                               double alpha_i, const double* B, int64_t ldb, double beta,
                               double beta_i, double* C, int64_t ldc, bool cplx,
                               hipStream_t stream);

// Wide blocks (real-view M or N above 128, up to kTsmmWideMax): the row band
// of A in LDS, B packed into the stream's workspace; same arguments as
// launch_tsmm_lds.
constexpr int kTsmmWideMax = 1024;
template <typename T>
skinny_status_t launch_tsmm_wide(skinny_handle_t h, size_t blockCount,
                                 int64_t M, int64_t N, int64_t K, const T* A,
                                 int64_t lda, T alpha, T alpha_i, const T* B,
                                 int64_t ldb, T beta, T beta_i, T* C,
                                 int64_t ldc, bool cplx, hipStream_t stream);

}  // namespace skinny
