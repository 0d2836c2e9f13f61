// FP64 TSMTTSM for callers that cannot allocate device memory.
//
// skinny_dtsmttsm keeps its split-K partials in the stream's workspace of a
// handle, which the library allocates (and grows) on demand.  This entry
// point runs the same plan and the same kernels with the workspace taken from
// a fixed device array, `scratch`, of kNoallocScratch bytes: arrival counters
// in the first kCounterBytes (zero when the code object is loaded, reset to
// zero by the kernel that uses them), split-K partials after them.  Each GPU
// holds its own copy of the array; its address on a device comes from
// hipGetSymbolAddress, taken once per device with that device current (the
// array's name in host code is not a device address).
//
// Every call on one GPU shares that GPU's array: calls that overlap in time
// on one GPU (different streams, or different host threads without ordering)
// race on the partials and must not be issued.
#include <mutex>

#include "internal.hpp"

namespace skinny {

constexpr size_t kNoallocPartials = SKINNY_NOALLOC_PARTIAL_BYTES;
constexpr size_t kNoallocScratch = kCounterBytes + kNoallocPartials;
constexpr int kNoallocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoallocScratch];

struct NoallocDevice {
  bool ready = false;
  void* ptr = nullptr;         // this device's copy of `scratch`
  skinny_handle_s plain;       // CU count only: no workspace, no tuned entries
};

static std::mutex g_noalloc_mu;
static NoallocDevice g_noalloc[kNoallocMaxDevices];

// The device's scratch address (and a plain handle for planning), set up on
// the first call on that device.  No allocation: hipGetSymbolAddress returns
// the address of the array the code object already holds.
static skinny_status_t noalloc_device(int dev, NoallocDevice** out) {
  if (dev < 0 || dev >= kNoallocMaxDevices)
    return fail(SKINNY_UNSUPPORTED, "skinny_dtsmttsm_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  NoallocDevice& d = g_noalloc[dev];
  if (!d.ready) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess)
      return hip_fail(e, "hipSetDevice");
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    int cus = 0;
    if (e == hipSuccess)
      e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "skinny_dtsmttsm_noalloc: scratch address");
    if (!p) return fail(SKINNY_HIP_ERROR, "skinny_dtsmttsm_noalloc: null scratch address");
    d.ptr = p;
    d.plain.device = dev;
    d.plain.cu_count = cus > 0 ? cus : 256;
    d.ready = true;
  }
  *out = &d;
  return SKINNY_OK;
}

}  // namespace skinny

using namespace skinny;

extern "C" {

skinny_status_t skinny_dtsmttsm_noalloc(skinny_handle_t h, size_t blockCount,
                                        int64_t M, int64_t N, int64_t K,
                                        const double* A, int64_t lda,
                                        double alpha, const double* B,
                                        int64_t ldb, double beta, double* C,
                                        int64_t ldc, skinny_stream_t stream) {
  static const char* fn = "skinny_dtsmttsm_noalloc";
  if (M < 0 || N < 0 || K < 0)
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": negative size");
  if (M > 0 && N > 0 && !C) return fail(SKINNY_INVALID_ARG, std::string(fn) + ": null C");
  if (M > 0 && N > 0 && K > 0 && (!A || !B))
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": null A or B");
  if (lda < (M > 1 ? M : 1) || ldb < (N > 1 ? N : 1) || ldc < (M > 1 ? M : 1))
    return fail(SKINNY_INVALID_ARG, std::string(fn) + ": leading dimension too small");
  if (M == 0 || N == 0) return SKINNY_OK;
  if (M > 1024 || N > 1024)
    return fail(SKINNY_UNSUPPORTED, std::string(fn) + ": M, N <= 1024 only");
  // the handle (optional) gives the device, the CU count and tuned grids; it
  // is never asked for a workspace, and no default handle is created
  int dev = 0;
  if (h) {
    dev = h->device;
  } else {
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  }
  NoallocDevice* d = nullptr;
  skinny_status_t s = noalloc_device(dev, &d);
  if (s != SKINNY_OK) return s;
  skinny_handle_t ph = h ? h : &d->plain;
  TsmttsmPlan p = plan_tsmttsm(ph, blockCount, M, N, K, 8, 8);
  if (p.scratch_bytes > kNoallocPartials) {
    if (blockCount > 0)
      return fail(SKINNY_INVALID_ARG,
                  std::string(fn) + ": blockCount " + std::to_string(blockCount) +
                      " needs " + std::to_string(p.scratch_bytes) +
                      " bytes of partials, more than the " +
                      std::to_string(kNoallocPartials) + " bytes the library holds");
    // the default grid: the largest split that fits
    size_t g = (size_t)p.G;
    while (g > 1 && p.scratch_bytes > kNoallocPartials) {
      g = g * 3 / 4;
      p = plan_tsmttsm(ph, g, M, N, K, 8, 8);
    }
    if (p.scratch_bytes > kNoallocPartials)
      return fail(SKINNY_INVALID_ARG, std::string(fn) + ": partials exceed the scratch");
  }
  Workspace w;
  w.base = static_cast<char*>(d->ptr);
  w.bytes = kNoallocScratch;
  return launch_tsmttsm_real<double>(ph, p, M, N, K, A, lda, alpha, B, ldb, beta,
                                     C, ldc, (hipStream_t)stream, &w);
}

}  // extern "C"
