// Handle, workspace and error plumbing of libskinny, plus the device-side
// counter RNG used to generate benchmark and test inputs in HBM.
#include <cstdio>
#include <cstring>

#include "internal.hpp"

namespace skinny {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

// why the process-wide default handle could not be created (empty if it was)
static std::string g_default_handle_error;

skinny_status_t no_handle(const std::string& fn) {
  return fail(SKINNY_HIP_ERROR, fn + ": no device handle (" +
                                    g_default_handle_error + ")");
}

skinny_status_t fail(skinny_status_t s, const std::string& msg) {
  set_error(msg);
  return s;
}

skinny_status_t hip_fail(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  set_error(m);
  return e == hipErrorOutOfMemory ? SKINNY_OOM : SKINNY_HIP_ERROR;
}

skinny_status_t acquire_workspace(skinny_handle_t h, hipStream_t stream,
                                  size_t scratch_bytes, Workspace* out) {
  const size_t total = kCounterBytes + scratch_bytes;
  std::lock_guard<std::mutex> lock(h->mu);
  Workspace& w = h->ws[stream];
  if (total > w.bytes) {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // grow geometrically so a sweep of shapes reallocates rarely
    size_t want = total < 2 * kCounterBytes ? 2 * kCounterBytes : total;
    if (w.bytes && want < 2 * w.bytes) want = 2 * w.bytes;
    void* p = nullptr;
    e = hipMalloc(&p, want);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(workspace)");
    e = hipMemset(p, 0, kCounterBytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      (void)hipFree(p);
      return hip_fail(e, "workspace counters");
    }
    if (w.base) h->retired.push_back(w.base);
    w.base = static_cast<char*>(p);
    w.bytes = want;
  }
  *out = w;
  return SKINNY_OK;
}

void release_workspace(skinny_handle_t h, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(h->mu);
  auto it = h->ws.find(stream);
  if (it == h->ws.end()) return;
  (void)hipStreamSynchronize(stream);
  if (it->second.base) (void)hipFree(it->second.base);
  h->ws.erase(it);
}

// ---- counter RNG (same function as the host generator) -------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static uint64_t host_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static uint64_t stream_key(uint64_t seed, uint64_t stream_id) {
  return host_mix64(seed ^ host_mix64(stream_id + 0x632BE59BD9B4E019ull));
}

template <int KIND>
__global__ void __launch_bounds__(256)
    fill_kernel(void* out, int64_t n, uint64_t key, int64_t offset, int lo,
                int hi) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint64_t span = (uint64_t)(hi - lo + 1);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    uint64_t b = mix64(key + (uint64_t)(offset + i) * 0xD1B54A32D192ED03ull);
    if (KIND == 0) {
      static_cast<double*>(out)[i] = (double)(b >> 11) * 0x1p-52 - 1.0;
    } else if (KIND == 1) {
      static_cast<float*>(out)[i] =
          (float)((double)(b >> 40) * 0x1p-23 - 1.0);
    } else {
      static_cast<double*>(out)[i] =
          (double)(lo + (int64_t)((b >> 32) % span));
    }
  }
}

template <int KIND>
static skinny_status_t fill(void* out, int64_t n, uint64_t seed,
                            uint64_t stream_id, int64_t offset, int lo, int hi,
                            skinny_stream_t stream) {
  if (n < 0 || (n > 0 && !out) || hi < lo)
    return fail(SKINNY_INVALID_ARG, "fill: bad arguments");
  if (n == 0) return SKINNY_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(fill_kernel<KIND>, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, out, n, stream_key(seed, stream_id),
                     offset, lo, hi);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "fill launch");
  return SKINNY_OK;
}

}  // namespace skinny

using namespace skinny;

extern "C" {

skinny_status_t skinny_create(skinny_handle_t* handle) {
  if (!handle) return fail(SKINNY_INVALID_ARG, "skinny_create: null handle");
  skinny_handle_t h = new skinny_handle_s();
  hipError_t e = hipGetDevice(&h->device);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(e, "hipGetDevice");
  }
  int cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            h->device);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(e, "hipDeviceGetAttribute");
  }
  h->cu_count = cus > 0 ? cus : 256;
  Workspace w;
  skinny_status_t s = acquire_workspace(h, nullptr, kCounterBytes, &w);
  if (s != SKINNY_OK) {
    delete h;
    return s;
  }
  *handle = h;
  return SKINNY_OK;
}

skinny_status_t skinny_destroy(skinny_handle_t h) {
  if (!h) return SKINNY_OK;
  (void)hipSetDevice(h->device);
  (void)hipDeviceSynchronize();
  for (auto& kv : h->ws)
    if (kv.second.base) (void)hipFree(kv.second.base);
  for (void* p : h->retired) (void)hipFree(p);
  delete h;
  return SKINNY_OK;
}

skinny_status_t skinny_reserve(skinny_handle_t h, size_t bytes) {
  return skinny_reserve_stream(h, bytes, nullptr);
}

skinny_status_t skinny_reserve_stream(skinny_handle_t h, size_t bytes,
                                      skinny_stream_t stream) {
  h = resolve(h);
  if (!h) return no_handle("reserve");
  Workspace w;
  return acquire_workspace(h, (hipStream_t)stream, bytes, &w);
}

skinny_handle_t skinny_default_handle(void) {
  static std::once_flag once;
  static skinny_handle_t h = nullptr;
  std::call_once(once, [] {
    if (skinny_create(&h) != SKINNY_OK) {
      h = nullptr;
      g_default_handle_error = g_last_error;
    }
  });
  return h;
}

const char* skinny_status_string(skinny_status_t s) {
  switch (s) {
    case SKINNY_OK: return "ok";
    case SKINNY_UNSUPPORTED: return "unsupported";
    case SKINNY_INVALID_ARG: return "invalid argument";
    case SKINNY_HIP_ERROR: return "hip error";
    case SKINNY_OOM: return "out of memory";
  }
  return "unknown";
}

const char* skinny_last_error(void) { return g_last_error.c_str(); }

int skinny_cu_count(skinny_handle_t h) {
  h = resolve(h);
  return h ? h->cu_count : 0;
}

skinny_status_t skinny_fill_uniform(double* out, int64_t n, uint64_t seed,
                                    uint64_t stream_id, int64_t offset,
                                    skinny_stream_t stream) {
  return fill<0>(out, n, seed, stream_id, offset, 0, 0, stream);
}

skinny_status_t skinny_fill_uniform_f(float* out, int64_t n, uint64_t seed,
                                      uint64_t stream_id, int64_t offset,
                                      skinny_stream_t stream) {
  return fill<1>(out, n, seed, stream_id, offset, 0, 0, stream);
}

skinny_status_t skinny_fill_int(double* out, int64_t n, uint64_t seed,
                                uint64_t stream_id, int64_t offset, int lo,
                                int hi, skinny_stream_t stream) {
  return fill<2>(out, n, seed, stream_id, offset, lo, hi, stream);
}

}  // extern "C"
