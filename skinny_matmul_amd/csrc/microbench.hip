// Microbenchmarks measuring this chip's own ceilings: HBM bandwidth for the
// access mixes of the tall-skinny products, the FP64 vector FMA and FP64 /
// INT8 matrix-core peaks.  Counterparts of microbenches/membw.cu (copy, scale,
// update, sum, triad, reduce at a 1 GiB footprint, membw.cu:19-164) and
// microbenches/roofline.cu (an FP64 multiply chain, roofline.cu:19-33),
// rewritten for gfx950: 16-byte accesses, working sets far above the 256 MB
// Infinity Cache, launch shapes from the CU count, and timing with HIP events
// on the caller's stream.
#include <algorithm>
#include <vector>

#include "internal.hpp"

namespace skinny {
namespace {

typedef double d2 __attribute__((ext_vector_type(2)));

// kind 0: read (sum), 1: copy, 2: triad a = b + s c, 3: write, 4: update a += s b
// U accesses of 16 bytes per array and thread in flight (grid-stride); NT
// issues the loads of b and c with the nontemporal hint.
template <int KIND, int U, bool NT>
__device__ __forceinline__ d2 mb_load(const d2* p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}

template <int KIND, int U, bool NT>
__global__ void __launch_bounds__(256)
    stream_kernel(d2* __restrict__ a, const d2* __restrict__ b,
                  const d2* __restrict__ c, int64_t n2, double s,
                  double* __restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  d2 acc = {0.0, 0.0};
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  auto one = [&](int64_t j, d2 vb, d2 vc, d2 va) {
    if (KIND == 0) {
      acc += vb;
    } else if (KIND == 1) {
      a[j] = vb;
    } else if (KIND == 2) {
      a[j] = vb + s * vc;
    } else if (KIND == 3) {
      a[j] = d2{s, s};
    } else {
      a[j] = va + s * vb;
    }
  };
  for (; i + (U - 1) * stride < n2; i += U * stride) {
    d2 vb[U], vc[U], va[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t j = i + u * stride;
      vb[u] = (KIND != 3) ? mb_load<KIND, U, NT>(&b[j]) : d2{0, 0};
      vc[u] = (KIND == 2) ? mb_load<KIND, U, NT>(&c[j]) : d2{0, 0};
      va[u] = (KIND == 4) ? a[j] : d2{0, 0};
    }
#pragma unroll
    for (int u = 0; u < U; u++) one(i + u * stride, vb[u], vc[u], va[u]);
  }
  for (; i < n2; i += stride)
    one(i, (KIND != 3) ? b[i] : d2{0, 0}, (KIND == 2) ? c[i] : d2{0, 0},
        (KIND == 4) ? a[i] : d2{0, 0});
  // keep the read-only loop alive without a store per element
  if (KIND == 0 && acc.x + acc.y == 12345.678) sink[threadIdx.x] = acc.x;
}

template <typename F>
float time_ms(F&& launch, hipStream_t st, int reps) {
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  launch();  // warm-up
  float best = 1e30f;
  for (int r = 0; r < reps; r++) {
    (void)hipEventRecord(e0, st);
    launch();
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// FP64 FMA chains: 8 independent chains per lane
__global__ void __launch_bounds__(256) fma_kernel(double* out, int iters,
                                                  double x) {
  double v[8];
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = x + j + threadIdx.x;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = __builtin_fma(v[j], 0.999999, 1e-7);
  }
  double s = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) s += v[j];
  if (s == 12345.678) out[threadIdx.x] = s;
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// FP64 matrix core: 4 independent accumulator chains per wave
__global__ void __launch_bounds__(256) mfma_f64_kernel(double* out, int iters,
                                                       double x) {
  v4d acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) acc[j] = v4d{0, 0, 0, 0};
  double a = x + threadIdx.x, b = x - threadIdx.x;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int j = 0; j < 4; j++)
      acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  if (s == 12345.678) out[threadIdx.x] = s;
}

// INT8 matrix core (32x32x32, int32 accumulate): 2 chains per wave
__global__ void __launch_bounds__(256) mfma_i8_kernel(int* out, int iters,
                                                      int x) {
  v16i acc[2];
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int r = 0; r < 16; r++) acc[j][r] = 0;
  v4i a = {x, x + 1, x + 2, (int)threadIdx.x};
  v4i b = {x - 1, x, (int)threadIdx.x, x + 3};
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int j = 0; j < 2; j++)
      acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[j], 0, 0, 0);
  }
  int s = 0;
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int r = 0; r < 16; r++) s += acc[j][r];
  if (s == 12345) out[threadIdx.x] = s;
}

}  // namespace
}  // namespace skinny

using namespace skinny;

extern "C" {

skinny_status_t skinny_mb_stream(int kind, size_t bytes_per_array, int reps,
                                 double* gbs, skinny_stream_t stream) {
  if (!gbs || kind < 0 || kind > 4 || bytes_per_array < 4096 || reps < 1)
    return fail(SKINNY_INVALID_ARG, "skinny_mb_stream: bad arguments");
  skinny_handle_t h = skinny_default_handle();
  if (!h) return no_handle("microbench");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n2 = (int64_t)(bytes_per_array / 16);
  const size_t nb = (size_t)n2 * 16;
  void *a = nullptr, *b = nullptr, *c = nullptr, *sink = nullptr;
  hipError_t e = hipMalloc(&a, nb);
  if (e == hipSuccess) e = hipMalloc(&b, nb);
  if (e == hipSuccess) e = hipMalloc(&c, nb);
  if (e == hipSuccess) e = hipMalloc(&sink, 4096);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, nb, st);
  if (e == hipSuccess) e = hipMemsetAsync(b, 0, nb, st);
  if (e == hipSuccess) e = hipMemsetAsync(c, 0, nb, st);
  if (e != hipSuccess) {
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(c);
    (void)hipFree(sink);
    return hip_fail(e, "skinny_mb_stream alloc");
  }
  // best over launch shapes: 1, 2, 4 or 8 workgroups per CU, 1 or 4
  // accesses in flight per thread, and loads with or without the
  // nontemporal hint (the best shape differs by kind: writes and copies peak
  // with few workgroups; reads peak at 2 per CU with the hint, which took a
  // read stream from 6.4 to 7.1 TB/s)
  float ms = 1e30f;
  for (int per : {1, 2, 4, 8}) {
    for (int unroll : {1, 4}) {
     for (int nt : {0, 1}) {
      if (nt && kind == 3) continue;  // no loads
      const int blocks = per * h->cu_count;
      auto launch = [&]() {
        d2* A = (d2*)a;
        const d2* B = (const d2*)b;
        const d2* Cc = (const d2*)c;
        double* sk = (double*)sink;
#define SK_MB1(k, u)                                                                          \
  if (nt)                                                                                     \
    hipLaunchKernelGGL((stream_kernel<k, u, true>), dim3(blocks), dim3(256), 0, st, A, B, Cc, \
                       n2, 1.5, sk);                                                          \
  else                                                                                        \
    hipLaunchKernelGGL((stream_kernel<k, u, false>), dim3(blocks), dim3(256), 0, st, A, B,    \
                       Cc, n2, 1.5, sk);
#define SK_MB(k)      \
  if (unroll == 1) {  \
    SK_MB1(k, 1)      \
  } else {            \
    SK_MB1(k, 4)      \
  }
        switch (kind) {
          case 0: SK_MB(0) break;
          case 1: SK_MB(1) break;
          case 2: SK_MB(2) break;
          case 3: SK_MB(3) break;
          default: SK_MB(4) break;
        }
#undef SK_MB
#undef SK_MB1
      };
      ms = std::min(ms, time_ms(launch, st, reps));
     }
    }
  }
  e = hipGetLastError();
  const double arrays[5] = {1, 2, 3, 1, 3};  // bytes moved per array size
  *gbs = arrays[kind] * (double)nb / (ms * 1e-3) / 1e9;
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(c);
  (void)hipFree(sink);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_stream launch");
  return SKINNY_OK;
}

// kind 0: FP64 vector FMA, 1: FP64 MFMA 16x16x4, 2: INT8 MFMA 32x32x32.
// Result in TFLOP/s (TOP/s for INT8).
skinny_status_t skinny_mb_peak(int kind, int reps, double* tflops,
                               skinny_stream_t stream) {
  if (!tflops || kind < 0 || kind > 2 || reps < 1)
    return fail(SKINNY_INVALID_ARG, "skinny_mb_peak: bad arguments");
  skinny_handle_t h = skinny_default_handle();
  if (!h) return no_handle("microbench");
  hipStream_t st = (hipStream_t)stream;
  void* out = nullptr;
  hipError_t e = hipMalloc(&out, 4096);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_peak alloc");
  const int blocks = 8 * h->cu_count;
  const int iters = kind == 0 ? 4096 : 1024;
  auto launch = [&]() {
    if (kind == 0)
      hipLaunchKernelGGL(fma_kernel, dim3(blocks), dim3(256), 0, st,
                         (double*)out, iters, 1.0);
    else if (kind == 1)
      hipLaunchKernelGGL(mfma_f64_kernel, dim3(blocks), dim3(256), 0, st,
                         (double*)out, iters, 1.0);
    else
      hipLaunchKernelGGL(mfma_i8_kernel, dim3(blocks), dim3(256), 0, st,
                         (int*)out, iters, 1);
  };
  float ms = time_ms(launch, st, reps);
  e = hipGetLastError();
  (void)hipFree(out);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_peak launch");
  const double threads = (double)blocks * 256.0;
  double ops;
  if (kind == 0)
    ops = threads * iters * 8 * 2;  // 8 chains, 2 flops per FMA
  else if (kind == 1)
    ops = threads / 64.0 * iters * 4 * (16.0 * 16 * 4 * 2);
  else
    ops = threads / 64.0 * iters * 2 * (32.0 * 32 * 32 * 2);
  *tflops = ops / (ms * 1e-3) / 1e12;
  return SKINNY_OK;
}

}  // extern "C"
