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

#include "grid_barrier.hpp"
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

// Copy variants (kind 5 of skinny_mb_stream): with chunk == 0 a grid-stride
// copy with U accesses of 16 bytes in flight per thread and the
// nontemporal hint on the loads (NTL) and/or the stores (NTS); with
// chunk > 0 each wave copies `chunk` consecutive elements per turn, U loads
// per lane at a time (longer read and write bursts per wave).
template <int U, bool NTL, bool NTS>
__global__ void __launch_bounds__(256)
    copy_variant_kernel(d2* __restrict__ a, const d2* __restrict__ b, int64_t n2, int64_t chunk) {
  auto ld = [&](int64_t j) { return NTL ? __builtin_nontemporal_load(&b[j]) : b[j]; };
  auto st = [&](int64_t j, d2 v) {
    if (NTS)
      __builtin_nontemporal_store(v, &a[j]);
    else
      a[j] = v;
  };
  if (chunk == 0) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += U * stride) {
      d2 v[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t j = i + u * stride;
        v[u] = j < n2 ? ld(j) : d2{0, 0};
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t j = i + u * stride;
        if (j < n2) st(j, v[u]);
      }
    }
    return;
  }
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  for (int64_t c0 = wave * chunk; c0 < n2; c0 += nwaves * chunk) {
    const int64_t cend = c0 + chunk < n2 ? c0 + chunk : n2;
    for (int64_t i = c0 + lane; i < cend; i += 64 * U) {
      d2 v[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t j = i + 64 * u;
        v[u] = j < cend ? ld(j) : d2{0, 0};
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t j = i + 64 * u;
        if (j < cend) st(j, v[u]);
      }
    }
  }
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

// With clk set, the first lane of workgroup 0 records the shader-clock
// counter (s_memtime) and the constant-rate real-time counter (s_memrealtime)
// at its start (idx 0) and end (idx 1): clk[idx], clk[2 + idx].
__device__ __forceinline__ void clock_stamp(unsigned long long* clk, int idx) {
  if (clk && blockIdx.x == 0 && threadIdx.x == 0) {
    clk[idx] = __builtin_amdgcn_s_memtime();
    clk[2 + idx] = __builtin_amdgcn_s_memrealtime();
  }
}

// FP64 FMA chains: 8 independent chains per lane
__global__ void __launch_bounds__(256) fma_kernel(double* out, int iters,
                                                  double x, unsigned long long* clk) {
  clock_stamp(clk, 0);
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
  clock_stamp(clk, 1);
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// FP64 matrix core: 4 independent accumulator chains per wave
__global__ void __launch_bounds__(256) mfma_f64_kernel(double* out, int iters,
                                                       double x, unsigned long long* clk) {
  clock_stamp(clk, 0);
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
  clock_stamp(clk, 1);
}

// INT8 matrix core (32x32x32, int32 accumulate): 2 chains per wave
__global__ void __launch_bounds__(256) mfma_i8_kernel(int* out, int iters,
                                                      int x, unsigned long long* clk) {
  clock_stamp(clk, 0);
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
  clock_stamp(clk, 1);
}

// FP64 matrix core, four-block 4x4x4 form: 16 independent chains per wave
__global__ void __launch_bounds__(256) mfma_f64_q4_kernel(double* out, int iters,
                                                          double x, unsigned long long* clk) {
  clock_stamp(clk, 0);
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; j++) acc[j] = 0.0;
  double a = x + threadIdx.x, b = x - threadIdx.x;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int j = 0; j < 16; j++)
      acc[j] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0;
#pragma unroll
  for (int j = 0; j < 16; j++) s += acc[j];
  if (s == 12345.678) out[threadIdx.x] = s;
  clock_stamp(clk, 1);
}

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// FP32 matrix core, three instruction shapes with independent chains per wave:
// F = 0: 16x16x4 (4 chains), F = 1: 32x32x2 (2 chains), F = 2: 4x4x1 with 16
// blocks (4 chains)
template <int F>
__global__ void __launch_bounds__(256) mfma_f32_kernel(float* out, int iters, float x,
                                                       unsigned long long* clk) {
  clock_stamp(clk, 0);
  const float a = x + threadIdx.x, b = x - threadIdx.x;
  float s = 0.0f;
  if constexpr (F == 1) {
    v16f acc[2];
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[j][r] = 0.0f;
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int j = 0; j < 2; j++)
        acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[j], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) s += acc[j][r];
  } else {
    v4f acc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) acc[j] = v4f{0, 0, 0, 0};
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[j] = F == 0 ? __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0)
                        : __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, acc[j], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  }
  if (s == 12345.678f) out[threadIdx.x] = s;
  clock_stamp(clk, 1);
}

// The idle clock: one wave reads both counters about `spin` real-time ticks
// apart (the loop ends once the real-time counter has advanced that far).
__global__ void clock_kernel(unsigned long long* clk, int spin) {
  if (threadIdx.x != 0) return;
  const unsigned long long c0 = __builtin_amdgcn_s_memtime();
  const unsigned long long r0 = __builtin_amdgcn_s_memrealtime();
  unsigned long long r1 = r0;
  while (r1 - r0 < (unsigned long long)spin) r1 = __builtin_amdgcn_s_memrealtime();
  clk[0] = c0;
  clk[1] = __builtin_amdgcn_s_memtime();
  clk[2] = r0;
  clk[3] = r1;
}

// ---- memory hierarchy: latency and bandwidth by working-set size ---------
// (the role of microbenches/instruction_rates.cu:30-148: LDS throughput and
// the L1 / L2 latency sweep, measured here for gfx950's L1, L2, Infinity
// Cache and HBM)

// Pointer chain over the 128-byte lines of a buffer: line l holds the word
// offset of line (l + P) mod n_lines, P odd and near 0.618 n_lines, so the
// chain visits every line (n_lines is a power of two) in a scattered order.
__global__ void __launch_bounds__(256)
    chain_init_kernel(unsigned* a, int64_t n_lines, int64_t P) {
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < n_lines;
       l += (int64_t)gridDim.x * 256) {
    const int64_t nxt = (l + P) & (n_lines - 1);
    a[l * 32] = (unsigned)(nxt * 32);
  }
}

// One lane follows the chain: every load's address is the previous load's
// value.  Word offsets stay below n_lines * 32 by construction.
__global__ void chase_kernel(const unsigned* a, unsigned start, int64_t steps,
                             unsigned* sink) {
  if (threadIdx.x != 0) return;
  unsigned idx = start;
  for (int64_t i = 0; i < steps; i++) idx = a[idx];
  sink[0] = idx;
}

// `passes` reads of the whole buffer by the whole grid (grid-stride, 4
// accesses of 16 bytes in flight per lane); a workgroup reads the same
// addresses on every pass, so a buffer that fits in a cache level is served
// from it after the first pass.
__global__ void __launch_bounds__(256)
    reread_kernel(const d2* __restrict__ b, int64_t n2, int passes,
                  double* __restrict__ sink) {
  d2 acc = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int p = 0; p < passes; p++) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += 4 * stride) {
      d2 v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t j = i + u * stride;
        v[u] = j < n2 ? b[j] : d2{0.0, 0.0};
      }
#pragma unroll
      for (int u = 0; u < 4; u++) acc += v[u];
    }
  }
  if (acc.x + acc.y == 12345.678) sink[threadIdx.x] = acc.x;
}

// LDS read bandwidth: 16-byte reads, lane l of wave w reading entry l (or
// 64 + l) of its wave's row, so the 64 lanes of a read are conflict-free.
// Eight independent chains per lane: each read's index adds a value from
// the previous read of the same chain (always 0), which keeps the compiler
// from hoisting the loop-invariant reads.
template <int ITERS>
__global__ void __launch_bounds__(256) lds_bw_kernel(unsigned* sink) {
  __shared__ uint4 buf[4][128];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  buf[w][l] = uint4{(unsigned)l, 1u, 2u, 3u};
  buf[w][64 + l] = uint4{(unsigned)l, 5u, 6u, 7u};
  __syncthreads();
  uint4 acc = {0, 0, 0, 0};
  unsigned off[8];
#pragma unroll
  for (int j = 0; j < 8; j++) off[j] = 0;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const uint4 v = buf[w][(j & 1) * 64 + l + off[j]];
      off[j] = v.y >> 8;  // 0: entries hold 1 or 5 here
      acc.x += v.x;
      acc.z += v.z;
      acc.w ^= v.w;
    }
  }
  if ((acc.x + acc.z + acc.w) == 0xdeadbeefu) sink[threadIdx.x] = acc.x;
}

// LDS latency: one lane follows a chain of dependent LDS reads.
__global__ void __launch_bounds__(64) lds_chase_kernel(int steps,
                                                       unsigned* sink) {
  __shared__ unsigned buf[1024];
  for (int i = threadIdx.x; i < 1024; i += 64) buf[i] = (unsigned)((i + 37) & 1023);
  __syncthreads();
  if (threadIdx.x != 0) return;
  unsigned idx = 0;
  for (int i = 0; i < steps; i++) idx = buf[idx & 1023];
  sink[0] = idx;
}

// Grid barriers back to back (the cost global_sync/sync.cu measures with its
// atomic arrival counter): n barriers, then the exit that resets the counters.
__global__ void __launch_bounds__(256)
    barrier_kernel(unsigned* bar, int n, unsigned spin_limit) {
  for (int i = 1; i <= n; i++) grid_barrier(bar, (unsigned)i * gridDim.x, spin_limit);
  grid_exit(bar);
}

}  // namespace
}  // namespace skinny

using namespace skinny;

extern "C" {

skinny_status_t skinny_mb_stream(int kind, size_t bytes_per_array, int reps,
                                 double* gbs, skinny_stream_t stream) {
  if (!gbs || kind < 0 || kind > 5 || bytes_per_array < 4096 || reps < 1)
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
  if (kind == 5) {
    // copy variants: nontemporal stores and/or loads (grid-stride, 4 in
    // flight), and per-wave bursts of 16 KiB, 64 KiB or 1 MiB (8 in flight)
    d2* A = (d2*)a;
    const d2* B = (const d2*)b;
    for (int per : {1, 2, 4, 8}) {
      const int blocks = per * h->cu_count;
      for (int v = 0; v < 4; v++) {
        auto launch = [&]() {
          if (v == 0)
            hipLaunchKernelGGL((copy_variant_kernel<4, false, true>), dim3(blocks), dim3(256), 0,
                               st, A, B, n2, (int64_t)0);
          else if (v == 1)
            hipLaunchKernelGGL((copy_variant_kernel<4, true, true>), dim3(blocks), dim3(256), 0,
                               st, A, B, n2, (int64_t)0);
          else if (v == 2)
            hipLaunchKernelGGL((copy_variant_kernel<4, true, false>), dim3(blocks), dim3(256), 0,
                               st, A, B, n2, (int64_t)0);
          else
            hipLaunchKernelGGL((copy_variant_kernel<4, false, false>), dim3(blocks), dim3(256), 0,
                               st, A, B, n2, (int64_t)0);
        };
        ms = std::min(ms, time_ms(launch, st, reps));
      }
      for (int64_t chunk : {(int64_t)1024, (int64_t)4096, (int64_t)65536}) {
        for (int ntl = 0; ntl < 2; ntl++) {
          auto launch = [&]() {
            if (ntl)
              hipLaunchKernelGGL((copy_variant_kernel<8, true, false>), dim3(blocks), dim3(256),
                                 0, st, A, B, n2, chunk);
            else
              hipLaunchKernelGGL((copy_variant_kernel<8, false, false>), dim3(blocks), dim3(256),
                                 0, st, A, B, n2, chunk);
          };
          ms = std::min(ms, time_ms(launch, st, reps));
        }
      }
    }
  }
  for (int per : {1, 2, 4, 8}) {
    if (kind == 5) break;
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
  const double arrays[6] = {1, 2, 3, 1, 3, 2};  // bytes moved per array size
  *gbs = arrays[kind] * (double)nb / (ms * 1e-3) / 1e9;
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(c);
  (void)hipFree(sink);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_stream launch");
  return SKINNY_OK;
}

// kind 0: latency (ns per dependent load) of a chain over `bytes` bytes;
// 1: aggregate read bandwidth (GB/s) of the whole GPU re-reading a buffer of
// `bytes` bytes; 2: LDS read bandwidth (GB/s, all CUs; bytes ignored);
// 3: LDS latency (ns per dependent read; bytes ignored).
skinny_status_t skinny_mb_cache(int kind, size_t bytes, int reps,
                                double* out, skinny_stream_t stream) {
  if (!out || kind < 0 || kind > 3 || reps < 1 ||
      ((kind == 0 || kind == 1) && (bytes < 4096 || bytes > ((size_t)16 << 30))))
    return fail(SKINNY_INVALID_ARG, "skinny_mb_cache: bad arguments");
  skinny_handle_t h = skinny_default_handle();
  if (!h) return no_handle("microbench");
  hipStream_t st = (hipStream_t)stream;
  void* sink = nullptr;
  hipError_t e = hipMalloc(&sink, 4096);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_cache alloc");
  void* buf = nullptr;
  if (kind <= 1) {
    // power-of-two number of 128-byte lines for the chain
    size_t nb = 128;
    while (nb * 2 <= bytes) nb *= 2;
    if (kind == 1) nb = bytes / 16 * 16;
    e = hipMalloc(&buf, nb);
    if (e != hipSuccess) {
      (void)hipFree(sink);
      return hip_fail(e, "skinny_mb_cache alloc");
    }
    bytes = nb;
  }
  float ms = 0.0f;
  double count = 0.0;
  if (kind == 0) {
    const int64_t n_lines = (int64_t)(bytes / 128);
    int64_t P = (int64_t)(0.6180339887 * (double)n_lines) | 1;
    if (P >= n_lines) P = n_lines > 1 ? n_lines - 1 : 1;
    if (n_lines == 1) P = 0;
    const int gb = (int)std::min<int64_t>((n_lines + 255) / 256, 1024);
    hipLaunchKernelGGL(chain_init_kernel, dim3(gb), dim3(256), 0, st, (unsigned*)buf,
                       n_lines, P);
    // Buffers that fit the 256 MiB Infinity Cache are first read by the
    // whole GPU and once along the chain (up to 64k lines), so the timed
    // chase finds them in the smallest level that holds them.  Larger ones
    // are not warmed, and each timed launch starts where the previous one
    // stopped, so its lines come from HBM.
    const int64_t steps = 1 << 16;
    if (bytes <= ((size_t)256 << 20)) {
      hipLaunchKernelGGL(reread_kernel, dim3(h->cu_count), dim3(256), 0, st,
                         (const d2*)buf, (int64_t)(bytes / 16), 1, (double*)sink);
      hipLaunchKernelGGL(chase_kernel, dim3(1), dim3(64), 0, st, (const unsigned*)buf, 0u,
                         std::min<int64_t>(n_lines, steps), (unsigned*)sink);
    }
    int64_t launch = 0;
    ms = time_ms([&]() {
      // first line of segment `launch` of the chain: line (launch steps P) mod n
      const int64_t line = (int64_t)(((unsigned __int128)(launch * steps) * (unsigned __int128)P) %
                                     (unsigned __int128)n_lines);
      const unsigned start = bytes <= ((size_t)256 << 20) ? 0u : (unsigned)(line * 32);
      launch++;
      hipLaunchKernelGGL(chase_kernel, dim3(1), dim3(64), 0, st, (const unsigned*)buf, start,
                         steps, (unsigned*)sink);
    }, st, reps);
    count = (double)steps;
    *out = (double)ms * 1e6 / count;
  } else if (kind == 1) {
    const int64_t n2 = (int64_t)(bytes / 16);
    // about 16 GiB of reads per launch (at least one pass)
    const int blocks = 4 * h->cu_count;
    int passes = (int)std::max<int64_t>(1, ((int64_t)16 << 30) / (int64_t)bytes);
    if (passes > 1 << 20) passes = 1 << 20;
    ms = time_ms([&]() {
      hipLaunchKernelGGL(reread_kernel, dim3(blocks), dim3(256), 0, st, (const d2*)buf, n2,
                         passes, (double*)sink);
    }, st, reps);
    *out = (double)bytes * passes / (ms * 1e-3) / 1e9;
  } else if (kind == 2) {
    constexpr int IT = 4096;
    const int blocks = 8 * h->cu_count;
    ms = time_ms([&]() {
      hipLaunchKernelGGL(lds_bw_kernel<IT>, dim3(blocks), dim3(256), 0, st, (unsigned*)sink);
    }, st, reps);
    *out = (double)blocks * 256 * IT * 8 * 16 / (ms * 1e-3) / 1e9;
  } else {
    const int steps = 1 << 16;
    ms = time_ms([&]() {
      hipLaunchKernelGGL(lds_chase_kernel, dim3(1), dim3(64), 0, st, steps, (unsigned*)sink);
    }, st, reps);
    *out = (double)ms * 1e6 / steps;
  }
  e = hipGetLastError();
  (void)hipFree(buf);
  (void)hipFree(sink);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_cache launch");
  return SKINNY_OK;
}

// Microseconds per grid barrier for a cooperative grid of per_cu workgroups
// of 256 threads per CU (1 <= per_cu <= the occupancy limit).
skinny_status_t skinny_mb_barrier(int per_cu, int reps, double* us,
                                  skinny_stream_t stream) {
  if (!us || per_cu < 1 || reps < 1)
    return fail(SKINNY_INVALID_ARG, "skinny_mb_barrier: bad arguments");
  skinny_handle_t h = skinny_default_handle();
  if (!h) return no_handle("microbench");
  int maxper = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&maxper, barrier_kernel, 256, 0);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_barrier occupancy");
  if (per_cu > maxper)
    return fail(SKINNY_UNSUPPORTED, "skinny_mb_barrier: more workgroups per CU than fit");
  hipStream_t st = (hipStream_t)stream;
  unsigned* bar = nullptr;
  e = hipMalloc(&bar, 4 * sizeof(unsigned));
  if (e == hipSuccess) e = hipMemsetAsync(bar, 0, 4 * sizeof(unsigned), st);
  if (e != hipSuccess) {
    (void)hipFree(bar);
    return hip_fail(e, "skinny_mb_barrier alloc");
  }
  const int grid = per_cu * h->cu_count;
  int n = 1000;
  unsigned spin_limit = 1u << 25;
  void* args[] = {&bar, &n, &spin_limit};
  hipError_t le = hipSuccess;
  const float ms = time_ms([&]() {
    const hipError_t x = hipLaunchCooperativeKernel((const void*)barrier_kernel, dim3(grid),
                                                    dim3(256), args, 0, st);
    if (x != hipSuccess) le = x;
  }, st, reps);
  unsigned flag = 0;
  e = le;
  if (e == hipSuccess) e = hipMemcpy(&flag, bar + 2, sizeof(unsigned), hipMemcpyDeviceToHost);
  (void)hipFree(bar);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_barrier launch");
  if (flag) return fail(SKINNY_HIP_ERROR, "skinny_mb_barrier: a barrier timed out");
  *us = (double)ms * 1e3 / n;
  return SKINNY_OK;
}

// The peak loop of `kind` (as skinny_mb_peak) on `blocks` workgroups, with
// the clock stamps written to clk when it is not null.
static void launch_peak(int kind, int blocks, int iters, void* out, unsigned long long* clk,
                        hipStream_t st) {
  switch (kind) {
    case 0: hipLaunchKernelGGL(fma_kernel, dim3(blocks), dim3(256), 0, st, (double*)out,
                               iters, 1.0, clk); break;
    case 1: hipLaunchKernelGGL(mfma_f64_kernel, dim3(blocks), dim3(256), 0, st, (double*)out,
                               iters, 1.0, clk); break;
    case 2: hipLaunchKernelGGL(mfma_i8_kernel, dim3(blocks), dim3(256), 0, st, (int*)out,
                               iters, 1, clk); break;
    case 3: hipLaunchKernelGGL(mfma_f64_q4_kernel, dim3(blocks), dim3(256), 0, st,
                               (double*)out, iters, 1.0, clk); break;
    case 4: hipLaunchKernelGGL(mfma_f32_kernel<0>, dim3(blocks), dim3(256), 0, st, (float*)out,
                               iters, 1.0f, clk); break;
    case 5: hipLaunchKernelGGL(mfma_f32_kernel<1>, dim3(blocks), dim3(256), 0, st, (float*)out,
                               iters, 1.0f, clk); break;
    default: hipLaunchKernelGGL(mfma_f32_kernel<2>, dim3(blocks), dim3(256), 0, st,
                                (float*)out, iters, 1.0f, clk); break;
  }
}

// kind 0: FP64 vector FMA, 1: FP64 MFMA 16x16x4, 2: INT8 MFMA 32x32x32,
// 3: FP64 MFMA 4x4x4 (four blocks), 4: FP32 MFMA 16x16x4, 5: FP32 MFMA
// 32x32x2, 6: FP32 MFMA 4x4x1 (16 blocks).  Result in TFLOP/s (TOP/s for
// INT8).
skinny_status_t skinny_mb_peak(int kind, int reps, double* tflops,
                               skinny_stream_t stream) {
  if (!tflops || kind < 0 || kind > 6 || reps < 1)
    return fail(SKINNY_INVALID_ARG, "skinny_mb_peak: bad arguments");
  skinny_handle_t h = skinny_default_handle();
  if (!h) return no_handle("microbench");
  hipStream_t st = (hipStream_t)stream;
  void* out = nullptr;
  hipError_t e = hipMalloc(&out, 4096);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_peak alloc");
  const int blocks = 8 * h->cu_count;
  const int iters = kind == 0 ? 4096 : 1024;
  auto launch = [&]() { launch_peak(kind, blocks, iters, out, nullptr, st); };
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
  else if (kind == 2)
    ops = threads / 64.0 * iters * 2 * (32.0 * 32 * 32 * 2);
  else if (kind == 3)
    ops = threads / 64.0 * iters * 16 * (4.0 * 4 * 4 * 4 * 2);
  else if (kind == 4)
    ops = threads / 64.0 * iters * 4 * (16.0 * 16 * 4 * 2);
  else if (kind == 5)
    ops = threads / 64.0 * iters * 2 * (32.0 * 32 * 2 * 2);
  else
    ops = threads / 64.0 * iters * 4 * (16.0 * 4 * 4 * 1 * 2);
  *tflops = ops / (ms * 1e-3) / 1e12;
  return SKINNY_OK;
}

// Shader clock (MHz) while the peak loop of `kind` (as skinny_mb_peak) runs,
// from the counters workgroup 0 of that loop reads at its start and end;
// kind -1: a one-wave spin of about 1 ms with nothing else running.
skinny_status_t skinny_mb_clock(int kind, double* mhz, skinny_stream_t stream) {
  if (!mhz || kind < -1 || kind > 6)
    return fail(SKINNY_INVALID_ARG, "skinny_mb_clock: bad arguments");
  skinny_handle_t h = skinny_default_handle();
  if (!h) return no_handle("microbench");
  int rate_khz = 0;
  hipError_t e = hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, h->device);
  if (e != hipSuccess) return hip_fail(e, "wall clock rate");
  if (rate_khz <= 0) return fail(SKINNY_HIP_ERROR, "skinny_mb_clock: no wall clock rate");
  hipStream_t st = (hipStream_t)stream;
  void* out = nullptr;
  unsigned long long* clk = nullptr;
  e = hipMalloc(&out, 4096);
  if (e == hipSuccess) e = hipMalloc(&clk, 32);
  if (e != hipSuccess) {
    (void)hipFree(out);
    return hip_fail(e, "skinny_mb_clock alloc");
  }
  const int blocks = 8 * h->cu_count;
  if (kind >= 0)
    launch_peak(kind, blocks, kind == 0 ? 16384 : 4096, out, clk, st);
  else
    hipLaunchKernelGGL(clock_kernel, dim3(1), dim3(64), 0, st, clk, rate_khz);  // 1 ms
  e = hipGetLastError();
  unsigned long long hc[4] = {0, 0, 0, 0};
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipMemcpy(hc, clk, 32, hipMemcpyDeviceToHost);
  (void)hipFree(out);
  (void)hipFree(clk);
  if (e != hipSuccess) return hip_fail(e, "skinny_mb_clock");
  const double ticks = (double)(hc[3] - hc[2]);
  if (ticks <= 0) return fail(SKINNY_HIP_ERROR, "skinny_mb_clock: no ticks");
  *mhz = (double)(hc[1] - hc[0]) / (ticks / (rate_khz * 1e3)) / 1e6;
  return SKINNY_OK;
}

}  // extern "C"
