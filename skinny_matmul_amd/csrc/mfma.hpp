// Matrix-core (MFMA) building blocks shared by the TSMTTSM and TSMM kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace skinny {

// Global-memory view of a pointer into device memory.  Pointers that are
// read from memory (the fields of a work-item record) are generic, and
// accesses through them compile to flat_* instructions, which need 64-bit
// vector addresses and count against both the vector-memory and the LDS
// wait counters; through this view they compile to global_* instructions.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// LDS view of a pointer into a workgroup's shared memory (ds_* instead of
// flat_* instructions when the pointer was passed through a call).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(3))) T* lmem(T* p) {
  return (__attribute__((address_space(3))) T*)p;
}

// Constant-memory view of a pointer to data that no wave writes during the
// kernel: uniform accesses through it compile to scalar loads (s_load, through
// the scalar data cache, into SGPRs) instead of per-lane vector loads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* cmem(const T* p) {
  return (const __attribute__((address_space(4))) T*)p;
}

// Hand-over of workgroup partials to the workgroup that sums them, within one
// launch.  Each XCD has its own L2, and a release at agent scope writes the
// XCD's L2 back (buffer_wbl2), which costs microseconds per workgroup when
// hundreds of them finish together.  Partials are instead written with
// agent-scope stores (global_store ... sc1), which write through the L2, so
// the only thing an arrival must ensure is that its stores have completed:
// published() waits for them (s_waitcnt vmcnt(0), which is also the second
// half of the release sequence) and keeps the compiler from moving memory
// accesses across it.  The summing workgroup still takes the acquire side
// (an L2 invalidate) before it reads.
template <typename T>
__device__ __forceinline__ void st_agent(T* p, T v) {
  __hip_atomic_store(gmem(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void published() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Streaming load with the nontemporal hint (global_load ... nt): inputs that
// are read once do not displace cached data.  On MI355X a read stream with
// the hint runs at about 7.1 TB/s against 6.4 TB/s without it.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(gmem(p));
}

// 16 x 16 x 4 MFMA, fp64 and fp32.  Operands: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16].  Result element v of lane l is
// D[drow(l / 16, v)][l % 16]; the fp64 form interleaves the rows (4 v + l/16),
// the fp32 form blocks them (4 (l/16) + v).  Both mappings were measured on
// gfx950 by the rectangular parity tests.
template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int drow(int g, int v) { return 4 * v + g; }
};

template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int drow(int g, int v) { return 4 * g + v; }
};

}  // namespace skinny
