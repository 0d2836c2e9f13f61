// Deterministic workgroup-wide sums of many partial results.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace skinny {

// Sum `count` partials of PS entries (src[q * PS + e]) with the whole
// workgroup: entry e is summed by a team of TPE threads (thread j of the team
// takes partials j, j + TPE, ...; the team is then combined by a fixed LDS
// tree), so a small output pays one round trip instead of `count`.  The order
// depends only on (PS, count): results are reproducible.  The sums go to
// dst (group level; agent-scope stores, read by another workgroup of the
// launch) or to `fin` in LDS (final level, PS <= ldsn / 2).  The
// LDS area holds ldsn elements: the first blockDim.x are the tree.
template <typename T>
__device__ inline void team_sum(const T* src, int count, int PS, T* dst, T* lds,
                         int ldsn, bool to_lds) {
  const int nthr = blockDim.x;  // a power of two, <= ldsn / 2
  int pad = 1;
  while (pad < PS) pad <<= 1;
  int tpe = nthr / pad;
  if (tpe < 1) tpe = 1;
  int cpow = 1;
  while (cpow < count) cpow <<= 1;
  if (tpe > cpow) tpe = cpow;
  const int per_pass = nthr / tpe;
  T* tree = lds;
  T* fin = lds + ldsn / 2;
  const int tid = threadIdx.x;
  const int j = tid % tpe;
  for (int e0 = 0; e0 < PS; e0 += per_pass) {
    const int e = e0 + tid / tpe;
    T s = T(0);
    if (e < PS) {
      const T* p = src + e;
      int q = j;
      for (; q + 3 * tpe < count; q += 4 * tpe) {
        const T v0 = p[(int64_t)q * PS];
        const T v1 = p[(int64_t)(q + tpe) * PS];
        const T v2 = p[(int64_t)(q + 2 * tpe) * PS];
        const T v3 = p[(int64_t)(q + 3 * tpe) * PS];
        s += v0;
        s += v1;
        s += v2;
        s += v3;
      }
      for (; q < count; q += tpe) s += p[(int64_t)q * PS];
    }
    tree[tid] = s;
    __syncthreads();
    for (int w = tpe >> 1; w > 0; w >>= 1) {
      if (j < w) tree[tid] += tree[tid + w];
      __syncthreads();
    }
    if (j == 0 && e < PS) {
      if (to_lds)
        fin[e] = tree[tid];
      else
        __hip_atomic_store((__attribute__((address_space(1))) T*)(dst + e), tree[tid],
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
}

// Same summation as team_sum, handing each finished entry to f(e, sum)
// (called by one thread per entry).  `tree` holds blockDim.x elements.
template <typename T, typename F>
__device__ inline void team_reduce(const T* src, int count, int PS, T* tree,
                                   F f) {
  const int nthr = blockDim.x;
  int pad = 1;
  while (pad < PS) pad <<= 1;
  int tpe = nthr / pad;
  if (tpe < 1) tpe = 1;
  int cpow = 1;
  while (cpow < count) cpow <<= 1;
  if (tpe > cpow) tpe = cpow;
  const int per_pass = nthr / tpe;
  const int tid = threadIdx.x;
  const int j = tid % tpe;
  for (int e0 = 0; e0 < PS; e0 += per_pass) {
    const int e = e0 + tid / tpe;
    T s = T(0);
    if (e < PS) {
      const T* p = src + e;
      int q = j;
      for (; q + 3 * tpe < count; q += 4 * tpe) {
        const T v0 = p[(int64_t)q * PS];
        const T v1 = p[(int64_t)(q + tpe) * PS];
        const T v2 = p[(int64_t)(q + 2 * tpe) * PS];
        const T v3 = p[(int64_t)(q + 3 * tpe) * PS];
        s += v0;
        s += v1;
        s += v2;
        s += v3;
      }
      for (; q < count; q += tpe) s += p[(int64_t)q * PS];
    }
    if (tpe > 1) {
      tree[tid] = s;
      __syncthreads();
      for (int w = tpe >> 1; w > 0; w >>= 1) {
        if (j < w) tree[tid] += tree[tid + w];
        __syncthreads();
      }
      s = tree[tid];
      __syncthreads();
    }
    if (j == 0 && e < PS) f(e, s);
  }
}

}  // namespace skinny
