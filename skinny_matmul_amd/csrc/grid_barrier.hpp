// Grid-wide barrier for cooperative launches (all workgroups resident).
//
// Thread 0 of each workgroup releases the workgroup's writes at agent scope,
// adds 1 to an arrival counter that only grows within a launch, and polls it
// until it reaches `target` (barrier number x grid size), then acquires.
// The role of the atomic arrival counter of global_sync/sync.cu:38-53.  A
// poll limit bounds the wait: past it the barrier sets bar[2] and lets the
// workgroup go on, so a launch whose workgroups are not all resident fails
// the call instead of hanging the device.  bar[0] (arrivals) and bar[1]
// (exits) are reset by the last workgroup to leave (grid_exit).
#pragma once

#include <hip/hip_runtime.h>

namespace skinny {

__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned target,
                                             unsigned spin_limit) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > spin_limit) {
        __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Leave the grid: the last workgroup out resets the arrival and exit counters
// for the next launch.
__device__ __forceinline__ void grid_exit(unsigned* bar) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned t = __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_AGENT);
    if (t == gridDim.x - 1) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&bar[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace skinny
