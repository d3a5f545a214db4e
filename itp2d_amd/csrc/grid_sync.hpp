// Device-side synchronisation shared by the cooperative kernels (block
// Jacobi, tridiagonalization). HIP device code only.
#pragma once
#include <hip/hip_runtime.h>

namespace itp {

// Grid-wide barrier of a cooperative launch: an arrival counter (zeroed
// before the launch) with agent-scope release/acquire fences. Barrier number
// k of every workgroup waits for k * gridDim.x arrivals. A workgroup that is
// not the last to arrive polls the counter at most `limit` times; past that
// it raises the error flag instead of waiting on. Returns true when the flag
// is raised (by this or any other workgroup): the caller then leaves the
// kernel, and every workgroup still waiting on a later barrier gives up after
// its own bounded wait, so the launch always drains.
__device__ __forceinline__ bool grid_barrier(unsigned* cnt, unsigned& epoch, unsigned* err, unsigned limit) {
  __shared__ int failed;
  __syncthreads();
  const unsigned target = ++epoch * gridDim.x;
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned arrived = atomicAdd(cnt, 1u) + 1u;
    if (arrived < target) {
      for (unsigned spins = 0;; spins++) {
        if (spins >= limit) {
          atomicOr(err, 1u);
          break;
        }
        if (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();
    failed = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
  }
  __syncthreads();
  return failed != 0;
}

}  // namespace itp
