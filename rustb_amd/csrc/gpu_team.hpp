// Workgroup "team" used by the team-generic device routines.
#pragma once

#include <hip/hip_runtime.h>

namespace rtb {

struct GpuTeam {
  int id, nt;
  __device__ GpuTeam() : id(threadIdx.x), nt(blockDim.x) {}
  __device__ void sync() { __syncthreads(); }

  // Sum over the workgroup, returned to every thread.  Deterministic: a fixed
  // butterfly inside each wave, then the per-wave partials in wave order.
  template <typename T>
  __device__ T reduce_sum(T x, T* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    const int wave = id >> 6, nw = (nt + 63) >> 6;
    if ((id & 63) == 0) red[wave] = x;
    __syncthreads();
    T s = T(0);
    for (int w = 0; w < nw; ++w) s += red[w];
    __syncthreads();
    return s;
  }
};

}  // namespace rtb
