// Workgroup "team" used by the team-generic device routines.
#pragma once

#include <hip/hip_runtime.h>

namespace rtb {

struct GpuTeam {
  static constexpr bool kPairs = true;   // lanes (2i, 2i+1) share rows
  static constexpr int kRedStride = 32;  // >= 2 x the waves of a workgroup
  int id, nt;
  __device__ GpuTeam() : id(threadIdx.x), nt(blockDim.x) {}
  __device__ void sync() { __syncthreads(); }

  // x + partner's x for the lane pair (2i, 2i+1); both lanes must be active.
  template <typename T>
  __device__ T pair_sum(T x, T*) {
    return x + __shfl_xor(x, 1, 64);
  }

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

  // The same sum without a trailing barrier: `red` must not be rewritten
  // until every thread has passed another barrier.
  template <typename T>
  __device__ T reduce_sum_ns(T x, T* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    const int wave = id >> 6, nw = (nt + 63) >> 6;
    if ((id & 63) == 0) red[wave] = x;
    __syncthreads();
    T s = T(0);
    for (int w = 0; w < nw; ++w) s += red[w];
    return s;
  }

  // Two sums at once (no trailing barrier), red holds interleaved pairs.
  template <typename T>
  __device__ void reduce_sum2_ns(T& x, T& y, T* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      x += __shfl_xor(x, off, 64);
      y += __shfl_xor(y, off, 64);
    }
    const int wave = id >> 6, nw = (nt + 63) >> 6;
    if ((id & 63) == 0) {
      red[2 * wave] = x;
      red[2 * wave + 1] = y;
    }
    __syncthreads();
    T sx = T(0), sy = T(0);
    for (int w = 0; w < nw; ++w) {
      sx += red[2 * w];
      sy += red[2 * w + 1];
    }
    x = sx;
    y = sy;
  }
};

// `nt` threads of a workgroup working on one matrix while the other threads
// of the workgroup work on others; sync() is the whole-workgroup barrier, so
// every sub-team must execute the same barrier sequence (idle sub-teams take
// id >= n and skip all per-index work).
struct GpuSubTeam {
  static constexpr bool kPairs = false;
  int id, nt;
  __device__ GpuSubTeam(int id_, int nt_) : id(id_), nt(nt_) {}
  __device__ void sync() { __syncthreads(); }
};

}  // namespace rtb
