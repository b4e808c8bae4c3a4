// Workgroup "team" used by the team-generic device routines.
#pragma once

#include <hip/hip_runtime.h>

namespace rtb {

// Sum over the 16 lanes of each DPP row, returned to every lane of the row:
// quad swaps (xor 1, xor 2), then half-row and row mirrors.  Register-to-
// register moves, no LDS round trips (cf. a shuffle through ds_bpermute).
template <int Ctrl>
__device__ inline float dpp_mov(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), Ctrl, 0xF, 0xF, false));
}
template <int Ctrl>
__device__ inline double dpp_mov(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), Ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), Ctrl, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <typename T>
__device__ inline T row_sum16(T x) {
  x += dpp_mov<0xB1>(x);   // quad_perm [1, 0, 3, 2]
  x += dpp_mov<0x4E>(x);   // quad_perm [2, 3, 0, 1]
  x += dpp_mov<0x141>(x);  // row_half_mirror
  x += dpp_mov<0x140>(x);  // row_mirror
  return x;
}
__device__ inline float lane_value(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ inline double lane_value(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
// Sum over a full wave of 64 active lanes, wave-uniform; fixed order.
template <typename T>
__device__ inline T wave_sum64(T x) {
  x = row_sum16(x);
  return (lane_value(x, 0) + lane_value(x, 16)) + (lane_value(x, 32) + lane_value(x, 48));
}

// Sum over the aligned groups of 8 lanes (quad swaps, then the half-row
// mirror), returned to every lane of the group.
template <typename T>
__device__ inline T group_sum8(T x) {
  x += dpp_mov<0xB1>(x);   // quad_perm [1, 0, 3, 2]
  x += dpp_mov<0x4E>(x);   // quad_perm [2, 3, 0, 1]
  x += dpp_mov<0x141>(x);  // row_half_mirror
  return x;
}

struct GpuTeam {
  static constexpr int kRedStride = 32;  // >= 2 x the waves of a workgroup
  int id, nt;
  __device__ GpuTeam() : id(threadIdx.x), nt(blockDim.x) {}
  __device__ void sync() { __syncthreads(); }
  // Orders the LDS accesses of the lanes of one wave (they execute in
  // lockstep; only the compiler's reordering has to be fenced).
  __device__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // Sum over the aligned group of G (8 or 16) consecutive lanes, to each of
  // them; every lane of the wave calls it.
  template <int G, typename T>
  __device__ T group_sum(T x) {
    static_assert(G == 8 || G == 16, "group size");
    if constexpr (G == 16) {
      return row_sum16(x);
    } else {
      return group_sum8(x);
    }
  }

  // Sum over the workgroup, returned to every thread, without a trailing
  // barrier: `red` must not be rewritten until every thread has passed
  // another barrier.  Deterministic: each wave in a fixed order, then the
  // per-wave partials in wave order.  Every lane of the workgroup calls it.
  template <typename T>
  __device__ T reduce_sum_ns(T x, T* red) {
    x = wave_sum64(x);
    const int wave = id >> 6, nw = (nt + 63) >> 6;
    if ((id & 63) == 0) red[wave] = x;
    __syncthreads();
    T s = T(0);
    for (int w = 0; w < nw; ++w) s += red[w];
    return s;
  }
};

// `nt` threads of a workgroup working on one matrix while the other threads
// of the workgroup work on others; sync() is the whole-workgroup barrier, so
// every sub-team must execute the same barrier sequence (idle sub-teams take
// id >= n and skip all per-index work).
struct GpuSubTeam {
  int id, nt;
  __device__ GpuSubTeam(int id_, int nt_) : id(id_), nt(nt_) {}
  __device__ void sync() { __syncthreads(); }
};

}  // namespace rtb
