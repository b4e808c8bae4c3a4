// Internal declarations shared by the librtb_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/rtb_hip.h"

namespace rtb {

// Thread-local error text behind rtb_last_error().
void set_error(const char* fmt, ...);
inline int fail(int code, const char* msg) {
  set_error("%s", msg);
  return code;
}

#define RTB_HIP(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      ::rtb::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,   \
                       __LINE__);                                                           \
      return RTB_ERR_DEVICE;                                                                \
    }                                                                                       \
  } while (0)

#define RTB_TRY(expr)        \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != RTB_OK) {     \
      return rc_;            \
    }                        \
  } while (0)

// Largest nsta the batched eigensolver handles (packed lower triangle of the
// FP64 matrix in LDS: n(n+1)/2 * 16 B <= 160 KB).
constexpr int kMaxEighN = 128;

}  // namespace rtb

// Device-resident model (a1 storage, model.rs:150-158), plus the small
// per-R / per-orbital tables the kernels read.
struct rtb_model {
  int dim = 0, spin = 0, norb = 0, nsta = 0, nR = 0;
  int device = 0;
  std::vector<int64_t> hamR;  // host copy [nR*dim]
  std::vector<double> orb;    // host copy [norb*dim]
  std::vector<double> lat;    // host copy [dim*dim]
  double det_lat = 0.0;
  double* d_ham = nullptr;    // [nR][n][n] complex (re, im)
  double* d_R = nullptr;      // [nR][3]   R as doubles (zero padded)
  double* d_Rcart = nullptr;  // [nR][3]   R . lat
  double* d_tau = nullptr;    // [n][3]    fractional tau, spin duplicated
  double* d_taucart = nullptr;  // [n][3]  tau . lat
};
