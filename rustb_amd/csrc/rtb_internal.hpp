// Internal declarations shared by the librtb_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <functional>
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

// ---- call workspaces (workspace.cpp) ----
// A slot is a non-blocking HIP stream (default priority) plus one device
// block that grows to the largest workspace a call on it has needed; calls
// lease a slot for their duration, so that steady-state calls allocate no
// device memory.  At most max_concurrent_calls() calls (RTB_MAX_CONCURRENT_CALLS,
// default 4) hold a slot on one device at once; further callers wait.
struct Slot;
int max_concurrent_calls();
class Lease {
 public:
  Lease();   // takes a slot on the current device (waits while all are busy)
  ~Lease();  // synchronises the slot's stream, then returns the slot
  Lease(const Lease&) = delete;
  Lease& operator=(const Lease&) = delete;
  int status() const { return rc_; }
  hipStream_t stream() const;
  Slot* slot() const { return s_; }
  // the slot's i-th timing event (created on first use, then reused)
  int event(size_t i, hipEvent_t* out);

 private:
  Slot* s_ = nullptr;
  int rc_ = RTB_OK;
};

// A call's device buffers, carved with a bump pointer (256-byte aligned)
// from a leased slot's block (which grows when a call needs more) or from a
// fixed buffer (no growth: a request beyond it is RTB_ERR_UNSUPPORTED).  A
// fixed buffer with a null base only counts the bytes requested (sizing).
class Work {
 public:
  explicit Work(Lease& l);
  Work(void* base, size_t cap);
  Work(const Work&) = delete;
  Work& operator=(const Work&) = delete;
  template <typename T>
  int alloc(T** out, size_t count) {
    void* p = nullptr;
    const int rc = take(count * sizeof(T), &p);
    *out = static_cast<T*>(p);
    return rc;
  }
  size_t used() const { return used_; }

 private:
  int take(size_t bytes, void** out);
  Slot* slot_ = nullptr;
  char* base_ = nullptr;
  size_t cap_ = 0, off_ = 0, used_ = 0;
};

// Event-pair timing of kernel launches on one stream, by stage (see
// rtb_stats), with the leased slot's reusable events; a Timer made without
// a lease times nothing.  collect() runs after the stream is synchronised
// and publishes the sums as the calling thread's rtb_last_stats().
void publish_stats(const rtb_stats& s);
class Timer {
 public:
  explicit Timer(Lease& l) : lease_(&l), st_(l.stream()) {}
  Timer() = default;
  Timer(const Timer&) = delete;
  Timer& operator=(const Timer&) = delete;
  int begin(int stage) {
    if (!lease_) return RTB_OK;
    Rec r{stage, nullptr, nullptr, 0, 0.0, 0.0};
    RTB_TRY(lease_->event(2 * recs_.size(), &r.a));
    RTB_TRY(lease_->event(2 * recs_.size() + 1, &r.b));
    RTB_HIP(hipEventRecord(r.a, st_));
    recs_.push_back(r);
    return RTB_OK;
  }
  int end(int launches, double flop = 0.0, double mflop = 0.0) {
    if (!lease_) return RTB_OK;
    recs_.back().launches = launches;
    recs_.back().flop = flop;
    recs_.back().mflop = mflop;
    RTB_HIP(hipEventRecord(recs_.back().b, st_));
    return RTB_OK;
  }
  // the eigenpair checks of the call: largest refined residual (eV) and the
  // number of k-points re-solved in FP64
  void note_check(double max_residual, size_t resolved) {
    max_res_ = max_residual;
    resolved_ = resolved;
  }
  int collect(size_t nk, int n) {
    rtb_stats s{};
    RTB_TRY(gather(nk, n, &s));
    publish_stats(s);
    return RTB_OK;
  }
  // the sums without publishing them
  int gather(size_t nk, int n, rtb_stats* out) {
    rtb_stats s{};
    s.nk = nk;
    s.nsta = n;
    s.max_residual = max_res_;
    s.resolved = resolved_;
    for (auto& r : recs_) {
      float ms = 0.f;
      RTB_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      s.ms[r.stage] += ms;
      s.launches[r.stage] += r.launches;
      s.flop[r.stage] += r.flop;
      s.mflop[r.stage] += r.mflop;
    }
    if (!recs_.empty()) {
      float ms = 0.f;
      RTB_HIP(hipEventElapsedTime(&ms, recs_.front().a, recs_.back().b));
      s.ms[RTB_STAGE_TOTAL] = ms;
      for (int i = 0; i < RTB_STAGE_TOTAL; ++i) {
        s.launches[RTB_STAGE_TOTAL] += s.launches[i];
        s.flop[RTB_STAGE_TOTAL] += s.flop[i];
        s.mflop[RTB_STAGE_TOTAL] += s.mflop[i];
      }
    }
    *out = s;
    return RTB_OK;
  }

 private:
  struct Rec {
    int stage;
    hipEvent_t a, b;
    int launches;
    double flop, mflop;
  };
  Lease* lease_ = nullptr;
  hipStream_t st_ = nullptr;
  std::vector<Rec> recs_;
  double max_res_ = 0.0;
  size_t resolved_ = 0;
};

// Algorithmic flop counts used in the stats.
inline double flop_zgemm(int n, size_t nb) { return 8.0 * n * n * static_cast<double>(n) * nb; }
inline double flop_zgemm_herm(int n, size_t nb) { return 4.0 * n * n * (n + 1.0) * nb; }
inline double flop_tridiag(int n, size_t nb) { return 16.0 / 3.0 * n * n * static_cast<double>(n) * nb; }
inline double flop_backtrans(int n, size_t nb) { return 8.0 * n * n * static_cast<double>(n) * nb; }
// Matrix-core operations the product kernel executes for nb n x n complex
// products: 64 x 64 output tiles (the lower triangle of them when herm), the
// k loop padded to 16, three real products (3M) of 2 flops per multiply-add.
inline double mflop_zgemm(int n, size_t nb, bool herm) {
  const double side = (n + 63) / 64, kpad = ((n + 15) / 16) * 16.0;
  const double tiles = herm ? side * (side + 1) / 2 : side * side;
  return tiles * 64.0 * 64.0 * kpad * 6.0 * static_cast<double>(nb);
}

// EnergyCut Hall conductivity, host part (energy_cut.cpp).  Band tracking:
// perm [nx ny][n] (band b of vertex i is raw band perm[i][b]) from the raw
// overlaps ovx [(nx-1) ny][n][n] = U_i^H U_{i+ny}, ovy [nx ny - 1][n][n] =
// U_i^H U_{i+1} (complex, or their squared moduli when squared).  The 2D
// Fermi cut of E [nk][n] and Im K [nk][n][n] (Ki, element stride ks) for
// sorted mu, fractional-coordinate measure, thermal width in eV.
// 3D (nz > 1, vertex (ix ny + iy) nz + iz): ovx pairs (i, i + ny nz), ovy
// (i, i + nz), ovz (i, i + 1), each indexed by its lower vertex; the cut is
// over tetrahedra.  A 2D mesh is nz = 1 (ovz unused).
void energy_cut_track(size_t nx, size_t ny, size_t nz, int n, const double* ovx, const double* ovy,
                      const double* ovz, bool squared, std::vector<int>& perm);
// The walk's schedule, which depends on the mesh only: per vertex the
// neighbours already visited when the walk assigns it (neighbour order) and a
// level one above theirs; the vertices grouped by level (walk order inside a
// level).  Vertex 0 alone is level 0.
struct TrackSchedule {
  std::vector<size_t> by_level;
  std::vector<size_t> level_start;  // [levels + 1]
  std::vector<unsigned char> ndep;  // [nk]
  std::vector<size_t> dep;          // [nk][6]
};
void energy_cut_schedule(size_t nx, size_t ny, size_t nz, TrackSchedule& ts);
void energy_cut_sigma(size_t nx, size_t ny, size_t nz, int dim, int n, const double* E, const double* Ki, size_t ks,
                      const int* perm, const double* mu, size_t n_mu, double width, double eta, double* out);
// The thermal convolution of a zero-width cut t0(mu, n_mu, out) (t0 itself
// for width 0).
int energy_cut_convolve(const double* mu, size_t n_mu, double width,
                        const std::function<int(const double*, size_t, double*)>& t0, double* out);

// Largest nsta of the register-resident eigensolver route (one workgroup
// holds the matrix in registers); above it, up to kMaxN, the two-stage route
// (dense -> band -> tridiagonal) with the matrix in global memory.
constexpr int kMaxEighN = 128;
constexpr int kMaxN = 256;

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
  bool herm = true;           // ham[-R] == ham[R]^H bit for bit, for every R
  double* d_rmat = nullptr;   // [nR][dim][n][n] complex position matrices (or null)
  double* d_ham = nullptr;    // [nR][n][n] complex (re, im)
  double* d_R = nullptr;      // [nR][3]   R as doubles (zero padded)
  double* d_Rcart = nullptr;  // [nR][3]   R . lat
  double* d_tau = nullptr;    // [n][3]    fractional tau, spin duplicated
  double* d_taucart = nullptr;  // [n][3]  tau . lat
};
