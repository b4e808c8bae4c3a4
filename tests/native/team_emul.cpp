// CPU emulation of the eigensolver workgroups: the team-generic device
// source (rustb_amd/csrc/eigh_core.hpp, eigh_reg.hpp) run by host threads
// with a barrier, in the slice shapes of the GPU dispatch.  Test-only: lets
// the indexing and numerics be checked on the host (with AddressSanitizer
// and ThreadSanitizer).
#include <barrier>
#include <thread>
#include <vector>

#include "../../rustb_amd/csrc/eigh_band.hpp"
#include "../../rustb_amd/csrc/eigh_core.hpp"
#include "../../rustb_amd/csrc/eigh_reg.hpp"

namespace {

struct CpuTeam {
  static constexpr int kRedStride = 2048;  // >= 2 x the largest emulated team
  int id, nt;
  std::barrier<>* bar;
  double* grp = nullptr;  // [nt] scratch of group_sum
  void sync() { bar->arrive_and_wait(); }
  // the lanes of a wave need only ordering on the GPU; here a full barrier
  void wave_sync() { sync(); }
  template <int G, typename T>
  T group_sum(T x) {
    grp[id] = static_cast<double>(x);
    sync();
    T s = T(0);
    const int g0 = id / G * G;
    for (int i = 0; i < G; ++i) s += static_cast<T>(grp[g0 + i]);
    sync();
    return s;
  }
  template <typename T>
  T reduce_sum_ns(T x, T* red) {
    red[id] = x;
    sync();
    T s = T(0);
    for (int i = 0; i < nt; ++i) s += red[i];
    return s;
  }
};

// Register-resident route: tridiag_reg and backtrans_reg with NB * S threads,
// the tridiagonal eigenpairs with the team-generic routines in between.
template <typename T, int NB, int S, int Sbt>
int run_reg(int n, const double* H, double* E, double* Z, int vectors, int rows_out) {
  using namespace rtb;
  constexpr int NT = NB * S, NTb = NB * Sbt;
  constexpr int SM = S > Sbt ? S : Sbt;
  std::vector<cplx<T>> xc(NB), vp(3 * NB), wp(NB), yp(2 * SM * NB);
  std::vector<T> red(2 * 2048), d(n), e(n), e2(n), Ev(n), cnt(n);
  std::vector<cplx<T>> Vg(static_cast<size_t>(n) * n), taug(n);
  std::vector<T> Y(static_cast<size_t>(n) * n), Sw(static_cast<size_t>(n) * n * 3), twist(n);
  std::vector<T> es(NB);
  RegSmem<T> sm{xc.data(), vp.data(), wp.data(), yp.data(), red.data(), es.data()};
  auto team = [](int nt, auto&& f) {
    std::barrier<> bar(nt);
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back([&, i] {
        CpuTeam tm{i, nt, &bar};
        f(tm);
      });
    for (auto& t : th) t.join();
  };
  team(NT, [&](auto& tm) { tridiag_reg<T, NB, S>(tm, n, H, sm, d.data(), e.data(), Vg.data(), taug.data()); });
  team(n, [&](auto& tm) {
    bisect_all<T>(tm, n, d.data(), e.data(), e2.data(), Ev.data(), cnt.data());
    if (vectors) tridiag_vectors<T>(tm, n, d.data(), e.data(), Ev.data(), Y.data(), n, Sw.data(), twist.data());
  });
  for (int j = 0; j < n; ++j) E[j] = static_cast<double>(Ev[j]);
  if (vectors)
    team(NTb, [&](auto& tm) {
      backtrans_reg<T, NB, Sbt>(tm, n, Y.data(), n, twist.data(), e.data(), Sw.data(), Vg.data(), taug.data(), sm, Z,
                                rows_out);
    });
  return 0;
}

// Two-stage route (eigh_band.hpp) with the team sizes of the GPU dispatch:
// band_reduce (256 threads), band_chase (256; FP32: band_chase_lds, 768), the
// tridiagonal eigenpairs,
// twisted_column, band_backtrans (W * S threads per block of W columns).
template <typename T, int W, int S>
int run_band(int n, const double* H, double* E, double* Z, int vectors, int rows_out) {
  using namespace rtb;
  const BandLayout L = band_layout(n);
  std::vector<cplx<T>> A(static_cast<size_t>(n) * n), band(L.total);
  std::vector<T> d(n), e(n), e2(n), Ev(n), cnt(n), twist(n), Y(static_cast<size_t>(n) * n),
      Sw(static_cast<size_t>(n) * n * 3);
  std::vector<cplx<T>> pV(static_cast<size_t>(n) * kBandLd), pY(static_cast<size_t>(n) * kBandLd),
      small(3 * kBandB * kBandB + kBandB), hv(kBandB), Ys(band_backtrans_team_elems<T, W>(n));
  std::vector<T> red(2 * 2048);
  std::vector<double> grp(2048);
  auto team = [&](int nt, auto&& f) {
    std::barrier<> bar(nt);
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back([&, i] {
        CpuTeam tm{i, nt, &bar, grp.data()};
        f(tm);
      });
    for (auto& t : th) t.join();
  };
  BandReduceSmem<T> sm{pV.data(), pY.data(), small.data(), small.data() + kBandB * kBandB,
                       small.data() + 2 * kBandB * kBandB, small.data() + 3 * kBandB * kBandB, red.data()};
  team(256, [&](auto& tm) { band_reduce<T>(tm, n, H, A.data(), band.data() + L.v1, band.data() + L.t1, sm); });
  if constexpr (sizeof(T) == 4) {  // the GPU's FP32 chase: band region in team memory, 768 threads
    std::vector<cplx<T>> AB(band_chase_team_elems(n));
    team(768, [&](auto& tm) {
      band_chase_lds<T>(tm, n, A.data(), d.data(), e.data(), band.data() + L.v2, band.data() + L.tau2, AB.data());
    });
  } else {
    team(256, [&](auto& tm) {
      band_chase<T>(tm, n, A.data(), d.data(), e.data(), band.data() + L.v2, band.data() + L.tau2, hv.data(),
                    red.data());
    });
  }
  team(n, [&](auto& tm) {
    bisect_all<T>(tm, n, d.data(), e.data(), e2.data(), Ev.data(), cnt.data());
    if (vectors) tridiag_vectors<T>(tm, n, d.data(), e.data(), Ev.data(), Y.data(), n, Sw.data(), twist.data());
  });
  for (int j = 0; j < n; ++j) E[j] = static_cast<double>(Ev[j]);
  if (!vectors) return 0;
  for (int j = 0; j < n; ++j) twisted_column<T>(n, e.data(), Sw.data(), n, j, twist[j], Y.data());
  for (int j0 = 0; j0 < n; j0 += W)
    team(W * S, [&](auto& tm) {
      band_backtrans<T, W, S>(tm, n, j0, Y.data(), n, band.data() + L.v1, band.data() + L.t1, band.data() + L.v2,
                              band.data() + L.tau2, Ys.data(), Z, rows_out, nullptr);
    });
  return 0;
}

}  // namespace

extern "C" int emul_eigh_band(int n, int fp32, const double* H, double* E, double* Z, int vectors, int rows_out) {
  if (n < 2 || n > 256) return -1;
  if (fp32) return run_band<float, 64, 16>(n, H, E, Z, vectors, rows_out);
  return run_band<double, 32, 16>(n, H, E, Z, vectors, rows_out);
}

extern "C" int emul_eigh_reg(int n, int fp32, const double* H, double* E, double* Z, int vectors, int rows_out) {
  if (n < 1 || n > 128) return -1;
  // slice shapes of the GPU dispatch (eigh.hip RegShape)
  if (fp32) {
    if (n <= 32) return run_reg<float, 32, 2, 2>(n, H, E, Z, vectors, rows_out);
    if (n <= 64) return run_reg<float, 64, 2, 2>(n, H, E, Z, vectors, rows_out);
    return run_reg<float, 128, 4, 2>(n, H, E, Z, vectors, rows_out);
  }
  if (n <= 32) return run_reg<double, 32, 2, 2>(n, H, E, Z, vectors, rows_out);
  if (n <= 64) return run_reg<double, 64, 2, 2>(n, H, E, Z, vectors, rows_out);
  return run_reg<double, 128, 4, 4>(n, H, E, Z, vectors, rows_out);
}
