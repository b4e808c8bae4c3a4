// CPU emulation of the eigensolver workgroups: the team-generic device
// source (rustb_amd/csrc/eigh_core.hpp, eigh_reg.hpp) run by host threads
// with a barrier, in the slice shapes of the GPU dispatch.  Test-only: lets
// the indexing and numerics be checked on the host (with AddressSanitizer
// and ThreadSanitizer).
#include <barrier>
#include <thread>
#include <vector>

#include "../../rustb_amd/csrc/eigh_core.hpp"
#include "../../rustb_amd/csrc/eigh_reg.hpp"

namespace {

struct CpuTeam {
  static constexpr int kRedStride = 2048;  // >= 2 x the largest emulated team
  int id, nt;
  std::barrier<>* bar;
  void sync() { bar->arrive_and_wait(); }
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

}  // namespace

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
