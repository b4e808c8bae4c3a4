// CPU emulation of one eigensolver workgroup: the team-generic device source
// (rustb_amd/csrc/eigh_core.hpp) run by `nt` host threads with a barrier.
// Test-only: lets the indexing and numerics be checked on the host (with
// AddressSanitizer) for any team size.
#include <barrier>
#include <thread>
#include <vector>

#include "../../rustb_amd/csrc/eigh_core.hpp"
#include "../../rustb_amd/csrc/eigh_reg.hpp"

namespace {

template <bool Pairs>
struct CpuTeam {
  static constexpr bool kPairs = Pairs;
  static constexpr int kRedStride = 2048;  // >= 2 x the largest emulated team
  int id, nt;
  std::barrier<>* bar;
  void sync() { bar->arrive_and_wait(); }
  template <typename T>
  T pair_sum(T x, T* red) {
    red[id] = x;
    sync();
    T y = x + red[id ^ 1];
    sync();
    return y;
  }
  template <typename T>
  T reduce_sum_ns(T x, T* red) {
    red[id] = x;
    sync();
    T s = T(0);
    for (int i = 0; i < nt; ++i) s += red[i];
    return s;
  }
  template <typename T>
  void reduce_sum2_ns(T& x, T& y, T* red) {
    red[2 * id] = x;
    red[2 * id + 1] = y;
    sync();
    T sx = T(0), sy = T(0);
    for (int i = 0; i < nt; ++i) {
      sx += red[2 * i];
      sy += red[2 * i + 1];
    }
    x = sx;
    y = sy;
  }
  template <typename T>
  T reduce_sum(T x, T* red) {
    red[id] = x;
    sync();
    T s = T(0);
    for (int i = 0; i < nt; ++i) s += red[i];
    sync();
    return s;
  }
};

template <typename T, bool Pairs>
int run(int n, int nt, const double* H, double* E, double* Z, int vectors) {
  using namespace rtb;
  std::vector<cplx<T>> P(static_cast<size_t>(n) * (n + 1) / 2), v(n), w(n);
  std::vector<T> d(n), e(n), e2(n), Ev(n), red(4 * 2048 + nt);
  std::vector<cplx<T>> Vg(static_cast<size_t>(n) * n), taug(n);
  std::vector<T> Y(static_cast<size_t>(n) * n), S(static_cast<size_t>(n) * n * 3), ysc(n);
  std::vector<cplx<T>> Zb(static_cast<size_t>(n) * n), vs(2 * n), part(nt > n ? nt : n), sv(n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c <= r; ++c) {
      const double* h = H + (static_cast<size_t>(r) * n + c) * 2;
      P[pidx(r, c)] = {static_cast<T>(h[0]), static_cast<T>(h[1])};
    }
  std::barrier<> bar(nt);
  auto body = [&](int id) {
    CpuTeam<Pairs> tm{id, nt, &bar};
    tridiagonalize<T>(tm, n, P.data(), v.data(), w.data(), d.data(), e.data(), red.data(), Vg.data(),
                      taug.data());
    bisect_all<T>(tm, n, d.data(), e.data(), e2.data(), Ev.data());
    if (vectors) {
      tridiag_vectors<T>(tm, n, d.data(), e.data(), Ev.data(), Y.data(), n, S.data(), ysc.data());
      for (int f = id; f < n * n; f += nt) Zb[f] = {Y[f] * ysc[f % n], T(0)};
      tm.sync();
      apply_q<T>(tm, n, n, Zb.data(), Vg.data(), taug.data(), vs.data(), part.data(), sv.data());
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nt; ++i) th.emplace_back(body, i);
  for (auto& t : th) t.join();
  for (int j = 0; j < n; ++j) E[j] = static_cast<double>(Ev[j]);
  if (vectors)
    for (int f = 0; f < n * n; ++f) {
      Z[2 * f] = static_cast<double>(Zb[f].re);
      Z[2 * f + 1] = static_cast<double>(Zb[f].im);
    }
  return 0;
}

// Register-resident route: tridiag_reg and backtrans_reg with NB * S threads,
// the tridiagonal eigenpairs with the team-generic routines in between.
template <typename T, int NB, int S, int Sbt>
int run_reg(int n, const double* H, double* E, double* Z, int vectors, int rows_out) {
  using namespace rtb;
  constexpr int NT = NB * S, NTb = NB * Sbt;
  constexpr int SM = S > Sbt ? S : Sbt;
  std::vector<cplx<T>> xc(NB), vp(3 * NB), wp(NB), yp(2 * SM * NB);
  std::vector<T> red(2 * 2048), d(n), e(n), e2(n), Ev(n), red2(4 * 2048 + n);
  std::vector<cplx<T>> Vg(static_cast<size_t>(n) * n), taug(n);
  std::vector<T> Y(static_cast<size_t>(n) * n), Sw(static_cast<size_t>(n) * n * 3), ysc(n);
  RegSmem<T> sm{xc.data(), vp.data(), wp.data(), yp.data(), red.data()};
  auto team = [](int nt, auto&& f) {
    std::barrier<> bar(nt);
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back([&, i] {
        CpuTeam<false> tm{i, nt, &bar};
        f(tm);
      });
    for (auto& t : th) t.join();
  };
  team(NT, [&](auto& tm) { tridiag_reg<T, NB, S>(tm, n, H, sm, d.data(), e.data(), Vg.data(), taug.data()); });
  team(n, [&](auto& tm) {
    bisect_all<T>(tm, n, d.data(), e.data(), e2.data(), Ev.data());
    if (vectors) tridiag_vectors<T>(tm, n, d.data(), e.data(), Ev.data(), Y.data(), n, Sw.data(), ysc.data());
  });
  for (int j = 0; j < n; ++j) E[j] = static_cast<double>(Ev[j]);
  if (vectors)
    team(NTb, [&](auto& tm) {
      backtrans_reg<T, NB, Sbt>(tm, n, Y.data(), n, ysc.data(), Vg.data(), taug.data(), sm, Z, rows_out);
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

extern "C" int emul_eigh(int n, int nt, int fp32, const double* H, double* E, double* Z, int vectors) {
  // even team sizes exercise the paired-row matvec used on the GPU
  if (nt % 2 == 0)
    return fp32 ? run<float, true>(n, nt, H, E, Z, vectors) : run<double, true>(n, nt, H, E, Z, vectors);
  return fp32 ? run<float, false>(n, nt, H, E, Z, vectors) : run<double, false>(n, nt, H, E, Z, vectors);
}
