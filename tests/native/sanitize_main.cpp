// Sanitizer driver for the host emulations of the device source: runs the
// eigensolver workgroups (team_emul.cpp: eigh_core.hpp, eigh_reg.hpp with
// host threads and a barrier) and the cluster Rayleigh-Ritz
// (refine_emul.cpp: refine_core.hpp) on seeded Hermitian matrices and checks
// residuals and orthonormality itself, so that it can be built with
// AddressSanitizer + UndefinedBehaviorSanitizer or ThreadSanitizer as a
// standalone program.  Exit status 0: every case within its bound.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "refine_emul.cpp"
#include "team_emul.cpp"

namespace {

using cd = std::complex<double>;

struct Rng {  // SplitMix64
  uint64_t s;
  double uniform() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  }
};

std::vector<cd> hermitian(int n, uint64_t seed) {
  Rng r{seed};
  std::vector<cd> H(static_cast<size_t>(n) * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      const cd v = i == j ? cd(r.uniform(), 0.0) : cd(r.uniform(), r.uniform());
      H[i * n + j] = v;
      H[j * n + i] = std::conj(v);
    }
  return H;
}

// max |H z_j - E_j z_j| and max |Z^H Z - I| for column eigenvectors Z
void check(int n, const std::vector<cd>& H, const std::vector<double>& E, const std::vector<cd>& Z, double& res,
           double& orth) {
  res = orth = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      cd s = -E[j] * Z[i * n + j];
      for (int q = 0; q < n; ++q) s += H[i * n + q] * Z[q * n + j];
      res = std::fmax(res, std::abs(s));
    }
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      cd s = a == b ? cd(-1.0, 0.0) : cd(0.0, 0.0);
      for (int q = 0; q < n; ++q) s += std::conj(Z[q * n + a]) * Z[q * n + b];
      orth = std::fmax(orth, std::abs(s));
    }
}

}  // namespace

int main() {
  int bad = 0;
  // sizes of each slice shape (32, 64, 128 rows) and the smallest ones
  const int sizes[2][4] = {{1, 2, 65, 128}, {17, 33, 100, 128}};
  for (int fp32 = 0; fp32 < 2; ++fp32)
    for (int n : sizes[fp32]) {
      const std::vector<cd> H = hermitian(n, 1000 + n);
      std::vector<double> E(n);
      std::vector<cd> Z(static_cast<size_t>(n) * n);
      const int rc = emul_eigh_reg(n, fp32, reinterpret_cast<const double*>(H.data()), E.data(),
                                   reinterpret_cast<double*>(Z.data()), 1, 0);
      double res = 0.0, orth = 0.0;
      check(n, H, E, Z, res, orth);
      const double tol_r = fp32 ? 2e-5 * n : 1e-13 * n, tol_o = fp32 ? 1e-4 : 1e-13 * n;
      const bool ok = rc == 0 && res <= tol_r && orth <= tol_o;
      std::printf("eigh n=%d fp32=%d residual %.2e orthonormality %.2e %s\n", n, fp32, res, orth, ok ? "ok" : "FAIL");
      bad += !ok;
    }
  // the two-stage route (eigh_band.hpp) at sizes with one, two and several
  // panels
  for (int fp32 = 0; fp32 < 2; ++fp32)
    for (int n : {fp32 ? 33 : 17, 50}) {
      const std::vector<cd> H = hermitian(n, 2000 + n);
      std::vector<double> E(n);
      std::vector<cd> Z(static_cast<size_t>(n) * n);
      const int rc = emul_eigh_band(n, fp32, reinterpret_cast<const double*>(H.data()), E.data(),
                                    reinterpret_cast<double*>(Z.data()), 1, 0);
      double res = 0.0, orth = 0.0;
      check(n, H, E, Z, res, orth);
      const double tol_r = fp32 ? 2e-5 * n : 1e-13 * n, tol_o = fp32 ? 1e-4 : 1e-13 * n;
      const bool ok = rc == 0 && res <= tol_r && orth <= tol_o;
      std::printf("band n=%d fp32=%d residual %.2e orthonormality %.2e %s\n", n, fp32, res, orth, ok ? "ok" : "FAIL");
      bad += !ok;
    }
  // cluster Rayleigh-Ritz: exact eigenvectors of a matrix with a degenerate
  // triple, mixed inside the triple, then restored
  {
    const int n = 12;
    std::vector<cd> Q = hermitian(n, 77);
    std::vector<double> E0(n);
    std::vector<cd> Z0(static_cast<size_t>(n) * n);
    emul_eigh_reg(n, 0, reinterpret_cast<const double*>(Q.data()), E0.data(), reinterpret_cast<double*>(Z0.data()),
                  1, 0);
    std::vector<double> w = {-2.0, -1.5, -1.0, 0.5, 0.5, 0.5, 1.0, 1.2, 1.4, 2.0, 2.5, 3.0};
    std::vector<cd> H(static_cast<size_t>(n) * n, cd(0.0, 0.0));  // H = Z0 diag(w) Z0^H
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        for (int q = 0; q < n; ++q) H[i * n + j] += Z0[i * n + q] * w[q] * std::conj(Z0[j * n + q]);
    std::vector<cd> Z = Z0;
    const double c = std::cos(0.3), s = std::sin(0.3);
    for (int i = 0; i < n; ++i) {  // rotate columns 3 and 4 into each other
      const cd a = Z[i * n + 3], b = Z[i * n + 4];
      Z[i * n + 3] = c * a + s * b;
      Z[i * n + 4] = -s * a + c * b;
    }
    std::vector<cd> HZ(static_cast<size_t>(n) * n, cd(0.0, 0.0));
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        for (int q = 0; q < n; ++q) HZ[i * n + j] += H[i * n + q] * Z[q * n + j];
    std::vector<double> E = w;
    emul_cluster_rr(n, reinterpret_cast<double*>(Z.data()), reinterpret_cast<const double*>(HZ.data()), E.data(),
                    1e-3);
    double res = 0.0, orth = 0.0;
    check(n, H, E, Z, res, orth);
    const bool ok = res <= 1e-13 && orth <= 1e-13;
    std::printf("cluster_rr residual %.2e orthonormality %.2e %s\n", res, orth, ok ? "ok" : "FAIL");
    bad += !ok;
  }
  return bad ? 1 : 0;
}
