// Reference-equivalent CPU path (test and CPU-baseline infrastructure only).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, as the checker / CPU baseline; the product never links it.
//
// Restates the reference's per-k algorithm in C++ with OpenMP over k (one
// k-point per thread, as rayon's par_iter over k does) and a host LAPACK
// (MKL sequential or single-threaded OpenBLAS) for zheev / zgemm:
//   gen_v_projected (velocity.rs:470-731): phases exp(i 2 pi R.k) not reduced,
//     zaxpy accumulation of H_lat and of each projected velocity in stored R
//     order, rows with w.R == 0 skipped, tau term + Atom gauge fused;
//   eigh (ndarray-linalg -> zheev 'V'; solve_ham.rs:110-129);
//   A = U^H J U, B = U^H v_b U with zgemm, Omega_n Kubo loop
//     (response/traits.rs:160-200).
// Eigenvalue-only mode follows solve_band_onek (solve_ham.rs:52-63, zheev 'N').
#include <dlfcn.h>
#include <omp.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

using cd = std::complex<double>;
typedef void (*zheev_t)(const char*, const char*, const int*, cd*, const int*, double*, cd*, const int*, double*,
                        int*, size_t, size_t);
typedef void (*zgemm_t)(const char*, const char*, const int*, const int*, const int*, const cd*, const cd*,
                        const int*, const cd*, const int*, const cd*, cd*, const int*, size_t, size_t);
zheev_t g_zheev = nullptr;
zgemm_t g_zgemm = nullptr;
std::string g_lib;

}  // namespace

extern "C" {

// Load LAPACK from `path`, symbols `<prefix>zheev_` / `<prefix>zgemm_`, and
// force single-threaded BLAS (threads come from OpenMP over k).
int cpu_ref_init(const char* path, const char* prefix) {
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
  std::string p(prefix ? prefix : "");
  g_zheev = reinterpret_cast<zheev_t>(dlsym(h, (p + "zheev_").c_str()));
  g_zgemm = reinterpret_cast<zgemm_t>(dlsym(h, (p + "zgemm_").c_str()));
  if (!g_zheev || !g_zgemm) return -2;
  if (auto f = reinterpret_cast<int (*)(int)>(dlsym(h, "MKL_Set_Threading_Layer"))) f(1);  // SEQUENTIAL
  if (auto f = reinterpret_cast<void (*)(int)>(dlsym(h, (p + "openblas_set_num_threads").c_str()))) f(1);
  g_lib = path;
  return 0;
}

const char* cpu_ref_lib() { return g_lib.c_str(); }

// mode 0: E only; mode 2: Omega + E (two projected directions, Atom gauge).
// ham [nR][n][n] complex; R [nR][3]; wR [2][nR]; tau [n][3]; wtau [2][n];
// k [nk][3].  Returns 0, or the first nonzero LAPACK info.
int cpu_ref_run(int mode, int n, int nR, const double* ham_, const double* R, const double* wR, const double* tau,
                const double* wtau, const double* k, long nk, double eta, int threads, double* omega, double* E) {
  if (!g_zheev || !g_zgemm) return -100;
  const cd* ham = reinterpret_cast<const cd*>(ham_);
  const size_t n2 = static_cast<size_t>(n) * n;
  const double two_pi = 2.0 * M_PI;
  int err = 0;
#pragma omp parallel num_threads(threads)
  {
    std::vector<cd> us(nR), H(n2), V0(n2), V1(n2), phi(n), U(n2), J(n2), T1(n2), A(n2), B(n2);
    std::vector<double> w(n), rwork(3 * n);
    int lwork = -1, info = 0;
    cd wq;
    const char* jobz = mode == 0 ? "N" : "V";
    g_zheev(jobz, "L", &n, H.data(), &n, w.data(), &wq, &lwork, rwork.data(), &info, 1, 1);
    lwork = static_cast<int>(wq.real());
    std::vector<cd> work(lwork > 1 ? lwork : 1);
#pragma omp for schedule(dynamic, 1)
    for (long ik = 0; ik < nk; ++ik) {
      const double* kk = k + ik * 3;
      for (int r = 0; r < nR; ++r) {
        double arg = two_pi * (R[r * 3] * kk[0] + R[r * 3 + 1] * kk[1] + R[r * 3 + 2] * kk[2]);
        us[r] = cd(std::cos(arg), std::sin(arg));
      }
      std::fill(H.begin(), H.end(), cd(0.0, 0.0));
      for (int r = 0; r < nR; ++r) {  // zaxpy in stored R order
        const cd u = us[r];
        const cd* h = ham + r * n2;
        for (size_t e = 0; e < n2; ++e) H[e] += u * h[e];
      }
      if (mode == 2) {
        cd* Vs[2] = {V0.data(), V1.data()};
        for (int p = 0; p < 2; ++p) {
          std::fill(Vs[p], Vs[p] + n2, cd(0.0, 0.0));
          for (int r = 0; r < nR; ++r) {
            const double rw = wR[p * nR + r];
            if (rw == 0.0) continue;
            const cd a = us[r] * cd(0.0, 1.0) * rw;
            const cd* h = ham + r * n2;
            for (size_t e = 0; e < n2; ++e) Vs[p][e] += a * h[e];
          }
        }
      }
      for (int o = 0; o < n; ++o) {
        double arg = two_pi * (tau[o * 3] * kk[0] + tau[o * 3 + 1] * kk[1] + tau[o * 3 + 2] * kk[2]);
        phi[o] = cd(std::cos(arg), std::sin(arg));
      }
      for (int m = 0; m < n; ++m) {
        const cd cm = std::conj(phi[m]);
        for (int c = 0; c < n; ++c) {
          const cd g = cm * phi[c];
          if (mode == 2) {
            for (int p = 0; p < 2; ++p) {
              cd* Vp = p == 0 ? V0.data() : V1.data();
              const double diff = wtau[p * n + c] - wtau[p * n + m];
              Vp[m * n + c] = (Vp[m * n + c] + cd(0.0, diff) * H[m * n + c]) * g;
            }
          }
        }
        for (int c = 0; c < n; ++c) H[m * n + c] *= cm * phi[c];
      }
      // the row-major buffer read column-major is H^T = conj(H): same spectrum,
      // eigenvectors W = conj(U) in its columns (ndarray-linalg's convention)
      int inf = 0;
      g_zheev(jobz, "L", &n, H.data(), &n, w.data(), work.data(), &lwork, rwork.data(), &inf, 1, 1);
      if (inf != 0) {
#pragma omp atomic write
        err = inf;
      }
      for (int j = 0; j < n; ++j) E[ik * n + j] = w[j];
      if (mode != 2) continue;
      // column-major operands: U[:, j] = conj(W[:, j]); J, v_b transposed in
      const cd one(1.0, 0.0), zero(0.0, 0.0);
      for (size_t e = 0; e < n2; ++e) U[e] = std::conj(H[e]);
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
          J[c * n + r] = V0[r * n + c];
          T1[c * n + r] = V1[r * n + c];
        }
      // A = U^H (J U); B = U^H (v_b U)   (column-major)
      g_zgemm("N", "N", &n, &n, &n, &one, J.data(), &n, U.data(), &n, &zero, A.data(), &n, 1, 1);
      g_zgemm("C", "N", &n, &n, &n, &one, U.data(), &n, A.data(), &n, &zero, J.data(), &n, 1, 1);
      g_zgemm("N", "N", &n, &n, &n, &one, T1.data(), &n, U.data(), &n, &zero, A.data(), &n, 1, 1);
      g_zgemm("C", "N", &n, &n, &n, &one, U.data(), &n, A.data(), &n, &zero, B.data(), &n, 1, 1);
      // J now holds A (column-major: A_bm = J[b + m n]); B_mb = B[m + b n]
      const double eta2 = eta * eta;
      for (int b = 0; b < n; ++b) {
        double acc = 0.0;
        for (int m = 0; m < n; ++m) {
          if (m == b) continue;
          const cd kbm = J[b + static_cast<size_t>(m) * n] * B[m + static_cast<size_t>(b) * n];
          const double d = w[b] - w[m];
          acc += -2.0 * kbm.imag() / (d * d + eta2);
        }
        omega[ik * n + b] = acc;
      }
    }
  }
  return err;
}

}  // extern "C"
