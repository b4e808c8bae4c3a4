// FP64 refinement routines shared by the GPU kernels and host-side checks.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define RTB_RHD __host__ __device__
#else
#define RTB_RHD
#endif

namespace rtb {

constexpr int kMaxCluster = 8;

struct C2 {
  double re, im;
};
struct Cd {  // layout of an interleaved (re, im) double pair
  double x, y;
};
RTB_RHD inline C2 cm(C2 a, C2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
RTB_RHD inline C2 cmc(C2 a, C2 b) { return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }  // conj(a) b

// Small-matrix part of the Rayleigh-Ritz step for one cluster of c columns:
// from the Gram matrix G = Zc^H Zc (buf[0]) and the projection
// M = Zc^H (H Zc) (buf[1]) computes X (buf[6]) such that Zc X are the
// orthonormal Ritz vectors, and the Ritz values ev[0..c) in ascending order.
// Returns false (nothing computed) when G is not positive definite.
// buf: 7 scratch c x c matrices (G and M are overwritten).
RTB_RHD inline bool cluster_rr_small(int c, C2 (*buf)[kMaxCluster][kMaxCluster], double* ev) {
  using std::sqrt;
  using std::fabs;
  using std::atan2;
  using std::cos;
  using std::sin;
  C2(*G)[kMaxCluster] = buf[0];
  C2(*M)[kMaxCluster] = buf[1];
  C2(*Li)[kMaxCluster] = buf[2];
  C2(*T)[kMaxCluster] = buf[3];
  C2(*P)[kMaxCluster] = buf[4];
  C2(*U)[kMaxCluster] = buf[5];
  C2(*X)[kMaxCluster] = buf[6];
  // Hermitian parts (M is Hermitian in exact arithmetic)
  for (int a = 0; a < c; ++a)
    for (int b = a; b < c; ++b) {
      C2 x = M[a][b], y = M[b][a];
      C2 hsym = {0.5 * (x.re + y.re), 0.5 * (x.im - y.im)};
      M[a][b] = hsym;
      M[b][a] = {hsym.re, -hsym.im};
    }
  // Cholesky G = L L^H (L lower), stored in G's lower triangle
  for (int a = 0; a < c; ++a) {
    double d = G[a][a].re;
    for (int q = 0; q < a; ++q) d -= G[a][q].re * G[a][q].re + G[a][q].im * G[a][q].im;
    if (!(d > 0.0)) return false;
    d = sqrt(d);
    G[a][a] = {d, 0.0};
    for (int b = a + 1; b < c; ++b) {
      C2 s = G[b][a];
      for (int q = 0; q < a; ++q) {
        C2 t = cm(G[b][q], {G[a][q].re, -G[a][q].im});
        s.re -= t.re;
        s.im -= t.im;
      }
      G[b][a] = {s.re / d, s.im / d};
    }
  }
  // Li = L^{-1}: solve column by column (lower triangular)
  for (int a = 0; a < c; ++a)
    for (int b = 0; b < c; ++b) Li[a][b] = {a == b ? 1.0 : 0.0, 0.0};
  for (int col = 0; col < c; ++col)
    for (int a = 0; a < c; ++a) {
      C2 s = Li[a][col];
      for (int q = 0; q < a; ++q) {
        C2 t = cm(G[a][q], Li[q][col]);
        s.re -= t.re;
        s.im -= t.im;
      }
      Li[a][col] = {s.re / G[a][a].re, s.im / G[a][a].re};
    }
  // P = L^{-1} M L^{-H}
  for (int a = 0; a < c; ++a)
    for (int b = 0; b < c; ++b) {
      C2 s = {0.0, 0.0};
      for (int q = 0; q < c; ++q) {
        C2 t = cm(Li[a][q], M[q][b]);
        s.re += t.re;
        s.im += t.im;
      }
      T[a][b] = s;
    }
  for (int a = 0; a < c; ++a)
    for (int b = 0; b < c; ++b) {
      C2 s = {0.0, 0.0};
      for (int q = 0; q < c; ++q) {
        C2 t = cm(T[a][q], {Li[b][q].re, -Li[b][q].im});
        s.re += t.re;
        s.im += t.im;
      }
      P[a][b] = s;
    }
  // cyclic complex Jacobi on P; accumulate U (P = U D U^H)
  for (int a = 0; a < c; ++a)
    for (int b = 0; b < c; ++b) U[a][b] = {a == b ? 1.0 : 0.0, 0.0};
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int a = 0; a < c; ++a) {
      dia += P[a][a].re * P[a][a].re;
      for (int b = a + 1; b < c; ++b) off += P[a][b].re * P[a][b].re + P[a][b].im * P[a][b].im;
    }
    if (off <= 1e-34 * dia || off == 0.0) break;
    for (int p = 0; p < c - 1; ++p)
      for (int q = p + 1; q < c; ++q) {
        const C2 apq = P[p][q];
        const double mag = sqrt(apq.re * apq.re + apq.im * apq.im);
        if (mag == 0.0) continue;
        const double app = P[p][p].re, aqq = P[q][q].re;
        // rotation annihilating P[p][q]: phase e = apq / |apq|
        const double theta = 0.5 * atan2(2.0 * mag, aqq - app);
        const double cs = cos(theta), sn = sin(theta);
        const C2 e = {apq.re / mag, apq.im / mag};
        // columns p, q: [p q] <- [p q] [[cs, sn e], [-sn conj(e), cs]]
        for (int a = 0; a < c; ++a) {
          const C2 xp = P[a][p], xq = P[a][q];
          const C2 t1 = cm(xq, {-sn * e.re, sn * e.im});  // xq * (-sn conj(e))
          const C2 t2 = cm(xp, {sn * e.re, sn * e.im});   // xp * (sn e)
          P[a][p] = {cs * xp.re + t1.re, cs * xp.im + t1.im};
          P[a][q] = {cs * xq.re + t2.re, cs * xq.im + t2.im};
          const C2 up = U[a][p], uq = U[a][q];
          const C2 s1 = cm(uq, {-sn * e.re, sn * e.im});
          const C2 s2 = cm(up, {sn * e.re, sn * e.im});
          U[a][p] = {cs * up.re + s1.re, cs * up.im + s1.im};
          U[a][q] = {cs * uq.re + s2.re, cs * uq.im + s2.im};
        }
        // rows p, q: conjugate transpose of the same rotation
        for (int b = 0; b < c; ++b) {
          const C2 yp = P[p][b], yq = P[q][b];
          const C2 t1 = cm({-sn * e.re, -sn * e.im}, yq);  // (-sn e) yq
          const C2 t2 = cm({sn * e.re, -sn * e.im}, yp);   // (sn conj(e)) yp
          P[p][b] = {cs * yp.re + t1.re, cs * yp.im + t1.im};
          P[q][b] = {cs * yq.re + t2.re, cs * yq.im + t2.im};
        }
      }
  }
  // sort eigenvalues ascending (selection) with U columns
  int order[kMaxCluster];
  for (int a = 0; a < c; ++a) order[a] = a;
  for (int a = 0; a < c; ++a)
    for (int b = a + 1; b < c; ++b)
      if (P[order[b]][order[b]].re < P[order[a]][order[a]].re) {
        int t = order[a];
        order[a] = order[b];
        order[b] = t;
      }
  // X = L^{-H} U: new cluster columns Zc X
  for (int a = 0; a < c; ++a)
    for (int b = 0; b < c; ++b) {
      C2 s = {0.0, 0.0};
      for (int q = 0; q < c; ++q) {
        C2 t = cm({Li[q][a].re, -Li[q][a].im}, U[q][order[b]]);
        s.re += t.re;
        s.im += t.im;
      }
      X[a][b] = s;
    }
  for (int b = 0; b < c; ++b) ev[b] = P[order[b]][order[b]].re;
  return true;
}

// Cluster boundaries: [j0, j1] with consecutive gaps <= thr, at most
// kMaxCluster members (a longer run is split).
RTB_RHD inline int cluster_end(int n, const double* Ek, int j0, double thr) {
  int j1 = j0;
  while (j1 + 1 < n && Ek[j1 + 1] - Ek[j1] <= thr && j1 + 1 - j0 < kMaxCluster) ++j1;
  return j1;
}

// Rayleigh-Ritz inside every cluster of one k-point (columns of Zk, HZ =
// H Z for unit columns), serial form: the n-length sums and the column
// rotation around cluster_rr_small.  buf: 7 scratch c x c matrices.
RTB_RHD inline void cluster_rr_one(int n, Cd* Zk, const Cd* Hk, double* Ek, double gap_rel,
                                   C2 (*buf)[kMaxCluster][kMaxCluster]) {
  using std::fabs;
  using std::fmax;
  double scale = 0.0;
  for (int j = 0; j < n; ++j) scale = fmax(scale, fabs(Ek[j]));
  const double thr = gap_rel * fmax(scale, 1.0);
  int j0 = 0;
  while (j0 < n) {
    const int j1 = cluster_end(n, Ek, j0, thr);
    const int c = j1 - j0 + 1;
    if (c >= 2) {
      for (int a = 0; a < c; ++a)
        for (int b = 0; b < c; ++b) {
          C2 g = {0.0, 0.0}, mm = {0.0, 0.0};
          for (int r = 0; r < n; ++r) {
            const Cd za = Zk[static_cast<size_t>(r) * n + j0 + a];
            const Cd zb = Zk[static_cast<size_t>(r) * n + j0 + b];
            const Cd hb = Hk[static_cast<size_t>(r) * n + j0 + b];
            C2 t = cmc({za.x, za.y}, {zb.x, zb.y});
            g.re += t.re;
            g.im += t.im;
            C2 u = cmc({za.x, za.y}, {hb.x, hb.y});
            mm.re += u.re;
            mm.im += u.im;
          }
          buf[0][a][b] = g;
          buf[1][a][b] = mm;
        }
      double ev[kMaxCluster];
      if (cluster_rr_small(c, buf, ev)) {
        C2(*X)[kMaxCluster] = buf[6];
        for (int r = 0; r < n; ++r) {
          C2 row[kMaxCluster];
          for (int a = 0; a < c; ++a) {
            const Cd z = Zk[static_cast<size_t>(r) * n + j0 + a];
            row[a] = {z.x, z.y};
          }
          for (int b = 0; b < c; ++b) {
            C2 s = {0.0, 0.0};
            for (int a = 0; a < c; ++a) {
              C2 t = cm(row[a], X[a][b]);
              s.re += t.re;
              s.im += t.im;
            }
            Zk[static_cast<size_t>(r) * n + j0 + b] = Cd{s.re, s.im};
          }
        }
        for (int b = 0; b < c; ++b) Ek[j0 + b] = ev[b];
      }
    }
    j0 = j1 + 1;
  }
}

}  // namespace rtb
