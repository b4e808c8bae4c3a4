// Two-stage reduction for the batched eigensolver at 128 < n <= 256, where a
// matrix no longer fits on chip (one FP64 matrix of order 256 is 1 MB):
//
//   1. dense -> band of half-bandwidth b (kBandB) by blocked Householder
//      panels: QR of the b columns below the band (zgeqr2 reflectors in
//      LDS), their compact WY form Q = I - V T V^H (zlarft), and the two-sided
//      update A22 <- Q^H A22 Q = A22 - V Z^H - Z V^H with Y = A22 V T,
//      Z = Y - 1/2 V (T^H V^H Y); the matrix stays in global memory and each
//      panel reads it twice and writes it once (band_reduce);
//   2. band -> real symmetric tridiagonal by Householder bulge chasing: sweep
//      i annihilates column i below its sub-diagonal with a reflector on rows
//      [i+1, i+b], then chases the bulge down b rows at a time, each step
//      annihilating the first column of the bulge; the fill a sweep leaves is
//      removed by the next one (band_chase);
//   3. the tridiagonal eigenpairs as for n <= 128 (bisection, twisted
//      factorisations, inverse iteration in clusters), with the eigenvectors
//      of T written out (twisted_column);
//   4. Z = Q1 Q2 Y: the chasing reflectors in reverse order, then the panels
//      in reverse order, on blocks of W columns of Y held in team memory
//      (band_backtrans).
//
// Team-generic like eigh_core.hpp: the same source runs on a GPU workgroup
// and on host threads for verification.  Teams provide id, nt, sync(),
// reduce_sum_ns(x, red), group_sum<G>(x) (the sum over the aligned group of
// G consecutive ids, returned to each of them) and wave_sync()
// (orders team-memory accesses among the threads of one wave).
#pragma once

#include "eigh_core.hpp"

namespace rtb {

constexpr int kBandB = 16;       // half-bandwidth of the intermediate band matrix
constexpr int kBandLd = kBandB + 1;  // padded row stride of the panels in team memory
constexpr int kMaxBandN = 256;   // largest order of the two-stage route

// Per-matrix storage of the reflectors (complex elements):
//   V1 [P][n][b]   panel p's V (rows 0..m_p - 1 used, m_p = n - (p + 1) b)
//   T1 [P][b][b]   panel p's T (upper triangular)
//   V2 [n - 1][K][b]  sweep i, step k: reflector on rows i + 1 + k b ..
//   tau2 [n - 1][K]
// band_backtrans's team memory: the column block Ys [n][W + pad], then two
// staging buffers of NB slots (reflectors or panel rows) of b + 1 complex
// elements (the b entries, then the reflector's tau).  The strides keep the
// lanes of one LDS access on distinct banks: thread (c, s) reads row s + S u
// of column c, so a row stride of 2 mod 16 complex (FP32) or odd (FP64,
// 16-byte elements) spreads the 8 rows x 2 columns of a 16-lane access; the
// 8 lanes of a column read staged rows 17 elements apart.
template <typename T>
struct BandBtStage {
  static constexpr int NB = sizeof(T) == 4 ? 64 : 32;
  static constexpr int pad = sizeof(T) == 4 ? 2 : 1;
};
template <typename T, int W>
RTB_HD inline size_t band_backtrans_team_elems(int n) {
  return static_cast<size_t>(n) * (W + BandBtStage<T>::pad) +
         2 * static_cast<size_t>(BandBtStage<T>::NB) * (kBandB + 1);
}

struct BandLayout {
  int n, P, K;
  size_t v1, t1, v2, tau2, total;
};

RTB_HD inline BandLayout band_layout(int n) {
  BandLayout L;
  L.n = n;
  L.P = n > kBandB ? (n - 1) / kBandB : 0;
  L.K = (n - 1 + kBandB - 1) / kBandB;
  if (L.K < 1) L.K = 1;
  L.v1 = 0;
  L.t1 = L.v1 + static_cast<size_t>(L.P) * n * kBandB;
  L.v2 = L.t1 + static_cast<size_t>(L.P) * kBandB * kBandB;
  L.tau2 = L.v2 + static_cast<size_t>(n > 1 ? n - 1 : 1) * L.K * kBandB;
  L.total = L.tau2 + static_cast<size_t>(n > 1 ? n - 1 : 1) * L.K;
  return L;
}

// zlarfg on (alpha, |x|^2 of the rest): H = I - tau v v^H with v[0] = 1,
// H^H (alpha, x) = (beta, 0); the rest of v is x * scale.
template <typename T>
struct Refl {
  cplx<T> tau, scale;
  T beta;
};
template <typename T>
RTB_HD inline Refl<T> make_reflector(cplx<T> alpha, T xnorm2) {
  Refl<T> h;
  h.tau = {T(0), T(0)};
  h.scale = {T(0), T(0)};
  h.beta = alpha.re;
  if (xnorm2 == T(0) && alpha.im == T(0)) return h;
  const T nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm2);
  h.beta = alpha.re >= T(0) ? -nrm : nrm;
  const T rb = T(1) / h.beta;
  h.tau = {(h.beta - alpha.re) * rb, -alpha.im * rb};
  const T ar = alpha.re - h.beta, ai = alpha.im;
  const T rden = T(1) / (ar * ar + ai * ai);
  h.scale = {ar * rden, -ai * rden};
  return h;
}

template <typename T>
RTB_HD inline void cfma(cplx<T>& acc, cplx<T> a, cplx<T> b) {  // acc += a b
  acc.re += a.re * b.re - a.im * b.im;
  acc.im += a.re * b.im + a.im * b.re;
}
template <typename T>
RTB_HD inline void cfmac(cplx<T>& acc, cplx<T> a, cplx<T> b) {  // acc += conj(a) b
  acc.re += a.re * b.re + a.im * b.im;
  acc.im += a.re * b.im - a.im * b.re;
}

// Team memory of band_reduce: two m x b complex panels (m <= n - b), three
// b x b matrices, b reflector scalars and reduction scratch of
// 2 * Team::kRedStride.
template <typename T>
struct BandReduceSmem {
  cplx<T>* V;    // [n - b][b + 1]  the panel, then its reflectors
  cplx<T>* Y;    // [n - b][b + 1]  A22 V, then Y, then Z
  cplx<T>* Tm;   // [b][b]
  cplx<T>* W;    // [b][b]
  cplx<T>* M;    // [b][b]
  cplx<T>* tau;  // [b]
  T* red;        // [2 * kRedStride]
};

// --------------------------------------------------------------------------
// Stage 1.  A (n x n complex, row-major, global) receives the Hermitian
// matrix of H's lower triangle (Re of the diagonal), as zheev('L') reads a
// row-major H, and is reduced in place to band form: entries more than b
// below or above the diagonal are zero on return.  V1, T1: the panels.
// Needs nt >= 16 (b - 1) (a group of 16 threads per panel column) and nt a
// multiple of 16.
// --------------------------------------------------------------------------
template <typename T, typename Team>
RTB_HD void band_reduce(Team& tm, int n, const double* H, cplx<T>* A, cplx<T>* V1, cplx<T>* T1,
                        BandReduceSmem<T> sm) {
  constexpr int b = kBandB;
  constexpr int ld = kBandLd;  // team-memory row stride of the panels (padded: no bank conflicts)
  const int nt = tm.nt, id = tm.id;
  for (size_t e = id; e < static_cast<size_t>(n) * n; e += nt) {
    const int r = static_cast<int>(e / n), c = static_cast<int>(e % n);
    cplx<T> x;
    if (r > c) {
      x = {static_cast<T>(H[e * 2]), static_cast<T>(H[e * 2 + 1])};
    } else if (r < c) {
      const size_t t = static_cast<size_t>(c) * n + r;
      x = {static_cast<T>(H[t * 2]), -static_cast<T>(H[t * 2 + 1])};
    } else {
      x = {static_cast<T>(H[e * 2]), T(0)};
    }
    A[e] = x;
  }
  tm.sync();
  const int P = n > b ? (n - 1) / b : 0;
  int red_sel = 0;
  for (int p = 0; p < P; ++p) {
    const int c0 = p * b, r0 = c0 + b, m = n - r0;
    cplx<T>* V1p = V1 + static_cast<size_t>(p) * n * b;
    // the panel A[r0 .. n-1][c0 .. c0+b-1] into team memory
    for (int e = id; e < m * b; e += nt) sm.V[(e / b) * ld + e % b] = A[static_cast<size_t>(r0 + e / b) * n + c0 + e % b];
    tm.sync();
    // QR of the panel (zgeqr2): reflector c annihilates V[c+1 .., c]; the
    // remaining columns get H_c^H = I - conj(tau) v v^H.  Thread group g of
    // 16 threads owns column c + 1 + g for the updates.
    const int g = id / 16, l = id % 16;
    for (int c = 0; c < b; ++c) {
      if (c >= m) {  // no rows left under this column
        if (id == 0) sm.tau[c] = {T(0), T(0)};
        continue;
      }
      T part = T(0);
      for (int i = c + 1 + id; i < m; i += nt) {
        const cplx<T> x = sm.V[i * ld + c];
        part += x.re * x.re + x.im * x.im;
      }
      T* red = sm.red + red_sel * Team::kRedStride;
      red_sel ^= 1;
      const T xn2 = tm.reduce_sum_ns(part, red);
      const Refl<T> h = make_reflector(sm.V[c * ld + c], xn2);
      tm.sync();  // every thread has read V[c][c]
      for (int i = c + 1 + id; i < m; i += nt) sm.V[i * ld + c] = cmul(sm.V[i * ld + c], h.scale);
      if (id == 0) {
        sm.V[c * ld + c] = {h.beta, T(0)};
        sm.tau[c] = h.tau;
      }
      tm.sync();
      const int c2 = c + 1 + g;
      const bool live = (h.tau.re != T(0) || h.tau.im != T(0)) && c2 < b;
      cplx<T> w = {T(0), T(0)};
      if (live) {
        for (int i = c + l; i < m; i += 16) {
          const cplx<T> v = i == c ? cplx<T>{T(1), T(0)} : sm.V[i * ld + c];
          cfmac(w, v, sm.V[i * ld + c2]);
        }
      }
      w.re = tm.template group_sum<16>(w.re);
      w.im = tm.template group_sum<16>(w.im);
      if (live) {
        const cplx<T> f = cmul(conjc(h.tau), w);  // conj(tau) v^H x
        for (int i = c + l; i < m; i += 16) {
          const cplx<T> v = i == c ? cplx<T>{T(1), T(0)} : sm.V[i * ld + c];
          const cplx<T> d = cmul(v, f);
          sm.V[i * ld + c2].re -= d.re;
          sm.V[i * ld + c2].im -= d.im;
        }
      }
      tm.sync();
    }
    // T of the compact WY form (zlarft, forward, columnwise):
    // T[c][c] = tau_c, T[0:c, c] = -tau_c T[0:c, 0:c] V[:, 0:c]^H v_c.
    // First W = V^H V (strictly upper part), one thread per pair.
    auto vat = [&](int i, int c) -> cplx<T> {
      return i < c ? cplx<T>{T(0), T(0)} : (i == c ? cplx<T>{T(1), T(0)} : sm.V[i * ld + c]);
    };
    for (int q = id; q < b * b; q += nt) {
      const int a = q / b, c = q % b;
      cplx<T> s = {T(0), T(0)};
      if (a < c)
        for (int i = c; i < m; ++i) cfmac(s, vat(i, a), vat(i, c));
      sm.W[q] = s;
    }
    tm.sync();
    if (id == 0) {
      for (int c = 0; c < b; ++c) {
        const cplx<T> tc = sm.tau[c];
        sm.Tm[c * b + c] = tc;
        for (int a = 0; a < c; ++a) {
          cplx<T> s = {T(0), T(0)};
          for (int q = a; q < c; ++q) cfma(s, sm.Tm[a * b + q], sm.W[q * b + c]);
          const cplx<T> t = cmul(tc, s);
          sm.Tm[a * b + c] = {-t.re, -t.im};
        }
        for (int a = c + 1; a < b; ++a) sm.Tm[a * b + c] = {T(0), T(0)};
      }
    }
    tm.sync();
    // the panel back into A: R (upper triangle of the top b x b block) and
    // zeros below, mirrored into the upper part; V and T to the panel store
    for (int e = id; e < m * b; e += nt) {
      const int i = e / b, c = e % b;
      const cplx<T> x = i <= c ? sm.V[i * ld + c] : cplx<T>{T(0), T(0)};
      A[static_cast<size_t>(r0 + i) * n + c0 + c] = x;
      A[static_cast<size_t>(c0 + c) * n + r0 + i] = conjc(x);
      V1p[e] = vat(i, c);
    }
    for (int q = id; q < b * b; q += nt) T1[static_cast<size_t>(p) * b * b + q] = sm.Tm[q];
    tm.sync();
    // V as a full m x b block (unit diagonal, zeros above) for the products
    for (int e = id; e < m * b; e += nt) sm.V[(e / b) * ld + e % b] = vat(e / b, e % b);
    tm.sync();
    // Y = A22 V: row i of Y from column i of A22 (Hermitian: conj(A22[k][i])),
    // read across threads i; all b columns per thread
    for (int i = id; i < m; i += nt) {
      cplx<T> acc[b];
#pragma unroll
      for (int c = 0; c < b; ++c) acc[c] = {T(0), T(0)};
      const cplx<T>* col = A + static_cast<size_t>(r0) * n + r0 + i;
      for (int k = 0; k < m; ++k) {
        const cplx<T> a = col[static_cast<size_t>(k) * n];
#pragma unroll
        for (int c = 0; c < b; ++c) cfmac(acc[c], a, sm.V[k * ld + c]);
      }
      // Y = (A22 V) T, T upper triangular
#pragma unroll
      for (int c = b - 1; c >= 0; --c) {
        cplx<T> s = {T(0), T(0)};
#pragma unroll
        for (int a = 0; a <= c; ++a) cfma(s, acc[a], sm.Tm[a * b + c]);
        sm.Y[i * ld + c] = s;
      }
    }
    tm.sync();
    // W = V^H Y, then M = T^H W = X^H A22 X (X = V T)
    for (int q = id; q < b * b; q += nt) {
      const int a = q / b, c = q % b;
      cplx<T> s = {T(0), T(0)};
      for (int i = a; i < m; ++i) cfmac(s, sm.V[i * ld + a], sm.Y[i * ld + c]);
      sm.W[q] = s;
    }
    tm.sync();
    for (int q = id; q < b * b; q += nt) {
      const int a = q / b, c = q % b;
      cplx<T> s = {T(0), T(0)};
      for (int r = 0; r <= a; ++r) cfmac(s, sm.Tm[r * b + a], sm.W[r * b + c]);
      sm.M[q] = s;
    }
    tm.sync();
    // Z = Y - 1/2 V M (in place)
    for (int e = id; e < m * b; e += nt) {
      const int i = e / b, c = e % b;
      cplx<T> s = {T(0), T(0)};
      for (int a = 0; a < b; ++a) cfma(s, sm.V[i * ld + a], sm.M[a * b + c]);
      sm.Y[i * ld + c].re -= T(0.5) * s.re;
      sm.Y[i * ld + c].im -= T(0.5) * s.im;
    }
    tm.sync();
    // A22 -= V Z^H + Z V^H, threads along the columns j (coalesced rows)
    for (int j = id; j < m; j += nt) {
      cplx<T> vj[b], zj[b];
#pragma unroll
      for (int c = 0; c < b; ++c) {
        vj[c] = conjc(sm.V[j * ld + c]);
        zj[c] = conjc(sm.Y[j * ld + c]);
      }
      for (int i = 0; i < m; ++i) {
        cplx<T> s = {T(0), T(0)};
#pragma unroll
        for (int c = 0; c < b; ++c) {
          cfma(s, sm.V[i * ld + c], zj[c]);
          cfma(s, sm.Y[i * ld + c], vj[c]);
        }
        cplx<T>& a = A[static_cast<size_t>(r0 + i) * n + r0 + j];
        a.re -= s.re;
        a.im -= s.im;
      }
    }
    tm.sync();
  }
}

// --------------------------------------------------------------------------
// Stage 2.  A (band form, full Hermitian storage) is reduced in place to a
// real symmetric tridiagonal: D[i] = A[i][i], E[i] = A[i+1][i] (E[n-1] = 0).
// Sweep i, step k: the reflector on rows [r0, r1] (r0 = i + 1 + k b) that
// annihilates column `col` below row r0 (col = i for k = 0, else the first
// row of the previous step) is applied from the left to columns
// [col, r1 + b] and from the right to rows [col, r1 + b]; stored in V2, tau2.
// hv: team memory of b complex.  nt a multiple of 16 (the right
// application takes 16 threads per row).
// --------------------------------------------------------------------------
template <typename T, typename Team>
RTB_HD void band_chase(Team& tm, int n, cplx<T>* A, T* D, T* E, cplx<T>* V2, cplx<T>* tau2, cplx<T>* hv, T* red) {
  constexpr int b = kBandB;
  const int nt = tm.nt, id = tm.id;
  const BandLayout L = band_layout(n);
  int red_sel = 0;
  for (int i = 0; i + 1 < n; ++i) {
    int col = i;
    for (int k = 0;; ++k) {
      const int r0 = i + 1 + k * b;
      if (r0 > n - 1) break;
      const int r1 = r0 + b - 1 < n - 1 ? r0 + b - 1 : n - 1;
      const int len = r1 - r0 + 1;
      // reflector from A[r0 .. r1][col]
      T part = T(0);
      for (int q = 1 + id; q < len; q += nt) {
        const cplx<T> x = A[static_cast<size_t>(r0 + q) * n + col];
        part += x.re * x.re + x.im * x.im;
      }
      T* rd = red + red_sel * Team::kRedStride;
      red_sel ^= 1;
      const T xn2 = tm.reduce_sum_ns(part, rd);
      const Refl<T> h = make_reflector(A[static_cast<size_t>(r0) * n + col], xn2);
      if (id < b) {
        cplx<T> v = {T(0), T(0)};
        if (id == 0) {
          v = {T(1), T(0)};
        } else if (id < len) {
          v = cmul(A[static_cast<size_t>(r0 + id) * n + col], h.scale);
        }
        hv[id] = v;
        V2[(static_cast<size_t>(i) * L.K + k) * b + id] = v;
      }
      if (id == 0) tau2[static_cast<size_t>(i) * L.K + k] = h.tau;
      tm.sync();
      const bool apply = h.tau.re != T(0) || h.tau.im != T(0);
      const int hi = r1 + b < n - 1 ? r1 + b : n - 1;
      if (apply) {
        // left: rows [r0, r1], columns [col, hi]: x <- x - conj(tau) v (v^H x)
        const cplx<T> ct = conjc(h.tau);
        for (int c = col + id; c <= hi; c += nt) {
          cplx<T> w = {T(0), T(0)};
          for (int q = 0; q < len; ++q) cfmac(w, hv[q], A[static_cast<size_t>(r0 + q) * n + c]);
          const cplx<T> f = cmul(ct, w);
          for (int q = 0; q < len; ++q) {
            const cplx<T> d = cmul(hv[q], f);
            cplx<T>& a = A[static_cast<size_t>(r0 + q) * n + c];
            a.re -= d.re;
            a.im -= d.im;
          }
        }
      }
      tm.sync();
      if (apply) {
        // right: rows [col, hi], columns [r0, r1]: y <- y - tau (y v) v^H;
        // 16 threads per row, one column each (contiguous in the row), the
        // row's dot product by a group sum
        const int g = id / 16, q = id % 16, rows_per = nt / 16;
        const cplx<T> vq = q < len ? hv[q] : cplx<T>{T(0), T(0)};
        for (int r0w = col; r0w <= hi; r0w += rows_per) {  // same trip count in every thread
          const int r = r0w + g;
          const bool in = r <= hi && q < len;
          cplx<T> a = {T(0), T(0)};
          if (in) a = A[static_cast<size_t>(r) * n + r0 + q];
          cplx<T> w = cmul(a, vq);
          w.re = tm.template group_sum<16>(w.re);
          w.im = tm.template group_sum<16>(w.im);
          if (in) {
            const cplx<T> d = cmul(cmul(w, h.tau), conjc(vq));
            A[static_cast<size_t>(r) * n + r0 + q] = {a.re - d.re, a.im - d.im};
          }
        }
      }
      tm.sync();
      col = r0;
    }
  }
  for (int j = id; j < n; j += nt) {
    D[j] = A[static_cast<size_t>(j) * n + j].re;
    E[j] = j + 1 < n ? A[static_cast<size_t>(j + 1) * n + j].re : T(0);
  }
  tm.sync();
}

// Stage 2 with the band region in team memory (the FP32 route: at n = 256 it
// is 135 KB): AB[r][d] holds A[r][r + d - 2b] for |d - 2b| <= 2b (the chase
// touches |c - r| <= 2b - 1), rows kChaseLd = 4b + 2 apart so the 16 rows a
// reflector spans fall on distinct banks.  The steps and formulas of
// band_chase; the team (nt a multiple of 16) works in groups of 16 threads:
// every group forms the reflector from the column (thread q of a group holds
// v_q), then for the left application each group takes a column of
// [col, hi] and its 16 rows, for the right one a row of [col, hi] and its 16
// columns, the dot products by group sums.  A[r][c] for |c - r| <= 2b is read
// from A as band_reduce left it.
constexpr int kChaseLd = 4 * kBandB + 2;
RTB_HD inline size_t band_chase_team_elems(int n) { return static_cast<size_t>(n) * kChaseLd; }

template <typename T, typename Team>
RTB_HD void band_chase_lds(Team& tm, int n, const cplx<T>* A, T* D, T* E, cplx<T>* V2, cplx<T>* tau2,
                           cplx<T>* AB) {
  constexpr int b = kBandB, W2 = 2 * kBandB, ND = 4 * kBandB + 1;
  const int nt = tm.nt, id = tm.id;
  const int g = id / 16, q = id % 16, groups = nt / 16;
  const BandLayout L = band_layout(n);
  auto at = [&](int r, int c) -> cplx<T>& { return AB[r * kChaseLd + (c - r + W2)]; };
  for (int e = id; e < n * ND; e += nt) {
    const int r = e / ND, d = e - r * ND, c = r + d - W2;
    AB[r * kChaseLd + d] = c >= 0 && c < n ? A[static_cast<size_t>(r) * n + c] : cplx<T>{T(0), T(0)};
  }
  tm.sync();
  for (int i = 0; i + 1 < n; ++i) {
    int col = i;
    for (int k = 0;; ++k) {
      const int r0 = i + 1 + k * b;
      if (r0 > n - 1) break;
      const int r1 = r0 + b - 1 < n - 1 ? r0 + b - 1 : n - 1;
      const int len = r1 - r0 + 1;
      const int hi = r1 + b < n - 1 ? r1 + b : n - 1;
      // reflector from A[r0 .. r1][col], formed by every group
      const cplx<T> x = q < len ? at(r0 + q, col) : cplx<T>{T(0), T(0)};
      const T part = q >= 1 ? x.re * x.re + x.im * x.im : T(0);
      const T xn2 = tm.template group_sum<16>(part);
      const Refl<T> h = make_reflector(at(r0, col), xn2);
      cplx<T> vq = {T(0), T(0)};
      if (q == 0)
        vq = {T(1), T(0)};
      else if (q < len)
        vq = cmul(x, h.scale);
      if (id < b) V2[(static_cast<size_t>(i) * L.K + k) * b + id] = vq;
      if (id == 0) tau2[static_cast<size_t>(i) * L.K + k] = h.tau;
      tm.sync();  // every group has read the column before it changes
      const bool apply = h.tau.re != T(0) || h.tau.im != T(0);
      if (apply) {
        // left: rows [r0, r1], columns [col, hi]: x <- x - conj(tau) v (v^H x)
        const cplx<T> ct = conjc(h.tau);
        for (int c0 = col; c0 <= hi; c0 += groups) {  // same trip count in every thread
          const int c = c0 + g;
          const bool in = c <= hi && q < len;
          cplx<T> a = {T(0), T(0)};
          if (in) a = at(r0 + q, c);
          cplx<T> w = {T(0), T(0)};
          cfmac(w, vq, a);
          w.re = tm.template group_sum<16>(w.re);
          w.im = tm.template group_sum<16>(w.im);
          if (in) {
            const cplx<T> d = cmul(vq, cmul(ct, w));
            at(r0 + q, c) = {a.re - d.re, a.im - d.im};
          }
        }
      }
      tm.sync();
      if (apply) {
        // right: rows [col, hi], columns [r0, r1]: y <- y - tau (y v) v^H
        for (int rr = col; rr <= hi; rr += groups) {  // same trip count in every thread
          const int r = rr + g;
          const bool in = r <= hi && q < len;
          cplx<T> a = {T(0), T(0)};
          if (in) a = at(r, r0 + q);
          cplx<T> w = cmul(a, vq);
          w.re = tm.template group_sum<16>(w.re);
          w.im = tm.template group_sum<16>(w.im);
          if (in) {
            const cplx<T> d = cmul(cmul(w, h.tau), conjc(vq));
            at(r, r0 + q) = {a.re - d.re, a.im - d.im};
          }
        }
      }
      tm.sync();
      col = r0;
    }
  }
  for (int j = id; j < n; j += nt) {
    D[j] = at(j, j).re;
    E[j] = j + 1 < n ? at(j + 1, j).re : T(0);
  }
  tm.sync();
}

// Column j of the eigenvectors of T (Y[i * ldy + j], unit) from its twisted
// factorisation (Ts = [D+ | D-], [2][n][ldy], as twisted_all stores them);
// twist < 0: a cluster member, already in Y.
template <typename T>
RTB_HD void twisted_column(int n, const T* e, const T* Ts, int ldy, int j, T twist, T* Y) {
  const int r = static_cast<int>(twist);
  if (r < 0) return;
  const T* Dp = Ts + j;
  const T* Dm = Ts + static_cast<size_t>(n) * ldy + j;
  T nrm2 = T(1), v = T(1);
  Y[static_cast<size_t>(r) * ldy + j] = T(1);
  // unrolled: the factor loads do not depend on v, so several are in flight
#pragma unroll 8
  for (int k = r - 1; k >= 0; --k) {
    v = -e[k] * recip(Dp[static_cast<size_t>(k) * ldy]) * v;
    Y[static_cast<size_t>(k) * ldy + j] = v;
    nrm2 += v * v;
  }
  v = T(1);
#pragma unroll 8
  for (int k = r + 1; k < n; ++k) {
    v = -e[k - 1] * recip(Dm[static_cast<size_t>(k) * ldy]) * v;
    Y[static_cast<size_t>(k) * ldy + j] = v;
    nrm2 += v * v;
  }
  const T inv = T(1) / sqrt(nrm2);
#pragma unroll 8
  for (int k = 0; k < n; ++k) Y[static_cast<size_t>(k) * ldy + j] *= inv;
}

// --------------------------------------------------------------------------
// Stage 4.  Columns [j0, j0 + W) of Z = Q1 Q2 Y for the real eigenvectors
// Y of T (Y[i * ldy + j]).  Team of W * S threads: thread (c, s) =
// (id / S, id % S) takes rows s, s + S, .. of each reflector's rows for
// column j0 + c; the S threads of a column are consecutive ids (one wave).
// Ys: team memory of band_backtrans_team_elems.  Z out as in backtrans_reg: rows_out = 0
// Zout[r][c], rows_out = 1 Zout[c][r]; gdiag (optional): squared norms of
// the stored columns (FP64).
// --------------------------------------------------------------------------
template <typename T, int W, int S, typename Team, typename TO>
RTB_HD void band_backtrans(Team& tm, int n, int j0, const T* Y, int ldy, const cplx<T>* V1, const cplx<T>* T1,
                           const cplx<T>* V2, const cplx<T>* tau2, cplx<T>* Ys, TO* Zout, int rows_out,
                           double* gdiag) {
  constexpr int b = kBandB;
  constexpr int Wp = W + BandBtStage<T>::pad;  // padded row stride of Ys
  static_assert(b % S == 0 && (S == 8 || S == 16), "rows per thread");
  const BandLayout L = band_layout(n);
  const int c = tm.id / S, s = tm.id % S;
  const int j = j0 + c;
  const bool live = c < W && j < n;
  for (int e = tm.id; e < n * W; e += tm.nt) {
    const int i = e / W, cc = e % W;
    Ys[i * Wp + cc] = {j0 + cc < n ? Y[static_cast<size_t>(i) * ldy + j0 + cc] : T(0), T(0)};
  }
  tm.sync();
  // The reflectors stream through team memory: a sequence of chunks, each
  // loaded by the whole team with coalesced reads into registers while the
  // previous chunk is applied from team memory (one memory latency per chunk
  // rather than one per reflector).  Chunks: first the chasing reflectors
  // (Q2) in batches of NB slots, slot r = i K + k holding sweep i, step k
  // (k < nsteps(i) used), taken in descending order (last sweep first, last
  // step first); then per panel of Q1 (last first) its rows in chunks of NB,
  // twice: for w = V^H Z, then for Z <- Z - V (T w).  Thread (c, s) keeps
  // rows s, s + S, .. of every reflector and panel in increasing order, so
  // each column's arithmetic is the same as applying them one by one.
  constexpr int NB = BandBtStage<T>::NB;
  constexpr int NT = W * S;  // team size
  constexpr int SL = b + 1;                       // staged slot: b entries, then tau
  constexpr int EV = (NB * SL + NT - 1) / NT;      // staging registers per thread
  constexpr int Q = b / S;                        // rows of a reflector per thread
  cplx<T>* Sv = Ys + static_cast<size_t>(n) * Wp;  // [2][NB][SL]
  auto nsteps = [&](int i) { return (n - 1 - i + b - 1) / b; };
  // whole sweeps per chasing chunk (a sweep's K slots are contiguous)
  const int spc = NB / L.K > 1 ? NB / L.K : 1;
  const int nq2 = n >= 2 ? (n - 1 + spc - 1) / spc : 0;
  auto prow = [&](int p) { return n - (p + 1) * b; };  // rows of panel p
  auto pch = [&](int p) { return (prow(p) + NB - 1) / NB; };
  int nchunks = nq2;
  for (int p = 0; p < L.P; ++p) nchunks += 2 * pch(p);
  struct Chunk {
    int kind;      // 0: chasing slots [lo, hi) (whole sweeps); 1, 2: panel p rows [lo, hi), first / second pass
    int p, lo, hi;
    bool first, last;  // first / last chunk of its pass
  };
  auto chunk = [&](int t) {
    Chunk ch{0, 0, 0, 0, false, false};
    if (t < nq2) {
      const int ihi = n - 1 - t * spc, ilo = ihi > spc ? ihi - spc : 0;
      ch.lo = ilo * L.K;
      ch.hi = ihi * L.K;
      return ch;
    }
    t -= nq2;
    for (int p = L.P - 1; p >= 0; --p) {
      const int c2 = pch(p);
      if (t < 2 * c2) {
        const int ci = t < c2 ? t : t - c2;
        ch.kind = t < c2 ? 1 : 2;
        ch.p = p;
        ch.lo = ci * NB;
        ch.hi = ch.lo + NB < prow(p) ? ch.lo + NB : prow(p);
        ch.first = ci == 0;
        ch.last = ci == c2 - 1;
        return ch;
      }
      t -= 2 * c2;
    }
    return ch;
  };
  cplx<T> rv[EV];
  auto load = [&](int t) {
    const Chunk ch = chunk(t);
    const cplx<T>* v = ch.kind == 0 ? V2 + static_cast<size_t>(ch.lo) * b
                                    : V1 + (static_cast<size_t>(ch.p) * n + ch.lo) * b;
    const int rows = ch.hi - ch.lo;
#pragma unroll
    for (int u = 0; u < EV; ++u) {
      const int e = tm.id + u * NT;
      const int r = e / SL, q = e - r * SL;
      cplx<T> x = {T(0), T(0)};
      if (r < rows) {
        if (q < b)
          x = v[r * b + q];
        else if (ch.kind == 0)
          x = tau2[ch.lo + r];
      }
      rv[u] = x;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int u = 0; u < EV; ++u) {
      const int e = tm.id + u * NT;
      if (e < NB * SL) Sv[buf * NB * SL + e] = rv[u];
    }
  };
  cplx<T> w[b];  // a panel's V^H Z, then T (V^H Z), for this column
  if (nchunks > 0) {
    load(0);
    store(0);
  }
  tm.sync();
  for (int t = 0; t < nchunks; ++t) {
    const int buf = t & 1;
    if (t + 1 < nchunks) load(t + 1);
    const Chunk ch = chunk(t);
    const cplx<T>* sv = Sv + buf * NB * SL;
    if (ch.kind == 0) {
      // sweep i's windows [i + 1 + k b, i + 1 + (k + 1) b) are disjoint, so
      // its reflectors commute: all window sums, then all updates (K
      // independent chains per thread instead of one)
      constexpr int KM = (kMaxBandN - 1 + b - 1) / b;  // steps per sweep at most
      const int ilo = ch.lo / L.K, ihi = ch.hi / L.K;
      for (int i = ihi - 1; i >= ilo; --i) {
        const int ki = nsteps(i);
        const cplx<T>* vs = sv + (i - ilo) * L.K * SL;  // slot (i, k) at vs + k SL
        cplx<T> x[KM];
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          x[k] = {T(0), T(0)};
          if (k < ki) {
            const cplx<T>* v = vs + k * SL;
            const int r0 = i + 1 + k * b;
            const int len = r0 + b - 1 < n - 1 ? b : n - r0;
#pragma unroll
            for (int u = 0; u < Q; ++u) {
              const int q = s + S * u;
              if (q < len) cfmac(x[k], v[q], Ys[(r0 + q) * Wp + c]);
            }
          }
        }
#pragma unroll
        for (int k = 0; k < KM; ++k)
          if (k < ki) {
            x[k].re = tm.template group_sum<S>(x[k].re);
            x[k].im = tm.template group_sum<S>(x[k].im);
          }
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          if (k < ki) {
            const cplx<T>* v = vs + k * SL;
            const cplx<T> tau = v[b];
            if (tau.re != T(0) || tau.im != T(0)) {  // team-uniform
              const cplx<T> f = cmul(tau, x[k]);
              const int r0 = i + 1 + k * b;
              const int len = r0 + b - 1 < n - 1 ? b : n - r0;
#pragma unroll
              for (int u = 0; u < Q; ++u) {
                const int q = s + S * u;
                if (q < len) {
                  const cplx<T> d = cmul(v[q], f);
                  Ys[(r0 + q) * Wp + c].re -= d.re;
                  Ys[(r0 + q) * Wp + c].im -= d.im;
                }
              }
            }
          }
        }
        tm.wave_sync();
      }
    } else {
      // Q1 panel p: Z <- Z - V (T (V^H Z)) on rows [r0, n)
      const int r0 = (ch.p + 1) * b;
      if (ch.kind == 1) {
        if (ch.first) {
#pragma unroll
          for (int a = 0; a < b; ++a) w[a] = {T(0), T(0)};
        }
        for (int i = ch.lo + s; i < ch.hi; i += S) {  // NB is a multiple of S: rows s mod S
          const cplx<T> z = Ys[(r0 + i) * Wp + c];
          const cplx<T>* vr = sv + (i - ch.lo) * SL;
#pragma unroll
          for (int a = 0; a < b; ++a) cfmac(w[a], vr[a], z);
        }
        if (ch.last) {
#pragma unroll
          for (int a = 0; a < b; ++a) {
            w[a].re = tm.template group_sum<S>(w[a].re);
            w[a].im = tm.template group_sum<S>(w[a].im);
          }
          // w <- T w in place: row a of the upper triangular T reads w[q >= a]
          const cplx<T>* Tp = T1 + static_cast<size_t>(ch.p) * b * b;
#pragma unroll
          for (int a = 0; a < b; ++a) {
            cplx<T> x = {T(0), T(0)};
#pragma unroll
            for (int q = a; q < b; ++q) cfma(x, Tp[a * b + q], w[q]);
            w[a] = x;
          }
        }
      } else {
        for (int i = ch.lo + s; i < ch.hi; i += S) {
          const cplx<T>* vr = sv + (i - ch.lo) * SL;
          cplx<T> d = {T(0), T(0)};
#pragma unroll
          for (int a = 0; a < b; ++a) cfma(d, vr[a], w[a]);
          Ys[(r0 + i) * Wp + c].re -= d.re;
          Ys[(r0 + i) * Wp + c].im -= d.im;
        }
        tm.wave_sync();
      }
    }
    if (t + 1 < nchunks) store(buf ^ 1);
    tm.sync();
  }
  // out
  double g2 = 0.0;
  if (live) {
    for (int i = s; i < n; i += S) {
      const cplx<T> z = Ys[i * Wp + c];
      const size_t o = rows_out ? static_cast<size_t>(j) * n + i : static_cast<size_t>(i) * n + j;
      Zout[o * 2] = static_cast<TO>(z.re);
      Zout[o * 2 + 1] = static_cast<TO>(z.im);
      g2 += static_cast<double>(z.re) * z.re + static_cast<double>(z.im) * z.im;
    }
  }
  g2 = tm.template group_sum<S>(g2);
  if (gdiag && live && s == 0) gdiag[j] = g2;
}

}  // namespace rtb
