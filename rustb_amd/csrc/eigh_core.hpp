// Batched Hermitian eigensolver core: one matrix per team (workgroup).
//
// Algorithm (the LAPACK zheevx route, restated for one workgroup per matrix):
//   1. Householder reduction of the Hermitian matrix to a real symmetric
//      tridiagonal T = Q^H A Q (zhetd2 with zlarfg reflectors, lower storage),
//      the matrix held as a packed lower triangle in LDS;
//   2. eigenvalues of T by Sturm-count bisection, one thread per eigenvalue
//      (dstebz's counting rule with the pivmin guard);
//   3. eigenvectors of T by inverse iteration with a partially pivoted
//      tridiagonal LU (dlagtf/dlagts style), one thread per eigenvalue, with
//      modified Gram-Schmidt inside clusters of close eigenvalues (dstein);
//   4. back-transformation Z = Q Y by the stored reflectors (zunmtr).
//
// Every routine is written against a `Team` (id, nt, sync(), reduce_sum()),
// so the same source runs on the GPU (a workgroup) and on the CPU (threads
// with a barrier) for host-side verification.  No wave intrinsics appear
// here; the GPU Team implements reduce_sum with them.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define RTB_HD __host__ __device__
#else
#define RTB_HD
#endif

namespace rtb {

#if !defined(__HIP_DEVICE_COMPILE__)
using std::fabs;
using std::fmax;
using std::fmin;
using std::sqrt;
#endif

template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
RTB_HD inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
RTB_HD inline cplx<T> cmulc(cplx<T> a, cplx<T> b) {  // conj(a) * b
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
template <typename T>
RTB_HD inline cplx<T> conjc(cplx<T> a) {
  return {a.re, -a.im};
}

template <typename T>
struct Lim;
template <>
struct Lim<double> {
  RTB_HD static constexpr double eps() { return DBL_EPSILON * 0.5; }  // LAPACK dlamch('E')
  RTB_HD static constexpr double safmin() { return DBL_MIN; }
};
template <>
struct Lim<float> {
  RTB_HD static constexpr float eps() { return FLT_EPSILON * 0.5f; }
  RTB_HD static constexpr float safmin() { return FLT_MIN; }
};

// packed lower triangle index, r >= c
RTB_HD inline int pidx(int r, int c) { return r * (r + 1) / 2 + c; }

// A(r, c) of the Hermitian matrix stored as a packed lower triangle
template <typename T>
RTB_HD inline cplx<T> herm_at(const cplx<T>* P, int r, int c) {
  return r >= c ? P[pidx(r, c)] : conjc(P[pidx(c, r)]);
}

// Shared-memory layout of one team's eigensolver workspace (element counts).
struct EighSmem {
  static RTB_HD size_t bytes(int n, int nt, size_t real_size) {
    size_t cx = 2 * real_size;
    size_t packed = static_cast<size_t>(n) * (n + 1) / 2;
    return packed * cx + 2 * static_cast<size_t>(n) * cx  // P, v, w
           + 4 * static_cast<size_t>(n) * real_size        // d, e, e2, E
           + static_cast<size_t>(nt + 64) * real_size;     // reduction scratch
  }
};

// --------------------------------------------------------------------------
// Stage 1: Householder tridiagonalization (zhetd2, UPLO = 'L').
//   P    : packed lower triangle of A (team-shared), destroyed
//   v, w : team-shared complex vectors of length n
//   d, e : team-shared outputs (T diagonal, sub-diagonal e[0..n-2])
//   red  : team-shared reduction scratch
//   Vg   : global [n-1][n] reflector vectors (v[r] = 0 for r <= j, v[j+1] = 1)
//   taug : global [n-1] reflector scalars
// Q = H(0) H(1) ... H(n-2), H(j) = I - tau_j v_j v_j^H, and Q^H A Q = T.
// --------------------------------------------------------------------------
template <typename T, typename Team>
RTB_HD void tridiagonalize(Team& tm, int n, cplx<T>* P, cplx<T>* v, cplx<T>* w, T* d, T* e,
                           T* red, cplx<T>* Vg, cplx<T>* taug) {
  // Four barriers per column: the norm reduction, the reflector broadcast,
  // the dot reduction (which also publishes p), and the end of the rank-2
  // update.  Each reduction has its own scratch, and every scratch is
  // rewritten only after at least one further barrier.
  const int id = tm.id, nt = tm.nt;
  T* red_norm = red;
  T* red_dot = red + Team::kRedStride;
  T* red_pair = red + 3 * Team::kRedStride;
  for (int j = 0; j + 1 < n; ++j) {
    // zlarfg on alpha = A(j+1, j), x = A(j+2:n, j)
    T part = T(0);
    for (int r = j + 2 + id; r < n; r += nt) {
      cplx<T> x = P[pidx(r, j)];
      part += x.re * x.re + x.im * x.im;
    }
    const T xnorm2 = tm.reduce_sum_ns(part, red_norm);
    const cplx<T> alpha = P[pidx(j + 1, j)];
    cplx<T> tau = {T(0), T(0)};
    cplx<T> scale = {T(0), T(0)};
    T beta = alpha.re;
    if (!(xnorm2 == T(0) && alpha.im == T(0))) {
      T nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm2);
      beta = alpha.re >= T(0) ? -nrm : nrm;
      tau.re = (beta - alpha.re) / beta;
      tau.im = -alpha.im / beta;
      // scale = 1 / (alpha - beta)
      T ar = alpha.re - beta, ai = alpha.im;
      T den = ar * ar + ai * ai;
      scale.re = ar / den;
      scale.im = -ai / den;
    }
    // reflector vector (rows > j) into v and global storage
    for (int r = id; r < n; r += nt) {
      cplx<T> vr = {T(0), T(0)};
      if (r == j + 1) {
        vr.re = T(1);
      } else if (r > j + 1) {
        vr = cmul(P[pidx(r, j)], scale);
      }
      v[r] = vr;
      if (Vg) Vg[static_cast<size_t>(j) * n + r] = vr;
    }
    if (id == 0) {
      d[j] = P[pidx(j, j)].re;
      e[j] = beta;
      if (taug) taug[j] = tau;
    }
    tm.sync();
    if (tau.re != T(0) || tau.im != T(0)) {
      // p = tau * A22 v (rows j+1..n-1) into w, and the partial of p^H v.
      // With Team::kPairs two threads share a row (even/odd columns) and
      // combine through pair_sum; the trip count is uniform across the team.
      cplx<T> dot = {T(0), T(0)};
      const int half = Team::kPairs ? (id & 1) : 0;
      const int cstep = Team::kPairs ? 2 : 1;
      const int rid = Team::kPairs ? (id >> 1) : id;
      const int rnt = Team::kPairs ? (nt >> 1) : nt;
      const int m = n - j - 1;
      for (int base = 0; base < m; base += rnt) {
        const int r = j + 1 + base + rid;
        const bool act = (base + rid < m) && (rid < rnt);
        cplx<T> s = {T(0), T(0)};
        if (act) {
          const cplx<T>* Prow = P + pidx(r, 0);
          for (int c = j + 1 + half; c <= r; c += cstep) {
            cplx<T> a = Prow[c];
            cplx<T> vc = v[c];
            s.re += a.re * vc.re - a.im * vc.im;
            s.im += a.re * vc.im + a.im * vc.re;
          }
          // same column parity class as the first loop (c = j+1+half mod 2)
          int c = r + 1;
          if (Team::kPairs && ((c - (j + 1 + half)) & 1) != 0) ++c;
          for (; c < n; c += cstep) {
            cplx<T> a = P[pidx(c, r)];  // A(r,c) = conj(A(c,r))
            cplx<T> vc = v[c];
            s.re += a.re * vc.re + a.im * vc.im;
            s.im += a.re * vc.im - a.im * vc.re;
          }
        }
        if (Team::kPairs) {
          s.re = tm.pair_sum(s.re, red_pair);
          s.im = tm.pair_sum(s.im, red_pair);
        }
        if (act && half == 0) {
          cplx<T> p = cmul(tau, s);
          w[r] = p;
          cplx<T> pv = cmulc(p, v[r]);
          dot.re += pv.re;
          dot.im += pv.im;
        }
      }
      T dre = dot.re, dim_ = dot.im;
      tm.reduce_sum2_ns(dre, dim_, red_dot);
      // alpha2 = -1/2 tau (p^H v); the update uses w = p + alpha2 v on the fly
      cplx<T> a2 = cmul(tau, cplx<T>{dre, dim_});
      a2.re *= T(-0.5);
      a2.im *= T(-0.5);
      // A22 -= v w^H + w v^H on the packed lower triangle (rows j+1..n-1),
      // row index tracked incrementally along each thread's strided walk
      const int f0 = pidx(j + 1, 0), f1 = n * (n + 1) / 2;
      int r = j + 1, rs = f0;  // current row and its packed start
      cplx<T> vr = v[r], wr = w[r];
      {
        cplx<T> t = cmul(a2, vr);
        wr.re += t.re;
        wr.im += t.im;
      }
      for (int f = f0 + id; f < f1; f += nt) {
        bool moved = false;
        while (f >= rs + r + 1) {
          rs += r + 1;
          ++r;
          moved = true;
        }
        if (moved) {
          vr = v[r];
          wr = w[r];
          cplx<T> t = cmul(a2, vr);
          wr.re += t.re;
          wr.im += t.im;
        }
        const int c = f - rs;
        if (c <= j) continue;
        cplx<T> vc = v[c], wc = w[c];
        {
          cplx<T> t = cmul(a2, vc);
          wc.re += t.re;
          wc.im += t.im;
        }
        cplx<T> t1 = cmul(vr, conjc(wc));
        cplx<T> t2 = cmul(wr, conjc(vc));
        P[f].re -= t1.re + t2.re;
        P[f].im -= t1.im + t2.im;
      }
      tm.sync();
    }
  }
  if (id == 0) {
    d[n - 1] = P[pidx(n - 1, n - 1)].re;
    if (n >= 1) e[n - 1] = T(0);
  }
  tm.sync();
}

// --------------------------------------------------------------------------
// Stage 2: Sturm-count bisection (dstebz rule).  Thread t computes the
// eigenvalues t, t+nt, ... ascending into E.  e2[i] = e[i]^2 must be filled.
// --------------------------------------------------------------------------
// 1/q: on the GPU the hardware reciprocal refined by two Newton steps (error
// well below the bisection tolerance); on the host a plain division.
RTB_HD inline double recip(double q) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(q);
  double e = fma(-q, r, 1.0);
  r = fma(r, e, r);
  e = fma(-q, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / q;
#endif
}
RTB_HD inline float recip(float q) {
#if defined(__HIP_DEVICE_COMPILE__)
  float r = __builtin_amdgcn_rcpf(q);
  float e = fmaf(-q, r, 1.0f);
  return fmaf(r, e, r);
#else
  return 1.0f / q;
#endif
}

template <typename T>
RTB_HD inline int sturm_count(int n, const T* d, const T* e2, T x, T pivmin) {
  int cnt = 0;
  T q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < T(0);
  for (int i = 1; i < n; ++i) {
    q = d[i] - x - e2[i - 1] * recip(q);
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < T(0);
  }
  return cnt;
}

template <typename T>
RTB_HD inline void gershgorin(int n, const T* d, const T* e, T& gl, T& gu) {
  gl = d[0];
  gu = d[0];
  for (int i = 0; i < n; ++i) {
    T off = (i > 0 ? fabs(e[i - 1]) : T(0)) + (i + 1 < n ? fabs(e[i]) : T(0));
    gl = fmin(gl, d[i] - off);
    gu = fmax(gu, d[i] + off);
  }
}

template <typename T, typename Team>
RTB_HD void bisect_all(Team& tm, int n, const T* d, const T* e, T* e2, T* E) {
  for (int i = tm.id; i < n; i += tm.nt) e2[i] = (i + 1 < n) ? e[i] * e[i] : T(0);
  tm.sync();
  T gl, gu;
  gershgorin(n, d, e, gl, gu);
  T emax2 = T(0);
  for (int i = 0; i + 1 < n; ++i) emax2 = fmax(emax2, e2[i]);
  const T eps = Lim<T>::eps();
  const T pivmin = Lim<T>::safmin() * fmax(T(1), emax2);
  const T tnorm = fmax(fabs(gl), fabs(gu));
  const T widen = T(2) * eps * tnorm * n + T(4) * pivmin;
  gl -= widen;
  gu += widen;
  const T atol = eps * tnorm;
  if (n <= 2) {
    // closed forms: (a+d)/2 -+ sqrt(((a-d)/2)^2 + e^2) (exact at e.g. [[0,3],[3,0]])
    if (tm.id == 0) {
      if (n == 1) {
        E[0] = d[0];
      } else {
        T mean = T(0.5) * (d[0] + d[1]), half = T(0.5) * (d[0] - d[1]);
        T r = sqrt(half * half + e[0] * e[0]);
        E[0] = mean - r;
        E[1] = mean + r;
      }
    }
    tm.sync();
    return;
  }
  for (int j = tm.id; j < n; j += tm.nt) {
    T lo = gl, hi = gu;
    for (int it = 0; it < 200; ++it) {
      T tol = atol + T(2) * eps * fmax(fabs(lo), fabs(hi));
      if (hi - lo <= tol) break;
      T mid = T(0.5) * (lo + hi);
      if (mid <= lo || mid >= hi) break;
      if (sturm_count(n, d, e2, mid, pivmin) <= j) {
        lo = mid;
      } else {
        hi = mid;
      }
    }
    E[j] = T(0.5) * (lo + hi);
  }
  tm.sync();
}

// --------------------------------------------------------------------------
// Stage 3: inverse iteration for the eigenvectors of T (dstein-like).
//   Y  : global [n][n], Y[i*ldy + j] = component i of eigenvector j (column-
//        major by eigenvector, coalesced across threads)
//   S  : global scratch [3][n][n] (pivoted-LU U rows, same layout as Y)
// --------------------------------------------------------------------------
RTB_HD inline double rtb_hash_uniform(uint32_t j, uint32_t i) {
  uint32_t h = j * 0x9E3779B1u ^ (i + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (static_cast<double>(h) + 0.5) * (2.0 / 4294967296.0) - 1.0;  // (-1, 1)
}

// One solve (T - lam I) z = b with partial pivoting; b in Y[:, j] on entry,
// z in Y[:, j] on exit, normalised to unit 2-norm.
template <typename T>
RTB_HD void invit_solve(int n, const T* d, const T* e, T lam, T tol, T* Y, int ldy, int j,
                        T* S) {
  const size_t nn = static_cast<size_t>(n) * ldy;
  T* U0 = S;
  T* U1 = S + nn;
  T* U2 = S + 2 * nn;
#define Yat(i) Y[static_cast<size_t>(i) * ldy + j]
#define SAt(arr, i) arr[static_cast<size_t>(i) * ldy + j]
  T a = d[0] - lam;
  T bsup = n > 1 ? e[0] : T(0);
  for (int k = 0; k + 1 < n; ++k) {
    T c = e[k];
    T an = d[k + 1] - lam;
    T bn = (k + 2 < n) ? e[k + 1] : T(0);
    T bk = Yat(k), bk1 = Yat(k + 1);
    if (fabs(a) >= fabs(c)) {
      T l = (a != T(0)) ? c / a : T(0);
      SAt(U0, k) = a;
      SAt(U1, k) = bsup;
      SAt(U2, k) = T(0);
      Yat(k + 1) = bk1 - l * bk;
      a = an - l * bsup;
      bsup = bn;
    } else {
      T l = a / c;
      SAt(U0, k) = c;
      SAt(U1, k) = an;
      SAt(U2, k) = bn;
      Yat(k) = bk1;
      Yat(k + 1) = bk - l * bk1;
      a = bsup - l * an;
      bsup = -l * bn;
    }
  }
  SAt(U0, n - 1) = a;
  // back substitution with tiny pivots perturbed to +-tol
  T znrm2 = T(0);
  T z1 = T(0), z2 = T(0);  // z[k+1], z[k+2]
  for (int k = n - 1; k >= 0; --k) {
    T u0 = SAt(U0, k);
    if (fabs(u0) < tol) u0 = (u0 >= T(0)) ? tol : -tol;
    T rhs = Yat(k);
    if (k + 1 < n) rhs -= SAt(U1, k) * z1;
    if (k + 2 < n) rhs -= SAt(U2, k) * z2;
    T z = rhs / u0;
    Yat(k) = z;
    z2 = z1;
    z1 = z;
    znrm2 += z * z;
  }
  // normalise (the solution can be ~1/tol large; rescale in two steps)
  T zmax = T(0);
  for (int k = 0; k < n; ++k) zmax = fmax(zmax, fabs(Yat(k)));
  if (zmax > T(0)) {
    T s = T(0);
    for (int k = 0; k < n; ++k) {
      T t = Yat(k) / zmax;
      s += t * t;
    }
    T inv = T(1) / (zmax * sqrt(s));
    for (int k = 0; k < n; ++k) Yat(k) *= inv;
  }
#undef Yat
#undef SAt
}

// Modified Gram-Schmidt of eigenvector jj against jj0..jj-1 (columns of Y).
template <typename T>
RTB_HD void mgs_column(int n, T* Y, int ldy, int jj0, int jj) {
  for (int ii = jj0; ii < jj; ++ii) {
    T dot = T(0);
    for (int k = 0; k < n; ++k) dot += Y[static_cast<size_t>(k) * ldy + ii] * Y[static_cast<size_t>(k) * ldy + jj];
    for (int k = 0; k < n; ++k) Y[static_cast<size_t>(k) * ldy + jj] -= dot * Y[static_cast<size_t>(k) * ldy + ii];
  }
  T s = T(0);
  for (int k = 0; k < n; ++k) {
    T t = Y[static_cast<size_t>(k) * ldy + jj];
    s += t * t;
  }
  T inv = s > T(0) ? T(1) / sqrt(s) : T(0);
  for (int k = 0; k < n; ++k) Y[static_cast<size_t>(k) * ldy + jj] *= inv;
}

// Eigenvector of an isolated eigenvalue lam by a twisted factorisation
// (Parlett & Dhillon, LAA 309 (2000) 121-151; Fernando, SIMAX 18 (1997)):
// T - lam I = L+ D+ L+^T (top-down) = U- D- U-^T (bottom-up); the twist
// index r minimises |gamma_k| = |D+_k + D-_k - (d_k - lam)|, and
//   z_r = 1,  z_k = -(e_k / D+_k) z_{k+1} (k < r),
//   z_k = -(e_{k-1} / D-_k) z_{k-1} (k > r)
// solves (T - lam I) z = gamma_r e_r: one solve, no iteration, residual
// |gamma_r| / ||z||.  Pivots below pivmin are replaced by -pivmin (the Sturm
// count rule).  S: [2][n] scratch columns (D+ and D-), same layout as Y.
// z is left unnormalised in Y; the return value is 1 / ||z||.
template <typename T>
RTB_HD T twisted_vector(int n, const T* d, const T* e, T lam, T pivmin, T* Y, int ldy, int j, T* S) {
  // The recurrences are serial; the scratch reads are not, so they are
  // issued kB at a time ahead of the arithmetic (one memory latency per
  // block instead of one per step).
  constexpr int kB = 16;
  const size_t nn = static_cast<size_t>(n) * ldy;
  T* Dp = S;
  T* Dm = S + nn;
#define Yat(i) Y[static_cast<size_t>(i) * ldy + j]
#define SAt(arr, i) arr[static_cast<size_t>(i) * ldy + j]
  T dp = d[0] - lam;
  if (fabs(dp) < pivmin) dp = -pivmin;
  SAt(Dp, 0) = dp;
  for (int k = 0; k + 1 < n; ++k) {
    dp = d[k + 1] - lam - e[k] * e[k] * recip(dp);
    if (fabs(dp) < pivmin) dp = -pivmin;
    SAt(Dp, k + 1) = dp;
  }
  T dm = d[n - 1] - lam;
  if (fabs(dm) < pivmin) dm = -pivmin;
  SAt(Dm, n - 1) = dm;
  int r = n - 1;
  T gmin = fabs(dp);  // gamma_{n-1} = D+_{n-1}
  for (int k0 = n - 2; k0 >= 0; k0 -= kB) {
    T pre[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) pre[u] = k0 - u >= 0 ? SAt(Dp, k0 - u) : T(0);
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        const T dk = d[k] - lam;
        dm = dk - e[k] * e[k] * recip(dm);
        if (fabs(dm) < pivmin) dm = -pivmin;
        SAt(Dm, k) = dm;
        const T g = fabs(pre[u] + dm - dk);
        if (g < gmin) {
          gmin = g;
          r = k;
        }
      }
    }
  }
  T z = T(1), nrm2 = T(1);
  Yat(r) = z;
  for (int k0 = r - 1; k0 >= 0; k0 -= kB) {
    T pre[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) pre[u] = k0 - u >= 0 ? SAt(Dp, k0 - u) : T(1);
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        z = -e[k] * recip(pre[u]) * z;
        Yat(k) = z;
        nrm2 += z * z;
      }
    }
  }
  z = T(1);
  for (int k0 = r + 1; k0 < n; k0 += kB) {
    T pre[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) pre[u] = k0 + u < n ? SAt(Dm, k0 + u) : T(1);
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 + u;
      if (k < n) {
        z = -e[k - 1] * recip(pre[u]) * z;
        Yat(k) = z;
        nrm2 += z * z;
      }
    }
  }
#undef Yat
#undef SAt
  return T(1) / sqrt(nrm2);
}

// Eigenvectors of T: eigenvalues farther than ortol * tnorm from both
// neighbours (ortol = 1e-3, dstein's cluster threshold) by the twisted
// factorisation; members of a cluster by inverse iteration (dstein: a
// pivoted LU solve from a fixed pseudo-random start, twice, each followed
// by modified Gram-Schmidt inside the cluster).  Column j of Y times ysc[j]
// is the unit eigenvector (twisted vectors are stored unnormalised).
template <typename T, typename Team>
RTB_HD void tridiag_vectors(Team& tm, int n, const T* d, const T* e, const T* E, T* Y, int ldy,
                            T* S, T* ysc) {
  T gl, gu;
  gershgorin(n, d, e, gl, gu);
  const T tnorm = fmax(fabs(gl), fabs(gu));
  const T eps = Lim<T>::eps();
  const T tol = fmax(eps * tnorm, Lim<T>::safmin() * T(1e4));
  const T ortol = T(1e-3) * tnorm;
  T emax2 = T(0);
  for (int i = 0; i + 1 < n; ++i) emax2 = fmax(emax2, e[i] * e[i]);
  const T pivmin = Lim<T>::safmin() * fmax(T(1), emax2);
  for (int pass = 0; pass < 2; ++pass) {
    for (int j = tm.id; j < n; j += tm.nt) {
      const bool clustered = (j > 0 && E[j] - E[j - 1] <= ortol) || (j + 1 < n && E[j + 1] - E[j] <= ortol);
      if (!clustered) {
        if (pass == 0) ysc[j] = twisted_vector(n, d, e, E[j], pivmin, Y, ldy, j, S);
        continue;
      }
      ysc[j] = T(1);
      if (pass == 0) {
        for (int i = 0; i < n; ++i)
          Y[static_cast<size_t>(i) * ldy + j] = static_cast<T>(rtb_hash_uniform(j, i));
      }
      invit_solve(n, d, e, E[j], tol, Y, ldy, j, S);
    }
    tm.sync();
    // cluster heads orthogonalise their members in order
    for (int j = tm.id; j < n; j += tm.nt) {
      bool head = (j == 0) || (E[j] - E[j - 1] > ortol);
      if (!head) continue;
      int last = j;
      while (last + 1 < n && E[last + 1] - E[last] <= ortol) ++last;
      for (int jj = j + 1; jj <= last; ++jj) mgs_column(n, Y, ldy, j, jj);
    }
    tm.sync();
  }
}

// --------------------------------------------------------------------------
// Stage 4: back-transformation of a block of columns, Z = Q Z,
//   Z    : team-shared complex [n][cb] (row r, column c), holds Y on entry
//   Vg/taug : reflectors from stage 1
//   vs   : team-shared complex [2][n] (double-buffered reflector)
//   part : team-shared complex [max(nt, cb)] (row-slice partial dot products)
//   s    : team-shared complex [cb]
// Per reflector: s_c = tau sum_r conj(v_r) Z[r][c] (row slices in parallel,
// then a fixed-order sum of the slices), Z[r][c] -= v_r s_c.
// --------------------------------------------------------------------------
template <typename T, typename Team>
RTB_HD void apply_q(Team& tm, int n, int cb, cplx<T>* Z, const cplx<T>* Vg, const cplx<T>* taug, cplx<T>* vs,
                    cplx<T>* part, cplx<T>* s) {
  const int S = tm.nt / cb > 0 ? tm.nt / cb : 1;  // row slices per column
  int buf = 0;
  if (n >= 2)
    for (int r = tm.id; r < n; r += tm.nt) vs[r] = Vg[static_cast<size_t>(n - 2) * n + r];
  tm.sync();
  for (int j = n - 2; j >= 0; --j) {
    const cplx<T>* v = vs + buf * n;
    const cplx<T> tau = taug[j];
    const bool active = !(tau.re == T(0) && tau.im == T(0));
    if (active) {
      for (int q = tm.id; q < S * cb; q += tm.nt) {
        const int c = q % cb, sl = q / cb;
        cplx<T> acc = {T(0), T(0)};
        for (int r = j + 1 + sl; r < n; r += S) {
          cplx<T> t = cmulc(v[r], Z[r * cb + c]);
          acc.re += t.re;
          acc.im += t.im;
        }
        part[q] = acc;
      }
    }
    tm.sync();
    if (active) {
      for (int c = tm.id; c < cb; c += tm.nt) {
        cplx<T> acc = {T(0), T(0)};
        for (int sl = 0; sl < S; ++sl) {
          acc.re += part[sl * cb + c].re;
          acc.im += part[sl * cb + c].im;
        }
        s[c] = cmul(tau, acc);
      }
    }
    if (j >= 1)  // prefetch the next reflector into the other buffer
      for (int r = tm.id; r < n; r += tm.nt) vs[(buf ^ 1) * n + r] = Vg[static_cast<size_t>(j - 1) * n + r];
    tm.sync();
    if (active) {
      const int tot = (n - j - 1) * cb;
      for (int f = tm.id; f < tot; f += tm.nt) {
        int r = j + 1 + f / cb, c = f % cb;
        cplx<T> t = cmul(v[r], s[c]);
        Z[r * cb + c].re -= t.re;
        Z[r * cb + c].im -= t.im;
      }
    }
    tm.sync();
    buf ^= 1;
  }
}

}  // namespace rtb
