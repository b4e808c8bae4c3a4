// Batched Hermitian eigensolver core: the tridiagonal stages, one matrix per
// team (workgroup or sub-team).
//
// Algorithm (the LAPACK zheevx route for the tridiagonal part):
//   1. Householder reduction to a real symmetric tridiagonal T = Q^H A Q
//      (zhetd2 with zlarfg reflectors): tridiag_reg in eigh_reg.hpp;
//   2. eigenvalues of T by Sturm-count bisection, one thread per eigenvalue
//      (dstebz's counting rule with the pivmin guard): bisect_all;
//   3. eigenvectors of T: for an isolated eigenvalue the top-down and
//      bottom-up factorisations of T - lam I and the twist index (the vector
//      itself is generated where it is consumed, in backtrans_reg); inside a
//      cluster, inverse iteration with a partially pivoted tridiagonal LU and
//      modified Gram-Schmidt (dstein): tridiag_vectors;
//   4. back-transformation Z = Q Y (zunmtr): backtrans_reg in eigh_reg.hpp.
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

// The brackets start from one multisection sweep: the team counts at the n - 1
// points splitting the Gershgorin interval into n equal parts (one point per
// thread, counts in cnt[n]), and eigenvalue j starts from the part holding
// it, log2(n) bisection steps shorter.
template <typename T, typename Team>
RTB_HD void bisect_all(Team& tm, int n, const T* d, const T* e, T* e2, T* E, T* cnt) {
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
  const T width = gu - gl;
  auto point = [&](int m) { return gl + width * static_cast<T>(m + 1) / static_cast<T>(n); };
  for (int m = tm.id; m + 1 < n; m += tm.nt) cnt[m] = static_cast<T>(sturm_count(n, d, e2, point(m), pivmin));
  tm.sync();
  for (int j = tm.id; j < n; j += tm.nt) {
    // first point whose count exceeds j (counts are non-decreasing)
    int a = 0, b = n - 1;
    while (a < b) {
      const int mid = (a + b) / 2;
      if (cnt[mid] > static_cast<T>(j)) {
        b = mid;
      } else {
        a = mid + 1;
      }
    }
    T lo = a > 0 ? point(a - 1) : gl, hi = a < n - 1 ? point(a) : gu;
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
// Stage 3: inverse iteration for the eigenvectors of T inside clusters
// (dstein-like).
//   Y  : global [n][ldy], Y[i*ldy + j] = component i of eigenvector j (column-
//        major by eigenvector, coalesced across threads)
//   S  : global scratch [3][n][ldy] (pivoted-LU U rows, same layout as Y)
// Each column walk is serial in k.  Its loads are issued kColBlock at a time
// ahead of the arithmetic, so a walk waits for memory once per block rather
// than once per step (a load is not issued ahead of the stores before it).
// --------------------------------------------------------------------------
constexpr int kColBlock = 8;

RTB_HD inline double rtb_hash_uniform(uint32_t j, uint32_t i) {
  uint32_t h = j * 0x9E3779B1u ^ (i + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (static_cast<double>(h) + 0.5) * (2.0 / 4294967296.0) - 1.0;  // (-1, 1)
}

// One solve (T - lam I) z = b with partial pivoting (the pivot choice as
// selects, not branches), z normalised to unit 2-norm into Y[:, j].  b is the
// pseudo-random start rtb_hash_uniform(j, i) when `seeded`, else Y[:, j].
template <typename T>
RTB_HD void invit_solve(int n, const T* d, const T* e, T lam, T tol, T* __restrict__ Y, int ldy, int j,
                        T* __restrict__ S, bool seeded) {
  constexpr int kB = kColBlock;
  const size_t nn = static_cast<size_t>(n) * ldy;
  T* __restrict__ U0 = S;
  T* __restrict__ U1 = S + nn;
  T* __restrict__ U2 = S + 2 * nn;
#define AT(arr, i) arr[static_cast<size_t>(i) * ldy + j]
  // forward elimination; yk is the running right-hand side of pivot row k
  T a = d[0] - lam;
  T bsup = n > 1 ? e[0] : T(0);
  T yk = seeded ? static_cast<T>(rtb_hash_uniform(j, 0)) : AT(Y, 0);
  for (int k0 = 0; k0 + 1 < n; k0 += kB) {
    T pre[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int i = k0 + u + 1;
      pre[u] = i < n ? (seeded ? static_cast<T>(rtb_hash_uniform(j, i)) : AT(Y, i)) : T(0);
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 + u;
      if (k + 1 < n) {
        const T c = e[k], an = d[k + 1] - lam, bn = k + 2 < n ? e[k + 1] : T(0), y1 = pre[u];
        const bool sw = fabs(a) < fabs(c);  // row k+1 becomes pivot row k
        const T piv = sw ? c : a, oth = sw ? a : c;
        const T l = piv != T(0) ? oth / piv : T(0);
        AT(U0, k) = piv;
        AT(U1, k) = sw ? an : bsup;
        AT(U2, k) = sw ? bn : T(0);
        AT(Y, k) = sw ? y1 : yk;
        yk = (sw ? yk : y1) - l * (sw ? y1 : yk);
        a = (sw ? bsup : an) - l * (sw ? an : bsup);
        bsup = sw ? -l * bn : bn;
      }
    }
  }
  AT(U0, n - 1) = a;
  AT(U1, n - 1) = T(0);
  AT(U2, n - 1) = T(0);
  AT(Y, n - 1) = yk;
  // back substitution with tiny pivots perturbed to +-tol; the 2-norm is
  // accumulated scaled (scl * sqrt(ssq)), as the solution can be ~1/tol large
  T z1 = T(0), z2 = T(0), scl = T(0), ssq = T(0);
  for (int k0 = n - 1; k0 >= 0; k0 -= kB) {
    T p0[kB], p1[kB], p2[kB], py[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 - u;
      p0[u] = k >= 0 ? AT(U0, k) : T(1);
      p1[u] = k >= 0 ? AT(U1, k) : T(0);
      p2[u] = k >= 0 ? AT(U2, k) : T(0);
      py[u] = k >= 0 ? AT(Y, k) : T(0);
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        T u0 = p0[u];
        if (fabs(u0) < tol) u0 = (u0 >= T(0)) ? tol : -tol;
        const T z = (py[u] - p1[u] * z1 - p2[u] * z2) / u0;
        AT(Y, k) = z;
        z2 = z1;
        z1 = z;
        const T az = fabs(z), big = fmax(az, scl);
        const T r = big > T(0) ? fmin(az, scl) / big : T(0);
        ssq = az > scl ? T(1) + ssq * r * r : ssq + r * r;
        scl = big;
      }
    }
  }
  const T inv = scl > T(0) ? T(1) / (scl * sqrt(ssq)) : T(0);
  for (int k0 = 0; k0 < n; k0 += kB) {
    T p[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) p[u] = k0 + u < n ? AT(Y, k0 + u) : T(0);
#pragma unroll
    for (int u = 0; u < kB; ++u)
      if (k0 + u < n) AT(Y, k0 + u) = p[u] * inv;
  }
#undef AT
}

// Modified Gram-Schmidt of eigenvector jj against jj0..jj-1 (columns of Y),
// then unit 2-norm.
template <typename T>
RTB_HD void mgs_column(int n, T* __restrict__ Y, int ldy, int jj0, int jj) {
  constexpr int kB = kColBlock;
#define AT(i, c) Y[static_cast<size_t>(i) * ldy + (c)]
  T s = T(0);
  for (int ii = jj0; ii < jj; ++ii) {
    T dot = T(0);
    for (int k0 = 0; k0 < n; k0 += kB) {
      T pa[kB], pb[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        pa[u] = k0 + u < n ? AT(k0 + u, ii) : T(0);
        pb[u] = k0 + u < n ? AT(k0 + u, jj) : T(0);
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) dot += pa[u] * pb[u];
    }
    s = T(0);
    for (int k0 = 0; k0 < n; k0 += kB) {
      T pa[kB], pb[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        pa[u] = k0 + u < n ? AT(k0 + u, ii) : T(0);
        pb[u] = k0 + u < n ? AT(k0 + u, jj) : T(0);
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        if (k0 + u < n) {
          const T t = pb[u] - dot * pa[u];
          AT(k0 + u, jj) = t;
          s += t * t;
        }
      }
    }
  }
  const T inv = s > T(0) ? T(1) / sqrt(s) : T(0);
  for (int k0 = 0; k0 < n; k0 += kB) {
    T p[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) p[u] = k0 + u < n ? AT(k0 + u, jj) : T(0);
#pragma unroll
    for (int u = 0; u < kB; ++u)
      if (k0 + u < n) AT(k0 + u, jj) = p[u] * inv;
  }
#undef AT
}

// Twisted factorisation for an isolated eigenvalue lam (Parlett & Dhillon,
// LAA 309 (2000) 121-151; Fernando, SIMAX 18 (1997)):
// T - lam I = L+ D+ L+^T (top-down) = U- D- U-^T (bottom-up); the twist
// index r minimises |gamma_k| = |D+_k + D-_k - (d_k - lam)|, and
//   z_r = 1,  z_k = -(e_k / D+_k) z_{k+1} (k < r),
//   z_k = -(e_{k-1} / D-_k) z_{k-1} (k > r)
// solves (T - lam I) z = gamma_r e_r: one solve, no iteration, residual
// |gamma_r| / ||z|| (twisted_z generates z).  Pivots below pivmin are
// replaced by -pivmin (the Sturm count rule).  D+ and D- are stored in
// S[0][k][j] and S[1][k][j] (column j, leading dimension ldy); returns r.
template <typename T>
RTB_HD int twisted_factor(int n, const T* d, const T* e, T lam, T pivmin, int ldy, int j, T* S) {
  // The recurrences are serial; the scratch reads are not, so they are
  // issued kB at a time ahead of the arithmetic (one memory latency per
  // block instead of one per step).
  constexpr int kB = 8;
  const size_t nn = static_cast<size_t>(n) * ldy;
  T* Dp = S;
  T* Dm = S + nn;
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
#undef SAt
  return r;
}

// Unit eigenvector from a twisted factorisation (see twisted_factor), all n
// components in order: for k < r the top-down chain, for k > r the
// bottom-up one.  Dp, Dm: D+ and D- of column j (stride ldy).
template <typename T>
RTB_HD void twisted_z(int n, const T* e, const T* Dp, const T* Dm, int ldy, int r, T* z) {
  T nrm2 = T(1), v = T(1);
  z[r] = T(1);
  for (int k = r - 1; k >= 0; --k) {
    v = -e[k] * recip(Dp[static_cast<size_t>(k) * ldy]) * v;
    z[k] = v;
    nrm2 += v * v;
  }
  v = T(1);
  for (int k = r + 1; k < n; ++k) {
    v = -e[k - 1] * recip(Dm[static_cast<size_t>(k) * ldy]) * v;
    z[k] = v;
    nrm2 += v * v;
  }
  const T inv = T(1) / sqrt(nrm2);
  for (int k = 0; k < n; ++k) z[k] *= inv;
}

// Thresholds shared by the eigenvector stages: eigenvalues closer than
// ortol = cluster_rel * tnorm to a neighbour form a cluster (dstein's
// cluster_rel is 1e-3); tol perturbs tiny LU pivots; pivmin guards the
// twisted pivots (the Sturm count rule).
template <typename T>
struct VecParams {
  T tol, ortol, pivmin;
};

template <typename T>
RTB_HD inline VecParams<T> vec_params(int n, const T* d, const T* e, T cluster_rel) {
  T gl, gu;
  gershgorin(n, d, e, gl, gu);
  const T tnorm = fmax(fabs(gl), fabs(gu));
  T emax2 = T(0);
  for (int i = 0; i + 1 < n; ++i) emax2 = fmax(emax2, e[i] * e[i]);
  return {fmax(Lim<T>::eps() * tnorm, Lim<T>::safmin() * T(1e4)), cluster_rel * tnorm,
          Lim<T>::safmin() * fmax(T(1), emax2)};
}

template <typename T>
RTB_HD inline bool in_cluster(int n, const T* E, int j, T ortol) {
  return (j > 0 && E[j] - E[j - 1] <= ortol) || (j + 1 < n && E[j + 1] - E[j] <= ortol);
}

// Eigenvector data of the isolated eigenvalues: the twisted factorisation
// (D+, D- in S[0..1], twist[j] = r); twist[j] = -1 marks a cluster member.
template <typename T, typename Team>
RTB_HD void twisted_all(Team& tm, int n, const T* d, const T* e, const T* E, int ldy, T* S, T* twist,
                        T cluster_rel) {
  const VecParams<T> p = vec_params(n, d, e, cluster_rel);
  for (int j = tm.id; j < n; j += tm.nt)
    twist[j] = in_cluster(n, E, j, p.ortol) ? T(-1) : static_cast<T>(twisted_factor(n, d, e, E[j], p.pivmin, ldy, j, S));
}

// Cluster members by inverse iteration (dstein: a pivoted LU solve from a
// fixed pseudo-random start, twice, each followed by modified Gram-Schmidt
// inside the cluster) into unit columns of Y; S[0..2] of those columns is
// scratch.  The other columns are not touched.
template <typename T, typename Team>
RTB_HD void cluster_vectors(Team& tm, int n, const T* d, const T* e, const T* E, T* Y, int ldy, T* S,
                            T cluster_rel) {
  const VecParams<T> p = vec_params(n, d, e, cluster_rel);
  for (int pass = 0; pass < 2; ++pass) {
    for (int j = tm.id; j < n; j += tm.nt)
      if (in_cluster(n, E, j, p.ortol)) invit_solve(n, d, e, E[j], p.tol, Y, ldy, j, S, pass == 0);
    tm.sync();
    // cluster heads orthogonalise their members in order
    for (int j = tm.id; j < n; j += tm.nt) {
      const bool head = (j == 0) || (E[j] - E[j - 1] > p.ortol);
      if (!head) continue;
      int last = j;
      while (last + 1 < n && E[last + 1] - E[last] <= p.ortol) ++last;
      for (int jj = j + 1; jj <= last; ++jj) mgs_column(n, Y, ldy, j, jj);
    }
    tm.sync();
  }
}

// Both eigenvector stages (S: [3][n][ldy] scratch).
template <typename T, typename Team>
RTB_HD void tridiag_vectors(Team& tm, int n, const T* d, const T* e, const T* E, T* Y, int ldy, T* S, T* twist,
                            T cluster_rel = T(1e-3)) {
  twisted_all(tm, n, d, e, E, ldy, S, twist, cluster_rel);
  cluster_vectors(tm, n, d, e, E, Y, ldy, S, cluster_rel);
}

}  // namespace rtb
