// Register-resident Householder tridiagonalization and back-transformation
// for one matrix of order n <= NB per team of NB * S threads.
//
// Thread t owns index r = t % NB and slice s = t / NB; it keeps C = NB / S
// complex entries in registers:
//   tridiag_reg     row r of A, columns c = s + S*q (q = 0..C-1)
//   backtrans_reg   column r of Z, rows    s + S*q
// Only vectors (the reflector, the matvec partials, w) pass through
// team-shared memory, so the O(n^3) work never touches LDS.  The arithmetic
// is the same zhetd2 / zunmtr recurrence as tridiagonalize() / apply_q() in
// eigh_core.hpp; the Hermitian matrix is kept in full (both triangles
// updated) so that every matvec and update is local to a thread's row.
//
// Cyclic slices (c = s + S*q) keep the S slices equally loaded while the
// active trailing block shrinks; work is skipped in blocks of 8 entries whose
// columns (rows) the current reflector does not touch.
#pragma once

#include "eigh_core.hpp"

namespace rtb {

template <typename T>
struct RegSmem {
  cplx<T>* xc;  // [NB]      column j of A, by row
  cplx<T>* vp;  // [3 * NB]  reflector(s), slice-major: vp[s*C + q] = v[s + S*q]
  cplx<T>* wp;  // [NB]      w, slice-major
  cplx<T>* yp;  // [2 * S * NB]  per-slice partial sums
  T* red;       // [2 * Team::kRedStride]
  T* es;        // [NB]      sub-diagonal of T (back-transformation)
};

// a[q] for a team-uniform q: a select chain the compiler keeps in registers
// (a dynamically indexed register array would be demoted to scratch memory).
template <int C, typename T>
RTB_HD inline cplx<T> reg_pick(const cplx<T> (&a)[C], int q) {
#if defined(__HIP_DEVICE_COMPILE__)
  T xr = a[0].re, xi = a[0].im;
#pragma unroll
  for (int i = 1; i < C; ++i) {
    const bool m = i == q;
    xr = m ? a[i].re : xr;
    xi = m ? a[i].im : xi;
  }
  return {xr, xi};
#else
  return a[q];
#endif
}

// Slice index of thread `id`.  With NB >= 64 every wave lies in one slice,
// so on the GPU the value is made wave-uniform (scalar), which turns the
// per-block skip tests into scalar branches.
template <int NB>
RTB_HD inline int slice_of(int id) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (NB >= 64) return __builtin_amdgcn_readfirstlane(id / NB);
#endif
  return id / NB;
}

template <typename T>
RTB_HD inline T fmadd(T a, T b, T c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return fma(a, b, c);
#else
  return std::fma(a, b, c);
#endif
}

// --------------------------------------------------------------------------
// Q^H A Q = T for the Hermitian A given as FP64 row-major Hm[n][n] (the lower
// triangle is read, as zhetrd with UPLO = 'L').  Outputs as tridiagonalize():
// Dg[0..n-1], Eg[0..n-2] (Eg[n-1] = 0), Vg[j][0..n-1], taug[j].
// --------------------------------------------------------------------------
template <typename T, int NB, int S, typename Team>
RTB_HD void tridiag_reg(Team& tm, int n, const double* Hm, RegSmem<T> sm, T* Dg, T* Eg, cplx<T>* Vg,
                        cplx<T>* taug) {
  constexpr int C = NB / S;
  constexpr int BL = C < 8 ? C : 8;
  constexpr int NBL = C / BL;
  static_assert(NB % S == 0 && C % BL == 0, "slice shape");
  const int r = tm.id % NB, s = slice_of<NB>(tm.id);
  cplx<T> a[C];
  {
    // A(r, c) = H[r][c] below the diagonal, conj(H[c][r]) above it and
    // Re H[r][r] on it: the Hermitian matrix of H's lower triangle, as
    // zheev('L') sees it (it ignores Im H[r][r]); the per-q addresses are
    // base + q * stride
    const int rr = r < n ? r : 0;
    const double* lo = Hm + (static_cast<size_t>(rr) * n + s) * 2;  // H[r][s + S q]
    const double* up = Hm + (static_cast<size_t>(s) * n + rr) * 2;  // H[s + S q][r]
    const size_t ustep = static_cast<size_t>(S) * n * 2;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int c = s + S * q;
      cplx<T> x = {T(0), T(0)};
      if (r < n && c < n) {
        if (r >= c) {
          const double* h = lo + q * (2 * S);
          x = {static_cast<T>(h[0]), r == c ? T(0) : static_cast<T>(h[1])};
        } else {
          const double* h = up + q * ustep;
          x = {static_cast<T>(h[0]), -static_cast<T>(h[1])};
        }
      }
      a[q] = x;
    }
  }
  T* red_n = sm.red;
  T* red_v = sm.red + Team::kRedStride;
  const int rslot = (r % S) * C + r / S;  // slice-major slot of index r
  // column `col` of A into xc (all rows), and the partial of |A(col+2:, col)|^2
  auto extract = [&](int col) -> T {
    T p = T(0);
    if (s == col % S) {
      const cplx<T> x = reg_pick<C, T>(a, col / S);
      sm.xc[r] = x;
      if (r >= col + 2) p = x.re * x.re + x.im * x.im;
    }
    return p;
  };
  T part = extract(0);
  for (int j = 0; j + 1 < n; ++j) {
    // B0: column j and its norm are visible
    const T xnorm2 = tm.reduce_sum_ns(part, red_n);
    const cplx<T> alpha = sm.xc[j + 1];
    const T djj = sm.xc[j].re;
    cplx<T> tau = {T(0), T(0)};
    cplx<T> scale = {T(0), T(0)};
    T beta = alpha.re;
    if (!(xnorm2 == T(0) && alpha.im == T(0))) {  // zlarfg
      T nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm2);
      beta = alpha.re >= T(0) ? -nrm : nrm;
      const T rb = recip(beta);
      tau.re = (beta - alpha.re) * rb;
      tau.im = -alpha.im * rb;
      T ar = alpha.re - beta, ai = alpha.im;
      const T rden = recip(ar * ar + ai * ai);
      scale.re = ar * rden;
      scale.im = -ai * rden;
    }
    if (s == 0) {
      cplx<T> v = {T(0), T(0)};
      if (r == j + 1) {
        v.re = T(1);
      } else if (r > j + 1) {
        v = cmul(sm.xc[r], scale);
      }
      sm.vp[rslot] = v;
      if (Vg && r < n) Vg[static_cast<size_t>(j) * n + r] = v;
    }
    if (tm.id == 0) {
      Dg[j] = djj;
      Eg[j] = beta;
      if (taug) taug[j] = tau;
    }
    tm.sync();  // B1: v visible
    if (tau.re != T(0) || tau.im != T(0)) {
      const cplx<T> vr = sm.vp[rslot];
      const cplx<T>* vs = sm.vp + s * C;
      cplx<T> y = {T(0), T(0)};
      if (r > j) {
#pragma unroll
        for (int b = 0; b < NBL; ++b) {
          if (s + S * (b * BL + BL - 1) > j) {
#pragma unroll
            for (int i = 0; i < BL; ++i) {
              const int q = b * BL + i;
              const cplx<T> vc = vs[q];
              y.re = fmadd(a[q].re, vc.re, y.re);
              y.re = fmadd(-a[q].im, vc.im, y.re);
              y.im = fmadd(a[q].re, vc.im, y.im);
              y.im = fmadd(a[q].im, vc.re, y.im);
            }
          }
        }
      }
      sm.yp[s * NB + r] = y;
      // B2: v^H A v (real for Hermitian A) and the partials are visible
      const T vav = tm.reduce_sum_ns(vr.re * y.re + vr.im * y.im, red_v);
      // w = p - 1/2 tau (p^H v) v with p = tau A v, and p^H v = conj(tau) v^H A v
      const T a2 = T(-0.5) * (tau.re * tau.re + tau.im * tau.im) * vav;
      cplx<T> wr = {T(0), T(0)};
      if (r > j) {
        cplx<T> ys = {T(0), T(0)};
#pragma unroll
        for (int t = 0; t < S; ++t) {
          ys.re += sm.yp[t * NB + r].re;
          ys.im += sm.yp[t * NB + r].im;
        }
        const cplx<T> p = cmul(tau, ys);
        wr = {p.re + a2 * vr.re, p.im + a2 * vr.im};
      }
      if (s == 0) sm.wp[rslot] = wr;
      tm.sync();  // B3: w visible
      if (r > j) {
        const cplx<T>* ws = sm.wp + s * C;
        // A -= v w^H + w v^H on this thread's entries of row r
#pragma unroll
        for (int b = 0; b < NBL; ++b) {
          if (s + S * (b * BL + BL - 1) > j) {
#pragma unroll
            for (int i = 0; i < BL; ++i) {
              const int q = b * BL + i;
              const cplx<T> vc = vs[q], wc = ws[q];
              T re = a[q].re, im = a[q].im;
              re = fmadd(-vr.re, wc.re, re);
              re = fmadd(-vr.im, wc.im, re);
              re = fmadd(-wr.re, vc.re, re);
              re = fmadd(-wr.im, vc.im, re);
              im = fmadd(-vr.im, wc.re, im);
              im = fmadd(vr.re, wc.im, im);
              im = fmadd(-wr.im, vc.re, im);
              im = fmadd(wr.re, vc.im, im);
              a[q] = {re, im};
            }
          }
        }
      }
    }
    part = extract(j + 1);
  }
  tm.sync();
  if (tm.id == 0 && n >= 1) {
    Dg[n - 1] = sm.xc[n - 1].re;
    Eg[n - 1] = T(0);
  }
}

// --------------------------------------------------------------------------
// Z = Q Y for the reflectors of tridiag_reg (zunmtr, SIDE = 'L', 'N'), Y the
// eigenvectors of T (column c = eigenvector c):
//   twist[c] >= 0: generated here from the twisted factorisation stored by
//                  tridiag_vectors (Ts = [D+ | D-], [2][n][ldy]) and the
//                  sub-diagonal e; every thread of column c walks all k
//                  (in step across the wave: coalesced reads) and keeps its
//                  rows;
//   twist[c] < 0:  read from the unit column c of Y ([n][ldy]).
// Z complex, written in type TO (FP64, or FP32 when T is: exact) as
//   rows_out = 0: Zout[r][c] (eigenvectors as columns)
//   rows_out = 1: Zout[c][r] (eigenvectors as rows)
// Per reflector j = n-2 .. 0: s_c = tau_j v_j^H Z[:, c], Z[:, c] -= v_j s_c.
// gdiag (optional, [n]): the squared 2-norms of the stored columns.
// --------------------------------------------------------------------------
template <typename T, int NB, int S, typename Team, typename TO>
RTB_HD void backtrans_reg(Team& tm, int n, const T* Y, int ldy, const T* twist, const T* e, const T* Ts,
                          const cplx<T>* Vg, const cplx<T>* taug, RegSmem<T> sm, TO* Zout, int rows_out,
                          double* gdiag = nullptr) {
  constexpr int C = NB / S;
  constexpr int BL = C < 8 ? C : 8;
  constexpr int NBL = C / BL;
  const int c = tm.id % NB, s = slice_of<NB>(tm.id);
  // e staged in team-shared memory: read per k by every thread, it would
  // otherwise occupy one scalar register per k
  for (int i = tm.id; i < n; i += tm.nt) sm.es[i] = e[i];
  tm.sync();
  const T* es = sm.es;
  cplx<T> z[C];
  {
    const int r = c < n ? static_cast<int>(twist[c]) : -1;
#pragma unroll
    for (int q = 0; q < C; ++q) z[q] = {T(0), T(0)};
    {
      // branch-free: every step computes the next chain value and keeps it
      // only inside the chain's range (k < r going up, r < k < n going down);
      // the scratch loads are scheduled kGenGroup register rows at a time.
      // Columns without a twist (r < 0) are overwritten from Y below.
      [[maybe_unused]] constexpr int kGenGroup = 4;  // unused in the host pass
      const int cc = c < n ? c : 0;
      const T* Dp = Ts + cc;
      const T* Dm = Ts + static_cast<size_t>(n) * ldy + cc;
      T nrm2 = T(1), v = T(1);
#pragma unroll
      for (int q = C - 1; q >= 0; --q) {  // k < r, descending
#pragma unroll
        for (int t = S - 1; t >= 0; --t) {
          const int k = t + S * q;
          if (k + 1 < NB) {
            const int kc = k < n ? k : n - 1;
            const T nv = -es[kc] * recip(Dp[static_cast<size_t>(kc) * ldy]) * v;
            const bool in = k < r;
            v = in ? nv : v;
            nrm2 += in ? nv * nv : T(0);
            if (t == s) z[q].re = in ? nv : (k == r ? T(1) : z[q].re);
          } else if (t == s) {
            z[q].re = k == r ? T(1) : z[q].re;
          }
        }
#if defined(__HIP_DEVICE_COMPILE__)
        if (q % kGenGroup == 0) __builtin_amdgcn_sched_barrier(0);  // bounded load look-ahead
#endif
      }
      v = T(1);
#pragma unroll
      for (int q = 0; q < C; ++q) {  // k > r, ascending
#pragma unroll
        for (int t = 0; t < S; ++t) {
          const int k = t + S * q;
          if (k >= 1) {
            const int kc = k < n ? k : n - 1;
            const int ke = kc >= 1 ? kc - 1 : 0;  // in range also for n = 1 (value unused there)
            const T nv = -es[ke] * recip(Dm[static_cast<size_t>(kc) * ldy]) * v;
            const bool in = r >= 0 && k > r && k < n;
            v = in ? nv : v;
            nrm2 += in ? nv * nv : T(0);
            if (t == s && in) z[q].re = nv;
          }
        }
#if defined(__HIP_DEVICE_COMPILE__)
        if (q % kGenGroup == kGenGroup - 1) __builtin_amdgcn_sched_barrier(0);
#endif
      }
      const T inv = T(1) / sqrt(nrm2);
#pragma unroll
      for (int q = 0; q < C; ++q) z[q].re *= inv;
    }
    if (r < 0 && c < n) {  // a cluster member: the unit column of Y
      const T* yb = Y + static_cast<size_t>(s) * ldy + c;  // Y[s + S q][c]
      const size_t ystep = static_cast<size_t>(S) * ldy;
#pragma unroll
      for (int q = 0; q < C; ++q) z[q].re = s + S * q < n ? yb[q * ystep] : T(0);
    }
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
    for (int q = 0; q < C; ++q) {
      // keep the zero imaginary parts opaque: otherwise the no-reflector path
      // is specialised with 2 * C constant doubles live at the stores
      asm volatile("" : "+v"(z[q].im));
    }
#endif
  }
  // reflector buffers rotate over three slots: a slot is rewritten two
  // barriers after its last read; the partial sums alternate over two.
  if (n >= 2 && tm.id < NB) {
    const int i = tm.id;
    sm.vp[(i % S) * C + i / S] =
        i < n ? Vg[static_cast<size_t>(n - 2) * n + i] : cplx<T>{T(0), T(0)};
  }
  tm.sync();
  for (int j = n - 2, k = 0; j >= 0; --j, ++k) {
    const cplx<T> tau = taug[j];
    cplx<T> nv = {T(0), T(0)};
    const bool stage = j >= 1 && tm.id < NB;
    if (stage && tm.id < n) nv = Vg[static_cast<size_t>(j - 1) * n + tm.id];
    const cplx<T>* vs = sm.vp + (k % 3) * NB + s * C;
    const bool act = !(tau.re == T(0) && tau.im == T(0));
    cplx<T> dot = {T(0), T(0)};
    if (act) {
#pragma unroll
      for (int b = 0; b < NBL; ++b) {
        if (s + S * (b * BL + BL - 1) > j) {
#pragma unroll
          for (int i = 0; i < BL; ++i) {
            const int q = b * BL + i;
            const cplx<T> v = vs[q];  // conj(v) * z
            dot.re = fmadd(v.re, z[q].re, dot.re);
            dot.re = fmadd(v.im, z[q].im, dot.re);
            dot.im = fmadd(v.re, z[q].im, dot.im);
            dot.im = fmadd(-v.im, z[q].re, dot.im);
          }
        }
      }
    }
    cplx<T>* dp = sm.yp + (k & 1) * S * NB;
    dp[s * NB + c] = dot;
    if (stage) {
      const int i = tm.id;
      sm.vp[((k + 1) % 3) * NB + (i % S) * C + i / S] = nv;
    }
    tm.sync();
    if (act) {
      cplx<T> acc = {T(0), T(0)};
#pragma unroll
      for (int t = 0; t < S; ++t) {
        acc.re += dp[t * NB + c].re;
        acc.im += dp[t * NB + c].im;
      }
      const cplx<T> w = cmul(tau, acc);
#pragma unroll
      for (int b = 0; b < NBL; ++b) {
        if (s + S * (b * BL + BL - 1) > j) {
#pragma unroll
          for (int i = 0; i < BL; ++i) {
            const int q = b * BL + i;
            const cplx<T> v = vs[q];  // z -= v * w
            T re = z[q].re, im = z[q].im;
            re = fmadd(-v.re, w.re, re);
            re = fmadd(v.im, w.im, re);
            im = fmadd(-v.re, w.im, im);
            im = fmadd(-v.im, w.re, im);
            z[q] = {re, im};
          }
        }
      }
    }
  }
  if (gdiag) {
    // squared 2-norm of column c of the stored Z, in FP64 (exact products of
    // the stored values), reduced over the S slices in shared memory
    double g = 0.0;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const double zr = static_cast<double>(static_cast<TO>(z[q].re));
      const double zi = static_cast<double>(static_cast<TO>(z[q].im));
      if (s + S * q < n) g += zr * zr + zi * zi;
    }
    tm.sync();  // the partial-sum buffer is free
    double* gp = reinterpret_cast<double*>(sm.yp);
    gp[s * NB + c] = g;
    tm.sync();
    if (s == 0 && c < n) {
      double t = 0.0;
      for (int u = 0; u < S; ++u) t += gp[u * NB + c];
      gdiag[c] = t;
    }
  }
  if (c < n) {
    // Zout element (r = s + S q, c): base + q * stride
    TO* zb = Zout + (rows_out ? static_cast<size_t>(c) * n + s : static_cast<size_t>(s) * n + c) * 2;
    const size_t zstep = rows_out ? static_cast<size_t>(S) * 2 : static_cast<size_t>(S) * n * 2;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (s + S * q < n) {
        TO* o = zb + q * zstep;
        o[0] = static_cast<TO>(z[q].re);
        o[1] = static_cast<TO>(z[q].im);
      }
    }
  }
}

}  // namespace rtb
