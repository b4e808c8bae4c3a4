// Bloch sum on a uniform k-mesh by axis-separable DFT passes (FP64).
//
// For k = (i1/N1, i2/N2, i3/N3) the phase factorises,
//   exp(i 2 pi k.R) = prod_d W_d^((i_d R_d) mod N_d),  W_d = exp(i 2 pi / N_d),
// so H(k) = sum_R1 W1^(i1 R1) sum_R2 W2^(i2 R2) sum_R3 W3^(i3 R3) h[R1,R2,R3].
// The hoppings are scattered once into a dense box [L1][L2][L3][n*n]
// (L_d = extent of R_d), then:
//   pass A (3D only): G_A[R1,R2][i3][e] = sum_R3 W3 h, and the R3-weighted G_A'
//   pass B:           G_B[R1][i2,i3][e] = sum_R2 W2 G_A (+ R2-weighted, + from G_A')
//   pass C (per chunk): H, and v_p = i sum_d (lat w_p)_d D_d with
//                     D_d = sum_R R_d e^{i2pi k.R} h_R, from sum_R1 W1 (R1-weighted) G_B*
// with the Atom gauge and tau term applied as in the dense kernel.  The roots
// W^m are exact-argument sincospi values (m = (i R) mod N in integers), so the
// result is the reference's sum re-associated, without rounding of k.R.
#include <algorithm>

#include "eigh.hpp"

namespace rtb {

namespace {

constexpr int kTW = 64;  // inner elements per tile

inline int imod(long long a, int n) {
  long long r = a % n;
  return static_cast<int>(r < 0 ? r + n : r);
}

__device__ inline int dmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

__global__ void copy_kernel(size_t count, const double2* __restrict__ x, double2* __restrict__ y) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g < count) y[g] = x[g];
}

__global__ void scatter_box_kernel(const double2* __restrict__ ham, const int* __restrict__ boxidx, int nR,
                                   size_t n2, double2* __restrict__ box) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= static_cast<size_t>(nR) * n2) return;
  const size_t r = g / n2, e = g % n2;
  box[static_cast<size_t>(boxidx[r]) * n2 + e] = ham[g];
}

// One DFT pass along an axis: X [outer][L][inner] -> Y_v [outer][N][inner],
//   Y_v[o][k][t] = sum_l w_v(l) W^((k (lo+l)) mod N) X[o][l][t],
//   w_v(l) = 1 (mode 0) or (lo + l) (mode 1).  Up to two variants per launch.
__global__ void __launch_bounds__(256) dft_axis_kernel(const double2* __restrict__ X, int L, int lo, size_t inner,
                                                       int N, int nv, int mode0, int mode1,
                                                       double2* __restrict__ Y0, double2* __restrict__ Y1) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* W = reinterpret_cast<double2*>(smem_raw);  // [N]
  double2* Xs = W + N;                                // [L][kTW]
  const size_t o = blockIdx.y;
  const size_t t0 = static_cast<size_t>(blockIdx.x) * kTW;
  const int tw = static_cast<int>(inner - t0 < static_cast<size_t>(kTW) ? inner - t0 : kTW);
  for (int m = threadIdx.x; m < N; m += 256) {
    double s, c;
    sincospi(2.0 * static_cast<double>(m) / static_cast<double>(N), &s, &c);
    W[m] = make_double2(c, s);
  }
  for (int q = threadIdx.x; q < L * kTW; q += 256) {
    const int l = q / kTW, t = q % kTW;
    Xs[q] = t < tw ? X[(o * L + l) * inner + t0 + t] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int t = threadIdx.x % kTW;
  if (t >= tw) return;
  for (int k = threadIdx.x / kTW; k < N; k += 256 / kTW) {
    double a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
    int m = dmod(k * lo, N);
    const int step = k % N;
    for (int l = 0; l < L; ++l) {
      const double2 w = W[m];
      const double2 x = Xs[l * kTW + t];
      const double tr = w.x * x.x - w.y * x.y, ti = w.x * x.y + w.y * x.x;
      const double wt = static_cast<double>(lo + l);
      if (mode0) {
        a0r += wt * tr;
        a0i += wt * ti;
      } else {
        a0r += tr;
        a0i += ti;
      }
      if (nv > 1) {
        if (mode1) {
          a1r += wt * tr;
          a1i += wt * ti;
        } else {
          a1r += tr;
          a1i += ti;
        }
      }
      m += step;
      if (m >= N) m -= N;
    }
    const size_t oi = (o * N + k) * inner + t0 + t;
    Y0[oi] = make_double2(a0r, a0i);
    if (nv > 1) Y1[oi] = make_double2(a1r, a1i);
  }
}

struct FinalArgs {
  int dim, n, L1, lo1, N1, N2, N3;
  size_t P;               // pairs (N2*N3)
  const double2* G0;      // [L1][P][n2]  (unweighted)
  const double2* G1;      // [L1][P][n2]  R2-weighted (dim >= 2)
  const double2* G2;      // [L1][P][n2]  R3-weighted (dim == 3)
  double alpha[2][3];     // (lat w_p)_d
  const double* tau;      // [n][3]
  const double* wtau;     // [2][n]
  int np;                 // 0 or 2
  int gauge;
  int combine;            // V0 + i V1 into one matrix per k
};

// Pass C for pairs [p0, p0 + gridDim.y): outputs in k1-innermost order.
// A workgroup owns kTW matrix elements of one pair for all N1 values of k1:
// the kTW columns of G0..G2 are staged in LDS once, and each thread forms
// kFB values of k1 per pass, so one LDS read of G serves kFB outputs.  The
// Atom-gauge factor conj(phi_m) phi_c = exp(i 2 pi (tau_c - tau_m).k) is one
// sincospi per element and k.
constexpr int kFB = 2;
constexpr int kTWF = 64;  // matrix elements per workgroup of the final pass

// NG: partial sums staged (G0 .. G_{NG-1}); VEL: the velocities are formed
// (d1 from G0, then G1, G2 for NG = 2, 3).
template <int NG, bool VEL>
__global__ void __launch_bounds__(256) dft_final_kernel(FinalArgs a, size_t p0, double2* __restrict__ H,
                                                        double2* __restrict__ V) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* W = reinterpret_cast<double2*>(smem_raw);  // [N1]
  double2* X0 = W + a.N1;                             // [L1][kTWF]
  double2* X1 = X0 + a.L1 * kTWF;
  double2* X2 = X1 + a.L1 * kTWF;
  const size_t n2 = static_cast<size_t>(a.n) * a.n;
  const size_t pl = blockIdx.y, pg = p0 + pl;
  const size_t t0 = static_cast<size_t>(blockIdx.x) * kTWF;
  const int tw = static_cast<int>(n2 - t0 < static_cast<size_t>(kTWF) ? n2 - t0 : kTWF);
  for (int m = threadIdx.x; m < a.N1; m += 256) {
    double s, c;
    sincospi(2.0 * static_cast<double>(m) / static_cast<double>(a.N1), &s, &c);
    W[m] = make_double2(c, s);
  }
  for (int q = threadIdx.x; q < a.L1 * kTWF; q += 256) {
    const int l = q / kTWF, t = q % kTWF;
    const size_t src = (static_cast<size_t>(l) * a.P + pg) * n2 + t0 + t;
    const bool ok = t < tw;
    X0[q] = ok ? a.G0[src] : make_double2(0.0, 0.0);
    if (NG > 1) X1[q] = ok ? a.G1[src] : make_double2(0.0, 0.0);
    if (NG > 2) X2[q] = ok ? a.G2[src] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int t = threadIdx.x % kTWF;
  if (t >= tw) return;
  const size_t e = t0 + t;
  const int mrow = static_cast<int>(e / a.n), mcol = static_cast<int>(e % a.n);
  const size_t i2 = pg / a.N3, i3 = pg % a.N3;
  const double k2 = a.dim >= 2 ? static_cast<double>(i2) / static_cast<double>(a.N2) : 0.0;
  const double k3 = a.dim >= 3 ? static_cast<double>(i3) / static_cast<double>(a.N3) : 0.0;
  const double* tm = a.tau + static_cast<size_t>(mrow) * 3;
  const double* tc = a.tau + static_cast<size_t>(mcol) * 3;
  const double dt1 = tc[0] - tm[0], dt23 = (tc[1] - tm[1]) * k2 + (tc[2] - tm[2]) * k3;
  double wd[2] = {0.0, 0.0};  // w_p . (tau_c - tau_m)
  for (int p = 0; p < a.np; ++p)
    wd[p] = a.wtau[static_cast<size_t>(p) * a.n + mcol] - a.wtau[static_cast<size_t>(p) * a.n + mrow];
  constexpr int kGroups = 256 / kTWF;
  for (int ib = (threadIdx.x / kTWF) * kFB; ib < a.N1; ib += kGroups * kFB) {
    double hr[kFB], hi[kFB], d1r[kFB], d1i[kFB], g1r[kFB], g1i[kFB], g2r[kFB], g2i[kFB];
    int m[kFB], step[kFB];
#pragma unroll
    for (int u = 0; u < kFB; ++u) {
      const int i1 = min(ib + u, a.N1 - 1);  // lanes past N1 repeat the last k1, unstored
      m[u] = dmod(i1 * a.lo1, a.N1);
      step[u] = i1;
      hr[u] = hi[u] = d1r[u] = d1i[u] = g1r[u] = g1i[u] = g2r[u] = g2i[u] = 0.0;
    }
    // unrolled so that the LDS reads of the next terms are issued while the
    // current ones are multiplied
#pragma unroll 4
    for (int l = 0; l < a.L1; ++l) {
      const double2 x = X0[l * kTWF + t];
      const double2 y = NG > 1 ? X1[l * kTWF + t] : make_double2(0.0, 0.0);
      const double2 z = NG > 2 ? X2[l * kTWF + t] : make_double2(0.0, 0.0);
      const double r1 = static_cast<double>(a.lo1 + l);
      double2 w[kFB];
#pragma unroll
      for (int u = 0; u < kFB; ++u) {
        w[u] = W[m[u]];
        m[u] += step[u];
        if (m[u] >= a.N1) m[u] -= a.N1;
      }
#pragma unroll
      for (int u = 0; u < kFB; ++u) {
        const double tr = w[u].x * x.x - w[u].y * x.y, ti = w[u].x * x.y + w[u].y * x.x;
        hr[u] += tr;
        hi[u] += ti;
        if (VEL) {
          d1r[u] += r1 * tr;
          d1i[u] += r1 * ti;
        }
        if (NG > 1) {
          g1r[u] += w[u].x * y.x - w[u].y * y.y;
          g1i[u] += w[u].x * y.y + w[u].y * y.x;
        }
        if (NG > 2) {
          g2r[u] += w[u].x * z.x - w[u].y * z.y;
          g2i[u] += w[u].x * z.y + w[u].y * z.x;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kFB; ++u) {
      const int i1 = ib + u;
      if (i1 >= a.N1) break;
      const double k1 = static_cast<double>(i1) / static_cast<double>(a.N1);
      double gr = 1.0, gi = 0.0;
      if (a.gauge == RTB_GAUGE_ATOM) {
        double arg = dt1 * k1 + dt23;
        arg -= rint(arg);
        sincospi(2.0 * arg, &gi, &gr);
      }
      const size_t kl = pl * a.N1 + i1;
      H[kl * n2 + e] = make_double2(hr[u] * gr - hi[u] * gi, hr[u] * gi + hi[u] * gr);
      double cr = 0.0, ci = 0.0;
      for (int p = 0; p < a.np; ++p) {
        const double sr = a.alpha[p][0] * d1r[u] + a.alpha[p][1] * g1r[u] + a.alpha[p][2] * g2r[u];
        const double si = a.alpha[p][0] * d1i[u] + a.alpha[p][1] * g1i[u] + a.alpha[p][2] * g2i[u];
        double xr = -si, xi = sr;  // i * sum
        if (a.gauge == RTB_GAUGE_ATOM) {
          xr += -wd[p] * hi[u];
          xi += wd[p] * hr[u];
          const double yr = xr * gr - xi * gi, yi = xr * gi + xi * gr;
          xr = yr;
          xi = yi;
        }
        if (a.combine) {
          cr += p == 0 ? xr : -xi;
          ci += p == 0 ? xi : xr;
        } else {
          V[(kl * 2 + p) * n2 + e] = make_double2(xr, xi);
        }
      }
      if (a.combine && a.np == 2) V[kl * n2 + e] = make_double2(cr, ci);
    }
  }
}

size_t axis_smem(int N, int L) { return sizeof(double2) * (static_cast<size_t>(N) + static_cast<size_t>(L) * kTW); }

int launch_axis(hipStream_t st, const double2* X, size_t outer, int L, int lo, size_t inner, int N, int nv,
                int mode0, int mode1, double2* Y0, double2* Y1) {
  const size_t tiles = (inner + kTW - 1) / kTW;
  for (size_t o0 = 0; o0 < outer; o0 += 65535) {
    const size_t oc = outer - o0 < 65535 ? outer - o0 : 65535;
    hipLaunchKernelGGL(dft_axis_kernel, dim3(static_cast<unsigned>(tiles), static_cast<unsigned>(oc)), dim3(256),
                       axis_smem(N, L), st, X + o0 * L * inner, L, lo, inner, N, nv, mode0, mode1,
                       Y0 + o0 * N * inner, Y1 ? Y1 + o0 * N * inner : nullptr);
  }
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

}  // namespace

int mesh_set_smem_limits() {
  const int lim = 160 * 1024;
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&dft_axis_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&dft_final_kernel<1, false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&dft_final_kernel<1, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&dft_final_kernel<2, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&dft_final_kernel<3, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  return RTB_OK;
}

int mesh_plan(const rtb_model* m, const size_t* kmesh, MeshPlan& p) {
  p.dim = m->dim;
  for (int d = 0; d < 3; ++d) {
    p.lo[d] = 0;
    p.L[d] = 1;
    p.N[d] = d < m->dim ? static_cast<int>(kmesh[d]) : 1;
  }
  for (int d = 0; d < m->dim; ++d) {
    long long lo = 0, hi = 0;
    for (int r = 0; r < m->nR; ++r) {
      lo = std::min<long long>(lo, m->hamR[r * m->dim + d]);
      hi = std::max<long long>(hi, m->hamR[r * m->dim + d]);
    }
    p.lo[d] = static_cast<int>(lo);
    p.L[d] = static_cast<int>(hi - lo + 1);
  }
  const size_t n2 = static_cast<size_t>(m->nsta) * m->nsta;
  p.box = static_cast<size_t>(p.L[0]) * p.L[1] * p.L[2];
  p.P = static_cast<size_t>(p.N[1]) * p.N[2];
  // element counts (complex) of the box and the pass intermediates
  p.box_elems = p.box * n2;
  p.ga_elems = p.dim == 3 ? 2 * static_cast<size_t>(p.L[0]) * p.L[1] * p.N[2] * n2 : 0;
  p.gb_elems = static_cast<size_t>(p.dim) * p.L[0] * p.P * n2;  // G0 (+G1, G2)
  // profitable when the box is not much sparser than the stored R set and the
  // passes fit comfortably next to the pipeline buffers
  const size_t bytes = 16 * (p.box_elems + p.ga_elems + p.gb_elems);
  p.ok = p.box <= 4 * static_cast<size_t>(m->nR) + 64 && bytes <= (size_t(96) << 30);
  for (int d = 0; d < 3; ++d) p.ok = p.ok && p.L[d] <= 32 && p.N[d] <= 4096;
  return RTB_OK;
}

void mesh_box_index(const rtb_model* m, const MeshPlan& p, std::vector<int>& idx) {
  idx.resize(static_cast<size_t>(m->nR));
  for (int r = 0; r < m->nR; ++r) {
    size_t f = 0;
    for (int d = 0; d < 3; ++d) {
      const int rd = d < m->dim ? static_cast<int>(m->hamR[r * m->dim + d]) : 0;
      f = f * p.L[d] + (rd - p.lo[d]);
    }
    idx[r] = static_cast<int>(f);
  }
}

// Passes A and B for the whole mesh (plus the box scatter).
//   work: >= box_elems + ga_elems complex scratch; G: gb_elems complex output
int mesh_prepare(hipStream_t st, const rtb_model* m, const MeshPlan& p, int* d_boxidx, double* work, double* G,
                 bool want_v) {
  std::vector<int> idx;
  mesh_box_index(m, p, idx);
  RTB_HIP(hipMemcpyAsync(d_boxidx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, st));
  return mesh_prepare_dev(st, m, p, d_boxidx, work, G, want_v);
}

int mesh_prepare_dev(hipStream_t st, const rtb_model* m, const MeshPlan& p, const int* d_boxidx, double* work,
                     double* G, bool want_v) {
  const size_t n2 = static_cast<size_t>(m->nsta) * m->nsta;
  double2* box = reinterpret_cast<double2*>(work);
  double2* GA = box + p.box_elems;
  double2* GB = reinterpret_cast<double2*>(G);
  RTB_TRY(zero_ints(st, reinterpret_cast<int*>(box), p.box_elems * (sizeof(double2) / sizeof(int))));
  const size_t tot = static_cast<size_t>(m->nR) * n2;
  hipLaunchKernelGGL(scatter_box_kernel, dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const double2*>(m->d_ham), d_boxidx, m->nR, n2, box);
  RTB_HIP(hipGetLastError());
  const size_t gsz = static_cast<size_t>(p.L[0]) * p.P * n2;  // one G array
  if (p.dim == 3) {
    // pass A over R3: [L1 L2][L3][n2] -> [L1 L2][N3][n2] (unweighted, R3-weighted)
    RTB_TRY(launch_axis(st, box, static_cast<size_t>(p.L[0]) * p.L[1], p.L[2], p.lo[2], n2, p.N[2],
                        want_v ? 2 : 1, 0, 1, GA, want_v ? GA + p.ga_elems / 2 : nullptr));
    // pass B over R2: [L1][L2][N3 n2] -> [L1][N2][N3 n2]
    RTB_TRY(launch_axis(st, GA, p.L[0], p.L[1], p.lo[1], static_cast<size_t>(p.N[2]) * n2, p.N[1],
                        want_v ? 2 : 1, 0, 1, GB, want_v ? GB + gsz : nullptr));
    if (want_v)
      RTB_TRY(launch_axis(st, GA + p.ga_elems / 2, p.L[0], p.L[1], p.lo[1], static_cast<size_t>(p.N[2]) * n2,
                          p.N[1], 1, 0, 0, GB + 2 * gsz, nullptr));
  } else if (p.dim == 2) {
    // pass B over R2: [L1][L2][n2] -> [L1][N2][n2]
    RTB_TRY(launch_axis(st, box, p.L[0], p.L[1], p.lo[1], n2, p.N[1], want_v ? 2 : 1, 0, 1, GB,
                        want_v ? GB + gsz : nullptr));
  } else {
    hipLaunchKernelGGL(copy_kernel, dim3(static_cast<unsigned>((p.box_elems + 255) / 256)), dim3(256), 0, st,
                       p.box_elems, box, GB);
    RTB_HIP(hipGetLastError());
  }
  return RTB_OK;
}

int mesh_final(hipStream_t st, const rtb_model* m, const MeshPlan& p, const double* G, const FourierArgs& a,
               const double* dirs, size_t p0, size_t np_pairs, double* H, double* V) {
  const size_t n2 = static_cast<size_t>(m->nsta) * m->nsta;
  const size_t gsz = static_cast<size_t>(p.L[0]) * p.P * n2;
  FinalArgs f;
  f.dim = m->dim;
  f.n = m->nsta;
  f.L1 = p.L[0];
  f.lo1 = p.lo[0];
  f.N1 = p.N[0];
  f.N2 = p.N[1];
  f.N3 = p.N[2];
  f.P = p.P;
  const double2* GB = reinterpret_cast<const double2*>(G);
  f.G0 = GB;
  f.G1 = m->dim >= 2 ? GB + gsz : nullptr;
  f.G2 = m->dim >= 3 ? GB + 2 * gsz : nullptr;
  f.np = a.np;
  f.combine = a.combine && a.np == 2;
  for (int q = 0; q < 2; ++q)
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      if (q < a.np && d < m->dim)
        for (int j = 0; j < m->dim; ++j) s += m->lat[d * m->dim + j] * dirs[q * m->dim + j];  // (lat w)_d
      f.alpha[q][d] = s;
    }
  f.tau = a.tau;
  f.wtau = a.wtau;
  f.gauge = a.gauge;
  const size_t smem = sizeof(double2) * (static_cast<size_t>(p.N[0]) + 3 * static_cast<size_t>(p.L[0]) * kTWF);
  const size_t tiles = (n2 + kTWF - 1) / kTWF;
  const size_t vk = f.combine ? 1 : 2;  // V matrices per k
  for (size_t q0 = 0; q0 < np_pairs; q0 += 65535) {
    const size_t qc = np_pairs - q0 < 65535 ? np_pairs - q0 : 65535;
    const dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(qc));
    double2* Hq = reinterpret_cast<double2*>(H) + q0 * p.N[0] * n2;
    double2* Vq = V ? reinterpret_cast<double2*>(V) + q0 * p.N[0] * vk * n2 : nullptr;
    if (!f.np) {
      hipLaunchKernelGGL((dft_final_kernel<1, false>), grid, dim3(256), smem, st, f, p0 + q0, Hq, Vq);
    } else if (f.dim == 1) {
      hipLaunchKernelGGL((dft_final_kernel<1, true>), grid, dim3(256), smem, st, f, p0 + q0, Hq, Vq);
    } else if (f.dim == 2) {
      hipLaunchKernelGGL((dft_final_kernel<2, true>), grid, dim3(256), smem, st, f, p0 + q0, Hq, Vq);
    } else {
      hipLaunchKernelGGL((dft_final_kernel<3, true>), grid, dim3(256), smem, st, f, p0 + q0, Hq, Vq);
    }
  }
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

}  // namespace rtb
