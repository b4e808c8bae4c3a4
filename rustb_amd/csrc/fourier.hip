// Bloch Fourier sum H(k) = sum_R e^{i 2 pi k.R} h_R and its projected
// k-derivatives (model_physics.rs:43-148, velocity.rs:470-731), FP64.
//
// Dense form: for a tile of TK k-points, every thread owns one matrix element
// e = (m, c) and accumulates, in stored R order, U_R h_R[e] for H and
// (w_p . R_cart) U_R h_R[e] for each projected direction p.  The phase table
// U[kk][r] of the tile is generated in LDS with the k.R argument reduced mod 1
// in FP64 before sincospi.  h_R[e] is read once per R and reused by the TK
// k-points of the tile (coalesced along e).  Epilogue: the Atom-gauge
// similarity conj(phi_m) phi_c and the tau term i (w.tau_c - w.tau_m) H_lat.
#include "eigh.hpp"

namespace rtb {

namespace {

constexpr int kRC = 32;     // R values staged per LDS round
constexpr int kBlock = 256; // threads per workgroup

__device__ inline void phase_of(double arg, double& c, double& s) {
  // exp(i 2 pi arg), arg reduced to [-1/2, 1/2] first
  double t = arg - rint(arg);
  sincospi(2.0 * t, &s, &c);
}

template <int TK, int NP>
__global__ void __launch_bounds__(kBlock) fourier_dense_kernel(FourierArgs a, const double* __restrict__ k,
                                                               int nk, double* __restrict__ H,
                                                               double* __restrict__ V) {
  __shared__ double Ur[TK][kRC];
  __shared__ double Ui[TK][kRC];
  __shared__ double W[NP > 0 ? NP : 1][kRC];
  __shared__ double Ks[TK][3];
  const int n2 = a.n * a.n;
  const int e = blockIdx.x * kBlock + threadIdx.x;
  const int k0 = blockIdx.y * TK;
  if (threadIdx.x < TK * 3) {
    int kk = threadIdx.x / 3, d = threadIdx.x % 3;
    Ks[kk][d] = (k0 + kk < nk) ? k[static_cast<size_t>(k0 + kk) * 3 + d] : 0.0;
  }
  double hr[TK], hi[TK];
  double vr[NP > 0 ? NP : 1][TK], vi[NP > 0 ? NP : 1][TK];
#pragma unroll
  for (int kk = 0; kk < TK; ++kk) {
    hr[kk] = 0.0;
    hi[kk] = 0.0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      vr[p][kk] = 0.0;
      vi[p][kk] = 0.0;
    }
  }
  const double2* ham2 = reinterpret_cast<const double2*>(a.ham);
  for (int R0 = 0; R0 < a.nR; R0 += kRC) {
    __syncthreads();
    const int rc = min(kRC, a.nR - R0);
    for (int q = threadIdx.x; q < TK * kRC; q += kBlock) {
      int kk = q / kRC, r = q % kRC;
      double c = 0.0, s = 0.0;
      if (r < rc && k0 + kk < nk) {
        const double* Rv = a.R + static_cast<size_t>(R0 + r) * 3;
        // same association as the reference: R0*k0 + R1*k1 + R2*k2
        double arg = Rv[0] * Ks[kk][0] + Rv[1] * Ks[kk][1] + Rv[2] * Ks[kk][2];
        phase_of(arg, c, s);
      }
      Ur[kk][r] = c;
      Ui[kk][r] = s;
    }
    for (int q = threadIdx.x; q < NP * kRC; q += kBlock) {
      int p = q / kRC, r = q % kRC;
      W[p][r] = r < rc ? a.wR[static_cast<size_t>(p) * a.nR + R0 + r] : 0.0;
    }
    __syncthreads();
    if (e < n2) {
      for (int r = 0; r < rc; ++r) {
        const double2 h = ham2[static_cast<size_t>(R0 + r) * n2 + e];
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
          const double ur = Ur[kk][r], ui = Ui[kk][r];
          const double tr = ur * h.x - ui * h.y;
          const double ti = ur * h.y + ui * h.x;
          hr[kk] += tr;
          hi[kk] += ti;
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            vr[p][kk] += W[p][r] * tr;
            vi[p][kk] += W[p][r] * ti;
          }
        }
      }
    }
  }
  if (e >= n2) return;
  const int m = e / a.n, c = e % a.n;
  const double* tm = a.tau + static_cast<size_t>(m) * 3;
  const double* tc = a.tau + static_cast<size_t>(c) * 3;
#pragma unroll
  for (int kk = 0; kk < TK; ++kk) {
    const int kg = k0 + kk;
    if (kg >= nk) break;
    double gr = 1.0, gi = 0.0;
    if (a.gauge == RTB_GAUGE_ATOM) {
      // conj(phi_m) phi_c = exp(i 2 pi k.(tau_c - tau_m)); each phase as in the
      // reference, exp(i 2 pi k.tau) per orbital, multiplied
      double cm, sm, cc, sc;
      phase_of(tm[0] * Ks[kk][0] + tm[1] * Ks[kk][1] + tm[2] * Ks[kk][2], cm, sm);
      phase_of(tc[0] * Ks[kk][0] + tc[1] * Ks[kk][1] + tc[2] * Ks[kk][2], cc, sc);
      gr = cm * cc + sm * sc;
      gi = cm * sc - sm * cc;
    }
    const size_t ho = (static_cast<size_t>(kg) * n2 + e) * 2;
    H[ho] = hr[kk] * gr - hi[kk] * gi;
    H[ho + 1] = hr[kk] * gi + hi[kk] * gr;
    double cr = 0.0, ci = 0.0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      // i * acc  (+ i (w.tau_c - w.tau_m) H_lat in the Atom gauge)
      double xr = -vi[p][kk], xi = vr[p][kk];
      if (a.gauge == RTB_GAUGE_ATOM) {
        const double diff = a.wtau[static_cast<size_t>(p) * a.n + c] - a.wtau[static_cast<size_t>(p) * a.n + m];
        xr += -diff * hi[kk];
        xi += diff * hr[kk];
        const double yr = xr * gr - xi * gi, yi = xr * gi + xi * gr;
        xr = yr;
        xi = yi;
      }
      if (NP == 2 && a.combine) {  // V0 + i V1
        cr += p == 0 ? xr : -xi;
        ci += p == 0 ? xi : xr;
      } else {
        const size_t vo = ((static_cast<size_t>(kg) * NP + p) * n2 + e) * 2;
        V[vo] = xr;
        V[vo + 1] = xi;
      }
    }
    if (NP == 2 && a.combine) {
      const size_t vo = (static_cast<size_t>(kg) * n2 + e) * 2;
      V[vo] = cr;
      V[vo + 1] = ci;
    }
  }
}

template <int TK, int NP>
int launch_dense(hipStream_t st, const FourierArgs& a, const double* k, size_t nk, double* H, double* V) {
  const int n2 = a.n * a.n;
  const size_t maxy = 65535;
  for (size_t off = 0; off < nk; off += maxy * TK) {
    size_t cnt = nk - off < maxy * TK ? nk - off : maxy * TK;
    dim3 grid((n2 + kBlock - 1) / kBlock, static_cast<unsigned>((cnt + TK - 1) / TK));
    hipLaunchKernelGGL((fourier_dense_kernel<TK, NP>), grid, dim3(kBlock), 0, st, a, k + off * 3,
                       static_cast<int>(cnt), H + off * static_cast<size_t>(n2) * 2,
                       V ? V + off * static_cast<size_t>(NP) * n2 * 2 : nullptr);
  }
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

__global__ void mesh_points_kernel(int dim, unsigned long long n1, unsigned long long n2,
                                   unsigned long long n3, unsigned long long k0,
                                   unsigned long long nk, double* __restrict__ k) {
  const unsigned long long i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= nk) return;
  unsigned long long f = k0 + i;
  // gen_kmesh (kpoints.rs:49-87): first axis outermost; k_d = i_d / N_d
  unsigned long long i3 = 0, i2 = 0, i1 = 0;
  if (dim == 3) {
    i3 = f % n3;
    f /= n3;
    i2 = f % n2;
    i1 = f / n2;
  } else if (dim == 2) {
    i2 = f % n2;
    i1 = f / n2;
  } else {
    i1 = f;
  }
  double* o = k + i * 3;
  o[0] = static_cast<double>(i1) / static_cast<double>(n1);
  o[1] = dim >= 2 ? static_cast<double>(i2) / static_cast<double>(n2) : 0.0;
  o[2] = dim >= 3 ? static_cast<double>(i3) / static_cast<double>(n3) : 0.0;
}

}  // namespace

int fourier_dense(hipStream_t st, const FourierArgs& a, const double* k, size_t nk, double* H, double* V) {
  if (nk == 0) return RTB_OK;
  switch (a.np) {
    case 0:
      return launch_dense<8, 0>(st, a, k, nk, H, nullptr);
    case 1:
      return launch_dense<8, 1>(st, a, k, nk, H, V);
    case 2:
      return launch_dense<8, 2>(st, a, k, nk, H, V);
    case 3:
      return launch_dense<8, 3>(st, a, k, nk, H, V);
    default:
      return fail(RTB_ERR_UNSUPPORTED, "at most 3 projected directions");
  }
}

// ---- the position-matrix term of the velocity (models with an r-matrix) ----
namespace {

// rproj[R][e] = sum_d w_d rmat[R][d][e] (e over the n^2 elements)
__global__ void rmatrix_project_kernel(size_t total, int dim, size_t n2, const double2* __restrict__ rmat, double w0,
                                       double w1, double w2, double2* __restrict__ out) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= total) return;
  const size_t r = g / n2, e = g % n2;
  const double w[3] = {w0, w1, w2};
  double re = 0.0, im = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double2 x = rmat[(r * dim + d) * n2 + e];
    re += w[d] * x.x;
    im += w[d] * x.y;
  }
  out[g] = make_double2(re, im);
}

__global__ void zero_diag_kernel(size_t nk, int n, double2* __restrict__ M) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= nk * n) return;
  const size_t k = g / n, i = g % n;
  M[k * n * n + i * n + i] = make_double2(0.0, 0.0);
}

// V[k][p] += i (T1 - T2)
__global__ void add_i_commutator_kernel(size_t total, size_t n2, int np, int p, const double2* __restrict__ T1,
                                        const double2* __restrict__ T2, double2* __restrict__ V) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= total) return;
  const size_t k = g / n2, e = g % n2;
  const double2 a = T1[g], b = T2[g];
  double2* v = V + (k * np + p) * n2 + e;
  const double cr = a.x - b.x, ci = a.y - b.y;
  *v = make_double2(v->x - ci, v->y + cr);
}

unsigned blocks_for(size_t total) { return static_cast<unsigned>((total + 255) / 256); }

}  // namespace

int rmatrix_velocity(hipStream_t st, const FourierArgs& a, const double* rmat, const double* dirs, const double* k,
                     size_t nk, const double* H, double* V, double* rproj, double* Rk, double* T1, double* T2) {
  if (nk == 0 || a.np == 0) return RTB_OK;
  const int n = a.n, dim = a.dim;
  const size_t n2 = static_cast<size_t>(n) * n;
  for (int p = 0; p < a.np; ++p) {
    double w[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; ++d) w[d] = dirs[p * dim + d];
    const size_t tot = static_cast<size_t>(a.nR) * n2;
    hipLaunchKernelGGL(rmatrix_project_kernel, dim3(blocks_for(tot)), dim3(256), 0, st, tot, dim, n2,
                       reinterpret_cast<const double2*>(rmat), w[0], w[1], w[2], reinterpret_cast<double2*>(rproj));
    RTB_HIP(hipGetLastError());
    // r_p(k): the same Bloch sum and gauge factor as H(k)
    FourierArgs b = a;
    b.np = 0;
    b.combine = 0;
    b.ham = rproj;
    RTB_TRY(fourier_dense(st, b, k, nk, Rk, nullptr));
    if (a.gauge == RTB_GAUGE_ATOM) {  // velocity.rs:395, 663: the diagonal of r_p is dropped
      hipLaunchKernelGGL(zero_diag_kernel, dim3(blocks_for(nk * n)), dim3(256), 0, st, nk, n,
                         reinterpret_cast<double2*>(Rk));
      RTB_HIP(hipGetLastError());
    }
    RTB_TRY(zgemm_batched(st, n, nk, H, n2, false, Rk, n2, T1, n2));
    RTB_TRY(zgemm_batched(st, n, nk, Rk, n2, false, H, n2, T2, n2));
    hipLaunchKernelGGL(add_i_commutator_kernel, dim3(blocks_for(nk * n2)), dim3(256), 0, st, nk * n2, n2, a.np, p,
                       reinterpret_cast<const double2*>(T1), reinterpret_cast<const double2*>(T2),
                       reinterpret_cast<double2*>(V));
    RTB_HIP(hipGetLastError());
  }
  return RTB_OK;
}

int mesh_points(hipStream_t st, int dim, const size_t* kmesh, size_t k0, size_t nk, double* k) {
  if (nk == 0) return RTB_OK;
  unsigned long long n1 = kmesh[0], n2 = dim >= 2 ? kmesh[1] : 1, n3 = dim >= 3 ? kmesh[2] : 1;
  hipLaunchKernelGGL(mesh_points_kernel, dim3(static_cast<unsigned>((nk + 255) / 256)), dim3(256), 0, st,
                     dim, n1, n2, n3, static_cast<unsigned long long>(k0), static_cast<unsigned long long>(nk), k);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

}  // namespace rtb
