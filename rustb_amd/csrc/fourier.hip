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
#include <algorithm>

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

// The epilogue of one k-point kv and element e = (m, c) of its matrices:
// the Atom-gauge factor conj(phi_m) phi_c, the tau term of the projected
// velocities, the stores.  ho: complex index of element e of H(k); vo: of
// element e of the first V matrix of k (NP matrices per k, or one holding
// V0 + i V1 with a.combine).
template <int NP>
__device__ inline void store_elem(const FourierArgs& a, const double* kv, int m, int c, double hr, double hi,
                                  const double* vr, const double* vi, double* __restrict__ H, size_t ho,
                                  double* __restrict__ V, size_t vo, size_t n2) {
  double gr = 1.0, gi = 0.0;
  if (a.gauge == RTB_GAUGE_ATOM) {
    // conj(phi_m) phi_c = exp(i 2 pi k.(tau_c - tau_m)); each phase as in the
    // reference, exp(i 2 pi k.tau) per orbital, multiplied
    const double* tm = a.tau + static_cast<size_t>(m) * 3;
    const double* tc = a.tau + static_cast<size_t>(c) * 3;
    double cm, sm, cc, sc;
    phase_of(tm[0] * kv[0] + tm[1] * kv[1] + tm[2] * kv[2], cm, sm);
    phase_of(tc[0] * kv[0] + tc[1] * kv[1] + tc[2] * kv[2], cc, sc);
    gr = cm * cc + sm * sc;
    gi = cm * sc - sm * cc;
  }
  H[2 * ho] = hr * gr - hi * gi;
  H[2 * ho + 1] = hr * gi + hi * gr;
  double cr = 0.0, ci = 0.0;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    // i * acc  (+ i (w.tau_c - w.tau_m) H_lat in the Atom gauge)
    double xr = -vi[p], xi = vr[p];
    if (a.gauge == RTB_GAUGE_ATOM) {
      const double diff = a.wtau[static_cast<size_t>(p) * a.n + c] - a.wtau[static_cast<size_t>(p) * a.n + m];
      xr += -diff * hi;
      xi += diff * hr;
      const double yr = xr * gr - xi * gi, yi = xr * gi + xi * gr;
      xr = yr;
      xi = yi;
    }
    if (NP == 2 && a.combine) {  // V0 + i V1
      cr += p == 0 ? xr : -xi;
      ci += p == 0 ? xi : xr;
    } else {
      const size_t o = vo + static_cast<size_t>(p) * n2;
      V[2 * o] = xr;
      V[2 * o + 1] = xi;
    }
  }
  if (NP == 2 && a.combine) {
    V[2 * vo] = cr;
    V[2 * vo + 1] = ci;
  }
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
#pragma unroll
  for (int kk = 0; kk < TK; ++kk) {
    const int kg = k0 + kk;
    if (kg >= nk) break;
    double pr[NP > 0 ? NP : 1], pi[NP > 0 ? NP : 1];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      pr[p] = vr[p][kk];
      pi[p] = vi[p][kk];
    }
    const size_t ho = static_cast<size_t>(kg) * n2 + e;
    const size_t vo = static_cast<size_t>(kg) * (NP == 2 && a.combine ? 1 : NP) * n2 + e;
    store_elem<NP>(a, Ks[kk], m, c, hr[kk], hi[kk], pr, pi, H, ho, V, vo, n2);
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

// ---- many models in one launch: a persistent work queue of tiles ----
//
// A tile is TK-deep k-point groups of one model on the uniform mesh (k-points
// generated from flat gen_kmesh indices, first axis outermost) times a block
// of matrix elements.  With n^2 <= 256 a workgroup holds G = min(256 / n^2,
// kQueueSlots / TK) groups of TK k-points over all n^2 elements (thread t:
// element t % n^2, group t / n^2), so small models fill the workgroup; above,
// G = 1 and the tile is one 256-element block.  Workgroups take tile indices from
// one atomic counter in the order of the table (the host sorts it by cost,
// longest first) until the table is exhausted, so every workgroup exits once
// the counter passes the tile count.  Per (R, element, k) the arithmetic is
// the dense kernel's: the same phases, stored R order and epilogue.
template <int TK>
__global__ void __launch_bounds__(kBlock) fourier_queue_kernel(const QueueModel* __restrict__ models,
                                                               const QueueTile* __restrict__ tiles, int ntiles,
                                                               unsigned* __restrict__ next) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* Ur = reinterpret_cast<double*>(smem_raw);  // [ks][kRC], ks = G * TK of the launch's largest G
  __shared__ double W[2][kRC];
  __shared__ double Ks[kQueueSlots][3];
  __shared__ int s_tile;
  for (;;) {
    __syncthreads();  // the previous tile's reads of the shared tables are done
    if (threadIdx.x == 0) s_tile = static_cast<int>(atomicAdd(next, 1u));
    __syncthreads();
    const int t = s_tile;
    if (t >= ntiles) return;
    const QueueTile tile = tiles[t];
    const QueueModel& md = models[tile.model];
    const FourierArgs& a = md.a;
    const int n2 = a.n * a.n;
    const int EB = n2 <= kBlock ? n2 : kBlock;
    const int G = n2 <= kBlock ? min(kBlock / n2, kQueueSlots / TK) : 1;
    const int ks = G * TK;
    double* Ui = Ur + static_cast<size_t>(ks) * kRC;
    const int grp = threadIdx.x / EB;
    const int e = tile.e0 + threadIdx.x % EB;
    const bool live = grp < G && e < n2;
    const int kn = tile.cnt;  // <= ks
    if (threadIdx.x < kn) {
      // gen_kmesh (kpoints.rs:49-87): first axis outermost; k_d = i_d / N_d
      long long f = tile.f0 + threadIdx.x;
      long long x[3] = {0, 0, 0};
      for (int d = a.dim - 1; d >= 0; --d) {
        x[d] = f % md.N[d];
        f /= md.N[d];
      }
      for (int d = 0; d < 3; ++d)
        Ks[threadIdx.x][d] = d < a.dim ? static_cast<double>(x[d]) / static_cast<double>(md.N[d]) : 0.0;
    }
    __syncthreads();
    double hr[TK], hi[TK], vr[2][TK], vi[2][TK];
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      hr[kk] = hi[kk] = 0.0;
      vr[0][kk] = vi[0][kk] = vr[1][kk] = vi[1][kk] = 0.0;
    }
    const double2* ham2 = reinterpret_cast<const double2*>(a.ham);
    for (int R0 = 0; R0 < a.nR; R0 += kRC) {
      __syncthreads();
      const int rc = min(kRC, a.nR - R0);
      for (int q = threadIdx.x; q < ks * kRC; q += kBlock) {
        const int kk = q / kRC, r = q % kRC;
        double c = 0.0, s = 0.0;
        if (r < rc && kk < kn) {
          const double* Rv = a.R + static_cast<size_t>(R0 + r) * 3;
          // same association as the reference: R0*k0 + R1*k1 + R2*k2
          const double arg = Rv[0] * Ks[kk][0] + Rv[1] * Ks[kk][1] + Rv[2] * Ks[kk][2];
          phase_of(arg, c, s);
        }
        Ur[q] = c;
        Ui[q] = s;
      }
      if (threadIdx.x < 2 * kRC) {
        const int p = threadIdx.x / kRC, r = threadIdx.x % kRC;
        W[p][r] = r < rc ? a.wR[static_cast<size_t>(p) * a.nR + R0 + r] : 0.0;
      }
      __syncthreads();
      if (live) {
        const double* Urg = Ur + static_cast<size_t>(grp) * TK * kRC;
        const double* Uig = Ui + static_cast<size_t>(grp) * TK * kRC;
        for (int r = 0; r < rc; ++r) {
          const double2 h = ham2[static_cast<size_t>(R0 + r) * n2 + e];
          const double w0 = W[0][r], w1 = W[1][r];
#pragma unroll
          for (int kk = 0; kk < TK; ++kk) {
            const double ur = Urg[kk * kRC + r], ui = Uig[kk * kRC + r];
            const double tr = ur * h.x - ui * h.y;
            const double ti = ur * h.y + ui * h.x;
            hr[kk] += tr;
            hi[kk] += ti;
            vr[0][kk] += w0 * tr;
            vi[0][kk] += w0 * ti;
            vr[1][kk] += w1 * tr;
            vi[1][kk] += w1 * ti;
          }
        }
      }
    }
    if (!live) continue;
    const int m = e / a.n, c = e % a.n;
    const size_t vper = a.combine ? 1 : 2;
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      const int kq = grp * TK + kk;
      if (kq >= kn) break;
      const double pr[2] = {vr[0][kk], vr[1][kk]}, pi[2] = {vi[0][kk], vi[1][kk]};
      store_elem<2>(a, Ks[kq], m, c, hr[kk], hi[kk], pr, pi, tile.H, static_cast<size_t>(kq) * n2 + e, tile.V,
                    static_cast<size_t>(kq) * vper * n2 + e, static_cast<size_t>(n2));
    }
  }
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

size_t fourier_queue_smem(int tk, int max_groups) {
  return 2 * sizeof(double) * static_cast<size_t>(tk) * max_groups * kRC;
}

int fourier_queue_groups(int n, int tk) {
  const int n2 = n * n;
  return n2 <= kBlock ? std::min(kBlock / n2, kQueueSlots / tk) : 1;
}

int fourier_queue(hipStream_t st, int tk, const QueueModel* models, const QueueTile* tiles, int ntiles,
                  int max_groups, unsigned* next, int workgroups) {
  if (ntiles == 0) return RTB_OK;
  if (tk < 1 || max_groups < 1 || max_groups * tk > kQueueSlots || workgroups < 1)
    return fail(RTB_ERR_INVALID_ARGUMENT, "work queue: bad launch shape");
  const size_t smem = fourier_queue_smem(tk, max_groups);
  switch (tk) {
    case 1:
      hipLaunchKernelGGL(fourier_queue_kernel<1>, dim3(workgroups), dim3(kBlock), smem, st, models, tiles, ntiles,
                         next);
      break;
    case 2:
      hipLaunchKernelGGL(fourier_queue_kernel<2>, dim3(workgroups), dim3(kBlock), smem, st, models, tiles, ntiles,
                         next);
      break;
    case 4:
      hipLaunchKernelGGL(fourier_queue_kernel<4>, dim3(workgroups), dim3(kBlock), smem, st, models, tiles, ntiles,
                         next);
      break;
    case 8:
      hipLaunchKernelGGL(fourier_queue_kernel<8>, dim3(workgroups), dim3(kBlock), smem, st, models, tiles, ntiles,
                         next);
      break;
    default:
      return fail(RTB_ERR_INVALID_ARGUMENT, "work queue: k-points per thread must be 1, 2, 4 or 8");
  }
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int fourier_queue_occupancy(int tk, int max_groups, int* per_cu) {
  const size_t smem = fourier_queue_smem(tk, max_groups);
  int b = 0;
  switch (tk) {
    case 1:
      RTB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fourier_queue_kernel<1>, kBlock, smem));
      break;
    case 2:
      RTB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fourier_queue_kernel<2>, kBlock, smem));
      break;
    case 4:
      RTB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fourier_queue_kernel<4>, kBlock, smem));
      break;
    case 8:
      RTB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fourier_queue_kernel<8>, kBlock, smem));
      break;
    default:
      return fail(RTB_ERR_INVALID_ARGUMENT, "work queue: k-points per thread must be 1, 2, 4 or 8");
  }
  *per_cu = std::max(b, 1);
  return RTB_OK;
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
