// FP64 refinement of FP32 eigenpairs (mixed form).
//
// One Ogita-Aishima step (Ogita & Aishima, Japan J. Indust. Appl. Math. 35,
// 2018: iterative refinement for symmetric eigendecompositions) on FP64
// matrix cores lifts the FP32 eigenvectors to FP64 accuracy outside
// clusters; then, with HZ = H Z in FP64:
//   * every column is renormalised and its eigenvalue replaced by the FP64
//     Rayleigh quotient  lambda_j = Re(z_j^H H z_j) / (z_j^H z_j), whose error
//     is quadratic in the FP32 eigenvector error;
//   * inside clusters of close eigenvalues (where the FP32 vectors mix the
//     members), an FP64 Rayleigh-Ritz step: orthonormalise the cluster's
//     columns (Cholesky of the Gram matrix), diagonalise the projected
//     Hermitian matrix by cyclic Jacobi, rotate the columns.
#include "eigh.hpp"
#include "refine_core.hpp"

namespace rtb {

namespace {


__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// One workgroup of NT threads (64, 128 or 256 by n) per k-point: first the
// Rayleigh quotients, column scales and residuals from the per-tile sums;
// then the threads test every eigenvalue gap in parallel; a k-point with a
// cluster forms H z' of all the cluster's columns in one pass over each row
// of HZ (a thread per row), the cluster's Gram and projection sums with the
// threads over rows (wave sums, then the waves' partials in wave order),
// thread 0 solves the c x c problem (LDS), and the threads rotate the
// cluster's columns row by row, forming the rotated columns' residuals on
// the way.  Column j of Z (and of HZ = H Z) times scale[j] is the unit
// vector; rotated columns are written unit.  Last, the k-point's check (see
// RefineCheck).
template <int NT>
__global__ void __launch_bounds__(NT) cluster_rr_kernel(int n, double2* __restrict__ Z,
                                                        const double2* __restrict__ HZ,
                                                        const float2* __restrict__ F,
                                                        const double* __restrict__ part, int tiles,
                                                        double* __restrict__ E, double* __restrict__ scale,
                                                        double gap_rel, RefineCheck chk) {
  constexpr int NW = NT / 64;
  constexpr int NP = kMaxCluster * kMaxCluster * 4;  // per-wave partials of the Gram / projection sums
  __shared__ C2 lds[7][kMaxCluster][kMaxCluster];
  __shared__ double2 hzc[kMaxCluster][kMaxN];  // refined H z' of the cluster's columns
  __shared__ double ev[kMaxCluster];
  __shared__ double s_res[kMaxN];  // residual of unit column j
  __shared__ double wred[NW][NP];
  __shared__ int solved;
  const size_t k = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double* Ek = E + k * n;
  double* sk = scale + k * n;
  // the maximum over the workgroup, to every thread
  auto block_max = [&](double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    if (lane == 0) wred[wave][0] = x;
    __syncthreads();
    double m = wred[0][0];
#pragma unroll
    for (int w = 1; w < NW; ++w) m = fmax(m, wred[w][0]);
    __syncthreads();
    return m;
  };
  {
    // Rayleigh quotients, column scales and residuals (against the
    // Ogita-Aishima quotients) from the per-tile sums of the correction
    // product's epilogue
    const double* pk = part + k * tiles * 3 * static_cast<size_t>(n);
    for (int j = tid; j < n; j += NT) {
      double num = 0.0, den = 0.0, res = 0.0;
      for (int t = 0; t < tiles; ++t) {
        num += pk[(3 * t) * static_cast<size_t>(n) + j];
        den += pk[(3 * t + 1) * static_cast<size_t>(n) + j];
        res += pk[(3 * t + 2) * static_cast<size_t>(n) + j];
      }
      Ek[j] = num / den;
      sk[j] = 1.0 / sqrt(den);
      s_res[j] = sqrt(res / den);
    }
    __threadfence_block();
    __syncthreads();
  }
  double m = 0.0;
  for (int j = tid; j < n; j += NT) m = fmax(m, fabs(Ek[j]));
  m = block_max(m);
  const double thr = gap_rel * fmax(m, 1.0);
  int close = 0;
  for (int j = tid; j + 1 < n; j += NT) close |= (Ek[j + 1] - Ek[j] <= thr);
  if (__syncthreads_or(close)) {
    double2* Zk = Z + k * n * n;
    const double2* Hk = HZ + k * n * n;
    const float2* Fk = F + k * n * n;
    int j0 = 0;
    while (j0 < n) {  // the same cluster sequence in every thread
      const int j1 = cluster_end(n, Ek, j0, thr);
      const int c = j1 - j0 + 1;
      // (a run of more than kMaxCluster close eigenvalues is split; the
      // residuals of the columns at a split fail the check)
      if (c >= 2) {
        // H z'_b = HZ (e_b + (F' - I)_b) for the cluster's columns, in FP64:
        // one pass over row r of HZ for all of them
        for (int r = tid; r < n; r += NT) {
          const double2* hr = Hk + static_cast<size_t>(r) * n;
          double2 acc[kMaxCluster];
#pragma unroll
          for (int b = 0; b < kMaxCluster; ++b) acc[b] = b < c ? hr[j0 + b] : make_double2(0.0, 0.0);
          for (int q = 0; q < n; ++q) {
            const double2 h = hr[q];
            const float2* fq = Fk + static_cast<size_t>(q) * n + j0;
#pragma unroll
            for (int b = 0; b < kMaxCluster; ++b)
              if (b < c) {
                const float2 f = fq[b];
                acc[b].x += h.x * f.x - h.y * f.y;
                acc[b].y += h.x * f.y + h.y * f.x;
              }
          }
#pragma unroll
          for (int b = 0; b < kMaxCluster; ++b)
            if (b < c) hzc[b][r] = acc[b];
        }
        __syncthreads();
        for (int p = 0; p < c * c; ++p) {
          const int a = p / c, b = p % c;
          double gr = 0.0, gi = 0.0, mr = 0.0, mi = 0.0;
          for (int r = tid; r < n; r += NT) {
            const double2 za = Zk[static_cast<size_t>(r) * n + j0 + a];
            const double2 zb = Zk[static_cast<size_t>(r) * n + j0 + b];
            const double2 hb = hzc[b][r];
            gr += za.x * zb.x + za.y * zb.y;  // conj(za) zb
            gi += za.x * zb.y - za.y * zb.x;
            mr += za.x * hb.x + za.y * hb.y;  // conj(za) hb
            mi += za.x * hb.y - za.y * hb.x;
          }
          gr = wave_sum(gr);
          gi = wave_sum(gi);
          mr = wave_sum(mr);
          mi = wave_sum(mi);
          if (lane == 0) {
            wred[wave][4 * p] = gr;
            wred[wave][4 * p + 1] = gi;
            wred[wave][4 * p + 2] = mr;
            wred[wave][4 * p + 3] = mi;
          }
        }
        __syncthreads();
        if (tid < c * c) {
          // sums over the unit columns s_a z_a, s_b z_b (H z_b scales with z_b)
          const int a = tid / c, b = tid % c;
          double gr = wred[0][4 * tid], gi = wred[0][4 * tid + 1], mr = wred[0][4 * tid + 2],
                 mi = wred[0][4 * tid + 3];
#pragma unroll
          for (int w = 1; w < NW; ++w) {
            gr += wred[w][4 * tid];
            gi += wred[w][4 * tid + 1];
            mr += wred[w][4 * tid + 2];
            mi += wred[w][4 * tid + 3];
          }
          const double sab = sk[j0 + a] * sk[j0 + b];
          lds[0][a][b] = {gr * sab, gi * sab};
          lds[1][a][b] = {mr * sab, mi * sab};
        }
        __syncthreads();
        if (tid == 0) solved = cluster_rr_small(c, lds, ev) ? 1 : 0;
        __syncthreads();
        if (solved) {
          double res2[kMaxCluster];
#pragma unroll
          for (int q = 0; q < kMaxCluster; ++q) res2[q] = 0.0;
          for (int r = tid; r < n; r += NT) {
            double2* zr = Zk + static_cast<size_t>(r) * n + j0;
            C2 row[kMaxCluster], hrow[kMaxCluster];
#pragma unroll
            for (int q = 0; q < kMaxCluster; ++q)
              if (q < c) {
                const double sq = sk[j0 + q];
                row[q] = {zr[q].x * sq, zr[q].y * sq};
                hrow[q] = {hzc[q][r].x * sq, hzc[q][r].y * sq};
              }
#pragma unroll
            for (int bb = 0; bb < kMaxCluster; ++bb) {
              if (bb < c) {
                C2 acc = {0.0, 0.0}, hacc = {0.0, 0.0};
#pragma unroll
                for (int q = 0; q < kMaxCluster; ++q) {
                  if (q < c) {
                    const C2 t = cm(row[q], lds[6][q][bb]);
                    const C2 u = cm(hrow[q], lds[6][q][bb]);
                    acc.re += t.re;
                    acc.im += t.im;
                    hacc.re += u.re;
                    hacc.im += u.im;
                  }
                }
                zr[bb] = make_double2(acc.re, acc.im);
                const double rr = fma(-ev[bb], acc.re, hacc.re), ri = fma(-ev[bb], acc.im, hacc.im);
                res2[bb] += rr * rr + ri * ri;
              }
            }
          }
#pragma unroll
          for (int bb = 0; bb < kMaxCluster; ++bb)
            if (bb < c) {
              const double x = wave_sum(res2[bb]);
              if (lane == 0) wred[wave][bb] = x;
            }
          __syncthreads();  // every thread has read the old scales
          if (tid < c) {
            double r2 = wred[0][tid];
#pragma unroll
            for (int w = 1; w < NW; ++w) r2 += wred[w][tid];
            Ek[j0 + tid] = ev[tid];
            sk[j0 + tid] = 1.0;  // the rotated columns are unit vectors
            s_res[j0 + tid] = sqrt(r2);
          }
        } else {
          for (int q = tid; q < c; q += NT) s_res[j0 + q] = HUGE_VAL;
        }
        __threadfence_block();
        __syncthreads();
      }
      j0 = j1 + 1;
    }
  }
  __threadfence_block();
  __syncthreads();
  // the check: every eigenvalue finite, every residual within the bound
  double mx = 0.0, rmax = 0.0;
  int bad = 0;
  for (int j = tid; j < n; j += NT) {
    const double e = Ek[j];
    bad |= !isfinite(e);
    mx = fmax(mx, fabs(e));
  }
  mx = block_max(mx);
  const double bound = chk.res_tol * fmax(1.0, mx);
  for (int j = tid; j < n; j += NT) {
    const double r = s_res[j];
    bad |= !(r <= bound);
    if (isfinite(r)) rmax = fmax(rmax, r);
  }
  rmax = block_max(rmax);
  const int any_bad = __syncthreads_or(bad);
  if (tid == 0) {
    if (any_bad && chk.fail_count) {
      const int slot = atomicAdd(chk.fail_count, 1);
      chk.fail_list[slot] = static_cast<int>(k);
    }
    if (chk.max_res) atomicMax(chk.max_res, static_cast<unsigned long long>(__double_as_longlong(rmax)));
  }
}

__global__ void count_nonfinite_kernel(int n, size_t nk, const double* __restrict__ E, int* __restrict__ count) {
  const size_t k = blockIdx.x * 4ull + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= nk) return;
  int bad = 0;
  for (int j = lane; j < n; j += 64) bad |= !isfinite(E[k * n + j]);
  if (__any(bad) && lane == 0) atomicAdd(count, 1);
}

__global__ void scatter_resolved_kernel(int n, const int* __restrict__ list, const double* __restrict__ Ef,
                                        const double2* __restrict__ Zf, double* __restrict__ E,
                                        double2* __restrict__ Z, double* __restrict__ scale) {
  const size_t i = blockIdx.x;
  const size_t k = static_cast<size_t>(list[i]);
  const size_t n2 = static_cast<size_t>(n) * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    E[k * n + j] = Ef[i * n + j];
    scale[k * n + j] = 1.0;
  }
  for (size_t e = threadIdx.x; e < n2; e += blockDim.x) Z[k * n2 + e] = Zf[i * n2 + e];
}

__global__ void zero_ints_kernel(int* p, size_t count) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i < count) p[i] = 0;
}

__global__ void transpose_kernel(int n, size_t nk, const double2* __restrict__ Z, const double* __restrict__ scale,
                                 double2* __restrict__ out) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  const size_t n2 = static_cast<size_t>(n) * n;
  if (g >= nk * n2) return;
  const size_t k = g / n2, e = g % n2;
  const size_t r = e / n, c = e % n;  // out[k][r][c] = Z[k][c][r] * scale[k][r]
  const double s = scale ? scale[k * n + r] : 1.0;
  const double2 z = Z[k * n2 + c * n + r];
  out[g] = make_double2(z.x * s, z.y * s);
}

}  // namespace

int transpose_batched(hipStream_t st, int n, size_t nk, const double* Z, const double* scale, double* out) {
  const size_t tot = nk * static_cast<size_t>(n) * n;
  if (tot == 0) return RTB_OK;
  hipLaunchKernelGGL(transpose_kernel, dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0, st, n, nk,
                     reinterpret_cast<const double2*>(Z), scale, reinterpret_cast<double2*>(out));
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

// One FP64 Ogita-Aishima refinement step of the FP32 eigenpairs, then Rayleigh
// quotients and Rayleigh-Ritz inside clusters:
//   HZ = H Z, S = Z^H HZ, G = Z^H Z, lt_i = S_ii / G_ii (G_ii exact, from the
//   back-transformation; G is consumed in its product's epilogue),
//   F'_ij = (S_ij - lt_j G_ij) / (lt_j - lt_i)  (i != j, |lt_i - lt_j| > thr)
//         = -G_ij / 2                           (i != j, inside a cluster)
//   F'_ii = 1 + (1 - G_ii) / 2,
//   Z' = Z + Z (F' - I), HZ' = HZ + HZ (F' - I)  (the FP32 vector error enters
//   squared).  F' - I is of the size of that error, so its two products run
//   on the FP32 matrix cores: their rounding (~1e-7 relative to a correction
//   already ~1e-6 or smaller) stays below the FP64 level of the sums.  HZ'
//   is never stored: its product's epilogue forms the Rayleigh sums
//   Re(z'^H HZ') and z'^H z' per column, and the cluster kernel rebuilds
//   H z' for the few columns inside clusters.
// Buffers (each [nk][n][n] complex): H (consumed; F' - I in FP32), Z (in,
// FP32), W1 (HZ), W3 (the Rayleigh sums); the refined eigenvectors are left in
// W2 as columns that become unit vectors when multiplied by scale[k][j]
// ([nk][n]; on entry the squared column norms of Z).
int refine_eigenpairs(hipStream_t st, Timer* tmr, int n, size_t nk, double* H, const float* Z, double* W1,
                      double* W2, double* W3, double* E, double* scale, double gap_rel, const RefineCheck& chk) {
  if (nk == 0) return RTB_OK;
  const size_t n2 = static_cast<size_t>(n) * n;
  // FP64 products: HZ, S = Z^H HZ, and G = Z^H Z consumed by its epilogue,
  // which writes F' - I (Z read from its exact FP32 storage; the exact
  // diagonal of G comes in `scale`)
  float* F = reinterpret_cast<float*>(H);  // F' - I into H (H is dead after HZ)
  if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_GEMM64));
  RTB_TRY(zgemm(st, n, nk, zop(H, n2), false, zop(Z, n2), W1, n2, false, nullptr));
  RTB_TRY(zgemm(st, n, nk, zop(Z, n2), true, zop(W1, n2), W2, n2, true, nullptr));
  RTB_TRY(zgemm_gram_oa(st, n, nk, Z, W2, scale, F, gap_rel, E));  // E <- the quotients lt
  if (tmr)
    RTB_TRY(tmr->end(3, flop_zgemm(n, nk) + 2 * flop_zgemm_herm(n, nk),
                     mflop_zgemm(n, nk, false) + 2 * mflop_zgemm(n, nk, true)));
  // Z' = Z + Z (F' - I) and HZ' = HZ + HZ (F' - I) in one kernel; HZ' is
  // consumed by its epilogue, which leaves the Rayleigh sums of the Z'
  // columns per 64-row tile in W3
  const int tiles = (n + 63) / 64;
  double* part = W3;
  if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_GEMM32));
  RTB_TRY(zgemm_corrections_f32(st, n, nk, Z, W1, F, E, W2, part));  // both products, one pass over F
  if (tmr) RTB_TRY(tmr->end(1, 2 * flop_zgemm(n, nk), 2 * mflop_zgemm(n, nk, false)));
  if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_REFINE));
  // a thread per row of the k-point's matrices
  auto rr = [&](auto kernel, int nt) {
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(nk)), dim3(nt), 0, st, n, reinterpret_cast<double2*>(W2),
                       reinterpret_cast<const double2*>(W1), reinterpret_cast<const float2*>(F), part, tiles, E,
                       scale, gap_rel, chk);
  };
  if (n <= 64)
    rr(cluster_rr_kernel<64>, 64);
  else if (n <= 128)
    rr(cluster_rr_kernel<128>, 128);
  else
    rr(cluster_rr_kernel<256>, 256);
  RTB_HIP(hipGetLastError());
  if (tmr) RTB_TRY(tmr->end(1));
  return RTB_OK;
}

int count_nonfinite(hipStream_t st, int n, size_t nk, const double* E, int* count) {
  if (nk == 0) return RTB_OK;
  hipLaunchKernelGGL(count_nonfinite_kernel, dim3(static_cast<unsigned>((nk + 3) / 4)), dim3(256), 0, st, n, nk, E,
                     count);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int scatter_resolved(hipStream_t st, int n, int cnt, const int* list, const double* Ef, const double* Zf, double* E,
                     double* Z, double* scale) {
  if (cnt <= 0) return RTB_OK;
  hipLaunchKernelGGL(scatter_resolved_kernel, dim3(static_cast<unsigned>(cnt)), dim3(256), 0, st, n, list, Ef,
                     reinterpret_cast<const double2*>(Zf), E, reinterpret_cast<double2*>(Z), scale);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int zero_ints(hipStream_t st, int* p, size_t count) {
  if (count == 0) return RTB_OK;
  hipLaunchKernelGGL(zero_ints_kernel, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, st, p, count);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

}  // namespace rtb
