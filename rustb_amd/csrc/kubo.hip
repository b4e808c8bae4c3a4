// Kubo band sum (response/traits.rs:160-200), occupation reductions
// (thermodynamics.rs:86-118, response/linear/mod.rs:204-221) and the
// Gaussian DOS (model_physics.rs:179-226), FP64.
#include "eigh.hpp"
#include "gpu_team.hpp"

namespace rtb {

namespace {

// helpers.rs:11-39 spin matrix S = sigma/2 (x) I_norb; J = (S v + v S) / 2,
// plus i w when w is given (the combined operator of the Kubo product)
__global__ void spin_current_kernel(int norb, size_t nb, int dir, const double2* __restrict__ v,
                                    size_t v_stride, const double2* __restrict__ w, double2* __restrict__ J) {
  const int n = 2 * norb;
  const size_t total = nb * n * n;
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= total) return;
  const size_t b = g / (static_cast<size_t>(n) * n);
  const int e = static_cast<int>(g % (static_cast<size_t>(n) * n));
  const int i = e / n, j = e % n;
  const double2* vb = v + b * v_stride;
  auto at = [&](int r, int c) { return vb[static_cast<size_t>(r) * n + c]; };
  double2 sv = make_double2(0.0, 0.0), vs = make_double2(0.0, 0.0);
  const int ip = i < norb ? i + norb : i - norb;  // spin partner of row i
  const int jp = j < norb ? j + norb : j - norb;
  if (dir == RTB_SPIN_Z) {
    double si = i < norb ? 0.5 : -0.5, sj = j < norb ? 0.5 : -0.5;
    double2 x = at(i, j);
    sv = make_double2(si * x.x, si * x.y);
    vs = make_double2(sj * x.x, sj * x.y);
  } else if (dir == RTB_SPIN_X) {
    double2 x = at(ip, j), y = at(i, jp);
    sv = make_double2(0.5 * x.x, 0.5 * x.y);
    vs = make_double2(0.5 * y.x, 0.5 * y.y);
  } else {  // Y: S[i][i+norb] = -i/2, S[i+norb][i] = +i/2
    double2 x = at(ip, j);
    double f = i < norb ? -0.5 : 0.5;  // (S v)_ij = f * i * v_{ip j}
    sv = make_double2(-f * x.y, f * x.x);
    double2 y = at(i, jp);
    double h = j < norb ? 0.5 : -0.5;  // (v S)_ij = v_{i jp} S_{jp j}; S_{j+norb, j} = +i/2
    vs = make_double2(-h * y.y, h * y.x);
  }
  double2 out = make_double2(0.5 * (sv.x + vs.x), 0.5 * (sv.y + vs.y));
  if (w) {
    const double2 x = w[b * v_stride + e];
    out.x -= x.y;
    out.y += x.x;
  }
  J[g] = out;
}

// Omega_j = sum_{m != j} (|C_mj|^2 - |C_jm|^2) / 2 / ((E_j - E_m)^2 + eta^2) for
// C = A + i B (A, B Hermitian): -2 Im(A_jm B_mj) = (|C_mj|^2 - |C_jm|^2) / 2,
// i.e. half the column sum minus the row sum of Q_rc = |C_rc|^2 /
// ((E_r - E_c)^2 + eta^2); the product kernel leaves those sums per 64 x 64
// tile (zgemm_kubo), summed here over the tiles.
__global__ void kubo_parts_kernel(int n, size_t nk, int tiles, const double* __restrict__ part,
                                  double* __restrict__ omega) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= nk * n) return;
  const size_t k = g / n;
  const int j = static_cast<int>(g % n);
  const double* pk = part + k * 2 * tiles * static_cast<size_t>(n);
  double col = 0.0, row = 0.0;
  for (int t = 0; t < tiles; ++t) {
    col += pk[static_cast<size_t>(t) * n + j];
    row += pk[static_cast<size_t>(tiles + t) * n + j];
  }
  omega[g] = 0.5 * (col - row);
}

__device__ inline double fermi(double e, double mu, double width) {
  // fermi_from_width (thermodynamics.rs:100-118)
  if (width == 0.0) return e < mu ? 1.0 : (e > mu ? 0.0 : 0.5);
  const double x = (e - mu) / width;
  if (x > 50.0) return 0.0;
  if (x < -50.0) return 1.0;
  return 1.0 / (1.0 + exp(x));
}

__device__ inline double width_of(double T) {
  return T <= 0.0 ? 0.0 : T * 8.617333262145e-5;  // thermodynamics.rs:10
}

__global__ void occupied_sum_kernel(int n, size_t nk, const double* __restrict__ omega,
                                    const double* __restrict__ E, double mu, double width,
                                    double* __restrict__ out) {
  const size_t k = blockIdx.x * 256ull + threadIdx.x;
  if (k >= nk) return;
  double s = 0.0;
  for (int b = 0; b < n; ++b) s += omega[k * n + b] * fermi(E[k * n + b], mu, width);
  out[k] = s;
}

// Layer-resolved occupied Berry curvature: per k-point and chemical
// potential mu_m, partial[k][m][l] = sum_j f(E_j, mu_m) Omega_j w_jl with
// w_jl = sum_{i in layer l} |U_ij|^2 for the unit eigenvector columns
// U = Z diag(scale) (scale may be null).  perm lists the rows of H by layer,
// rows of layer l at perm[bounds[l] .. bounds[l+1]).  One workgroup of 256
// threads per k-point, thread j per band (coalesced column walks); per-wave
// sums by DPP, then the waves in order.  n_mu * n_layers <= kLayerOut.
constexpr int kLayerOut = 1024;
__global__ void __launch_bounds__(256) layer_partials_kernel(int n, const double2* __restrict__ Z,
                                                            const double* __restrict__ scale,
                                                            const double* __restrict__ E,
                                                            const double* __restrict__ omega,
                                                            const double* __restrict__ mu, int n_mu, double width,
                                                            const int* __restrict__ perm,
                                                            const int* __restrict__ bounds, int n_layers,
                                                            double* __restrict__ partial) {
  __shared__ double wred[4][kLayerOut];
  const size_t k = blockIdx.x;
  const size_t nn = static_cast<size_t>(n) * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double2* Zk = Z + k * nn;
  for (int j0 = 0; j0 < n; j0 += 256) {
    const int j = j0 + tid;
    const bool live = j < n;
    const double sj = live && scale ? scale[k * n + j] : 1.0;
    const double om = live ? omega[k * n + j] : 0.0;
    const double ej = live ? E[k * n + j] : 0.0;
    for (int l = 0; l < n_layers; ++l) {
      double w = 0.0;
      if (live)
        for (int q = bounds[l]; q < bounds[l + 1]; ++q) {
          const double2 z = Zk[static_cast<size_t>(perm[q]) * n + j];
          w += z.x * z.x + z.y * z.y;
        }
      w *= sj * sj * om;
      for (int m = 0; m < n_mu; ++m) {
        const double v = wave_sum64(live ? w * fermi(ej, mu[m], width) : 0.0);
        if (lane == 0) wred[wave][m * n_layers + l] = (j0 == 0 ? 0.0 : wred[wave][m * n_layers + l]) + v;
      }
    }
  }
  __syncthreads();
  for (int o = tid; o < n_mu * n_layers; o += 256)
    partial[k * n_mu * n_layers + o] = (wred[0][o] + wred[1][o]) + (wred[2][o] + wred[3][o]);
}

constexpr int kHallK = 256;  // k-points per hall partial block

// deriv: weights -df/dE = f (1 - f) / width (fermi_derivative_from_width,
// thermodynamics.rs:122-126) instead of f
__global__ void hall_partials_kernel(int n, size_t nk, const double* __restrict__ omega,
                                     const double* __restrict__ E, const double* __restrict__ mu,
                                     int n_mu, double width, int deriv, double* __restrict__ partial) {
  // block = (k-block, mu); 256 threads over k, then a fixed tree
  __shared__ double red[256];
  const int imu = blockIdx.y;
  const size_t k = blockIdx.x * static_cast<size_t>(kHallK) + threadIdx.x;
  double s = 0.0;
  if (k < nk) {
    const double m = mu[imu];
    for (int b = 0; b < n; ++b) {
      const double f = fermi(E[k * n + b], m, width);
      s += omega[k * n + b] * (deriv ? f * (1.0 - f) / width : f);
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[static_cast<size_t>(blockIdx.x) * n_mu + imu] = red[0];
}

// Per k-point and chemical potential: out[k][m] = sum_n omega[k][n] f(E[k][n], mu_m)
// (the band sum of hall_partials, kept per k-point).
__global__ void kterm_kernel(int n, size_t nk, const double* __restrict__ omega, const double* __restrict__ E,
                             const double* __restrict__ mu, int n_mu, double width, double* __restrict__ out) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= nk * n_mu) return;
  const size_t k = g / n_mu;
  const double m = mu[g % n_mu];
  double s = 0.0;
  for (int b = 0; b < n; ++b) s += omega[k * n + b] * fermi(E[k * n + b], m, width);
  out[g] = s;
}

// out[s][m] = sum_{k in [start[s], start[s + 1])} term[k][m]: one workgroup per
// (segment, mu), strided partial sums then a fixed tree.
__global__ void __launch_bounds__(256) segment_sums_kernel(const double* __restrict__ term,
                                                           const long long* __restrict__ start, int n_mu,
                                                           double* __restrict__ out) {
  __shared__ double red[256];
  const int sg = blockIdx.x, m = blockIdx.y;
  const long long a = start[sg], b = start[sg + 1];
  double acc = 0.0;
  for (long long k = a + threadIdx.x; k < b; k += 256) acc += term[k * n_mu + m];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[static_cast<size_t>(sg) * n_mu + m] = red[0];
}

// out[j] = sum_r partial[r][j] for r < n_rows: one workgroup per column,
// strided partial sums then a fixed tree (for many rows).
__global__ void __launch_bounds__(256) column_sums_kernel(const double* __restrict__ partial, size_t n_rows, int m,
                                                          double* __restrict__ out) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (size_t r = threadIdx.x; r < n_rows; r += 256) acc += partial[r * m + j];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

constexpr int kDosE = 256;   // energies per block
constexpr int kDosEig = 2048; // eigenvalues per block

__global__ void dos_partials_kernel(size_t ne, const double* __restrict__ ev, double Emin, double dE, int En,
                                    double sigma0, double pi0, double* __restrict__ partial) {
  __shared__ double sev[kDosEig];
  const int j = blockIdx.x * kDosE + threadIdx.x;
  const size_t e0 = static_cast<size_t>(blockIdx.y) * kDosEig;
  const int cnt = static_cast<int>(ne - e0 < static_cast<size_t>(kDosEig) ? ne - e0 : kDosEig);
  for (int q = threadIdx.x; q < cnt; q += kDosE) sev[q] = ev[e0 + q];
  __syncthreads();
  if (j >= En) return;
  // linspace: a + step * i (ndarray::Array1::linspace)
  const double Ej = Emin + dE * j;
  double acc = 0.0;
  for (int q = 0; q < cnt; ++q) {
    const double A = (Ej - sev[q]) * sigma0;
    acc += exp(-A * A / 2.0) * sigma0 * pi0;
  }
  partial[static_cast<size_t>(blockIdx.y) * En + j] = acc;
}

__global__ void reduce_partials_kernel(const double* __restrict__ partial, int n_blk, int m, double scale,
                                       double* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= m) return;
  double s = 0.0;
  for (int b = 0; b < n_blk; ++b) s += partial[static_cast<size_t>(b) * m + j];
  out[j] = s * scale;
}

// Omega_n = sum_{m != n} -2 Im(A_nm B_mn) / ((E_n - E_m)^2 + eta^2) from A
// and B as the reference forms them (berry_curvature_at_impl,
// response/traits.rs:160-200), with no assumption that A and B are
// Hermitian; with `metric`, also g_n = sum_{m != n} Re(A_nm B_mn) / (...),
// the real part of the quantum geometric tensor (quantum_geometry.rs:168-186).  One k-point per workgroup (n <= 128): 32 x 32 tiles of A (row
// tile I, column tile J) and of B (tile (J, I)) through LDS, the products
// summed along each row of A.
constexpr int kOmTile = 32;
__global__ void __launch_bounds__(256) kubo_omega_exact_kernel(int n, const double2* __restrict__ A,
                                                               const double2* __restrict__ B,
                                                               const double* __restrict__ E, double eta,
                                                               double* __restrict__ omega,
                                                               double* __restrict__ metric) {
  __shared__ double2 sa[kOmTile][kOmTile + 1];
  __shared__ double2 sb[kOmTile][kOmTile + 1];
  __shared__ double se[kMaxN];
  const size_t k = blockIdx.x;
  const size_t nn = static_cast<size_t>(n) * n;
  A += k * nn;
  B += k * nn;
  const int tx = threadIdx.x % kOmTile, ty = threadIdx.x / kOmTile;  // 32 x 8
  for (int i = threadIdx.x; i < n; i += blockDim.x) se[i] = E[k * n + i];
  const int T = (n + kOmTile - 1) / kOmTile;
  const double eta2 = eta * eta;
  for (int I = 0; I < T; ++I) {
    double part[kOmTile / 8] = {};
    double gpart[kOmTile / 8] = {};
    for (int J = 0; J < T; ++J) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kOmTile / 8; ++q) {
        const int r = ty + 8 * q;
        const int ar = I * kOmTile + r, ac = J * kOmTile + tx;  // A[ar][ac]
        sa[r][tx] = ar < n && ac < n ? A[static_cast<size_t>(ar) * n + ac] : make_double2(0.0, 0.0);
        const int br = J * kOmTile + r, bc = I * kOmTile + tx;  // B[br][bc]
        sb[r][tx] = br < n && bc < n ? B[static_cast<size_t>(br) * n + bc] : make_double2(0.0, 0.0);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kOmTile / 8; ++q) {
        const int r = ty + 8 * q;
        const int row = I * kOmTile + r, col = J * kOmTile + tx;
        if (row < n && col < n && row != col) {
          const double2 a = sa[r][tx];  // A[row][col]
          const double2 b = sb[tx][r];  // B[col][row]
          const double im = a.x * b.y + a.y * b.x;
          const double d = se[row] - se[col];
          const double den = d * d + eta2;
          part[q] += -2.0 * im / den;
          if (metric) gpart[q] += (a.x * b.x - a.y * b.y) / den;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kOmTile / 8; ++q) {
      double v = part[q], g = gpart[q];
      for (int o = kOmTile / 2; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, kOmTile);
        g += __shfl_xor(g, o, kOmTile);
      }
      const int row = I * kOmTile + ty + 8 * q;
      if (tx == 0 && row < n) {
        omega[k * n + row] = v;
        if (metric) metric[k * n + row] = g;
      }
    }
  }
}

// Berry curvature dipole kernel per band (berry_curvature_dipole_n_onek,
// response/nonlinear/mod.rs:140-162): Omega^{c,f1}_n Re(A3_nn), and with
// `sym` the mean with Omega^{c,f2}_n Re(A2_nn) (FieldSymmetry::Symmetrized,
// mod.rs:284-300); A2 = U^H v_f1 U, A3 = U^H v_f2 U.
__global__ void dipole_combine_kernel(int n, size_t nk, const double2* __restrict__ A2,
                                      const double2* __restrict__ A3, const double* __restrict__ om1,
                                      const double* __restrict__ om2, int sym, double* __restrict__ out) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= nk * n) return;
  const size_t k = g / n;
  const int b = static_cast<int>(g % n);
  const size_t d = k * n * n + static_cast<size_t>(b) * n + b;
  const double first = om1[g] * A3[d].x;
  out[g] = sym ? (first + om2[g] * A2[d].x) * 0.5 : first;
}

}  // namespace

int spin_current(hipStream_t st, int norb, size_t nb, int spin_dir, const double* v, size_t v_stride,
                 const double* w, double* J) {
  const size_t n = 2 * static_cast<size_t>(norb);
  const size_t total = nb * n * n;
  if (total == 0) return RTB_OK;
  hipLaunchKernelGGL(spin_current_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st,
                     norb, nb, spin_dir, reinterpret_cast<const double2*>(v), v_stride / 2,
                     reinterpret_cast<const double2*>(w), reinterpret_cast<double2*>(J));
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int kubo_omega_from_parts(hipStream_t st, int n, size_t nk, const double* part, double* omega) {
  const size_t total = nk * n;
  if (total == 0) return RTB_OK;
  hipLaunchKernelGGL(kubo_parts_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, n, nk,
                     (n + 63) / 64, part, omega);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int kubo_omega_exact(hipStream_t st, int n, size_t nk, const double* A, const double* B, const double* E, double eta,
                     double* omega, double* metric) {
  if (nk == 0) return RTB_OK;
  if (n < 1 || n > kMaxN) return fail(RTB_ERR_UNSUPPORTED, "the band sum supports 1 <= nsta <= 256");
  if (nk > 0x7fffffffu) return fail(RTB_ERR_UNSUPPORTED, "too many k-points in one launch");
  hipLaunchKernelGGL(kubo_omega_exact_kernel, dim3(static_cast<unsigned>(nk)), dim3(256), 0, st, n,
                     reinterpret_cast<const double2*>(A), reinterpret_cast<const double2*>(B), E, eta, omega, metric);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int occupied_sum(hipStream_t st, int n, size_t nk, const double* omega, const double* E, double mu, double T,
                 double* out) {
  if (nk == 0) return RTB_OK;
  hipLaunchKernelGGL(occupied_sum_kernel, dim3(static_cast<unsigned>((nk + 255) / 256)), dim3(256), 0, st, n, nk,
                     omega, E, mu, T <= 0.0 ? 0.0 : T * 8.617333262145e-5, out);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int layer_partials(hipStream_t st, int n, size_t nk, const double* Z, const double* scale, const double* E,
                   const double* omega, const double* mu, int n_mu, double T, const int* perm, const int* bounds,
                   int n_layers, double* partial) {
  if (nk == 0) return RTB_OK;
  if (static_cast<long long>(n_mu) * n_layers > kLayerOut)
    return fail(RTB_ERR_UNSUPPORTED, "n_mu * n_layers above 1024");
  hipLaunchKernelGGL(layer_partials_kernel, dim3(static_cast<unsigned>(nk)), dim3(256), 0, st, n,
                     reinterpret_cast<const double2*>(Z), scale, E, omega, mu, n_mu,
                     T <= 0.0 ? 0.0 : T * 8.617333262145e-5, perm, bounds,
                     n_layers, partial);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int dipole_combine(hipStream_t st, int n, size_t nk, const double* A2, const double* A3, const double* om1,
                   const double* om2, bool sym, double* out) {
  const size_t total = nk * n;
  if (total == 0) return RTB_OK;
  hipLaunchKernelGGL(dipole_combine_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, n, nk,
                     reinterpret_cast<const double2*>(A2), reinterpret_cast<const double2*>(A3), om1, om2,
                     sym ? 1 : 0, out);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

namespace {
// Im K[b][i][j] = Im(A[b][i][j] B[b][j][i]) (compute_velocity_kernel,
// response/primitives.rs:106-111; only Im K enters Omega): one thread per
// element.
__global__ void cut_vertex_kernel(size_t total, int n, const double2* __restrict__ A, const double2* __restrict__ B,
                                  double* __restrict__ Ki) {
  const size_t g = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const size_t n2 = static_cast<size_t>(n) * n;
  const size_t b = g / n2, r = g % n2;
  const size_t i = r / n, j = r % n;
  const double2 a = A[g];
  const double2 c = B[b * n2 + j * n + i];
  Ki[g] = a.x * c.y + a.y * c.x;
}

// x[g] = |z[g]|^2 (complex z, real x; may not alias).
__global__ void abs2_kernel(size_t total, const double2* __restrict__ z, double* __restrict__ x) {
  const size_t g = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const double2 v = z[g];
  x[g] = v.x * v.x + v.y * v.y;
}
}  // namespace

int energy_cut_vertex(hipStream_t st, int n, size_t nk, const double* A, const double* B, double* Ki) {
  const size_t total = nk * n * n;
  if (total == 0) return RTB_OK;
  hipLaunchKernelGGL(cut_vertex_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, total, n,
                     reinterpret_cast<const double2*>(A), reinterpret_cast<const double2*>(B), Ki);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int abs2(hipStream_t st, size_t total, const double* z, double* x) {
  if (total == 0) return RTB_OK;
  hipLaunchKernelGGL(abs2_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, total,
                     reinterpret_cast<const double2*>(z), x);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int hall_blocks(size_t nk) { return static_cast<int>((nk + kHallK - 1) / kHallK); }

int hall_partials(hipStream_t st, int n, size_t nk, const double* omega, const double* E, const double* mu,
                  int n_mu, double T, double* partial, int n_blk, bool deriv) {
  if (nk == 0) return RTB_OK;
  hipLaunchKernelGGL(hall_partials_kernel, dim3(static_cast<unsigned>(n_blk), static_cast<unsigned>(n_mu)),
                     dim3(256), 0, st, n, nk, omega, E, mu, n_mu, T <= 0.0 ? 0.0 : T * 8.617333262145e-5,
                     deriv ? 1 : 0, partial);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int kterms(hipStream_t st, int n, size_t nk, const double* omega, const double* E, const double* mu, int n_mu,
           double T, double* out) {
  const size_t total = nk * static_cast<size_t>(n_mu);
  if (total == 0) return RTB_OK;
  hipLaunchKernelGGL(kterm_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, n, nk, omega,
                     E, mu, n_mu, T <= 0.0 ? 0.0 : T * 8.617333262145e-5, out);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int segment_sums(hipStream_t st, const double* term, const long long* start, int n_seg, int n_mu, double* out) {
  if (n_seg == 0 || n_mu == 0) return RTB_OK;
  if (n_mu > 65535) return fail(RTB_ERR_UNSUPPORTED, "at most 65535 chemical potentials");
  hipLaunchKernelGGL(segment_sums_kernel, dim3(static_cast<unsigned>(n_seg), static_cast<unsigned>(n_mu)), dim3(256),
                     0, st, term, start, n_mu, out);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int column_sums(hipStream_t st, const double* partial, size_t n_rows, int m, double* out) {
  if (m == 0) return RTB_OK;
  hipLaunchKernelGGL(column_sums_kernel, dim3(static_cast<unsigned>(m)), dim3(256), 0, st, partial, n_rows, m, out);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int dos_blocks(size_t ne) { return static_cast<int>((ne + kDosEig - 1) / kDosEig); }

int dos_partials(hipStream_t st, size_t ne, const double* evals, double Emin, double dE, int En, double sigma,
                 double* partial, int n_blk) {
  if (ne == 0) return RTB_OK;
  const double sigma0 = 1.0 / sigma;
  const double pi0 = 1.0 / sqrt(2.0 * M_PI);
  hipLaunchKernelGGL(dos_partials_kernel, dim3(static_cast<unsigned>((En + kDosE - 1) / kDosE),
                                               static_cast<unsigned>(n_blk)),
                     dim3(kDosE), 0, st, ne, evals, Emin, dE, En, sigma0, pi0, partial);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int reduce_partials(hipStream_t st, const double* partial, int n_blk, int m, double scale, double* out) {
  if (m == 0) return RTB_OK;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(static_cast<unsigned>((m + 255) / 256)), dim3(256), 0, st,
                     partial, n_blk, m, scale, out);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

}  // namespace rtb
