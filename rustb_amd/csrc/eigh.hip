// Batched Hermitian eigensolver kernels (gfx950).
//
//   eigh_tridiag_reg_kernel    one workgroup of NB * S threads per matrix:
//                              register-resident Householder reduction
//   eigh_trieig_kernel         one thread per eigenvalue, 256/n matrices per
//                              workgroup: bisection, then the twisted
//                              factorisations of the isolated eigenvalues
//   eigh_cluster_kernel        the same layout: inverse iteration for the
//                              members of eigenvalue clusters (a workgroup
//                              without clusters returns at once)
//   eigh_backtrans_reg_kernel  one workgroup per matrix: the eigenvectors of
//                              T (generated from the twisted factorisations)
//                              and Z = Q Y, register-resident
#include <algorithm>

#include "eigh.hpp"

#include "eigh_band.hpp"
#include "eigh_core.hpp"
#include "eigh_reg.hpp"
#include "gpu_team.hpp"

namespace rtb {

namespace {

// Register-resident tridiagonalization: one workgroup of NB * S threads per
// matrix (n <= NB), each thread holding NB / S entries of one row.
template <typename T, int NB, int S>
__global__ void __launch_bounds__(NB * S) eigh_tridiag_reg_kernel(int n, const double* __restrict__ H,
                                                                  T* __restrict__ Dg, T* __restrict__ Eg,
                                                                  cplx<T>* __restrict__ Vg,
                                                                  cplx<T>* __restrict__ taug) {
  __shared__ cplx<T> xc[NB], vp[NB], wp[NB], yp[S * NB];
  __shared__ T red[2 * GpuTeam::kRedStride];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  const size_t nn = static_cast<size_t>(n) * n;
  RegSmem<T> sm{xc, vp, wp, yp, red, nullptr};
  tridiag_reg<T, NB, S>(tm, n, H + mat * nn * 2, sm, Dg + mat * n, Eg + mat * n, Vg ? Vg + mat * nn : nullptr,
                        taug ? taug + mat * n : nullptr);
}

// Register-resident back-transformation Z = Q Y: one workgroup per matrix,
// each thread holding NB / S entries of one column of Z.  twist / e: the
// trieig outputs ([nmat][n] each), Ts: its scratch ([nmat][3][n][n]).  The
// FP32 kernels are held to 3 waves per SIMD (168 registers): the few spills
// this costs sit in the eigenvector generation before the reflector loop,
// and the third workgroup per CU hides the per-reflector barrier (FP32,
// n = 128: ~10% faster than 217 registers at 2 waves).
template <typename T, int NB, int S, typename TO>
__global__ void __launch_bounds__(NB * S) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1, 8))) eigh_backtrans_reg_kernel(int n, const T* __restrict__ Y,
                                                                    const T* __restrict__ twist,
                                                                    const T* __restrict__ e,
                                                                    const T* __restrict__ Ts,
                                                                    const cplx<T>* __restrict__ Vg,
                                                                    const cplx<T>* __restrict__ taug,
                                                                    TO* __restrict__ Zout, int rows_out,
                                                                    double* __restrict__ gdiag) {
  __shared__ cplx<T> vp[3 * NB], yp[2 * S * NB];
  __shared__ T es[NB];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  const size_t nn = static_cast<size_t>(n) * n;
  RegSmem<T> sm{nullptr, vp, nullptr, yp, nullptr, es};
  backtrans_reg<T, NB, S>(tm, n, Y + mat * nn, n, twist + mat * n, e + mat * n, Ts + mat * nn * 3, Vg + mat * nn,
                          taug + mat * n, sm, Zout + mat * nn * 2, rows_out, gdiag ? gdiag + mat * n : nullptr);
}

// Slice counts (threads per row or column = S).  Tridiagonalisation at
// n = 128: 4 slices (32 complex entries per thread, 8 waves per matrix, 4
// waves per SIMD) beat 2 slices (64 entries, 2 waves per SIMD) by ~20% in
// FP32; the back-transformation is indifferent and keeps 2 in FP32.
template <typename T, int NB>
struct RegShape;
template <> struct RegShape<float, 32> { static constexpr int S = 2, Sbt = 2; };
template <> struct RegShape<float, 64> { static constexpr int S = 2, Sbt = 2; };
template <> struct RegShape<float, 128> { static constexpr int S = 4, Sbt = 2; };
template <> struct RegShape<double, 32> { static constexpr int S = 2, Sbt = 2; };
template <> struct RegShape<double, 64> { static constexpr int S = 2, Sbt = 2; };
template <> struct RegShape<double, 128> { static constexpr int S = 4, Sbt = 4; };

template <typename T, int NB>
void launch_tridiag_reg(hipStream_t st, int n, size_t nmat, const double* H, T* Dg, T* Eg, cplx<T>* Vg,
                        cplx<T>* taug) {
  constexpr int S = RegShape<T, NB>::S;
  hipLaunchKernelGGL((eigh_tridiag_reg_kernel<T, NB, S>), dim3(static_cast<unsigned>(nmat)), dim3(NB * S), 0, st, n,
                     H, Dg, Eg, Vg, taug);
}

// nb: the block rows (32, 64 or 128, at least n), or 0 for the smallest
// that holds n.
inline int reg_block(int n, int nb) { return nb > 0 ? nb : (n <= 32 ? 32 : (n <= 64 ? 64 : 128)); }

template <typename T>
void tridiag_reg_dispatch(hipStream_t st, int n, size_t nmat, const double* H, T* Dg, T* Eg, cplx<T>* Vg,
                          cplx<T>* taug, int nb) {
  nb = reg_block(n, nb);
  if (nb == 32) {
    launch_tridiag_reg<T, 32>(st, n, nmat, H, Dg, Eg, Vg, taug);
  } else if (nb == 64) {
    launch_tridiag_reg<T, 64>(st, n, nmat, H, Dg, Eg, Vg, taug);
  } else {
    launch_tridiag_reg<T, 128>(st, n, nmat, H, Dg, Eg, Vg, taug);
  }
}

template <typename T, int NB, typename TO>
void launch_backtrans_reg(hipStream_t st, int n, size_t nmat, const T* Y, const T* DE, const T* Ts,
                          const cplx<T>* Vg, const cplx<T>* taug, TO* Z, int rows_out, double* gdiag) {
  constexpr int S = RegShape<T, NB>::Sbt;
  hipLaunchKernelGGL((eigh_backtrans_reg_kernel<T, NB, S, TO>), dim3(static_cast<unsigned>(nmat)), dim3(NB * S), 0, st,
                     n, Y, DE + 2 * nmat * n, DE + nmat * n, Ts, Vg, taug, Z, rows_out, gdiag);
}

template <typename T, typename TO>
void backtrans_reg_dispatch(hipStream_t st, int n, size_t nmat, const T* Y, const T* DE, const T* Ts,
                            const cplx<T>* Vg, const cplx<T>* taug, TO* Z, int rows_out, double* gdiag, int nb) {
  nb = reg_block(n, nb);
  if (nb == 32) {
    launch_backtrans_reg<T, 32, TO>(st, n, nmat, Y, DE, Ts, Vg, taug, Z, rows_out, gdiag);
  } else if (nb == 64) {
    launch_backtrans_reg<T, 64, TO>(st, n, nmat, Y, DE, Ts, Vg, taug, Z, rows_out, gdiag);
  } else {
    launch_backtrans_reg<T, 128, TO>(st, n, nmat, Y, DE, Ts, Vg, taug, Z, rows_out, gdiag);
  }
}

// ---- two-stage route (128 < n <= 256), see eigh_band.hpp ----
constexpr int kBandReduceThreads = 256;
// FP64 stage 2 team: 16 groups of 16, the right update of a step's 3b rows
// in three passes (each a dependent global round trip).  64 threads (12
// passes) took 29 ms per 1024 matrices at n = 256, 256 threads 23.7 ms; 128
// and 768 were slower (barriers over more waves, fewer matrices resident).
constexpr int kBandChaseThreads = 256;

template <typename T>
size_t band_reduce_smem(int n) {
  const size_t pan = static_cast<size_t>(n - kBandB) * kBandLd * sizeof(cplx<T>);
  return 2 * pan + 3 * kBandB * kBandB * sizeof(cplx<T>) + kBandB * sizeof(cplx<T>) +
         2 * GpuTeam::kRedStride * sizeof(T);
}

// Dense -> band, one workgroup per matrix (matrix in global memory).
template <typename T>
__global__ void __launch_bounds__(kBandReduceThreads) band_reduce_kernel(int n, const double* __restrict__ H,
                                                                        cplx<T>* __restrict__ A,
                                                                        cplx<T>* __restrict__ band) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  const size_t nn = static_cast<size_t>(n) * n;
  const BandLayout L = band_layout(n);
  cplx<T>* base = reinterpret_cast<cplx<T>*>(smem_raw);
  const size_t pan = static_cast<size_t>(n - kBandB) * kBandLd;
  BandReduceSmem<T> sm;
  sm.V = base;
  sm.Y = base + pan;
  sm.Tm = sm.Y + pan;
  sm.W = sm.Tm + kBandB * kBandB;
  sm.M = sm.W + kBandB * kBandB;
  sm.tau = sm.M + kBandB * kBandB;
  sm.red = reinterpret_cast<T*>(sm.tau + kBandB);
  cplx<T>* bm = band + mat * L.total;
  band_reduce<T>(tm, n, H + mat * nn * 2, A + mat * nn, bm + L.v1, bm + L.t1, sm);
}

// Band -> tridiagonal, one workgroup per matrix (band in global memory).
template <typename T>
__global__ void __launch_bounds__(kBandChaseThreads) band_chase_kernel(int n, cplx<T>* __restrict__ A,
                                                                      T* __restrict__ Dg, T* __restrict__ Eg,
                                                                      cplx<T>* __restrict__ band) {
  __shared__ cplx<T> hv[kBandB];
  __shared__ T red[2 * GpuTeam::kRedStride];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  const size_t nn = static_cast<size_t>(n) * n;
  const BandLayout L = band_layout(n);
  cplx<T>* bm = band + mat * L.total;
  band_chase<T>(tm, n, A + mat * nn, Dg + mat * n, Eg + mat * n, bm + L.v2, bm + L.tau2, hv, red);
}

// FP32 stage 2 with the band region in LDS (one workgroup of 768 threads
// per matrix).
constexpr int kBandChaseLdsThreads = 768;  // 48 groups of 16: one pass over the 3b columns (rows) of a step
__global__ void __launch_bounds__(kBandChaseLdsThreads) band_chase_lds_kernel(int n, const cplx<float>* __restrict__ A,
                                                                            float* __restrict__ Dg,
                                                                            float* __restrict__ Eg,
                                                                            cplx<float>* __restrict__ band) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  const size_t nn = static_cast<size_t>(n) * n;
  const BandLayout L = band_layout(n);
  cplx<float>* bm = band + mat * L.total;
  band_chase_lds<float>(tm, n, A + mat * nn, Dg + mat * n, Eg + mat * n, bm + L.v2, bm + L.tau2,
                        reinterpret_cast<cplx<float>*>(smem_raw));
}

// The eigenvectors of T of the isolated eigenvalues written into Y (the
// cluster members are there already): one thread per (matrix, column).
template <typename T>
__global__ void __launch_bounds__(256) twisted_columns_kernel(int n, size_t nmat, const T* __restrict__ Eg,
                                                             const T* __restrict__ S, const T* __restrict__ TW,
                                                             T* __restrict__ Y) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= nmat * n) return;
  const size_t mat = g / n;
  const int j = static_cast<int>(g % n);
  const size_t nn = static_cast<size_t>(n) * n;
  twisted_column<T>(n, Eg + mat * n, S + mat * nn * 3, n, j, TW[mat * n + j], Y + mat * nn);
}

// Z = Q1 Q2 Y for W columns per workgroup (W * S threads, Y block and the
// reflector staging buffers in LDS).
template <typename T>
struct BandBt;
// Threads per column S: one row of each chasing reflector per thread (FP32:
// 2489 -> 2400 ms per slab pass against 8; FP64: 65 ms per 1024 matrices at
// n = 256 against 72 ms with 8)
template <> struct BandBt<float> { static constexpr int W = 64, S = 16; };
template <> struct BandBt<double> { static constexpr int W = 32, S = 16; };

template <typename T, typename TO>
__global__ void __launch_bounds__(BandBt<T>::W * BandBt<T>::S) band_backtrans_kernel(
    int n, const T* __restrict__ Y, const cplx<T>* __restrict__ band, TO* __restrict__ Z, int rows_out,
    double* __restrict__ gdiag) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int W = BandBt<T>::W, S = BandBt<T>::S;
  GpuTeam tm;
  const size_t mat = blockIdx.y;
  const size_t nn = static_cast<size_t>(n) * n;
  const BandLayout L = band_layout(n);
  const cplx<T>* bm = band + mat * L.total;
  band_backtrans<T, W, S>(tm, n, blockIdx.x * W, Y + mat * nn, n, bm + L.v1, bm + L.t1, bm + L.v2, bm + L.tau2,
                          reinterpret_cast<cplx<T>*>(smem_raw), Z + mat * nn * 2, rows_out,
                          gdiag ? gdiag + mat * n : nullptr);
}

template <typename T>
size_t band_backtrans_smem(int n) {
  return band_backtrans_team_elems<T, BandBt<T>::W>(n) * sizeof(cplx<T>);
}

constexpr int kTeBlock = 256;
// Relative eigenvalue gap below which eigenvectors are formed by inverse
// iteration inside the cluster (dstein's threshold).
constexpr double kClusterRel = 1e-3;

// Tridiagonal eigenvalues (+ the twisted factorisations of the isolated ones
// when VEC): sub-team of n threads per matrix, 256 / n matrices per
// workgroup.  TW: [nmat][n] twist indices (-1: cluster member).
template <typename T, bool VEC>
__global__ void __launch_bounds__(kTeBlock) eigh_trieig_kernel(int n, size_t nmat, const T* __restrict__ Dg,
                                                              const T* __restrict__ Eg, double* __restrict__ Eout,
                                                              T* __restrict__ S, T* __restrict__ TW) {
  __shared__ T sh[5 * kTeBlock];
  const int per = kTeBlock / n;  // matrices per workgroup (n <= 128)
  const int sub = threadIdx.x / n;
  const size_t mat = static_cast<size_t>(blockIdx.x) * per + sub;
  const bool live = sub < per && mat < nmat;
  // idle threads take id >= n and sub-team 0's LDS (read-only for them)
  const int so = live ? sub : 0;
  GpuSubTeam tm(live ? static_cast<int>(threadIdx.x % n) : n, n);
  T* d = sh + so * 5 * n;
  T* e = d + n;
  T* e2 = e + n;
  T* E = e2 + n;
  T* cnt = E + n;
  if (live) {
    d[tm.id] = Dg[mat * n + tm.id];
    e[tm.id] = Eg[mat * n + tm.id];
  }
  __syncthreads();
  bisect_all<T>(tm, n, d, e, e2, E, cnt);
  if (live) Eout[mat * n + tm.id] = static_cast<double>(E[tm.id]);
  if (VEC) {
    const size_t mm = live ? mat : 0;
    twisted_all<T>(tm, n, d, e, E, n, S + mm * static_cast<size_t>(n) * n * 3, TW + mm * n, T(kClusterRel));
  }
}

// Eigenvectors of the cluster members (inverse iteration), same layout as
// eigh_trieig_kernel; a workgroup whose matrices have no cluster returns at
// once.  The eigenvalues are read back from Eval (exact: they were T).
template <typename T>
__global__ void __launch_bounds__(kTeBlock) eigh_cluster_kernel(int n, size_t nmat, const T* __restrict__ Dg,
                                                               const T* __restrict__ Eg, const double* __restrict__ Eval,
                                                               const T* __restrict__ TW, T* __restrict__ Y,
                                                               T* __restrict__ S) {
  __shared__ T sh[3 * kTeBlock];
  const int per = kTeBlock / n;
  const int sub = threadIdx.x / n;
  const size_t mat = static_cast<size_t>(blockIdx.x) * per + sub;
  const bool live = sub < per && mat < nmat;
  const int so = live ? sub : 0;
  GpuSubTeam tm(live ? static_cast<int>(threadIdx.x % n) : n, n);
  const bool member = live && TW[mat * n + tm.id] < T(0);
  if (!__syncthreads_or(member)) return;
  T* d = sh + so * 3 * n;
  T* e = d + n;
  T* E = e + n;
  if (live) {
    d[tm.id] = Dg[mat * n + tm.id];
    e[tm.id] = Eg[mat * n + tm.id];
    E[tm.id] = static_cast<T>(Eval[mat * n + tm.id]);
  }
  __syncthreads();
  const size_t mm = live ? mat : 0;
  cluster_vectors<T>(tm, n, d, e, E, Y + mm * static_cast<size_t>(n) * n, n, S + mm * static_cast<size_t>(n) * n * 3,
                     T(kClusterRel));
}

template <typename T>
int launch_trieig(hipStream_t st, int n, size_t nmat, const T* Dg, const T* Eg, double* E, bool vectors, T* Y, T* S,
                  T* TW) {
  const int per = kTeBlock / n;
  const unsigned blocks = static_cast<unsigned>((nmat + per - 1) / per);
  if (vectors) {
    hipLaunchKernelGGL((eigh_trieig_kernel<T, true>), dim3(blocks), dim3(kTeBlock), 0, st, n, nmat, Dg, Eg, E, S,
                       TW);
    RTB_HIP(hipGetLastError());
    hipLaunchKernelGGL((eigh_cluster_kernel<T>), dim3(blocks), dim3(kTeBlock), 0, st, n, nmat, Dg, Eg, E, TW, Y, S);
  } else {
    hipLaunchKernelGGL((eigh_trieig_kernel<T, false>), dim3(blocks), dim3(kTeBlock), 0, st, n, nmat, Dg, Eg, E, S,
                       TW);
  }
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

// The matrix the reference's eigh(UPLO::Lower) diagonalises for a row-major
// H (solve_ham.rs:122, response/traits.rs:168): ndarray-linalg hands the
// buffer to LAPACK as column-major, so LAPACK's lower triangle is the upper
// triangle of H.  Its eigvalsh_v(UPLO::Upper) (solve_ham.rs:61,
// ndarray_lapack.rs:412-448) passes the raw buffer, reading the lower one.
// For an H that is Hermitian only to rounding (e.g. read from a Wannier90
// file) the two differ, so the vector paths of such models first rebuild, in
// place,
// H[i][j] = conj(H[j][i]) for i > j and Im H[i][i] = 0 (what zheev uses of
// the diagonal).  One 32 x 32 tile pair (I >= J) per workgroup: the upper
// tile (J, I) is read by rows into LDS and written, conjugate-transposed,
// by rows into the lower tile (I, J); matrices are strided over grid.y.
constexpr int kHzTile = 32;
__global__ void __launch_bounds__(256) hermitize_from_upper_kernel(int n, size_t nmat, double* __restrict__ H) {
  __shared__ double2 t[kHzTile][kHzTile + 1];
  int p = blockIdx.x, I = 0;
  while (p > I) {
    p -= I + 1;
    ++I;
  }
  const int J = p;
  const int tx = threadIdx.x % kHzTile, ty = threadIdx.x / kHzTile;  // 32 x 8
  for (size_t mat = blockIdx.y; mat < nmat; mat += gridDim.y) {
    double2* A = reinterpret_cast<double2*>(H) + mat * n * n;
    for (int r = ty; r < kHzTile; r += 8) {
      const int gi = J * kHzTile + r, gj = I * kHzTile + tx;
      if (gi < n && gj < n) t[r][tx] = A[static_cast<size_t>(gi) * n + gj];
    }
    __syncthreads();
    for (int r = ty; r < kHzTile; r += 8) {
      const int gi = I * kHzTile + r, gj = J * kHzTile + tx;
      if (gi < n && gj < n) {
        double2* dst = A + static_cast<size_t>(gi) * n + gj;
        if (gi > gj) {
          const double2 v = t[tx][r];
          *dst = make_double2(v.x, -v.y);
        } else if (gi == gj) {
          *dst = make_double2(t[tx][r].x, 0.0);
        }
      }
    }
    __syncthreads();
  }
}

int hermitize_from_upper(hipStream_t st, int n, size_t nmat, double* H) {
  if (nmat == 0 || n < 1) return RTB_OK;
  const int T = (n + kHzTile - 1) / kHzTile;
  const unsigned gy = static_cast<unsigned>(std::min<size_t>(nmat, 65535));
  hipLaunchKernelGGL(hermitize_from_upper_kernel, dim3(T * (T + 1) / 2, gy), dim3(256), 0, st, n, nmat, H);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

template <typename T>
int eigh_launch(hipStream_t st, Timer* tmr, int n, size_t nmat, double* H, double* E, bool vectors, T* Vg,
                T* taug, T* Y, T* S, T* DE, bool from_upper, T* band, int nb) {
  if (n < 1 || n > kMaxBandN) return fail(RTB_ERR_UNSUPPORTED, "eigensolver supports 1 <= nsta <= 256");
  if (!eigh_block_valid(n, nb)) return fail(RTB_ERR_INVALID_ARGUMENT, "eigensolver block rows must be 0 or 32, 64, 128 >= nsta");
  if (n > kMaxEighN && !band) return fail(RTB_ERR_INVALID_ARGUMENT, "nsta above 128 needs the band buffers");
  if (nmat == 0) return RTB_OK;
  if (from_upper) {
    if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_REDUCE));
    RTB_TRY(hermitize_from_upper(st, n, nmat, H));
    if (tmr) RTB_TRY(tmr->end(1));
  }
  T* Dg = DE;
  T* Eg = DE + nmat * n;
  T* TW = DE + 2 * nmat * n;
  if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_TRIDIAG));
  if (n <= kMaxEighN) {
    tridiag_reg_dispatch<T>(st, n, nmat, H, Dg, Eg, reinterpret_cast<cplx<T>*>(Vg),
                            reinterpret_cast<cplx<T>*>(taug), nb);
    RTB_HIP(hipGetLastError());
    if (tmr) RTB_TRY(tmr->end(1, flop_tridiag(n, nmat)));
  } else {
    // the matrix itself is reduced in the Vg buffer (n x n complex)
    cplx<T>* A = reinterpret_cast<cplx<T>*>(Vg);
    cplx<T>* bnd = reinterpret_cast<cplx<T>*>(band);
    hipLaunchKernelGGL((band_reduce_kernel<T>), dim3(static_cast<unsigned>(nmat)), dim3(kBandReduceThreads),
                       band_reduce_smem<T>(n), st, n, H, A, bnd);
    RTB_HIP(hipGetLastError());
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL(band_chase_lds_kernel, dim3(static_cast<unsigned>(nmat)), dim3(kBandChaseLdsThreads),
                         band_chase_team_elems(n) * sizeof(cplx<float>), st, n, A, Dg, Eg, bnd);
    else
      hipLaunchKernelGGL((band_chase_kernel<T>), dim3(static_cast<unsigned>(nmat)), dim3(kBandChaseThreads), 0,
                         st, n, A, Dg, Eg, bnd);
    RTB_HIP(hipGetLastError());
    if (tmr) RTB_TRY(tmr->end(2, flop_tridiag(n, nmat)));
  }
  if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_TRIEIG));
  RTB_TRY(launch_trieig<T>(st, n, nmat, Dg, Eg, E, vectors, Y, S, TW));
  if (vectors && n > kMaxEighN) {
    const size_t tot = nmat * n;
    hipLaunchKernelGGL((twisted_columns_kernel<T>), dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0,
                       st, n, nmat, Eg, S, TW, Y);
    RTB_HIP(hipGetLastError());
  }
  if (tmr) RTB_TRY(tmr->end(vectors ? (n > kMaxEighN ? 3 : 2) : 1));
  return RTB_OK;
}

template <typename T, typename TO>
int eigh_backtransform(hipStream_t st, int n, size_t nmat, const T* Y, const T* DE, const T* S, const T* Vg,
                       const T* taug, TO* Z, bool rows_out, double* gdiag, const T* band, int nb) {
  if (nmat == 0) return RTB_OK;
  if (!eigh_block_valid(n, nb)) return fail(RTB_ERR_INVALID_ARGUMENT, "eigensolver block rows must be 0 or 32, 64, 128 >= nsta");
  if (n <= kMaxEighN) {
    backtrans_reg_dispatch<T, TO>(st, n, nmat, Y, DE, S, reinterpret_cast<const cplx<T>*>(Vg),
                                  reinterpret_cast<const cplx<T>*>(taug), Z, rows_out ? 1 : 0, gdiag, nb);
  } else {
    if (!band) return fail(RTB_ERR_INVALID_ARGUMENT, "nsta above 128 needs the band buffers");
    constexpr int W = BandBt<T>::W, S2 = BandBt<T>::S;
    const unsigned gx = static_cast<unsigned>((n + W - 1) / W);
    for (size_t m0 = 0; m0 < nmat; m0 += 65535) {
      const size_t cnt = std::min<size_t>(65535, nmat - m0);
      const size_t nn = static_cast<size_t>(n) * n;
      hipLaunchKernelGGL((band_backtrans_kernel<T, TO>), dim3(gx, static_cast<unsigned>(cnt)), dim3(W * S2),
                         band_backtrans_smem<T>(n), st, n, Y + m0 * nn,
                         reinterpret_cast<const cplx<T>*>(band) + m0 * band_layout(n).total, Z + m0 * nn * 2,
                         rows_out ? 1 : 0, gdiag ? gdiag + m0 * n : nullptr);
    }
  }
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

}  // namespace

int eigh_launch_fp32(hipStream_t st, Timer* tmr, int n, size_t nmat, double* H, double* E, bool vectors,
                     float* Vg, float* taug, float* Y, float* S, float* DE, bool from_upper, float* band, int nb) {
  return eigh_launch<float>(st, tmr, n, nmat, H, E, vectors, Vg, taug, Y, S, DE, from_upper, band, nb);
}

int eigh_backtransform_fp32(hipStream_t st, int n, size_t nmat, const float* Y, const float* DE, const float* S,
                            const float* Vg, const float* taug, float* Z, bool rows_out, double* gdiag,
                            const float* band, int nb) {
  return eigh_backtransform<float, float>(st, n, nmat, Y, DE, S, Vg, taug, Z, rows_out, gdiag, band, nb);
}

int eigh_launch_fp64(hipStream_t st, Timer* tmr, int n, size_t nmat, double* H, double* E, bool vectors,
                     double* Vg, double* taug, double* Y, double* S, double* DE, bool from_upper, double* band,
                     int nb) {
  return eigh_launch<double>(st, tmr, n, nmat, H, E, vectors, Vg, taug, Y, S, DE, from_upper, band, nb);
}

int eigh_backtransform_fp64(hipStream_t st, int n, size_t nmat, const double* Y, const double* DE, const double* S,
                            const double* Vg, const double* taug, double* Z, bool rows_out, const double* band,
                            int nb) {
  return eigh_backtransform<double, double>(st, n, nmat, Y, DE, S, Vg, taug, Z, rows_out, nullptr, band, nb);
}

bool eigh_block_valid(int n, int nb) {
  if (nb == 0) return true;
  return n <= kMaxEighN && (nb == 32 || nb == 64 || nb == 128) && nb >= n;
}

int hermitize_upper(hipStream_t st, int n, size_t nmat, double* H) { return hermitize_from_upper(st, n, nmat, H); }

size_t eigh_band_elems(int n) { return n > kMaxEighN ? band_layout(n).total * 2 : 0; }

int eigh_set_smem_limits() {
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&band_chase_lds_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(band_chase_team_elems(kMaxBandN) * sizeof(cplx<float>))));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&band_reduce_kernel<float>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(band_reduce_smem<float>(kMaxBandN))));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&band_reduce_kernel<double>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(band_reduce_smem<double>(kMaxBandN))));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&band_backtrans_kernel<float, float>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(band_backtrans_smem<float>(kMaxBandN))));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&band_backtrans_kernel<double, double>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(band_backtrans_smem<double>(kMaxBandN))));
  return RTB_OK;
}

}  // namespace rtb
