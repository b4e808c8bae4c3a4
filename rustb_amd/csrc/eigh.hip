// Batched Hermitian eigensolver kernels (gfx950).
//
//   eigh_tridiag_kernel     one workgroup per matrix: Householder reduction
//                           with the packed lower triangle in LDS
//   eigh_trieig_kernel      one thread per eigenvalue, 256/n matrices per
//                           workgroup: bisection (+ inverse iteration)
//   eigh_backtransform_kernel  Z = Q Y per block of columns
#include "eigh.hpp"

#include "eigh_core.hpp"
#include "eigh_reg.hpp"
#include "gpu_team.hpp"

namespace rtb {

namespace {

template <typename T>
struct TriSmem {
  cplx<T>* P;
  cplx<T>* v;
  cplx<T>* w;
  T *d, *e, *red;
};

template <typename T>
__device__ TriSmem<T> carve_tri(void* base, int n) {
  TriSmem<T> s;
  char* p = static_cast<char*>(base);
  s.P = reinterpret_cast<cplx<T>*>(p);
  p += sizeof(cplx<T>) * (static_cast<size_t>(n) * (n + 1) / 2);
  s.v = reinterpret_cast<cplx<T>*>(p);
  p += sizeof(cplx<T>) * n;
  s.w = reinterpret_cast<cplx<T>*>(p);
  p += sizeof(cplx<T>) * n;
  s.d = reinterpret_cast<T*>(p);
  s.e = s.d + n;
  s.red = s.e + n;
  return s;
}

// H: [nmat][n][n] complex FP64 (row-major; the lower triangle is read).
// Dg, Eg: [nmat][n] tridiagonal (diagonal, sub-diagonal); Vg, taug reflectors.
template <typename T>
__global__ void __launch_bounds__(256) eigh_tridiag_kernel(int n, const double* __restrict__ H, T* __restrict__ Dg,
                                                           T* __restrict__ Eg, cplx<T>* __restrict__ Vg,
                                                           cplx<T>* __restrict__ taug) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  TriSmem<T> s = carve_tri<T>(smem_raw, n);
  const double* Hm = H + mat * static_cast<size_t>(n) * n * 2;
  // lower triangle into packed storage, rows walked incrementally
  const int np = n * (n + 1) / 2;
  {
    int r = 0, rs = 0;
    for (int f = tm.id; f < np; f += tm.nt) {
      while (f >= rs + r + 1) {
        rs += r + 1;
        ++r;
      }
      const int c = f - rs;
      const double2 h = reinterpret_cast<const double2*>(Hm)[static_cast<size_t>(r) * n + c];
      s.P[f].re = static_cast<T>(h.x);
      s.P[f].im = static_cast<T>(h.y);
    }
  }
  tm.sync();
  cplx<T>* Vm = Vg ? Vg + mat * static_cast<size_t>(n) * n : nullptr;
  cplx<T>* tu = taug ? taug + mat * static_cast<size_t>(n) : nullptr;
  tridiagonalize<T>(tm, n, s.P, s.v, s.w, s.d, s.e, s.red, Vm, tu);
  for (int i = tm.id; i < n; i += tm.nt) {
    Dg[mat * n + i] = s.d[i];
    Eg[mat * n + i] = s.e[i];
  }
}

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
  RegSmem<T> sm{xc, vp, wp, yp, red};
  tridiag_reg<T, NB, S>(tm, n, H + mat * nn * 2, sm, Dg + mat * n, Eg + mat * n, Vg ? Vg + mat * nn : nullptr,
                        taug ? taug + mat * n : nullptr);
}

// Register-resident back-transformation Z = Q Y: one workgroup per matrix,
// each thread holding NB / S entries of one column of Z.
template <typename T, int NB, int S>
__global__ void __launch_bounds__(NB * S) eigh_backtrans_reg_kernel(int n, const T* __restrict__ Y,
                                                                    const T* __restrict__ ysc,
                                                                    const cplx<T>* __restrict__ Vg,
                                                                    const cplx<T>* __restrict__ taug,
                                                                    double* __restrict__ Zout, int rows_out) {
  __shared__ cplx<T> vp[3 * NB], yp[2 * S * NB];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  const size_t nn = static_cast<size_t>(n) * n;
  RegSmem<T> sm{nullptr, vp, nullptr, yp, nullptr};
  backtrans_reg<T, NB, S>(tm, n, Y + mat * nn, n, ysc + mat * n, Vg + mat * nn, taug + mat * n, sm, Zout + mat * nn * 2,
                          rows_out);
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

template <typename T>
void tridiag_reg_dispatch(hipStream_t st, int n, size_t nmat, const double* H, T* Dg, T* Eg, cplx<T>* Vg,
                          cplx<T>* taug) {
  if (n <= 32) {
    launch_tridiag_reg<T, 32>(st, n, nmat, H, Dg, Eg, Vg, taug);
  } else if (n <= 64) {
    launch_tridiag_reg<T, 64>(st, n, nmat, H, Dg, Eg, Vg, taug);
  } else {
    launch_tridiag_reg<T, 128>(st, n, nmat, H, Dg, Eg, Vg, taug);
  }
}

template <typename T, int NB>
void launch_backtrans_reg(hipStream_t st, int n, size_t nmat, const T* Y, const T* ysc, const cplx<T>* Vg,
                          const cplx<T>* taug, double* Z, int rows_out) {
  constexpr int S = RegShape<T, NB>::Sbt;
  hipLaunchKernelGGL((eigh_backtrans_reg_kernel<T, NB, S>), dim3(static_cast<unsigned>(nmat)), dim3(NB * S), 0, st,
                     n, Y, ysc, Vg, taug, Z, rows_out);
}

template <typename T>
void backtrans_reg_dispatch(hipStream_t st, int n, size_t nmat, const T* Y, const T* ysc, const cplx<T>* Vg,
                            const cplx<T>* taug, double* Z, int rows_out) {
  if (n <= 32) {
    launch_backtrans_reg<T, 32>(st, n, nmat, Y, ysc, Vg, taug, Z, rows_out);
  } else if (n <= 64) {
    launch_backtrans_reg<T, 64>(st, n, nmat, Y, ysc, Vg, taug, Z, rows_out);
  } else {
    launch_backtrans_reg<T, 128>(st, n, nmat, Y, ysc, Vg, taug, Z, rows_out);
  }
}

constexpr int kTeBlock = 256;

// Tridiagonal eigenvalues (+ eigenvectors): sub-team of n threads per matrix.
// With VEC, the per-eigenvector scale of Y (see tridiag_vectors) overwrites
// the diagonal Dg once it has been read.
template <typename T, bool VEC>
__global__ void __launch_bounds__(kTeBlock) eigh_trieig_kernel(int n, size_t nmat, T* __restrict__ Dg,
                                                              const T* __restrict__ Eg, double* __restrict__ Eout,
                                                              T* __restrict__ Y, T* __restrict__ S) {
  __shared__ T sh[4 * kTeBlock];
  const int per = kTeBlock / n;  // matrices per workgroup (n <= 128)
  const int sub = threadIdx.x / n;
  const size_t mat = static_cast<size_t>(blockIdx.x) * per + sub;
  const bool live = sub < per && mat < nmat;
  // idle threads take id >= n and sub-team 0's LDS (read-only for them)
  const int so = live ? sub : 0;
  GpuSubTeam tm(live ? static_cast<int>(threadIdx.x % n) : n, n);
  T* d = sh + so * 4 * n;
  T* e = d + n;
  T* e2 = e + n;
  T* E = e2 + n;
  if (live) {
    d[tm.id] = Dg[mat * n + tm.id];
    e[tm.id] = Eg[mat * n + tm.id];
  }
  __syncthreads();
  bisect_all<T>(tm, n, d, e, e2, E);
  if (live) Eout[mat * n + tm.id] = static_cast<double>(E[tm.id]);
  if (VEC) {
    const size_t mm = live ? mat : 0;
    T* Ym = Y + mm * static_cast<size_t>(n) * n;
    T* Sm = S + mm * static_cast<size_t>(n) * n * 3;
    tridiag_vectors<T>(tm, n, d, e, E, Ym, n, Sm, Dg + mm * n);
  }
}

// Z[:, c0:c0+cb] = Q Y[:, c0:c0+cb]; output complex FP64.
//   rows_out = 0: Z[mat][r][c] (eigenvectors as columns)
//   rows_out = 1: Z[mat][c][r] (eigenvectors as rows, solve_all layout)
template <typename T>
__global__ void __launch_bounds__(256) eigh_backtransform_kernel(int n, int cb, const T* __restrict__ Y,
                                                                 const cplx<T>* __restrict__ Vg,
                                                                 const cplx<T>* __restrict__ taug,
                                                                 double* __restrict__ Zout, int rows_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  const int c0 = blockIdx.y * cb;
  const int w = min(cb, n - c0);
  cplx<T>* Z = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* vs = Z + static_cast<size_t>(n) * cb;
  cplx<T>* part = vs + 2 * n;
  cplx<T>* sv = part + tm.nt;
  const T* Ym = Y + mat * static_cast<size_t>(n) * n;
  for (int f = tm.id; f < n * cb; f += tm.nt) {
    int r = f / cb, c = f % cb;
    T y = c < w ? Ym[static_cast<size_t>(r) * n + c0 + c] : T(0);
    Z[f] = {y, T(0)};
  }
  tm.sync();
  apply_q<T>(tm, n, cb, Z, Vg + mat * static_cast<size_t>(n) * n, taug + mat * static_cast<size_t>(n), vs, part,
             sv);
  double* Zm = Zout + mat * static_cast<size_t>(n) * n * 2;
  for (int f = tm.id; f < n * w; f += tm.nt) {
    int r, c;
    if (rows_out) {  // coalesce along r
      c = f / n;
      r = f % n;
    } else {
      r = f / w;
      c = f % w;
    }
    cplx<T> z = Z[r * cb + c];
    size_t o = rows_out ? (static_cast<size_t>(c0 + c) * n + r) : (static_cast<size_t>(r) * n + c0 + c);
    reinterpret_cast<double2*>(Zm)[o] = make_double2(static_cast<double>(z.re), static_cast<double>(z.im));
  }
}

size_t tri_smem(int n, size_t rs) {
  return sizeof(double) * 0 + (static_cast<size_t>(n) * (n + 1) / 2 + 2 * n) * 2 * rs + 2 * n * rs +
         (256 + 64) * rs;
}

int bt_cb(int n) { return n <= 64 ? n : 32; }
size_t bt_smem(int n, int cb, size_t rs) { return 2 * rs * (static_cast<size_t>(n) * cb + 2 * n + 256 + cb); }

}  // namespace

int eigh_set_smem_limits() {
  // up to 160 KB of dynamic LDS per workgroup on gfx950
  const int lim = 160 * 1024;
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&eigh_tridiag_kernel<double>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&eigh_backtransform_kernel<double>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&eigh_tridiag_kernel<float>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&eigh_backtransform_kernel<float>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  return RTB_OK;
}

// FP32 route of the mixed form: same kernels instantiated for float.
int eigh_launch_fp32(hipStream_t st, Timer* tmr, int n, size_t nmat, const double* H, double* E, bool vectors,
                     float* Vg, float* taug, float* Y, float* S, float* DE) {
  if (n < 1 || n > kMaxEighN) return fail(RTB_ERR_UNSUPPORTED, "eigensolver supports 1 <= nsta <= 128");
  if (nmat == 0) return RTB_OK;
  float* Dg = DE;
  float* Eg = DE + nmat * n;
  if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_TRIDIAG));
  tridiag_reg_dispatch<float>(st, n, nmat, H, Dg, Eg, reinterpret_cast<cplx<float>*>(Vg),
                              reinterpret_cast<cplx<float>*>(taug));
  RTB_HIP(hipGetLastError());
  if (tmr) RTB_TRY(tmr->end(1, flop_tridiag(n, nmat)));
  if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_TRIEIG));
  const int per = kTeBlock / n;
  const unsigned blocks = static_cast<unsigned>((nmat + per - 1) / per);
  if (vectors) {
    hipLaunchKernelGGL((eigh_trieig_kernel<float, true>), dim3(blocks), dim3(kTeBlock), 0, st, n, nmat, Dg, Eg, E,
                       Y, S);
  } else {
    hipLaunchKernelGGL((eigh_trieig_kernel<float, false>), dim3(blocks), dim3(kTeBlock), 0, st, n, nmat, Dg, Eg,
                       E, Y, S);
  }
  RTB_HIP(hipGetLastError());
  if (tmr) RTB_TRY(tmr->end(1));
  return RTB_OK;
}

int eigh_backtransform_fp32(hipStream_t st, int n, size_t nmat, const float* Y, const float* DE, const float* Vg,
                            const float* taug, double* Z, bool rows_out) {
  if (nmat == 0) return RTB_OK;
  backtrans_reg_dispatch<float>(st, n, nmat, Y, DE, reinterpret_cast<const cplx<float>*>(Vg),
                                reinterpret_cast<const cplx<float>*>(taug), Z, rows_out ? 1 : 0);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int eigh_launch_fp64(hipStream_t st, Timer* tmr, int n, size_t nmat, const double* H, double* E, bool vectors,
                     double* Vg, double* taug, double* Y, double* S, double* DE) {
  if (n < 1 || n > kMaxEighN) return fail(RTB_ERR_UNSUPPORTED, "eigensolver supports 1 <= nsta <= 128");
  if (nmat == 0) return RTB_OK;
  double* Dg = DE;
  double* Eg = DE + nmat * n;
  if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_TRIDIAG));
  tridiag_reg_dispatch<double>(st, n, nmat, H, Dg, Eg, reinterpret_cast<cplx<double>*>(Vg),
                               reinterpret_cast<cplx<double>*>(taug));
  RTB_HIP(hipGetLastError());
  if (tmr) RTB_TRY(tmr->end(1, flop_tridiag(n, nmat)));
  if (tmr) RTB_TRY(tmr->begin(RTB_STAGE_TRIEIG));
  const int per = kTeBlock / n;
  const unsigned blocks = static_cast<unsigned>((nmat + per - 1) / per);
  if (vectors) {
    hipLaunchKernelGGL((eigh_trieig_kernel<double, true>), dim3(blocks), dim3(kTeBlock), 0, st, n, nmat, Dg, Eg, E,
                       Y, S);
  } else {
    hipLaunchKernelGGL((eigh_trieig_kernel<double, false>), dim3(blocks), dim3(kTeBlock), 0, st, n, nmat, Dg, Eg,
                       E, Y, S);
  }
  RTB_HIP(hipGetLastError());
  if (tmr) RTB_TRY(tmr->end(1));
  return RTB_OK;
}

int eigh_backtransform_fp64(hipStream_t st, int n, size_t nmat, const double* Y, const double* DE,
                            const double* Vg, const double* taug, double* Z, bool rows_out) {
  if (nmat == 0) return RTB_OK;
  backtrans_reg_dispatch<double>(st, n, nmat, Y, DE, reinterpret_cast<const cplx<double>*>(Vg),
                                 reinterpret_cast<const cplx<double>*>(taug), Z, rows_out ? 1 : 0);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

}  // namespace rtb
