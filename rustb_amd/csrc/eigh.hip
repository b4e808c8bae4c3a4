// Batched Hermitian eigensolver kernels (gfx950): one workgroup per matrix.
#include "eigh.hpp"

#include "eigh_core.hpp"
#include "gpu_team.hpp"

namespace rtb {

namespace {

template <typename T>
struct Stage1Smem {
  cplx<T>* P;
  cplx<T>* v;
  cplx<T>* w;
  T *d, *e, *e2, *E, *red;
};

template <typename T>
__device__ Stage1Smem<T> carve(void* base, int n) {
  Stage1Smem<T> s;
  char* p = static_cast<char*>(base);
  s.P = reinterpret_cast<cplx<T>*>(p);
  p += sizeof(cplx<T>) * (static_cast<size_t>(n) * (n + 1) / 2);
  s.v = reinterpret_cast<cplx<T>*>(p);
  p += sizeof(cplx<T>) * n;
  s.w = reinterpret_cast<cplx<T>*>(p);
  p += sizeof(cplx<T>) * n;
  s.d = reinterpret_cast<T*>(p);
  s.e = s.d + n;
  s.e2 = s.e + n;
  s.E = s.e2 + n;
  s.red = s.E + n;
  return s;
}

// H: [nmat][n][n] complex FP64 (row-major); reads the lower triangle.
template <typename T, bool VEC>
__global__ void __launch_bounds__(256) eigh_stage1_kernel(int n, const double* __restrict__ H,
                                                          double* __restrict__ Eout,
                                                          cplx<T>* __restrict__ Vg,
                                                          cplx<T>* __restrict__ taug,
                                                          T* __restrict__ Y, T* __restrict__ S) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  Stage1Smem<T> s = carve<T>(smem_raw, n);
  const double* Hm = H + mat * static_cast<size_t>(n) * n * 2;
  // lower triangle into packed storage
  const int np = n * (n + 1) / 2;
  for (int f = tm.id; f < np; f += tm.nt) {
    int r = static_cast<int>((sqrtf(8.0f * f + 1.0f) - 1.0f) * 0.5f);
    if ((r + 1) * (r + 2) / 2 <= f) ++r;
    if (r * (r + 1) / 2 > f) --r;
    int c = f - r * (r + 1) / 2;
    const double* h = Hm + (static_cast<size_t>(r) * n + c) * 2;
    s.P[f].re = static_cast<T>(h[0]);
    s.P[f].im = static_cast<T>(h[1]);
  }
  tm.sync();
  cplx<T>* Vm = Vg ? Vg + mat * static_cast<size_t>(n) * n : nullptr;
  cplx<T>* tm_tau = taug ? taug + mat * static_cast<size_t>(n) : nullptr;
  tridiagonalize<T>(tm, n, s.P, s.v, s.w, s.d, s.e, s.red, Vm, tm_tau);
  bisect_all<T>(tm, n, s.d, s.e, s.e2, s.E);
  for (int j = tm.id; j < n; j += tm.nt) Eout[mat * n + j] = static_cast<double>(s.E[j]);
  if (VEC) {
    T* Ym = Y + mat * static_cast<size_t>(n) * n;
    T* Sm = S + mat * static_cast<size_t>(n) * n * 3;
    tridiag_vectors<T>(tm, n, s.d, s.e, s.E, Ym, n, Sm);
  }
}

// Z[:, c0:c0+cb] = Q Y[:, c0:c0+cb]; output complex FP64.
//   rows_out = false: Z[mat][r][c] (eigenvectors as columns)
//   rows_out = true : Z[mat][c][r] (eigenvectors as rows, solve_all layout)
template <typename T>
__global__ void __launch_bounds__(256) eigh_backtransform_kernel(int n, int cb, const T* __restrict__ Y,
                                                                 const cplx<T>* __restrict__ Vg,
                                                                 const cplx<T>* __restrict__ taug,
                                                                 double* __restrict__ Zout,
                                                                 int rows_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  GpuTeam tm;
  const size_t mat = blockIdx.x;
  const int c0 = blockIdx.y * cb;
  const int w = min(cb, n - c0);
  cplx<T>* Z = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* vs = Z + static_cast<size_t>(n) * cb;
  cplx<T>* sv = vs + n;
  const T* Ym = Y + mat * static_cast<size_t>(n) * n;
  for (int f = tm.id; f < n * cb; f += tm.nt) {
    int r = f / cb, c = f % cb;
    T y = c < w ? Ym[static_cast<size_t>(r) * n + c0 + c] : T(0);
    Z[f] = {y, T(0)};
  }
  tm.sync();
  apply_q<T>(tm, n, cb, Z, Vg + mat * static_cast<size_t>(n) * n, taug + mat * static_cast<size_t>(n),
             vs, sv);
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
    Zm[2 * o] = static_cast<double>(z.re);
    Zm[2 * o + 1] = static_cast<double>(z.im);
  }
}

}  // namespace

size_t eigh_stage1_smem(int n, int real_size) {
  return EighSmem::bytes(n, 256, static_cast<size_t>(real_size));
}

int eigh_launch_fp64(hipStream_t st, int n, size_t nmat, const double* H, double* E, bool vectors,
                     double* Vg, double* taug, double* Y, double* S) {
  if (n < 1 || n > kMaxEighN) return fail(RTB_ERR_UNSUPPORTED, "eigensolver supports 1 <= nsta <= 128");
  if (nmat == 0) return RTB_OK;
  const size_t sm = eigh_stage1_smem(n, sizeof(double));
  dim3 grid(static_cast<unsigned>(nmat));
  if (vectors) {
    hipLaunchKernelGGL((eigh_stage1_kernel<double, true>), grid, dim3(256), sm, st, n, H, E,
                       reinterpret_cast<cplx<double>*>(Vg), reinterpret_cast<cplx<double>*>(taug), Y, S);
  } else {
    hipLaunchKernelGGL((eigh_stage1_kernel<double, false>), grid, dim3(256), sm, st, n, H, E,
                       reinterpret_cast<cplx<double>*>(Vg), reinterpret_cast<cplx<double>*>(taug), Y, S);
  }
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int eigh_backtransform_fp64(hipStream_t st, int n, size_t nmat, const double* Y, const double* Vg,
                            const double* taug, double* Z, bool rows_out) {
  if (nmat == 0) return RTB_OK;
  const int cb = n <= 64 ? n : 32;
  const size_t sm = sizeof(cplx<double>) * (static_cast<size_t>(n) * cb + n + cb);
  dim3 grid(static_cast<unsigned>(nmat), static_cast<unsigned>((n + cb - 1) / cb));
  hipLaunchKernelGGL((eigh_backtransform_kernel<double>), grid, dim3(256), sm, st, n, cb, Y,
                     reinterpret_cast<const cplx<double>*>(Vg), reinterpret_cast<const cplx<double>*>(taug),
                     Z, rows_out ? 1 : 0);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int eigh_set_smem_limits() {
  // allow up to 160 KB of dynamic LDS per workgroup on gfx950
  const int lim = 160 * 1024;
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&eigh_stage1_kernel<double, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&eigh_stage1_kernel<double, false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&eigh_backtransform_kernel<double>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  return RTB_OK;
}

}  // namespace rtb
