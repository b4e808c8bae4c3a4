// Batched complex matrix products on the matrix cores (gfx950).
//
// zgemm3m_kernel: C = op(A) B in FP64 arithmetic for a batch of n x n
// complex matrices (n <= 128 in practice, any n works), row-major
// interleaved (re, im).  Each operand is stored in FP64 or FP32 (values exact
// in FP32: half the memory traffic); C is FP64, or consumed by an epilogue
// (the Ogita-Aishima correction from the Gram product, the Kubo band-sum
// weights from the velocity product).  zgemm3m_dual_kernel: the two FP32
// correction products of the refinement.  One workgroup computes one
// 64 x 64 output tile; its four waves own 32 x 32 quarters (2 x 2 blocks of
// the 16 x 16 x 4 MFMA).
//
//   * 3M product: Re = Ar Br - Ai Bi, Im = (Ar + Ai)(Br + Bi) - Ar Br - Ai Bi,
//     three real MFMAs per block and k-step instead of four;
//   * the next 16-wide k-chunk is loaded into registers while the current
//     one is multiplied from LDS;
//   * herm = 1: op(A) B is Hermitian (X^H Y with X^H Y = Y^H X, e.g. U^H J U);
//     only tiles on or below the diagonal are computed and the off-diagonal
//     ones are also stored conjugate-transposed;
//   * an optional per-matrix vector s scales the result to diag(s) C diag(s)
//     (products with unnormalised eigenvector columns U, unit after U diag(s));
//   * workgroup w runs on XCD w % 8; the tiles of one matrix are given to
//     workgroups on the same XCD so that they share its L2.
//
// MFMA operand layout (lane l): A[l % 16][l / 16], B[l / 16][l % 16];
// accumulator register r of lane l holds D[4r + l / 16][l % 16] (FP64) or
// D[4(l / 16) + r][l % 16] (FP32).
#include "eigh.hpp"
#include "gpu_team.hpp"

namespace rtb {

namespace {

constexpr int kTile = 64;  // output tile side
constexpr int kKc = 16;    // k-chunk
constexpr int kPad = 1;    // LDS row padding (doubles)
constexpr int kLd = kTile + kPad;
constexpr int kXcd = 8;

struct TileMap {
  int tiles;       // tiles per matrix
  int side;        // tiles per matrix side
  int herm;
};

// Tile t of a side x side grid: row-major, or the lower triangle row by row.
__device__ inline void tile_coords(int t, int side, int herm, int& ti, int& tj) {
  if (!herm) {
    ti = t / side;
    tj = t % side;
    return;
  }
  ti = 0;
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  tj = t - ti * (ti + 1) / 2;
}

// Arithmetic type of the products: FP64 (v_mfma_f64_16x16x4f64) or FP32
// (v_mfma_f32_16x16x4f32).  Operands are staged in LDS in that type; output
// and addend are FP64.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ v4 mac(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  // row of accumulator register r in lane group fk = lane / 16
  static __device__ int row(int r, int fk) { return 4 * r + fk; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ v4 mac(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ int row(int r, int fk) { return 4 * fk + r; }
};

// C = [D +] op(A) B, optionally scaled to diag(s) op(A) B diag(s) (s: n per
// matrix); D may alias A (it is read only), C aliases nothing.  The
// waves-per-SIMD attribute (FP64 2, FP32 4) bounds the registers so that the
// software-pipelined fragments fit next to the accumulators without spills.
template <typename V>
__device__ inline V zero2() {
  V v;
  v.x = 0;
  v.y = 0;
  return v;
}

// Optional epilogue of the Gram product G = Z^H Z of the Ogita-Aishima step:
// G itself is not stored; each element becomes the FP32 correction
//   F'_ij - I_ij = (S_ij - lt_j G_ij) / (lt_j - lt_i)   (i != j, |lt_i - lt_j| > thr)
//                = -G_ij / 2                           (i != j, inside a cluster)
//                = (1 - gd_i) / 2                      (i == j)
// from S = Z^H H Z (already stored), the exact diagonal gd of G and the
// Rayleigh quotients lt_i = S_ii / gd_i; thr = gap_rel * max(1, max |lt|).
// The tile at (0, 0) also writes the lt_i to lt ([nmat][n]) for the
// residuals of the correction products.
struct OaEpilogue {
  const double2* S = nullptr;  // [nmat][n][n], stride sC
  const double* gd = nullptr;  // [nmat][n]
  float2* F = nullptr;         // [nmat][n][n], stride sC
  double gap_rel = 0.0;
  double* lt = nullptr;        // [nmat][n]
};

// Optional epilogue of the Kubo product C = U^H (v_a + i v_b) U: C is not
// stored; with Q_ij = |C_ij|^2 / ((E_i - E_j)^2 + eta^2) (Q_ii = 0), the tile
// writes its column sums of Q (over its rows) to part[mat][0][tile row][n]
// and its row sums (over its columns) to part[mat][1][tile column][n];
// Omega_j = (sum of column sums - sum of row sums) / 2.
struct KuboEpilogue {
  const double* E = nullptr;  // [nmat][n]
  double eta2 = 0.0;
  double* part = nullptr;     // [nmat][2][tiles per side][n]
};

// SA, SB: storage of A (and D) and B, double2 or float2.
template <typename T, bool ADD, typename SA, typename SB>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 2 : 4, sizeof(T) == 8 ? 2 : 4)))
zgemm3m_kernel(int n, size_t nmat, TileMap tm, const SA* __restrict__ A, size_t sA, int conj_a,
               const SB* __restrict__ B, size_t sB, const SA* D, size_t sD, double2* __restrict__ C,
               size_t sC, const double* __restrict__ sc, OaEpilogue oa, KuboEpilogue kb) {
  typedef typename Mfma<T>::v4 v4;
  __shared__ double s_lt[kMaxN];  // Rayleigh quotients (OA epilogue)
  __shared__ double s_red[4];
  __shared__ double s_rq[2][kTile];  // Kubo column sums by wave row
  __shared__ double s_kr[2][kTile];     // Kubo row sums by wave column
  __shared__ T sA3[3][kKc][kLd];  // op(A) tile, k-major: re, im, re + im
  __shared__ T sB3[3][kKc][kLd];  // B tile, k-major
  // XCD-aware decomposition of the workgroup index
  const size_t g = blockIdx.x;
  const size_t loc = g / kXcd;
  const size_t mat = (loc / tm.tiles) * kXcd + g % kXcd;
  if (mat >= nmat) return;
  int ti, tj;
  tile_coords(static_cast<int>(loc % tm.tiles), tm.side, tm.herm, ti, tj);
  A += mat * sA;
  B += mat * sB;
  if (C) C += mat * sC;  // null when an epilogue consumes the product
  if (ADD) D += mat * sD;
  if (sc) sc += mat * n;
  const int row0 = ti * kTile, col0 = tj * kTile;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double thr = 0.0;
  if (oa.F) {  // workgroup-uniform
    oa.S += mat * sC;
    oa.gd += mat * n;
    oa.F += mat * sC;
    double m = 0.0;
    const bool put_lt = oa.lt && ti == 0 && tj == 0;
    for (int i = tid; i < n; i += 256) {
      const double l = oa.S[static_cast<size_t>(i) * n + i].x / oa.gd[i];
      s_lt[i] = l;
      if (put_lt) oa.lt[mat * n + i] = l;
      m = fmax(m, fabs(l));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (lane == 0) s_red[wave] = m;
    __syncthreads();
    thr = oa.gap_rel * fmax(fmax(fmax(s_red[0], s_red[1]), fmax(s_red[2], s_red[3])), 1.0);
  }
  auto oa_entry = [&](int i, int j, double gr, double gi, double2 sij) -> float2 {
    double fr, fi;
    if (i == j) {
      fr = 0.5 * (1.0 - oa.gd[i]);
      fi = 0.0;
    } else if (fabs(s_lt[i] - s_lt[j]) > thr) {
      const double inv = 1.0 / (s_lt[j] - s_lt[i]);
      fr = (sij.x - s_lt[j] * gr) * inv;
      fi = (sij.y - s_lt[j] * gi) * inv;
    } else {
      fr = -0.5 * gr;
      fi = -0.5 * gi;
    }
    return make_float2(static_cast<float>(fr), static_cast<float>(fi));
  };
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;

  // global -> register staging: 4 elements of op(A) and 4 of B per thread
  SA pa[4];
  SB pb[4];
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = tid + 256 * u;
      SA v = zero2<SA>();
      if (conj_a) {  // op(A)[r][k] = conj(A[k][r]); lanes along r
        const int kk = q / kTile, i = q % kTile;
        const int r = row0 + i, k = k0 + kk;
        if (r < n && k < n) v = A[static_cast<size_t>(k) * n + r];  // conjugated when staged
      } else {  // lanes along k within a row
        const int i = q / kKc, kk = q % kKc;
        const int r = row0 + i, k = k0 + kk;
        if (r < n && k < n) v = A[static_cast<size_t>(r) * n + k];
      }
      pa[u] = v;
      const int kk = q / kTile, j = q % kTile;
      const int k = k0 + kk, c = col0 + j;
      pb[u] = (k < n && c < n) ? B[static_cast<size_t>(k) * n + c] : zero2<SB>();
    }
  };
  auto store_chunk = [&]() {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = tid + 256 * u;
      int i, kk;
      if (conj_a) {
        kk = q / kTile;
        i = q % kTile;
      } else {
        i = q / kKc;
        kk = q % kKc;
      }
      const T ar = static_cast<T>(pa[u].x);
      const T ai = static_cast<T>(conj_a ? -pa[u].y : pa[u].y);
      sA3[0][kk][i] = ar;
      sA3[1][kk][i] = ai;
      sA3[2][kk][i] = ar + ai;
      const int kb = q / kTile, j = q % kTile;
      const T br = static_cast<T>(pb[u].x), bi = static_cast<T>(pb[u].y);
      sB3[0][kb][j] = br;
      sB3[1][kb][j] = bi;
      sB3[2][kb][j] = br + bi;
    }
  };

  v4 acc[3][2][2];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[p][a][b] = v4{T(0), T(0), T(0), T(0)};

  const int fr = lane & 15, fk = lane >> 4;
  // a wave whose 32 x 32 quadrant lies outside the matrix (small n, edge
  // tiles) stages its share of the tiles but issues no MFMAs: its outputs
  // are never stored, and the matrix cores of its SIMD serve other waves
  const bool quad_live = __builtin_amdgcn_readfirstlane(row0 + wr < n && col0 + wc < n ? 1 : 0) != 0;
  load_chunk(0);
  for (int k0 = 0; k0 < n; k0 += kKc) {
    __syncthreads();  // previous chunk fully consumed
    store_chunk();
    __syncthreads();
    if (k0 + kKc < n) load_chunk(k0 + kKc);  // in flight during the MFMAs
    if (!quad_live) continue;
    // the fragments of k-step s + 1 are read from LDS while the 12 MFMAs of
    // k-step s issue (two fragment register sets)
    T av[2][3][2], bv[2][3][2];
    auto frag = [&](int set, int kk) {
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          av[set][p][a] = sA3[p][kk + fk][wr + 16 * a + fr];
          bv[set][p][a] = sB3[p][kk + fk][wc + 16 * a + fr];
        }
    };
    frag(0, 0);
#pragma unroll
    for (int ks = 0; ks < kKc / 4; ++ks) {
      const int cur = ks & 1;
      if (ks + 1 < kKc / 4) frag(cur ^ 1, 4 * (ks + 1));
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            acc[p][a][b] = Mfma<T>::mac(av[cur][p][a], bv[cur][p][b], acc[p][a][b]);
    }
  }
  const bool mirror = tm.herm && ti != tj;
  double kb_col[2] = {0.0, 0.0};  // Kubo: per b column of this lane
  double kb_row[2][4] = {};       // Kubo: per (a, r) row of this lane
  double kb_er[2][4] = {}, kb_ec[2] = {};  // eigenvalues of this lane's rows and columns
  if (kb.part) {
    kb.E += mat * n;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wr + 16 * a + Mfma<T>::row(r, fk);
        kb_er[a][r] = i < n ? kb.E[i] : 0.0;
      }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = col0 + wc + 16 * b + fr;
      kb_ec[b] = j < n ? kb.E[j] : 0.0;
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wr + 16 * a + Mfma<T>::row(r, fk);
        const int j = col0 + wc + 16 * b + fr;
        if (i < n && j < n) {
          const double p1 = acc[0][a][b][r], p2 = acc[1][a][b][r], p3 = acc[2][a][b][r];
          double re = p1 - p2, im = p3 - p1 - p2;
          if (sc) {
            const double f = sc[i] * sc[j];
            re *= f;
            im *= f;
          }
          const size_t o = static_cast<size_t>(i) * n + j;
          if (kb.part) {  // the Kubo band-sum weights instead of C
            if (i != j) {
              // 1 / (d^2 + eta^2): hardware reciprocal and two Newton steps
              const double d = kb_er[a][r] - kb_ec[b];
              const double den = fma(d, d, kb.eta2);
              double rc = __builtin_amdgcn_rcp(den);
              rc = fma(rc, fma(-den, rc, 1.0), rc);
              rc = fma(rc, fma(-den, rc, 1.0), rc);
              const double q = (re * re + im * im) * rc;
              kb_col[b] += q;  // column j
              kb_row[a][r] += q;
            }
            continue;
          }
          if (oa.F) {  // the Ogita-Aishima correction instead of G (S Hermitian)
            const double2 sij = oa.S[o];
            oa.F[o] = oa_entry(i, j, re, im, sij);
            if (mirror)
              oa.F[static_cast<size_t>(j) * n + i] = oa_entry(j, i, re, -im, make_double2(sij.x, -sij.y));
            continue;
          }
          if (ADD) {
            const SA d = D[o];
            re += static_cast<double>(d.x);
            im += static_cast<double>(d.y);
          }
          if (C) {
            C[o] = make_double2(re, im);
            if (mirror) C[static_cast<size_t>(j) * n + i] = make_double2(re, -im);
          }
        }
      }
  if (kb.part) {  // workgroup-uniform: column and row sums of Q over the tile
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      kb_col[b] += __shfl_xor(kb_col[b], 16, 64);
      kb_col[b] += __shfl_xor(kb_col[b], 32, 64);
      if (fk == 0) s_rq[wave >> 1][wc + 16 * b + fr] = kb_col[b];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // the 16 lanes of a DPP row hold the 16 columns of one row
        const double x = row_sum16(kb_row[a][r]);
        if (fr == 0) s_kr[wave & 1][wr + 16 * a + Mfma<T>::row(r, fk)] = x;
      }
    __syncthreads();
    if (tid < kTile) {
      const size_t base = mat * 2 * static_cast<size_t>(tm.side);
      if (col0 + tid < n) kb.part[(base + ti) * n + col0 + tid] = s_rq[0][tid] + s_rq[1][tid];
      if (row0 + tid < n) kb.part[(base + tm.side + tj) * n + row0 + tid] = s_kr[0][tid] + s_kr[1][tid];
    }
  }
}

// Both correction products of the refinement in one pass over F' - I:
// Z' = Z + Z F and (HZ)' = HZ + HZ F on the FP32 matrix cores (Z, F in FP32
// storage, HZ in FP64), one 64 x 64 tile of each per workgroup sharing the
// staged F tile.  Z' is stored (FP64); (HZ)' is consumed in the epilogue:
// the Rayleigh sums Re(z'^H (HZ)') and z'^H z' and the squared residual
// |(HZ)' - lt_j z'|^2 of the tile's columns over its rows go to
// part[mat][tile row][3][n] (lt: [nmat][n], the Ogita-Aishima Rayleigh
// quotients).  Same lane layouts and pipelining as zgemm3m_kernel.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
zgemm3m_dual_kernel(int n, size_t nmat, TileMap tm, const float2* __restrict__ Z, const double2* __restrict__ HZ,
                    const float2* __restrict__ F, const double* __restrict__ lt, double2* __restrict__ Zp,
                    double* __restrict__ part) {
  typedef Mfma<float>::v4 v4;
  __shared__ float sA3[3][kKc][kLd];  // Z tile, k-major: re, im, re + im
  __shared__ float sH3[3][kKc][kLd];  // HZ tile (rounded to FP32)
  __shared__ float sB3[3][kKc][kLd];  // F tile
  __shared__ double s_rq[2][3][kTile];
  const size_t g = blockIdx.x;
  const size_t loc = g / kXcd;
  const size_t mat = (loc / tm.tiles) * kXcd + g % kXcd;
  if (mat >= nmat) return;
  int ti, tj;
  tile_coords(static_cast<int>(loc % tm.tiles), tm.side, 0, ti, tj);
  const size_t n2 = static_cast<size_t>(n) * n;
  Z += mat * n2;
  HZ += mat * n2;
  F += mat * n2;
  Zp += mat * n2;
  const int row0 = ti * kTile, col0 = tj * kTile;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  float2 pz[4], pf[4];
  double2 ph[4];
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = tid + 256 * u;
      {  // lanes along k within a row
        const int i = q / kKc, kk = q % kKc;
        const int r = row0 + i, k = k0 + kk;
        const bool ok = r < n && k < n;
        pz[u] = ok ? Z[static_cast<size_t>(r) * n + k] : zero2<float2>();
        ph[u] = ok ? HZ[static_cast<size_t>(r) * n + k] : zero2<double2>();
      }
      const int kk = q / kTile, j = q % kTile;
      const int k = k0 + kk, c = col0 + j;
      pf[u] = (k < n && c < n) ? F[static_cast<size_t>(k) * n + c] : zero2<float2>();
    }
  };
  auto store_chunk = [&]() {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = tid + 256 * u;
      const int i = q / kKc, kk = q % kKc;
      sA3[0][kk][i] = pz[u].x;
      sA3[1][kk][i] = pz[u].y;
      sA3[2][kk][i] = pz[u].x + pz[u].y;
      const float hr = static_cast<float>(ph[u].x), hi = static_cast<float>(ph[u].y);
      sH3[0][kk][i] = hr;
      sH3[1][kk][i] = hi;
      sH3[2][kk][i] = hr + hi;
      const int kb = q / kTile, j = q % kTile;
      sB3[0][kb][j] = pf[u].x;
      sB3[1][kb][j] = pf[u].y;
      sB3[2][kb][j] = pf[u].x + pf[u].y;
    }
  };
  v4 acz[3][2][2], ach[3][2][2];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        acz[p][a][b] = v4{0.f, 0.f, 0.f, 0.f};
        ach[p][a][b] = v4{0.f, 0.f, 0.f, 0.f};
      }
  const int fr = lane & 15, fk = lane >> 4;
  // quadrants outside the matrix issue no MFMAs (as in zgemm3m_kernel)
  const bool quad_live = __builtin_amdgcn_readfirstlane(row0 + wr < n && col0 + wc < n ? 1 : 0) != 0;
  load_chunk(0);
  for (int k0 = 0; k0 < n; k0 += kKc) {
    __syncthreads();
    store_chunk();
    __syncthreads();
    if (k0 + kKc < n) load_chunk(k0 + kKc);
    if (!quad_live) continue;
    // fragments of k-step s + 1 read from LDS while the 24 MFMAs of k-step s
    // issue (two register sets)
    float az[2][3][2], ah[2][3][2], bv[2][3][2];
    auto frag = [&](int set, int kk) {
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          az[set][p][a] = sA3[p][kk + fk][wr + 16 * a + fr];
          ah[set][p][a] = sH3[p][kk + fk][wr + 16 * a + fr];
          bv[set][p][a] = sB3[p][kk + fk][wc + 16 * a + fr];
        }
    };
    frag(0, 0);
#pragma unroll
    for (int ks = 0; ks < kKc / 4; ++ks) {
      const int cur = ks & 1;
      if (ks + 1 < kKc / 4) frag(cur ^ 1, 4 * (ks + 1));
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            acz[p][a][b] = Mfma<float>::mac(az[cur][p][a], bv[cur][p][b], acz[p][a][b]);
            ach[p][a][b] = Mfma<float>::mac(ah[cur][p][a], bv[cur][p][b], ach[p][a][b]);
          }
    }
  }
  double num[2] = {0.0, 0.0}, den[2] = {0.0, 0.0}, res[2] = {0.0, 0.0}, ltc[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int j = col0 + wc + 16 * b + fr;
    ltc[b] = j < n ? lt[mat * n + j] : 0.0;
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wr + 16 * a + Mfma<float>::row(r, fk);
        const int j = col0 + wc + 16 * b + fr;
        if (i < n && j < n) {
          const size_t o = static_cast<size_t>(i) * n + j;
          const float2 z = Z[o];
          const double2 h = HZ[o];
          double p1 = acz[0][a][b][r], p2 = acz[1][a][b][r], p3 = acz[2][a][b][r];
          const double zr = z.x + (p1 - p2), zi = z.y + (p3 - p1 - p2);
          p1 = ach[0][a][b][r];
          p2 = ach[1][a][b][r];
          p3 = ach[2][a][b][r];
          const double hr = h.x + (p1 - p2), hi = h.y + (p3 - p1 - p2);
          Zp[o] = make_double2(zr, zi);
          num[b] += zr * hr + zi * hi;  // Re(conj(z') h')
          den[b] += zr * zr + zi * zi;
          const double rr = fma(-ltc[b], zr, hr), ri = fma(-ltc[b], zi, hi);
          res[b] += rr * rr + ri * ri;
        }
      }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    num[b] += __shfl_xor(num[b], 16, 64);
    num[b] += __shfl_xor(num[b], 32, 64);
    den[b] += __shfl_xor(den[b], 16, 64);
    den[b] += __shfl_xor(den[b], 32, 64);
    res[b] += __shfl_xor(res[b], 16, 64);
    res[b] += __shfl_xor(res[b], 32, 64);
    if (fk == 0) {
      s_rq[wave >> 1][0][wc + 16 * b + fr] = num[b];
      s_rq[wave >> 1][1][wc + 16 * b + fr] = den[b];
      s_rq[wave >> 1][2][wc + 16 * b + fr] = res[b];
    }
  }
  __syncthreads();
  if (tid < kTile && col0 + tid < n) {
    double* pt = part + (mat * tm.side + ti) * 3 * static_cast<size_t>(n);
    pt[col0 + tid] = s_rq[0][0][tid] + s_rq[1][0][tid];
    pt[n + col0 + tid] = s_rq[0][1][tid] + s_rq[1][1][tid];
    pt[2 * n + col0 + tid] = s_rq[0][2][tid] + s_rq[1][2][tid];
  }
}

template <typename T, bool ADD, typename SA, typename SB>
int zgemm_launch(hipStream_t st, int n, size_t nb, const ZOperand& A, bool conj_a, const ZOperand& B,
                 const ZOperand& D, double* C, size_t sC, bool herm, const double* sc,
                 const OaEpilogue& oa = OaEpilogue{}, const KuboEpilogue& kb = KuboEpilogue{}) {
  if (nb == 0 || n <= 0) return RTB_OK;
  TileMap tm;
  tm.side = (n + kTile - 1) / kTile;
  tm.herm = herm ? 1 : 0;
  tm.tiles = herm ? tm.side * (tm.side + 1) / 2 : tm.side * tm.side;
  const SA* pA = static_cast<const SA*>(A.p);
  const SB* pB = static_cast<const SB*>(B.p);
  const SA* pD = static_cast<const SA*>(D.p);
  // grid: kXcd interleaved streams of (matrix group, tile); one launch per
  // slice of at most 2^24 workgroups (grid size in work-items < 2^32)
  const size_t per_launch = (static_cast<size_t>(1) << 24) / (kXcd * tm.tiles) * kXcd;
  for (size_t off = 0; off < nb; off += per_launch) {
    const size_t cnt = nb - off < per_launch ? nb - off : per_launch;
    const size_t groups = (cnt + kXcd - 1) / kXcd;
    const size_t wgs = groups * tm.tiles * kXcd;
    hipLaunchKernelGGL((zgemm3m_kernel<T, ADD, SA, SB>), dim3(static_cast<unsigned>(wgs)), dim3(256), 0, st, n, cnt,
                       tm, pA + off * A.stride, A.stride, conj_a ? 1 : 0, pB + off * B.stride, B.stride,
                       pD ? pD + off * D.stride : nullptr, D.stride,
                       C ? reinterpret_cast<double2*>(C) + off * sC : nullptr, sC, sc ? sc + off * n : nullptr,
                       OaEpilogue{oa.S ? oa.S + off * sC : nullptr, oa.gd ? oa.gd + off * n : nullptr,
                                  oa.F ? oa.F + off * sC : nullptr, oa.gap_rel, oa.lt ? oa.lt + off * n : nullptr},
                       KuboEpilogue{kb.E ? kb.E + off * n : nullptr, kb.eta2,
                                    kb.part ? kb.part + off * tm.side * 2 * static_cast<size_t>(n) : nullptr});
  }
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

}  // namespace

int zgemm(hipStream_t st, int n, size_t nb, ZOperand A, bool conj_a, ZOperand B, double* C, size_t sC, bool herm,
          const double* s) {
  const ZOperand none{};
  if (!A.f32 && !B.f32) return zgemm_launch<double, false, double2, double2>(st, n, nb, A, conj_a, B, none, C, sC, herm, s);
  if (!A.f32 && B.f32) return zgemm_launch<double, false, double2, float2>(st, n, nb, A, conj_a, B, none, C, sC, herm, s);
  if (A.f32 && !B.f32) return zgemm_launch<double, false, float2, double2>(st, n, nb, A, conj_a, B, none, C, sC, herm, s);
  return zgemm_launch<double, false, float2, float2>(st, n, nb, A, conj_a, B, none, C, sC, herm, s);
}

int zgemm_gram_oa(hipStream_t st, int n, size_t nb, const float* Z, const double* S, const double* gd, float* F,
                  double gap_rel, double* lt) {
  if (n > kMaxN) return fail(RTB_ERR_UNSUPPORTED, "zgemm_gram_oa supports n <= 256");
  const size_t n2 = static_cast<size_t>(n) * n;
  OaEpilogue oa;
  oa.S = reinterpret_cast<const double2*>(S);
  oa.gd = gd;
  oa.F = reinterpret_cast<float2*>(F);
  oa.gap_rel = gap_rel;
  oa.lt = lt;
  return zgemm_launch<double, false, float2, float2>(st, n, nb, zop(Z, n2), true, zop(Z, n2), ZOperand{}, nullptr,
                                                     n2, true, nullptr, oa);
}

int zgemm_kubo(hipStream_t st, int n, size_t nb, const double* U, const double* T1, const double* s,
               const double* E, double eta, double* part) {
  const size_t n2 = static_cast<size_t>(n) * n;
  KuboEpilogue kb;
  kb.E = E;
  kb.eta2 = eta * eta;
  kb.part = part;
  return zgemm_launch<double, false, double2, double2>(st, n, nb, zop(U, n2), true, zop(T1, n2), ZOperand{},
                                                       nullptr, n2, false, s, OaEpilogue{}, kb);
}

int zgemm_corrections_f32(hipStream_t st, int n, size_t nb, const float* Z, const double* HZ, const float* F,
                          const double* lt, double* Zp, double* part) {
  if (nb == 0 || n <= 0) return RTB_OK;
  TileMap tm;
  tm.side = (n + kTile - 1) / kTile;
  tm.herm = 0;
  tm.tiles = tm.side * tm.side;
  const size_t n2 = static_cast<size_t>(n) * n;
  const size_t per_launch = (static_cast<size_t>(1) << 24) / (kXcd * tm.tiles) * kXcd;
  for (size_t off = 0; off < nb; off += per_launch) {
    const size_t cnt = nb - off < per_launch ? nb - off : per_launch;
    const size_t wgs = (cnt + kXcd - 1) / kXcd * tm.tiles * kXcd;
    hipLaunchKernelGGL(zgemm3m_dual_kernel, dim3(static_cast<unsigned>(wgs)), dim3(256), 0, st, n, cnt, tm,
                       reinterpret_cast<const float2*>(Z) + off * n2, reinterpret_cast<const double2*>(HZ) + off * n2,
                       reinterpret_cast<const float2*>(F) + off * n2, lt + off * n,
                       reinterpret_cast<double2*>(Zp) + off * n2, part + off * tm.side * 3 * static_cast<size_t>(n));
  }
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int zgemm_batched(hipStream_t st, int n, size_t nb, const double* A, size_t sA, bool conj_a, const double* B,
                  size_t sB, double* C, size_t sC, bool herm) {
  return zgemm(st, n, nb, zop(A, sA), conj_a, zop(B, sB), C, sC, herm, nullptr);
}

int zgemm_batched_scaled(hipStream_t st, int n, size_t nb, const double* A, size_t sA, bool conj_a, const double* B,
                         size_t sB, double* C, size_t sC, bool herm, const double* s) {
  return zgemm(st, n, nb, zop(A, sA), conj_a, zop(B, sB), C, sC, herm, s);
}

}  // namespace rtb
