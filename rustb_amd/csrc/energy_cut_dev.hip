// The EnergyCut Fermi cut on the device: vertex Omega in tracked band order,
// then the zero-width cut of every (cell, simplex, band) for a list of
// chemical potentials (integrate_fermi_cut_2d_t0 / _3d_t0,
// src/response/energy_cut.rs:1231-1345, 1650-1797).  The per-simplex rules
// are energy_cut_core.hpp, the same code as the host part (energy_cut.cpp).
//
// One lane per (cell, simplex, band), lanes of a wave on consecutive bands.
// For each mu a lane settles the empty and full cases itself; the few that
// are cut (mu strictly inside the band's range on the simplex) are taken by
// the whole wave, two at a time for n <= 32: the 32 or 64 lanes of a group
// split the band sum over m of each quadrature point and add their shares in
// a fixed butterfly, so the loads of E and Im K rows are coalesced instead of
// one lane walking n bands.  The 64 lanes' contributions are then summed by
// DPP in a fixed order into a per-wave partial, and the partials are summed
// per mu in a fixed order (column_sums): the result does not depend on
// scheduling.
#include <limits>

#include "eigh.hpp"
#include "energy_cut_core.hpp"
#include "gpu_team.hpp"

namespace rtb {

namespace {

__global__ void cut_omega_kernel(cut::View V, size_t nk, double eta, double* __restrict__ om) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g >= nk * static_cast<size_t>(V.n)) return;
  om[g] = cut::vertex_omega(V, g / V.n, static_cast<int>(g % V.n), eta);
}

// simplex s of a cell with band b, item g = (cell * S + s) * n + b
template <int DIM>
struct Simplex {
  static constexpr int NV = DIM + 1;
  size_t vs[NV];
  double c[NV][DIM];
  double e_v[NV], om_v[NV];
  double measure, scale;
  int b;
};

template <int DIM>
__device__ inline void load_simplex(const cut::View& V, const double* __restrict__ om, size_t g, size_t nx, size_t ny,
                                    size_t nz, Simplex<DIM>& x) {
  constexpr int S = DIM == 2 ? 4 : 5;  // simplices per cell
  const int n = V.n;
  const size_t cell = g / (static_cast<size_t>(S) * n);
  const int s = static_cast<int>((g / n) % S);
  x.b = static_cast<int>(g % n);
  if constexpr (DIM == 2) {
    cut::cell_triangle(cell / ny, cell % ny, nx, ny, s, x.vs, x.c);
    x.measure = cut::tri_area(x.c);
    const double half_vol = (1.0 / static_cast<double>(nx)) * (1.0 / static_cast<double>(ny)) / 4.0;
    x.scale = half_vol / x.measure;
  } else {
    double vol;
    cut::cell_tetrahedron(cell / (ny * nz), (cell / nz) % ny, cell % nz, nx, ny, nz, s, x.vs, x.c, &vol);
    x.measure = cut::tet_vol(x.c[0], x.c[1], x.c[2], x.c[3]);
    x.scale = vol / x.measure;
  }
  for (int q = 0; q < Simplex<DIM>::NV; ++q) {
    x.e_v[q] = cut::band(V, x.vs[q], x.b);
    x.om_v[q] = om[x.vs[q] * n + x.b];
  }
}

// Lane t's share of the band sum of eval_berry_band_at_lam_buf
// (kernel.rs:138-182): the terms m = t, t + G, ..; the G lanes' shares add up
// to the sum.  The lane's KM terms (E and Im K behind the tracking
// permutation) are gathered once per cut item, and every quadrature point of
// the item is evaluated from registers.
template <int NV, int KM>
struct CachedLaneBandSum {
  double eb[NV];
  double em[KM][NV], km[KM][NV];
  bool use[KM];
  double eta2;
  __device__ CachedLaneBandSum(const cut::View& M, const size_t* vs, int b, double eta, int t, int G) {
    eta2 = eta * eta;
#pragma unroll
    for (int v = 0; v < NV; ++v) eb[v] = cut::band(M, vs[v], b);
#pragma unroll
    for (int u = 0; u < KM; ++u) {
      const int m = t + G * u;
      use[u] = m < M.n && m != b;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        em[u][v] = use[u] ? cut::band(M, vs[v], m) : 0.0;
        km[u][v] = use[u] ? cut::kim(M, vs[v], b, m) : 0.0;
      }
    }
  }
  template <int NVV>
  __device__ double at(const double* lam) const {
    static_assert(NVV == NV, "vertex count");
    double en = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v)
      if (lam[v] != 0.0) en += eb[v] * lam[v];
    double im = 0.0;
#pragma unroll
    for (int u = 0; u < KM; ++u) {
      if (!use[u]) continue;
      double e = 0.0, k = 0.0;
#pragma unroll
      for (int v = 0; v < NV; ++v)
        if (lam[v] != 0.0) e += em[u][v] * lam[v];
#pragma unroll
      for (int v = 0; v < NV; ++v)
        if (lam[v] != 0.0) k += km[u][v] * lam[v];
      const double de = en - e;
      const double denom = de * de + eta2;
      if (denom < 1e-30) continue;
      im += k / denom;
    }
    return -2.0 * im;
  }
};

// KM: terms per lane, ceil(n / G) with G = 32 lanes per item for n <= 32,
// 64 above.
template <int DIM, int KM>
__global__ void __launch_bounds__(256) cut_kernel(cut::View V, const double* __restrict__ om, size_t nx, size_t ny,
                                                  size_t nz, const double* __restrict__ mu, int n_mu, double eta,
                                                  double* __restrict__ partial) {
  constexpr int S = DIM == 2 ? 4 : 5;
  constexpr int NV = DIM + 1;
  const int n = V.n;
  const int G = n <= 32 ? 32 : 64;  // lanes per cut item
  const int lane = threadIdx.x & 63;
  const int t = lane & (G - 1);
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  const size_t g0 = g - lane;  // the wave's first item
  const size_t total = nx * ny * nz * S * static_cast<size_t>(n);
  bool live = g < total;
  Simplex<DIM> x{};
  double full_val = 0.0, e_min = 0.0, e_max = 0.0;
  if (live) {
    load_simplex<DIM>(V, om, g, nx, ny, nz, x);
    live = !(x.measure < cut::kEps);  // degenerate simplices of ragged meshes
    e_min = e_max = x.e_v[0];
    double osum = x.om_v[0];
    for (int q = 1; q < NV; ++q) {
      e_min = cut::dmin(e_min, x.e_v[q]);
      e_max = cut::dmax(e_max, x.e_v[q]);
      osum += x.om_v[q];
    }
    full_val = x.measure * osum / static_cast<double>(NV);
  }
  double* out = partial + (g / 64) * static_cast<size_t>(n_mu);
  for (int im = 0; im < n_mu; ++im) {
    const double m = mu[im];
    double v = 0.0;
    bool need = false;
    if (live) {
      // the partition points of the host loop: mu >= e_max - eps is the full
      // range, mu <= e_min + eps below it is empty, the rest is cut
      if (!(m < e_max - cut::kEps))
        v = x.scale * full_val;
      else
        need = !(m <= e_min + cut::kEps);
    }
    unsigned long long mask = __ballot(need);
    while (mask) {  // wave-uniform: every lane runs every pass
      const int l0 = __builtin_ctzll(mask);
      mask &= mask - 1;
      int l1 = -1;
      if (G == 32 && mask) {
        l1 = __builtin_ctzll(mask);
        mask &= mask - 1;
      }
      const int src = (G == 32 && lane >= 32) ? l1 : l0;  // this group's item
      double r = 0.0;
      if (src >= 0) {
        Simplex<DIM> y;
        load_simplex<DIM>(V, om, g0 + src, nx, ny, nz, y);
        const CachedLaneBandSum<NV, KM> f(V, y.vs, y.b, eta, t, G);
        const double one = t == 0 ? 1.0 : 0.0;
        if constexpr (DIM == 2)
          r = cut::hybrid_tri_by(y.c, y.e_v, y.om_v, m, f, one);
        else
          r = cut::hybrid_tet_by(y.c, y.e_v, y.om_v, m, f, one);
      }
      for (int o = G / 2; o > 0; o >>= 1) r += __shfl_xor(r, o);
      // group 0 (lanes 0..G-1) holds item l0's rule, group 1 item l1's; the
      // owners may sit in either half
      const double r0 = __shfl(r, 0), r1 = __shfl(r, 32);
      if (lane == l0) v = x.scale * r0;
      if (lane == l1) v = x.scale * r1;
    }
    v = wave_sum64(v);
    if (lane == 0) out[im] = v;
  }
}

// ---- band tracking (global_band_track, tracking.rs:123-163) ----
// One workgroup per vertex of a level: the squared overlaps of its scheduled
// neighbours (rows in the neighbour's tracked order) summed in neighbour
// order and divided by their count into M, as the host walk forms them;
// then greedy_assign (tracking.rs:44-68): rows by descending row maximum
// (stable), each taking its largest unassigned column, the lowest column
// among equal values.  A row whose overall best column (same scan rule) is
// still free takes it, which is what the scan over the free columns returns;
// otherwise wave 0 scans the free columns.  The same operations on the same
// values as the host walk, so the permutations are identical.  M [n][n] sits
// in LDS up to kTrackLdsMaxN bands and above in a global slice per
// workgroup of the launch (n = 256: 512 KiB); there a wave builds each row
// and ranks its first kTrackTop columns in registers, a row whose best
// column is taken takes the first free one of them, and only when all are
// taken does wave 0 scan the row in global memory.
constexpr int kTrackMaxDep = 6;  // visited neighbours of a 3D vertex at most
static_assert(sizeof(TrackEntry) == 8 * sizeof(unsigned), "schedule entry layout");
static_assert(kTrackDevMaxN <= 256, "the greedy's assigned-column mask is four 64-bit words");

// LDS bytes: M when it is in LDS, then rmax [n], best/order/bsort [n] each,
// the neighbours' permutations [kTrackMaxDep][n] and, with M in global
// memory, each row's first kTrackTop columns [n][kTrackTop]
constexpr int kTrackTop = 4;
size_t track_smem(int n) {
  const bool lds = n <= kTrackLdsMaxN;
  const size_t m = lds ? static_cast<size_t>(n) * n * sizeof(double) : 0;
  return m + n * sizeof(double) + (3 + kTrackMaxDep + (lds ? 0 : kTrackTop)) * static_cast<size_t>(n) * sizeof(int);
}

__global__ void identity_perm_kernel(size_t total, int n, int* __restrict__ perm) {
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  if (g < total) perm[g] = static_cast<int>(g % n);
}

// The overlap sum over the neighbours divided by their count: x / 1, x * 0.5
// and x * 0.25 are the correctly rounded x / c for c = 1, 2, 4 (exact
// scalings), so only 3, 5 and 6 take the division.
__device__ inline double by_count(double x, int cnt, double dc) {
  return cnt == 1 ? x : (cnt == 2 ? x * 0.5 : (cnt == 4 ? x * 0.25 : x / dc));
}

// kTrackThreads per vertex with M in LDS (one wave of 64 was slower at
// n = 32: the matrix build loses its parallelism); with M in global memory
// kTrackThreadsWide, one row per wave in flight: the build waits on global
// loads, and a level holds few vertices.
constexpr int kTrackThreads = 256;
constexpr int kTrackThreadsWide = 1024;
template <bool kLdsM>
__global__ void __launch_bounds__(kLdsM ? kTrackThreads : kTrackThreadsWide) track_kernel(const double* __restrict__ ovx, const double* __restrict__ ovy,
                                                   const double* __restrict__ ovz, size_t ny, size_t nz, int n,
                                                   const TrackEntry* __restrict__ sched, double* __restrict__ mwork,
                                                   int* __restrict__ perm) {
  constexpr int NT = kLdsM ? kTrackThreads : kTrackThreadsWide;
  extern __shared__ double smem[];
  const int nn = n * n;
  const size_t n2 = static_cast<size_t>(nn);
  double* M = kLdsM ? smem : mwork + blockIdx.x * n2;  // [n][n]
  double* rmax = kLdsM ? smem + n2 : smem;               // [n]
  int* best = reinterpret_cast<int*>(rmax + n);          // [n] each row's best column over all columns
  int* order = best + n;                                 // [n] rows by descending maximum
  int* bsort = order + n;                                // [n] best[order[s]]
  int* pv = bsort + n;                                   // [kTrackMaxDep][n] the neighbours' permutations
  int* top = pv + kTrackMaxDep * n;                      // [n][kTrackTop] (M in global memory only)
  const TrackEntry te = sched[blockIdx.x];              // uniform: one scalar load
  const size_t nb = te.v;
  const int cnt = static_cast<int>(te.cnt);
  if (cnt == 0) return;  // vertex 0 keeps the identity
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < kTrackMaxDep; ++r)
    if (r < cnt)
      for (int a = tid; a < n; a += NT) pv[r * n + a] = perm[static_cast<size_t>(te.dep[r]) * n + a];
  // each neighbour's overlap block and orientation; the pair's axis from its
  // index step, x first (with ny = 1 the x and y steps coincide and only x
  // pairs exist)
  const double* base[kTrackMaxDep];
  bool lower[kTrackMaxDep];  // nb is the pair's lower vertex: its raw band is the row
#pragma unroll
  for (int r = 0; r < kTrackMaxDep; ++r) {
    base[r] = nullptr;
    lower[r] = false;
    if (r < cnt) {
      const size_t vn = te.dep[r];
      const size_t lo = nb < vn ? nb : vn, step = (nb < vn ? vn : nb) - lo;
      const double* axis = step == ny * nz ? ovx : (step == nz ? ovy : ovz);
      base[r] = axis + lo * n2;
      lower[r] = nb == lo;
    }
  }
  __syncthreads();
  const double dc = static_cast<double>(cnt);
  // entry (a, b) of M: the neighbours' values in neighbour order
  const auto entry = [&](int a, int b) {
    double x = 0.0;
#pragma unroll
    for (int r = 0; r < kTrackMaxDep; ++r) {
      if (r < cnt) {
        const size_t pa = static_cast<size_t>(pv[r * n + a]);
        const double y = lower[r] ? base[r][static_cast<size_t>(b) * n + pa] : base[r][pa * n + b];
        x = r == 0 ? y : x + y;
      }
    }
    return by_count(x, cnt, dc);
  };
  if constexpr (kLdsM) {
    for (int e = tid; e < nn; e += NT) {
      const int a = e / n;
      M[e] = entry(a, e - a * n);
    }
    __syncthreads();
    // each row's maximum (fmax from -1: NaN dropped) and best column, the
    // lowest among equal values: a row per thread, the scan rotated to
    // spread the LDS banks
    for (int i = tid; i < n; i += NT) {
      double bv = -1.0;
      int bj = 0;
      for (int q = 0; q < n; ++q) {
        const int j = q + i < n ? q + i : q + i - n;
        const double v = M[i * n + j];
        if (v > bv || (v == bv && j < bj)) {
          bv = v;
          bj = j;
        }
      }
      rmax[i] = bv;
      best[i] = bj;
    }
  } else {
    // a row per wave, each lane its columns b = lane + 64 c: the row is
    // built, stored, and ranked from registers.  The columns in the order
    // of the host scan (larger value first, then lower column; NaN never
    // ranked) are found kTrackTop times over, each round among the columns
    // after the previous round's; the lanes' (value, column) pairs merge by
    // that order, and a lane or row with no candidate gives (-1, 0), which
    // every candidate (>= 0) beats.  Round 0 is the row's maximum and best
    // column as the host scan gives them.
    const int wv = tid >> 6, ln = tid & 63;
    for (int a = wv; a < n; a += NT / 64) {
      double xv[kTrackDevMaxN / 64];
#pragma unroll
      for (int c = 0; c < kTrackDevMaxN / 64; ++c) {
        const int b = ln + 64 * c;
        xv[c] = __builtin_nan("");
        if (b < n) {
          xv[c] = entry(a, b);
          M[static_cast<size_t>(a) * n + b] = xv[c];
        }
      }
      double pvl = 0.0;
      int pcl = -1;
#pragma unroll
      for (int k = 0; k < kTrackTop; ++k) {
        double bv = -1.0;
        int bj = 0;
#pragma unroll
        for (int c = 0; c < kTrackDevMaxN / 64; ++c) {
          const double x = xv[c];
          const int b = ln + 64 * c;
          const bool after = k == 0 || x < pvl || (x == pvl && b > pcl);
          if (after && x > bv) {
            bv = x;
            bj = b;
          }
        }
        for (int o = 32; o > 0; o >>= 1) {
          const double ov = __shfl_xor(bv, o);
          const int oj = __shfl_xor(bj, o);
          if (ov > bv || (ov == bv && oj < bj)) {
            bv = ov;
            bj = oj;
          }
        }
        if (ln == 0) {
          if (k == 0) {
            rmax[a] = bv;
            best[a] = bj;
          }
          top[a * kTrackTop + k] = bv >= 0.0 ? bj : -1;  // -1: no further candidate
        }
        pvl = bv;
        pcl = bj;
      }
    }
  }
  __syncthreads();
  // the stable descending sort as ranks
  for (int i = tid; i < n; i += NT) {
    const double x = rmax[i];
    int rank = 0;
    for (int k = 0; k < n; ++k) {
      const double y = rmax[k];
      rank += (y > x || (y == x && k < i)) ? 1 : 0;
    }
    order[rank] = i;
    bsort[rank] = best[i];
  }
  __syncthreads();
  if (tid >= 64) return;
  const int lane = tid;
  // assigned columns 64 c .. 64 c + 63 in word c (the same in every lane);
  // NW named words rather than an array indexed at run time
  constexpr int NW = kLdsM ? kTrackLdsMaxN / 64 : kTrackDevMaxN / 64;
  static_assert(NW == 2 || NW == 4, "two or four assigned-column words");
  unsigned long long tk0 = 0, tk1 = 0, tk2 = 0, tk3 = 0;
  const auto word = [&](int c) {
    if constexpr (NW == 2) return c == 0 ? tk0 : tk1;
    return c == 0 ? tk0 : (c == 1 ? tk1 : (c == 2 ? tk2 : tk3));
  };
  const auto taken = [&](int j) { return ((word(j >> 6) >> (j & 63)) & 1ull) != 0; };
  int* out = perm + nb * n;
  int r_next = order[0], j_next = bsort[0];  // read one step ahead
  for (int s = 0; s < n; ++s) {
    const int r = r_next;
    int j = j_next;
    if (s + 1 < n) {
      r_next = order[s + 1];
      j_next = bsort[s + 1];
    }
    bool scan = taken(j);  // the same in every lane
    if constexpr (!kLdsM) {
      // the first free column of the row's ranked list (entry 0 is j) is the
      // free columns' maximum; the full scan only when the list runs out
      if (scan) {
#pragma unroll
        for (int k = 1; k < kTrackTop; ++k) {
          const int c = top[r * kTrackTop + k];
          if (c < 0) break;
          if (!taken(c)) {
            j = c;
            scan = false;
            break;
          }
        }
      }
    }
    if (scan) {  // the scan over the free columns
      // no candidate: (-1, 0), as the host scan's start; every real
      // candidate is >= 0 and NaN never wins
      double bv = -1.0;
      int bj = 0;
#pragma unroll
      for (int c = 0; c < NW; ++c) {
        const int jj = lane + 64 * c;
        if (jj < n) {
          const bool t = (word(c) >> lane) & 1ull;
          const double v = M[static_cast<size_t>(r) * n + jj];
          if (!t && v > bv) {
            bv = v;
            bj = jj;
          }
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o);
        const int oj = __shfl_xor(bj, o);
        if (ov > bv || (ov == bv && oj < bj)) {
          bv = ov;
          bj = oj;
        }
      }
      j = bj;
    }
    const unsigned long long bit = 1ull << (j & 63);
    if constexpr (NW == 2) {
      if (j < 64)
        tk0 |= bit;
      else
        tk1 |= bit;
    } else {
      switch (j >> 6) {
        case 0: tk0 |= bit; break;
        case 1: tk1 |= bit; break;
        case 2: tk2 |= bit; break;
        default: tk3 |= bit; break;
      }
    }
    if (lane == 0) out[r] = j;
  }
}

}  // namespace

int energy_cut_omega(hipStream_t st, int n, size_t nk, const double* E, const double* Ki, const int* perm, double eta,
                     double* om) {
  const size_t total = nk * static_cast<size_t>(n);
  if (total == 0) return RTB_OK;
  const cut::View V{n, E, Ki, 1, perm};
  hipLaunchKernelGGL(cut_omega_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, V, nk,
                     eta, om);
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

size_t energy_cut_waves(int dim, size_t nx, size_t ny, size_t nz, int n) {
  const size_t total = nx * ny * nz * (dim == 2 ? 4 : 5) * static_cast<size_t>(n);
  return (total + 255) / 256 * 4;
}

namespace {

template <int DIM>
int launch_cut_dim(hipStream_t st, int n, unsigned blocks, const cut::View& V, const double* om, size_t nx, size_t ny,
                   size_t nz, const double* mu, int c, double eta, double* partial) {
  // the kernel's lanes per item: 32 for n <= 32, else 64; KM = ceil(n / G)
  const int G = n <= 32 ? 32 : 64;
  const int km = (n + G - 1) / G;
  if (km <= 1)
    hipLaunchKernelGGL((cut_kernel<DIM, 1>), dim3(blocks), dim3(256), 0, st, V, om, nx, ny, nz, mu, c, eta, partial);
  else if (km == 2)
    hipLaunchKernelGGL((cut_kernel<DIM, 2>), dim3(blocks), dim3(256), 0, st, V, om, nx, ny, nz, mu, c, eta, partial);
  else if (km <= 4)
    hipLaunchKernelGGL((cut_kernel<DIM, 4>), dim3(blocks), dim3(256), 0, st, V, om, nx, ny, nz, mu, c, eta, partial);
  else
    return fail(RTB_ERR_UNSUPPORTED, "the Fermi cut is for n <= 256");
  RTB_HIP(hipGetLastError());
  return RTB_OK;
}

int launch_cut(hipStream_t st, int dim, int n, unsigned blocks, const cut::View& V, const double* om, size_t nx,
               size_t ny, size_t nz, const double* mu, int c, double eta, double* partial) {
  if (dim == 2) return launch_cut_dim<2>(st, n, blocks, V, om, nx, ny, nz, mu, c, eta, partial);
  return launch_cut_dim<3>(st, n, blocks, V, om, nx, ny, nz, mu, c, eta, partial);
}

}  // namespace

int energy_cut_t0(hipStream_t st, int dim, size_t nx, size_t ny, size_t nz, int n, const double* E, const double* Ki,
                  const int* perm, const double* om, const double* mu, size_t n_mu, double eta, double* partial,
                  size_t chunk, double* out) {
  if (dim != 2 && dim != 3) return fail(RTB_ERR_INVALID_DIMENSION, "the cut is for 2D and 3D meshes");
  if (chunk == 0) return fail(RTB_ERR_INVALID_ARGUMENT, "zero mu chunk");
  const size_t waves = energy_cut_waves(dim, nx, ny, nz, n);
  const unsigned blocks = static_cast<unsigned>(waves / 4);
  const cut::View V{n, E, Ki, 1, perm};
  for (size_t m0 = 0; m0 < n_mu; m0 += chunk) {
    const int c = static_cast<int>(std::min(chunk, n_mu - m0));
    RTB_TRY(launch_cut(st, dim, n, blocks, V, om, nx, ny, nz, mu + m0, c, eta, partial));
    RTB_TRY(column_sums(st, partial, waves, c, out + m0));
  }
  return RTB_OK;
}

int energy_cut_track_smem_limits() {
  RTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&track_kernel<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(track_smem(kTrackLdsMaxN))));
  return RTB_OK;
}

int energy_cut_track_dev(hipStream_t st, size_t nx, size_t ny, size_t nz, int n, const double* ovx,
                         const double* ovy, const double* ovz, const TrackEntry* sched,
                         const std::vector<size_t>& level_start, double* mwork, size_t mwork_blocks, int* perm) {
  if (n < 1 || n > kTrackDevMaxN) return fail(RTB_ERR_UNSUPPORTED, "device band tracking is for n <= 256");
  const bool lds = n <= kTrackLdsMaxN;
  if (!lds && (!mwork || mwork_blocks == 0))
    return fail(RTB_ERR_INVALID_ARGUMENT, "device band tracking above 128 bands needs its matrix workspace");
  const size_t nk = nx * ny * nz;
  if (nk == 0) return RTB_OK;
  const size_t total = nk * static_cast<size_t>(n);
  hipLaunchKernelGGL(identity_perm_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, total,
                     n, perm);
  RTB_HIP(hipGetLastError());
  // level 0 is vertex 0 alone (no neighbours visited before it); above
  // kTrackLdsMaxN a level goes in launches of at most mwork_blocks vertices,
  // one workspace slice each
  const size_t step = lds ? std::numeric_limits<unsigned>::max() : mwork_blocks;
  for (size_t l = 1; l + 1 < level_start.size(); ++l) {
    for (size_t a = level_start[l]; a < level_start[l + 1]; a += step) {
      const size_t b = std::min(level_start[l + 1], a + step);
      if (lds)
        hipLaunchKernelGGL(track_kernel<true>, dim3(static_cast<unsigned>(b - a)), dim3(kTrackThreads), track_smem(n),
                           st, ovx, ovy, ovz, ny, nz, n, sched + a, nullptr, perm);
      else
        hipLaunchKernelGGL(track_kernel<false>, dim3(static_cast<unsigned>(b - a)), dim3(kTrackThreadsWide),
                           track_smem(n), st, ovx, ovy, ovz, ny, nz, n, sched + a, mwork, perm);
      RTB_HIP(hipGetLastError());
    }
  }
  return RTB_OK;
}

}  // namespace rtb
