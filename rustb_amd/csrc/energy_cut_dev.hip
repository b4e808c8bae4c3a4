// The EnergyCut Fermi cut on the device: vertex Omega in tracked band order,
// then the zero-width cut of every (cell, simplex, band) for a list of
// chemical potentials (integrate_fermi_cut_2d_t0 / _3d_t0,
// src/response/energy_cut.rs:1231-1345, 1650-1797).  The per-simplex
// arithmetic is energy_cut_core.hpp, the same code as the host part
// (energy_cut.cpp).
//
// One lane per (cell, simplex, band), lanes of a wave on consecutive bands;
// for each mu the 64 lanes' contributions are summed by DPP in a fixed order
// into a per-wave partial, and the partials are summed per mu in a fixed
// order (column_sums): the result does not depend on scheduling.
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

template <int DIM>
__global__ void __launch_bounds__(256) cut_kernel(cut::View V, const double* __restrict__ om, size_t nx, size_t ny,
                                                  size_t nz, const double* __restrict__ mu, int n_mu, double eta,
                                                  double* __restrict__ partial) {
  constexpr int S = DIM == 2 ? 4 : 5;  // simplices per cell
  constexpr int NV = DIM + 1;
  const int n = V.n;
  const size_t g = blockIdx.x * 256ull + threadIdx.x;
  const size_t total = nx * ny * nz * S * static_cast<size_t>(n);
  bool live = g < total;
  size_t vs[NV] = {};
  double c[NV][DIM] = {};
  double e_v[NV] = {}, om_v[NV] = {};
  double scale = 0.0, full_val = 0.0, e_min = 0.0, e_max = 0.0;
  int b = 0;
  if (live) {
    const size_t cell = g / (static_cast<size_t>(S) * n);
    const int s = static_cast<int>((g / n) % S);
    b = static_cast<int>(g % n);
    double measure = 0.0;
    if constexpr (DIM == 2) {
      size_t vv[4][3];
      double cc[4][3][2];
      cut::cell_triangles(cell / ny, cell % ny, nx, ny, vv, cc);
      for (int q = 0; q < 3; ++q) {
        vs[q] = vv[s][q];
        c[q][0] = cc[s][q][0];
        c[q][1] = cc[s][q][1];
      }
      measure = cut::tri_area(c);
      const double half_vol = (1.0 / static_cast<double>(nx)) * (1.0 / static_cast<double>(ny)) / 4.0;
      scale = half_vol / measure;
    } else {
      size_t vv[5][4];
      double cc[5][4][3], vol[5];
      cut::cell_tetrahedra(cell / (ny * nz), (cell / nz) % ny, cell % nz, nx, ny, nz, vv, cc, vol);
      for (int q = 0; q < 4; ++q) {
        vs[q] = vv[s][q];
        for (int d = 0; d < 3; ++d) c[q][d] = cc[s][q][d];
      }
      measure = cut::tet_vol(c[0], c[1], c[2], c[3]);
      scale = vol[s] / measure;
    }
    live = !(measure < cut::kEps);  // degenerate simplices of ragged meshes
    for (int q = 0; q < NV; ++q) {
      e_v[q] = cut::band(V, vs[q], b);
      om_v[q] = om[vs[q] * n + b];
    }
    e_min = e_max = e_v[0];
    double osum = om_v[0];
    for (int q = 1; q < NV; ++q) {
      e_min = cut::dmin(e_min, e_v[q]);
      e_max = cut::dmax(e_max, e_v[q]);
      osum += om_v[q];
    }
    full_val = measure * osum / static_cast<double>(NV);
  }
  const int lane = threadIdx.x & 63;
  double* out = partial + (g / 64) * static_cast<size_t>(n_mu);
  for (int im = 0; im < n_mu; ++im) {
    const double m = mu[im];
    double v = 0.0;
    if (live) {
      // the partition points of the host loop: mu >= e_max - eps is the full
      // range, mu <= e_min + eps below it is empty, the rest is cut
      if (!(m < e_max - cut::kEps)) {
        v = scale * full_val;
      } else if (!(m <= e_min + cut::kEps)) {
        if constexpr (DIM == 2)
          v = scale * cut::hybrid_tri(V, c, vs, e_v, om_v, m, eta, b);
        else
          v = scale * cut::hybrid_tet(V, c, vs, e_v, om_v, m, eta, b);
      }
    }
    v = wave_sum64(v);
    if (lane == 0) out[im] = v;
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
    if (dim == 2)
      hipLaunchKernelGGL(cut_kernel<2>, dim3(blocks), dim3(256), 0, st, V, om, nx, ny, nz, mu + m0, c, eta, partial);
    else
      hipLaunchKernelGGL(cut_kernel<3>, dim3(blocks), dim3(256), 0, st, V, om, nx, ny, nz, mu + m0, c, eta, partial);
    RTB_HIP(hipGetLastError());
    RTB_TRY(column_sums(st, partial, waves, c, out + m0));
  }
  return RTB_OK;
}

}  // namespace rtb
