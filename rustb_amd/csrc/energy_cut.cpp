// EnergyCut Hall conductivity in two and three dimensions, host part: band
// tracking over the k-mesh and the per-triangle (per-tetrahedron) Fermi cut,
// from vertex data the device forms (band energies, K^{ab}_nm = v^a_nm v^b_mn,
// and the raw eigenvector overlaps of mesh neighbours).  Follows
// hall_conductivity, Integration::EnergyCut (src/response/linear/mod.rs:226-268),
// global_band_track (src/response/tracking.rs:25-224), integrate_fermi_cut_2d
// (src/response/energy_cut.rs:1102-1409) and integrate_fermi_cut_3d
// (src/response/energy_cut.rs:1446-1850, build_tetrahedra_3d tracking.rs:344-412).
//
// Vertex data stays in the order the eigensolver produced; tracking keeps a
// permutation per vertex (band n of vertex i is raw band perm[i][n]) instead
// of moving matrices.
#include <algorithm>
#include <cmath>
#include <deque>
#include <thread>
#include <vector>

#include "rtb_internal.hpp"

#pragma clang fp contract(off)

namespace rtb {
namespace {

constexpr double kCutEps = 1e-12;  // ENERGY_CUT_EPS (energy_cut.rs:47)
constexpr double kXCut = 18.0;     // FERMI_X_CUT (energy_cut.rs:48)
constexpr int kXSteps = 72;        // FERMI_X_STEPS (energy_cut.rs:49)
// degree-2 symmetric 3-point rule, barycentric (quadrature.rs TRI_QUAD_PTS_3)
constexpr double kQ[3][3] = {{2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                             {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0},
                             {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}};
constexpr double kQW = 1.0 / 3.0;
// degree-2 symmetric 4-point rule for the tetrahedron (quadrature.rs
// TET_QUAD_PTS_4, weights 1/4 each)
constexpr double kTA = 0.5854101966249685, kTB = 0.1381966011250105;
constexpr double kTQ[4][4] = {{kTA, kTB, kTB, kTB}, {kTB, kTA, kTB, kTB}, {kTB, kTB, kTA, kTB}, {kTB, kTB, kTB, kTA}};
constexpr double kTW = 0.25;
// the five tetrahedra of a cube on its 0-7 body diagonal, corners numbered
// x + 2 y + 4 z, and their volume fractions (tracking.rs:8-23)
constexpr int kCubeTets[5][4] = {{0, 1, 2, 4}, {3, 1, 2, 7}, {5, 1, 4, 7}, {6, 2, 4, 7}, {1, 2, 4, 7}};
constexpr double kTetVolFactor[5] = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0};

struct Mesh {
  size_t nx, ny;
  int n;
  const double* E;    // [nk][n]
  const double* Ki;   // Im K [nk][n][n] at stride ks (2: inside complex K, 1: Im K alone)
  size_t ks;
  const double* ovx;  // [(nx-1) ny][n][n] complex, U_i^H U_{i+ny}
  const double* ovy;  // [nk-1][n][n] complex, U_i^H U_{i+1}
  const int* perm;    // [nk][n]
  const double* om;   // [nk][n] vertex Omega in tracked band order (or null)
  size_t nz = 1;      // 3D meshes (vertex index (ix ny + iy) nz + iz)
};

// Overlaps |<u_a(i)|u_b(j)>|^2 of mesh neighbours (build_overlap_matrix,
// tracking.rs:25-42) come from the raw products U_i^H U_j; the reference's
// evec is conj(U), which leaves each modulus unchanged.

// neighbour_indices (tracking.rs:165-224): +x, -x, +y, -y (+z, -z), not
// periodic; a 2D mesh is nz = 1.
int neighbours(size_t i, size_t nx, size_t ny, size_t nz, size_t out[6]) {
  const size_t ix = i / (ny * nz), iy = (i / nz) % ny, iz = i % nz;
  int c = 0;
  if (ix + 1 < nx) out[c++] = i + ny * nz;
  if (ix > 0) out[c++] = i - ny * nz;
  if (iy + 1 < ny) out[c++] = i + nz;
  if (iy > 0) out[c++] = i - nz;
  if (iz + 1 < nz) out[c++] = i + 1;
  if (iz > 0) out[c++] = i - 1;
  return c;
}

// greedy_assign (tracking.rs:44-68).
void greedy_assign(int n, const std::vector<double>& ov, std::vector<int>& perm, std::vector<char>& assigned,
                   std::vector<std::pair<int, double>>& rows) {
  rows.clear();
  for (int i = 0; i < n; ++i) {
    double best = -1.0;
    for (int j = 0; j < n; ++j) best = std::fmax(best, ov[static_cast<size_t>(i) * n + j]);
    rows.emplace_back(i, best);
  }
  std::stable_sort(rows.begin(), rows.end(), [](const auto& x, const auto& y) { return y.second < x.second; });
  std::fill(assigned.begin(), assigned.end(), 0);
  for (const auto& r : rows) {
    int best_j = 0;
    double best_val = -1.0;
    for (int j = 0; j < n; ++j) {
      const double v = ov[static_cast<size_t>(r.first) * n + j];
      if (!assigned[j] && v > best_val) {
        best_val = v;
        best_j = j;
      }
    }
    assigned[best_j] = 1;
    perm[r.first] = best_j;
  }
}

// Only Im K enters Omega (kernel.rs:54-80, 138-182).
double kim(const Mesh& M, size_t v, int a, int b) {
  const size_t n = M.n;
  return M.Ki[M.ks * (v * n * n + static_cast<size_t>(M.perm[v * n + a]) * n + M.perm[v * n + b])];
}
double band(const Mesh& M, size_t v, int a) { return M.E[v * M.n + M.perm[v * M.n + a]]; }

// eval_berry_kernel (kernel.rs:54-80), Omega only.
void vertex_omega(const Mesh& M, size_t v, double eta, double* om) {
  const double eta2 = eta * eta;
  for (int i = 0; i < M.n; ++i) {
    double im = 0.0;
    for (int m = 0; m < M.n; ++m) {
      if (m == i) continue;
      const double de = band(M, v, i) - band(M, v, m);
      const double denom = de * de + eta2;
      if (denom < 1e-30) continue;
      im += kim(M, v, i, m) / denom;
    }
    om[i] = -2.0 * im;
  }
}

// eval_berry_band_at_lam_buf (kernel.rs:138-182) over NV vertices.
template <int NV>
double band_at_lam(const Mesh& M, const size_t vs[NV], int n, const double lam[NV], double eta,
                   std::vector<double>& e, std::vector<double>& ki) {
  std::fill(e.begin(), e.end(), 0.0);
  std::fill(ki.begin(), ki.end(), 0.0);
  for (int v = 0; v < NV; ++v) {
    if (lam[v] == 0.0) continue;
    for (int m = 0; m < M.n; ++m) e[m] += band(M, vs[v], m) * lam[v];
  }
  for (int v = 0; v < NV; ++v) {
    if (lam[v] == 0.0) continue;
    for (int m = 0; m < M.n; ++m) ki[m] += kim(M, vs[v], n, m) * lam[v];
  }
  const double eta2 = eta * eta;
  double im = 0.0;
  for (int m = 0; m < M.n; ++m) {
    if (m == n) continue;
    const double de = e[n] - e[m];
    const double denom = de * de + eta2;
    if (denom < 1e-30) continue;
    im += ki[m] / denom;
  }
  return -2.0 * im;
}

double tri_area(const double c[3][2]) {  // triangle_area (energy_cut.rs:69-78)
  return 0.5 * std::fabs((c[1][0] - c[0][0]) * (c[2][1] - c[0][1]) - (c[2][0] - c[0][0]) * (c[1][1] - c[0][1]));
}

// sub_tri_area_2d with bary_to_phys_2d (energy_cut.rs:1102-1116).
double sub_area(const double c[3][2], const double t[3][3]) {
  double p[3][2];
  for (int q = 0; q < 3; ++q)
    for (int d = 0; d < 2; ++d) p[q][d] = t[q][0] * c[0][d] + t[q][1] * c[1][d] + t[q][2] * c[2][d];
  return 0.5 * std::fabs((p[1][0] - p[0][0]) * (p[2][1] - p[0][1]) - (p[2][0] - p[0][0]) * (p[1][1] - p[0][1]));
}

// clip_triangle (energy_cut.rs:1120-1169): the occupied part E <= mu of the
// linearly interpolated triangle as 0-2 sub-triangles (barycentric rows).
int clip(const double e[3], double mu, double out[2][3][3]) {
  const double eps = 1e-12;
  const double e_min = std::fmin(std::fmin(e[0], e[1]), e[2]);
  const double e_max = std::fmax(std::fmax(e[0], e[1]), e[2]);
  for (int s = 0; s < 2; ++s)
    for (int q = 0; q < 3; ++q)
      for (int d = 0; d < 3; ++d) out[s][q][d] = 0.0;
  if (mu <= e_min + eps) return 0;
  if (mu >= e_max - eps) {
    for (int q = 0; q < 3; ++q) out[0][q][q] = 1.0;
    return 1;
  }
  int idx[3] = {0, 1, 2};
  std::stable_sort(idx, idx + 3, [&](int i, int j) { return e[i] < e[j]; });
  const int a = idx[0], b = idx[1], c = idx[2];
  auto cut = [&](int i, int j, double* lam) {
    const double de = e[j] - e[i];
    double t = 0.5;
    if (!(std::fabs(de) < eps)) {
      t = (mu - e[i]) / de;
      t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    }
    lam[i] = 1.0 - t;
    lam[j] = t;
  };
  if (mu <= e[b] + eps) {
    out[0][0][a] = 1.0;
    cut(a, b, out[0][1]);
    cut(a, c, out[0][2]);
    return 1;
  }
  out[0][0][a] = 1.0;
  out[0][1][b] = 1.0;
  cut(a, c, out[0][2]);
  out[1][0][b] = 1.0;
  cut(a, c, out[1][1]);
  cut(b, c, out[1][2]);
  return 2;
}

// triangle_occupied_hybrid (energy_cut.rs:1184-1228).
double hybrid(const Mesh& M, const double c[3][2], const size_t vs[3], const double e_v[3], const double om_v[3],
              double mu, double eta, int n, std::vector<double>& e, std::vector<double>& ki) {
  const double e_min = std::fmin(std::fmin(e_v[0], e_v[1]), e_v[2]);
  const double e_max = std::fmax(std::fmax(e_v[0], e_v[1]), e_v[2]);
  const double area = tri_area(c);
  if (mu <= e_min + kCutEps) return 0.0;
  if (mu >= e_max - kCutEps) return area * (om_v[0] + om_v[1] + om_v[2]) / 3.0;
  double subs[2][3][3];
  const int ns = clip(e_v, mu, subs);
  double total = 0.0;
  for (int s = 0; s < ns; ++s) {
    const double sa = sub_area(c, subs[s]);
    if (sa < 1e-30) continue;
    for (int iq = 0; iq < 3; ++iq) {
      double lam[3];
      for (int d = 0; d < 3; ++d)  // combine_bary (energy_cut.rs:1172-1178)
        lam[d] = kQ[iq][0] * subs[s][0][d] + kQ[iq][1] * subs[s][1][d] + kQ[iq][2] * subs[s][2][d];
      total += sa * kQW * band_at_lam<3>(M, vs, n, lam, eta, e, ki);
    }
  }
  return total;
}

// integrate_fermi_cut_2d_t0 (energy_cut.rs:1231-1345) for the cells
// idx in [i0, i1), over both diagonals' triangles (build_triangles_2d_diagavg,
// tracking.rs:284-342), into acc [n_mu].
void cut_t0_cells(const Mesh& M, size_t i0, size_t i1, const double* mu, size_t n_mu, double eta, double* acc) {
  const double* om_all = M.om;
  const size_t nx = M.nx, ny = M.ny;
  const double inv_nx = 1.0 / static_cast<double>(nx), inv_ny = 1.0 / static_cast<double>(ny);
  const double half_vol = inv_nx * inv_ny / 4.0;
  const int n = M.n;
  std::vector<double> e(n), ki(n);
  for (size_t idx = i0; idx < i1; ++idx) {
    const size_t iy = idx % ny, ix = idx / ny;
    const size_t ixp = (ix + 1) % nx, iyp = (iy + 1) % ny;
    const size_t i00 = ix * ny + iy, i10 = ixp * ny + iy, i11 = ixp * ny + iyp, i01 = ix * ny + iyp;
    auto coord_of = [&](size_t i, double* c) {
      const double fx = static_cast<double>(ix), fy = static_cast<double>(iy);
      if (i == i00) {
        c[0] = fx * inv_nx, c[1] = fy * inv_ny;
      } else if (i == i10) {
        c[0] = (fx + 1.0) * inv_nx, c[1] = fy * inv_ny;
      } else if (i == i11) {
        c[0] = (fx + 1.0) * inv_nx, c[1] = (fy + 1.0) * inv_ny;
      } else {
        c[0] = fx * inv_nx, c[1] = (fy + 1.0) * inv_ny;
      }
    };
    const size_t tris[4][3] = {{i00, i10, i01}, {i11, i10, i01}, {i00, i10, i11}, {i00, i11, i01}};
    for (const auto& tv : tris) {
      double c[3][2];
      for (int q = 0; q < 3; ++q) coord_of(tv[q], c[q]);
      const double area = tri_area(c);
      if (area < kCutEps) continue;
      const double scale = half_vol / area;
      // eval_berry_kernel of the three vertices (energy_cut.rs:1268-1285),
      // formed once per vertex beforehand
      const double* o0 = om_all + tv[0] * n;
      const double* o1 = om_all + tv[1] * n;
      const double* o2 = om_all + tv[2] * n;
      for (int b = 0; b < n; ++b) {
        const double e_v[3] = {band(M, tv[0], b), band(M, tv[1], b), band(M, tv[2], b)};
        const double om_v[3] = {o0[b], o1[b], o2[b]};
        const double e_min = std::fmin(std::fmin(e_v[0], e_v[1]), e_v[2]);
        const double e_max = std::fmax(std::fmax(e_v[0], e_v[1]), e_v[2]);
        const double full_val = area * (om_v[0] + om_v[1] + om_v[2]) / 3.0;
        const size_t i_partial =
            std::partition_point(mu, mu + n_mu, [&](double x) { return x <= e_min + kCutEps; }) - mu;
        const size_t i_full = std::partition_point(mu, mu + n_mu, [&](double x) { return x < e_max - kCutEps; }) - mu;
        for (size_t im = i_partial; im < i_full; ++im)
          acc[im] += scale * hybrid(M, c, tv, e_v, om_v, mu[im], eta, b, e, ki);
        if (i_full < n_mu) {
          const double add = scale * full_val;
          for (size_t im = i_full; im < n_mu; ++im) acc[im] += add;
        }
      }
    }
  }
}

// ---- three dimensions: tetrahedra ----
double tet_vol(const double p0[3], const double p1[3], const double p2[3], const double p3[3]) {
  // tet_vol_from_pts (energy_cut.rs:1448-1458): |det| / 6, first-row expansion
  const double det = (p1[0] - p0[0]) * ((p2[1] - p0[1]) * (p3[2] - p0[2]) - (p2[2] - p0[2]) * (p3[1] - p0[1])) -
                     (p1[1] - p0[1]) * ((p2[0] - p0[0]) * (p3[2] - p0[2]) - (p2[2] - p0[2]) * (p3[0] - p0[0])) +
                     (p1[2] - p0[2]) * ((p2[0] - p0[0]) * (p3[1] - p0[1]) - (p2[1] - p0[1]) * (p3[0] - p0[0]));
  return std::fabs(det) / 6.0;
}

// sub_tet_vol_3d with bary_to_phys_3d (energy_cut.rs:1461-1488).
double sub_tet_vol(const double c[4][3], const double t[4][4]) {
  double p[4][3];
  for (int q = 0; q < 4; ++q)
    for (int d = 0; d < 3; ++d)
      p[q][d] = t[q][0] * c[0][d] + t[q][1] * c[1][d] + t[q][2] * c[2][d] + t[q][3] * c[3][d];
  return tet_vol(p[0], p[1], p[2], p[3]);
}

// sub_tet_k_quad (energy_cut.rs:1503-1526) with combine_bary_4 (:1492-1500).
double sub_tet_quad(const Mesh& M, const double c[4][3], const size_t vs[4], const double sub[4][4], int n, double eta,
                    std::vector<double>& e, std::vector<double>& ki) {
  const double sv = sub_tet_vol(c, sub);
  if (sv < 1e-30) return 0.0;
  double total = 0.0;
  for (int iq = 0; iq < 4; ++iq) {
    double lam[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < 4; ++k)
      for (int i = 0; i < 4; ++i) lam[i] += kTQ[iq][k] * sub[k][i];
    total += sv * kTW * band_at_lam<4>(M, vs, n, lam, eta, e, ki);
  }
  return total;
}

// tetrahedron_occupied_hybrid (energy_cut.rs:1529-1640): 0 when empty, the
// vertex-mean Omega times the volume when full, else the 4-point rule on the
// occupied sub-tetrahedra (1, 2 or 3 vertices below mu; with 3 below, the full
// tetrahedron minus the one above).
double hybrid_tet(const Mesh& M, const double c[4][3], const size_t vs[4], const double e_v[4], const double om_v[4],
                  double mu, double eta, int n, std::vector<double>& e, std::vector<double>& ki) {
  const double eps = kCutEps;
  const double full_vol = tet_vol(c[0], c[1], c[2], c[3]);
  if (full_vol < eps) return 0.0;
  const double e_min = std::fmin(std::fmin(e_v[0], e_v[1]), std::fmin(e_v[2], e_v[3]));
  const double e_max = std::fmax(std::fmax(e_v[0], e_v[1]), std::fmax(e_v[2], e_v[3]));
  if (mu <= e_min + eps) return 0.0;
  const double full_val = full_vol * (om_v[0] + om_v[1] + om_v[2] + om_v[3]) / 4.0;
  if (mu >= e_max - eps) return full_val;
  int idx[4] = {0, 1, 2, 3};
  std::stable_sort(idx, idx + 4, [&](int i, int j) { return e_v[i] < e_v[j]; });
  const int a = idx[0], b = idx[1], cc = idx[2], d = idx[3];
  auto unit = [](int i, double* l) {
    for (int q = 0; q < 4; ++q) l[q] = 0.0;
    l[i] = 1.0;
  };
  auto cut = [&](int i, int j, double* l) {
    for (int q = 0; q < 4; ++q) l[q] = 0.0;
    const double de = e_v[j] - e_v[i];
    double t = 0.5;
    if (!(std::fabs(de) < eps)) {
      t = (mu - e_v[i]) / de;
      t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    }
    l[i] = 1.0 - t;
    l[j] = t;
  };
  double s[4][4];
  if (mu <= e_v[b] + eps) {
    unit(a, s[0]), cut(a, b, s[1]), cut(a, cc, s[2]), cut(a, d, s[3]);
    return sub_tet_quad(M, c, vs, s, n, eta, e, ki);
  }
  if (mu <= e_v[cc] + eps) {  // three sub-tetrahedra through the p_ac - p_bd diagonal
    double total = 0.0;
    unit(a, s[0]), unit(b, s[1]), cut(a, cc, s[2]), cut(b, d, s[3]);
    total += sub_tet_quad(M, c, vs, s, n, eta, e, ki);
    unit(a, s[0]), cut(a, cc, s[1]), cut(a, d, s[2]), cut(b, d, s[3]);
    total += sub_tet_quad(M, c, vs, s, n, eta, e, ki);
    unit(b, s[0]), cut(b, cc, s[1]), cut(b, d, s[2]), cut(a, cc, s[3]);
    total += sub_tet_quad(M, c, vs, s, n, eta, e, ki);
    return total;
  }
  unit(d, s[0]), cut(a, d, s[1]), cut(b, d, s[2]), cut(cc, d, s[3]);
  return full_val - sub_tet_quad(M, c, vs, s, n, eta, e, ki);
}

// integrate_fermi_cut_3d_t0 (energy_cut.rs:1650-1797) for the cells idx in
// [i0, i1) (ix outermost), over the five tetrahedra of build_tetrahedra_3d,
// into acc [n_mu].
void cut_t0_cells_3d(const Mesh& M, size_t i0, size_t i1, const double* mu, size_t n_mu, double eta, double* acc) {
  const size_t nx = M.nx, ny = M.ny, nz = M.nz;
  const double inv_nx = 1.0 / static_cast<double>(nx), inv_ny = 1.0 / static_cast<double>(ny),
               inv_nz = 1.0 / static_cast<double>(nz);
  const double cube_vol = inv_nx * inv_ny * inv_nz;
  const int n = M.n;
  std::vector<double> e(n), ki(n);
  for (size_t idx = i0; idx < i1; ++idx) {
    const size_t iz = idx % nz, iy = (idx / nz) % ny, ix = idx / (ny * nz);
    const size_t ixp = (ix + 1) % nx, iyp = (iy + 1) % ny, izp = (iz + 1) % nz;
    size_t cid[8];
    double corner[8][3];
    for (int q = 0; q < 8; ++q) {
      const size_t x = q & 1 ? ixp : ix, y = q & 2 ? iyp : iy, z = q & 4 ? izp : iz;
      cid[q] = (x * ny + y) * nz + z;
      corner[q][0] = static_cast<double>(ix + (q & 1 ? 1 : 0)) * inv_nx;
      corner[q][1] = static_cast<double>(iy + (q & 2 ? 1 : 0)) * inv_ny;
      corner[q][2] = static_cast<double>(iz + (q & 4 ? 1 : 0)) * inv_nz;
    }
    for (int t = 0; t < 5; ++t) {
      size_t vs[4];
      double c[4][3];
      for (int q = 0; q < 4; ++q) {
        vs[q] = cid[kCubeTets[t][q]];
        for (int d = 0; d < 3; ++d) c[q][d] = corner[kCubeTets[t][q]][d];
      }
      const double vol = tet_vol(c[0], c[1], c[2], c[3]);
      if (vol < kCutEps) continue;
      const double scale = cube_vol * kTetVolFactor[t] / vol;
      for (int b = 0; b < n; ++b) {
        const double e_v[4] = {band(M, vs[0], b), band(M, vs[1], b), band(M, vs[2], b), band(M, vs[3], b)};
        const double om_v[4] = {M.om[vs[0] * n + b], M.om[vs[1] * n + b], M.om[vs[2] * n + b], M.om[vs[3] * n + b]};
        const double e_min = std::fmin(std::fmin(e_v[0], e_v[1]), std::fmin(e_v[2], e_v[3]));
        const double e_max = std::fmax(std::fmax(e_v[0], e_v[1]), std::fmax(e_v[2], e_v[3]));
        const double full_val = vol * (om_v[0] + om_v[1] + om_v[2] + om_v[3]) / 4.0;
        const size_t i_partial =
            std::partition_point(mu, mu + n_mu, [&](double x) { return x <= e_min + kCutEps; }) - mu;
        const size_t i_full = std::partition_point(mu, mu + n_mu, [&](double x) { return x < e_max - kCutEps; }) - mu;
        for (size_t im = i_partial; im < i_full; ++im)
          acc[im] += scale * hybrid_tet(M, c, vs, e_v, om_v, mu[im], eta, b, e, ki);
        if (i_full < n_mu) {
          const double add = scale * full_val;
          for (size_t im = i_full; im < n_mu; ++im) acc[im] += add;
        }
      }
    }
  }
}

size_t host_threads(size_t items) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return std::min<size_t>(std::min<size_t>(hw, 16), std::max<size_t>(1, items / 64));
}

// The zero-width cut summed over all cells: fixed row blocks on up to 16
// threads, partial sums added in block order (deterministic).
void cut_t0(const Mesh& M, int dim, const double* mu, size_t n_mu, double eta, double* out) {
  const size_t cells = M.nx * M.ny * M.nz;
  const size_t nt = host_threads(cells);
  std::vector<std::vector<double>> part(nt, std::vector<double>(n_mu, 0.0));
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      const size_t c0 = cells * t / nt, c1 = cells * (t + 1) / nt;
      if (dim == 3)
        cut_t0_cells_3d(M, c0, c1, mu, n_mu, eta, part[t].data());
      else
        cut_t0_cells(M, c0, c1, mu, n_mu, eta, part[t].data());
    });
  for (auto& x : th) x.join();
  for (size_t i = 0; i < n_mu; ++i) {
    double s = 0.0;
    for (size_t t = 0; t < nt; ++t) s += part[t][i];
    out[i] = s;
  }
}

double fermi_window_x(double x) {  // energy_cut.rs:59-66
  if (std::fabs(x) > 50.0) return 0.0;
  const double ex = std::exp(x);
  return ex / ((1.0 + ex) * (1.0 + ex));
}

}  // namespace

// global_band_track (tracking.rs:123-163): breadth-first from vertex 0; a new
// vertex takes greedy_assign of its overlaps with its visited neighbours
// (summed, then divided by their count).
void energy_cut_track(size_t nx, size_t ny, size_t nz, int n, const double* ovx, const double* ovy,
                      const double* ovz, bool squared, std::vector<int>& perm) {
  const size_t zs = squared ? 1 : 2;  // doubles per overlap element
  const size_t nk = nx * ny * nz;
  perm.resize(nk * n);
  for (size_t i = 0; i < nk; ++i)
    for (int b = 0; b < n; ++b) perm[i * n + b] = b;
  if (nk <= 1) return;
  std::vector<char> visited(nk, 0), assigned(n);
  std::vector<double> ov(static_cast<size_t>(n) * n);
  std::vector<int> p(n);
  std::vector<std::pair<int, double>> rows;
  std::deque<size_t> queue;
  visited[0] = 1;
  queue.push_back(0);
  size_t nbs[6], vns[6];
  while (!queue.empty()) {
    const size_t cur = queue.front();
    queue.pop_front();
    const int cn = neighbours(cur, nx, ny, nz, nbs);
    for (int q = 0; q < cn; ++q) {
      const size_t nb = nbs[q];
      if (visited[nb]) continue;
      int cnt = 0;
      const int vc = neighbours(nb, nx, ny, nz, vns);
      for (int r = 0; r < vc; ++r) {
        const size_t vn = vns[r];
        if (!visited[vn]) continue;
        // ov[a][b] = |<tracked band a of vn | raw band b of nb>|^2: element
        // (raw row r, column b) of U_vn^H U_nb sits at base + 2 (r rs + b cs);
        // the pair's axis from its index step (x first: with ny = 1 the x
        // and y steps coincide and only x neighbours exist)
        const size_t n2 = static_cast<size_t>(n) * n;
        const size_t lo = std::min(nb, vn), step = std::max(nb, vn) - lo;
        const double* ov_axis = step == ny * nz ? ovx : (step == nz ? ovy : ovz);
        const double* base = ov_axis + zs * lo * n2;
        size_t rs = n, cs = 1;
        if (nb == lo) rs = 1, cs = n;
        for (int a = 0; a < n; ++a) {
          const double* row = base + zs * static_cast<size_t>(perm[vn * n + a]) * rs;
          double* o = ov.data() + static_cast<size_t>(a) * n;
          for (int b = 0; b < n; ++b) {
            const double* z = row + zs * b * cs;
            const double x = squared ? z[0] : z[0] * z[0] + z[1] * z[1];
            o[b] = cnt == 0 ? x : o[b] + x;
          }
        }
        ++cnt;
      }
      if (cnt > 0) {
        for (double& o : ov) o /= static_cast<double>(cnt);
        greedy_assign(n, ov, p, assigned, rows);
        for (int b = 0; b < n; ++b) perm[nb * n + b] = p[b];
      }
      visited[nb] = 1;
      queue.push_back(nb);
    }
  }
}

// integrate_fermi_cut_2d / _3d (energy_cut.rs:1355-1409, 1798-1850): the
// zero-width cut, or its convolution with the Fermi window for width > 0.
// Fractional measure; a 2D mesh is nz = 1.
void energy_cut_sigma(size_t nx, size_t ny, size_t nz, int dim, int n, const double* E, const double* Ki, size_t ks,
                      const int* perm, const double* mu, size_t n_mu, double width, double eta, double* out) {
  // vertex Omega of every vertex in its tracked band order, once
  const size_t nk = nx * ny * nz;
  std::vector<double> om(nk * n);
  {
    const Mesh P{nx, ny, n, E, Ki, ks, nullptr, nullptr, perm, nullptr, nz};
    const size_t nt = host_threads(nk);
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t)
      th.emplace_back([&, t] {
        for (size_t v = nk * t / nt; v < nk * (t + 1) / nt; ++v) vertex_omega(P, v, eta, om.data() + v * n);
      });
    for (auto& x : th) x.join();
  }
  const Mesh M{nx, ny, n, E, Ki, ks, nullptr, nullptr, perm, om.data(), nz};
  if (width == 0.0) {
    cut_t0(M, dim, mu, n_mu, eta, out);
    return;
  }
  const double beta = 1.0 / width;
  const double x_max = 12.0;
  double mu_min = INFINITY, mu_max = -INFINITY;
  for (size_t i = 0; i < n_mu; ++i) {
    mu_min = std::fmin(mu_min, mu[i]);
    mu_max = std::fmax(mu_max, mu[i]);
  }
  const double mu_lo = mu_min - x_max / beta;
  const double mu_hi = mu_max + x_max / beta;
  const double dmu_fine = 0.1 / beta;
  const size_t n_ext = static_cast<size_t>(std::ceil((mu_hi - mu_lo) / dmu_fine)) + 1;
  // ndarray's linspace: a + i (b - a) / (n - 1)
  const double step = n_ext > 1 ? (mu_hi - mu_lo) / static_cast<double>(n_ext - 1) : 0.0;
  std::vector<double> mu_ext(n_ext), s0(n_ext);
  for (size_t i = 0; i < n_ext; ++i) mu_ext[i] = mu_lo + step * static_cast<double>(i);
  cut_t0(M, dim, mu_ext.data(), n_ext, eta, s0.data());
  const double dx = 2.0 * kXCut / static_cast<double>(kXSteps);
  for (size_t i = 0; i < n_mu; ++i) {
    double sum = 0.0;
    for (int iq = 0; iq < kXSteps; ++iq) {
      const double x = -kXCut + (static_cast<double>(iq) + 0.5) * dx;
      const double w = fermi_window_x(x);
      const double i_f = (mu[i] + x / beta - mu_lo) / dmu_fine;
      const double fl = std::floor(i_f);
      const size_t i_lo = fl > 0.0 ? static_cast<size_t>(fl) : 0;
      const size_t i_hi = std::min(i_lo + 1, n_ext - 1);
      if (i_hi > i_lo) {
        const double t = i_f - static_cast<double>(i_lo);
        const double val = s0[i_lo] + t * (s0[i_hi] - s0[i_lo]);
        sum += dx * w * val;
      }
    }
    out[i] = sum;
  }
}

}  // namespace rtb

int rtb_energy_cut_host(size_t nx, size_t ny, int nsta, const double* E, const double* K, const double* ovx,
                        const double* ovy, const double* mu, size_t n_mu, double T, double eta, double det,
                        double* sigma) {
  using namespace rtb;
  if (nx == 0 || ny == 0 || nsta <= 0) return fail(RTB_ERR_INVALID_KMESH, "kmesh and nsta must be positive");
  if (!E || !K || !mu || !sigma || (nx > 1 && !ovx) || (nx * ny > 1 && !ovy))
    return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  if (n_mu == 0) return fail(RTB_ERR_INVALID_RESPONSE, "chemical_potentials: must contain at least one value");
  for (size_t i = 0; i + 1 < n_mu; ++i)  // validate_sorted (config.rs:291-303)
    if (mu[i] > mu[i + 1]) return fail(RTB_ERR_INVALID_RESPONSE, "mu: must be sorted in ascending order");
  std::vector<int> perm;
  energy_cut_track(nx, ny, 1, nsta, ovx, ovy, nullptr, false, perm);
  const double width = T <= 0.0 ? 0.0 : T * 8.617333262145e-5;  // thermodynamics.rs:10, 50-59
  energy_cut_sigma(nx, ny, 1, 2, nsta, E, K + 1, 2, perm.data(), mu, n_mu, width, eta, sigma);
  for (size_t i = 0; i < n_mu; ++i) sigma[i] /= det;  // linear/mod.rs:267
  return RTB_OK;
}

int rtb_energy_cut_host_3d(size_t nx, size_t ny, size_t nz, int nsta, const double* E, const double* K,
                           const double* ovx, const double* ovy, const double* ovz, const double* mu, size_t n_mu,
                           double T, double eta, double det, double* sigma) {
  using namespace rtb;
  if (nx == 0 || ny == 0 || nz == 0 || nsta <= 0) return fail(RTB_ERR_INVALID_KMESH, "kmesh and nsta must be positive");
  const size_t nk = nx * ny * nz;
  if (!E || !K || !mu || !sigma || (nx > 1 && !ovx) || (nk > nz && ny > 1 && !ovy) || (nz > 1 && !ovz))
    return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  if (n_mu == 0) return fail(RTB_ERR_INVALID_RESPONSE, "chemical_potentials: must contain at least one value");
  for (size_t i = 0; i + 1 < n_mu; ++i)  // validate_sorted (config.rs:291-303)
    if (mu[i] > mu[i + 1]) return fail(RTB_ERR_INVALID_RESPONSE, "mu: must be sorted in ascending order");
  std::vector<int> perm;
  energy_cut_track(nx, ny, nz, nsta, ovx, ovy, ovz, false, perm);
  const double width = T <= 0.0 ? 0.0 : T * 8.617333262145e-5;  // thermodynamics.rs:10, 50-59
  energy_cut_sigma(nx, ny, nz, 3, nsta, E, K + 1, 2, perm.data(), mu, n_mu, width, eta, sigma);
  for (size_t i = 0; i < n_mu; ++i) sigma[i] /= det;  // linear/mod.rs:267
  return RTB_OK;
}
