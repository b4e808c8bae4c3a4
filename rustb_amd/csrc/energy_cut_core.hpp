// The per-simplex Fermi cut of the EnergyCut Hall conductivity, shared by the
// host part (energy_cut.cpp) and the device kernel (energy_cut_dev.hip): the same
// operations in the same order on both, so that the host part reproduces the
// reference's arithmetic step by step.  Follows triangle_occupied_hybrid and
// tetrahedron_occupied_hybrid (src/response/energy_cut.rs:1102-1228,
// 1446-1640), eval_berry_kernel and eval_berry_band_at_lam_buf
// (src/response/kernel.rs:54-80, 138-182), the quadrature rules of
// src/response/quadrature.rs and the cells of build_triangles_2d_diagavg /
// build_tetrahedra_3d (src/response/tracking.rs:8-23, 284-412).
#pragma once

#include <cmath>
#include <cstddef>

#if defined(__HIP__)
#define RTB_CUT __host__ __device__
#else
#define RTB_CUT
#endif

#pragma clang fp contract(off)

namespace rtb {
namespace cut {

constexpr double kEps = 1e-12;  // ENERGY_CUT_EPS (energy_cut.rs:47)
// degree-2 symmetric 3-point rule of the triangle (TRI_QUAD_PTS_3, weights 1/3)
constexpr double kQ3[3][3] = {{2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                              {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0},
                              {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}};
constexpr double kW3 = 1.0 / 3.0;
// degree-2 symmetric 4-point rule of the tetrahedron (TET_QUAD_PTS_4, weights 1/4)
constexpr double kTA = 0.5854101966249685, kTB = 0.1381966011250105;
constexpr double kQ4[4][4] = {{kTA, kTB, kTB, kTB}, {kTB, kTA, kTB, kTB}, {kTB, kTB, kTA, kTB}, {kTB, kTB, kTB, kTA}};
constexpr double kW4 = 0.25;
// the five tetrahedra of a cube on its 0-7 body diagonal (corners numbered
// x + 2 y + 4 z) and their volume fractions
constexpr int kCubeTets[5][4] = {{0, 1, 2, 4}, {3, 1, 2, 7}, {5, 1, 4, 7}, {6, 2, 4, 7}, {1, 2, 4, 7}};
constexpr double kTetVolFactor[5] = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0};

// Vertex data in the eigensolver's band order, read through the tracking
// permutation (band a of vertex v is raw band perm[v][a]).
struct View {
  int n;
  const double* E;   // [nk][n]
  const double* Ki;  // Im K [nk][n][n] at stride ks
  size_t ks;
  const int* perm;  // [nk][n]
};

RTB_CUT inline double band(const View& M, size_t v, int a) { return M.E[v * M.n + M.perm[v * M.n + a]]; }
RTB_CUT inline double kim(const View& M, size_t v, int a, int b) {
  const size_t n = M.n;
  return M.Ki[M.ks * (v * n * n + static_cast<size_t>(M.perm[v * n + a]) * n + M.perm[v * n + b])];
}
RTB_CUT inline double dmin(double a, double b) { return a < b ? a : (b < a ? b : a); }
RTB_CUT inline double dmax(double a, double b) { return a > b ? a : (b > a ? b : a); }

// eval_berry_kernel (kernel.rs:54-80), Omega of tracked band i at vertex v.
RTB_CUT inline double vertex_omega(const View& M, size_t v, int i, double eta) {
  const double eta2 = eta * eta;
  double im = 0.0;
  for (int m = 0; m < M.n; ++m) {
    if (m == i) continue;
    const double de = band(M, v, i) - band(M, v, m);
    const double denom = de * de + eta2;
    if (denom < 1e-30) continue;
    im += kim(M, v, i, m) / denom;
  }
  return -2.0 * im;
}

// eval_berry_band_at_lam_buf (kernel.rs:138-182) over NV vertices: E and row
// n of Im K interpolated at the barycentric point lam, then the band sum; each
// interpolated value is formed where it is used (the same sums as with the
// reference's buffers).
template <int NV>
RTB_CUT inline double band_at_lam(const View& M, const size_t* vs, int n, const double* lam, double eta) {
  double en = 0.0;
  for (int v = 0; v < NV; ++v)
    if (lam[v] != 0.0) en += band(M, vs[v], n) * lam[v];
  const double eta2 = eta * eta;
  double im = 0.0;
  for (int m = 0; m < M.n; ++m) {
    if (m == n) continue;
    double em = 0.0, km = 0.0;
    for (int v = 0; v < NV; ++v)
      if (lam[v] != 0.0) em += band(M, vs[v], m) * lam[v];
    for (int v = 0; v < NV; ++v)
      if (lam[v] != 0.0) km += kim(M, vs[v], n, m) * lam[v];
    const double de = en - em;
    const double denom = de * de + eta2;
    if (denom < 1e-30) continue;
    im += km / denom;
  }
  return -2.0 * im;
}

// e[i] for a run-time i, as a chain of selects, so the small arrays here stay
// in registers on the device.  The values are made opaque to the optimizer
// first (an empty asm statement on each, no instruction), else it folds the
// selects of loads back into one load at a run-time index, which keeps the
// array in scratch memory.
template <int N>
RTB_CUT inline double pick(const double* e, int i) {
  double v[N];
  for (int q = 0; q < N; ++q) {
    v[q] = e[q];
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(v[q]));
#endif
  }
  double x = v[0];
  for (int q = 1; q < N; ++q) x = q == i ? v[q] : x;
  return x;
}

// l = the unit vector e_i (N entries)
template <int N>
RTB_CUT inline void unit_vec(double* l, int i) {
  for (int q = 0; q < N; ++q) l[q] = q == i ? 1.0 : 0.0;
}

// indices of e[0..N) in ascending order, ties in index order (a stable sort):
// compare-exchanges of adjacent positions only (N rounds of odd-even
// transposition), each swapping on a strict <, keep equal keys in index order
template <int N>
RTB_CUT inline void sort_idx(const double* e, int* idx) {
  double k[N];
  int o[N];
  for (int i = 0; i < N; ++i) {
    k[i] = e[i];
    o[i] = i;
  }
  for (int round = 0; round < N; ++round)
    for (int i = round % 2; i + 1 < N; i += 2)
      if (k[i + 1] < k[i]) {
        const double tk = k[i];
        k[i] = k[i + 1];
        k[i + 1] = tk;
        const int to = o[i];
        o[i] = o[i + 1];
        o[i + 1] = to;
      }
  for (int i = 0; i < N; ++i) idx[i] = o[i];
}

// the point of edge (i, j) where the linear E crosses mu, as barycentrics
template <int N>
RTB_CUT inline void cut_point(const double* e, int i, int j, double mu, double* l) {
  const double ei = pick<N>(e, i);
  const double de = pick<N>(e, j) - ei;
  double t = 0.5;
  if (!(fabs(de) < kEps)) {
    t = (mu - ei) / de;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  }
  for (int q = 0; q < N; ++q) l[q] = q == i ? 1.0 - t : (q == j ? t : 0.0);
}

// ---- triangles ----
RTB_CUT inline double tri_area(const double c[3][2]) {  // triangle_area (energy_cut.rs:69-78)
  return 0.5 * fabs((c[1][0] - c[0][0]) * (c[2][1] - c[0][1]) - (c[2][0] - c[0][0]) * (c[1][1] - c[0][1]));
}

// sub_tri_area_2d with bary_to_phys_2d (energy_cut.rs:1102-1116).
RTB_CUT inline double sub_area(const double c[3][2], const double t[3][3]) {
  double p[3][2];
  for (int q = 0; q < 3; ++q)
    for (int d = 0; d < 2; ++d) p[q][d] = t[q][0] * c[0][d] + t[q][1] * c[1][d] + t[q][2] * c[2][d];
  return 0.5 * fabs((p[1][0] - p[0][0]) * (p[2][1] - p[0][1]) - (p[2][0] - p[0][0]) * (p[1][1] - p[0][1]));
}

// clip_triangle (energy_cut.rs:1120-1169): the occupied part E <= mu as 0-2
// sub-triangles (barycentric rows).
RTB_CUT inline int clip(const double e[3], double mu, double out[2][3][3]) {
  const double e_min = dmin(dmin(e[0], e[1]), e[2]);
  const double e_max = dmax(dmax(e[0], e[1]), e[2]);
  for (int s = 0; s < 2; ++s)
    for (int q = 0; q < 3; ++q)
      for (int d = 0; d < 3; ++d) out[s][q][d] = 0.0;
  if (mu <= e_min + kEps) return 0;
  if (mu >= e_max - kEps) {
    for (int q = 0; q < 3; ++q) out[0][q][q] = 1.0;
    return 1;
  }
  int idx[3];
  sort_idx<3>(e, idx);
  const int a = idx[0], b = idx[1], c = idx[2];
  if (mu <= pick<3>(e, b) + kEps) {
    unit_vec<3>(out[0][0], a);
    cut_point<3>(e, a, b, mu, out[0][1]);
    cut_point<3>(e, a, c, mu, out[0][2]);
    return 1;
  }
  unit_vec<3>(out[0][0], a);
  unit_vec<3>(out[0][1], b);
  cut_point<3>(e, a, c, mu, out[0][2]);
  unit_vec<3>(out[1][0], b);
  cut_point<3>(e, a, c, mu, out[1][1]);
  cut_point<3>(e, b, c, mu, out[1][2]);
  return 2;
}

// The band sum of eval_berry_band_at_lam_buf in the reference's order, for the
// hybrid rules below.  Those take any evaluator `f` whose f.at<NV>(lam) is the
// band sum at the barycentric point lam, or one lane's share of it when the
// sum over m is split across lanes; `one` weighs the rules' constant terms
// (1 here, 1 on one lane and 0 on the others of a split sum), so that the
// lanes' results add up to the rule.
struct SeqBandSum {
  View M;
  const size_t* vs;
  int n;
  double eta;
  template <int NV>
  RTB_CUT double at(const double* lam) const {
    return band_at_lam<NV>(M, vs, n, lam, eta);
  }
};

// triangle_occupied_hybrid (energy_cut.rs:1184-1228): 0 when empty, the
// vertex-mean Omega times the area when full, else the 3-point rule on each
// clipped sub-triangle.
template <class F>
RTB_CUT inline double hybrid_tri_by(const double c[3][2], const double e_v[3], const double om_v[3], double mu,
                                    const F& f, double one) {
  const double e_min = dmin(dmin(e_v[0], e_v[1]), e_v[2]);
  const double e_max = dmax(dmax(e_v[0], e_v[1]), e_v[2]);
  const double area = tri_area(c);
  if (mu <= e_min + kEps) return 0.0;
  if (mu >= e_max - kEps) return one * (area * (om_v[0] + om_v[1] + om_v[2]) / 3.0);
  double subs[2][3][3];
  const int ns = clip(e_v, mu, subs);
  double total = 0.0;
#pragma unroll
  for (int s = 0; s < 2; ++s) {  // unrolled: subs is indexed at compile time
    if (s >= ns) break;
    const double sa = sub_area(c, subs[s]);
    if (sa < 1e-30) continue;
    for (int iq = 0; iq < 3; ++iq) {
      double lam[3];
      for (int d = 0; d < 3; ++d)  // combine_bary (energy_cut.rs:1172-1178)
        lam[d] = kQ3[iq][0] * subs[s][0][d] + kQ3[iq][1] * subs[s][1][d] + kQ3[iq][2] * subs[s][2][d];
      total += sa * kW3 * f.template at<3>(lam);
    }
  }
  return total;
}

RTB_CUT inline double hybrid_tri(const View& M, const double c[3][2], const size_t vs[3], const double e_v[3],
                                 const double om_v[3], double mu, double eta, int n) {
  return hybrid_tri_by(c, e_v, om_v, mu, SeqBandSum{M, vs, n, eta}, 1.0);
}

// ---- tetrahedra ----
// tet_vol_from_pts (energy_cut.rs:1448-1458): |det| / 6, first-row expansion.
RTB_CUT inline double tet_vol(const double p0[3], const double p1[3], const double p2[3], const double p3[3]) {
  const double det = (p1[0] - p0[0]) * ((p2[1] - p0[1]) * (p3[2] - p0[2]) - (p2[2] - p0[2]) * (p3[1] - p0[1])) -
                     (p1[1] - p0[1]) * ((p2[0] - p0[0]) * (p3[2] - p0[2]) - (p2[2] - p0[2]) * (p3[0] - p0[0])) +
                     (p1[2] - p0[2]) * ((p2[0] - p0[0]) * (p3[1] - p0[1]) - (p2[1] - p0[1]) * (p3[0] - p0[0]));
  return fabs(det) / 6.0;
}

// sub_tet_vol_3d with bary_to_phys_3d (energy_cut.rs:1461-1488).
RTB_CUT inline double sub_tet_vol(const double c[4][3], const double t[4][4]) {
  double p[4][3];
  for (int q = 0; q < 4; ++q)
    for (int d = 0; d < 3; ++d)
      p[q][d] = t[q][0] * c[0][d] + t[q][1] * c[1][d] + t[q][2] * c[2][d] + t[q][3] * c[3][d];
  return tet_vol(p[0], p[1], p[2], p[3]);
}

// sub_tet_k_quad (energy_cut.rs:1503-1526) with combine_bary_4 (:1492-1500).
template <class F>
RTB_CUT inline double sub_tet_quad(const double c[4][3], const double sub[4][4], const F& f) {
  const double sv = sub_tet_vol(c, sub);
  if (sv < 1e-30) return 0.0;
  double total = 0.0;
  for (int iq = 0; iq < 4; ++iq) {
    double lam[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < 4; ++k)
      for (int i = 0; i < 4; ++i) lam[i] += kQ4[iq][k] * sub[k][i];
    total += sv * kW4 * f.template at<4>(lam);
  }
  return total;
}

// tetrahedron_occupied_hybrid (energy_cut.rs:1529-1640): 0 when empty, the
// vertex-mean Omega times the volume when full, else the 4-point rule on the
// occupied sub-tetrahedra (1, 2 or 3 vertices below mu; with 3 below, the full
// tetrahedron minus the one above).
template <class F>
RTB_CUT inline double hybrid_tet_by(const double c[4][3], const double e_v[4], const double om_v[4], double mu,
                                    const F& f, double one) {
  const double full_vol = tet_vol(c[0], c[1], c[2], c[3]);
  if (full_vol < kEps) return 0.0;
  const double e_min = dmin(dmin(e_v[0], e_v[1]), dmin(e_v[2], e_v[3]));
  const double e_max = dmax(dmax(e_v[0], e_v[1]), dmax(e_v[2], e_v[3]));
  if (mu <= e_min + kEps) return 0.0;
  const double full_val = full_vol * (om_v[0] + om_v[1] + om_v[2] + om_v[3]) / 4.0;
  if (mu >= e_max - kEps) return one * full_val;
  int idx[4];
  sort_idx<4>(e_v, idx);
  const int a = idx[0], b = idx[1], cc = idx[2], d = idx[3];
  double s[4][4];
  auto unit = [](int i, double* l) { unit_vec<4>(l, i); };
  if (mu <= pick<4>(e_v, b) + kEps) {
    unit(a, s[0]);
    cut_point<4>(e_v, a, b, mu, s[1]);
    cut_point<4>(e_v, a, cc, mu, s[2]);
    cut_point<4>(e_v, a, d, mu, s[3]);
    return sub_tet_quad(c, s, f);
  }
  if (mu <= pick<4>(e_v, cc) + kEps) {  // three sub-tetrahedra through the p_ac - p_bd diagonal
    double total = 0.0;
    unit(a, s[0]);
    unit(b, s[1]);
    cut_point<4>(e_v, a, cc, mu, s[2]);
    cut_point<4>(e_v, b, d, mu, s[3]);
    total += sub_tet_quad(c, s, f);
    unit(a, s[0]);
    cut_point<4>(e_v, a, cc, mu, s[1]);
    cut_point<4>(e_v, a, d, mu, s[2]);
    cut_point<4>(e_v, b, d, mu, s[3]);
    total += sub_tet_quad(c, s, f);
    unit(b, s[0]);
    cut_point<4>(e_v, b, cc, mu, s[1]);
    cut_point<4>(e_v, b, d, mu, s[2]);
    cut_point<4>(e_v, a, cc, mu, s[3]);
    total += sub_tet_quad(c, s, f);
    return total;
  }
  unit(d, s[0]);
  cut_point<4>(e_v, a, d, mu, s[1]);
  cut_point<4>(e_v, b, d, mu, s[2]);
  cut_point<4>(e_v, cc, d, mu, s[3]);
  return one * full_val - sub_tet_quad(c, s, f);
}

RTB_CUT inline double hybrid_tet(const View& M, const double c[4][3], const size_t vs[4], const double e_v[4],
                                 const double om_v[4], double mu, double eta, int n) {
  return hybrid_tet_by(c, e_v, om_v, mu, SeqBandSum{M, vs, n, eta}, 1.0);
}

// ---- cells ----
// Triangle t of 2D cell (ix, iy) over both diagonals
// (build_triangles_2d_diagavg, tracking.rs:284-342): vertices and fractional
// coordinates; each triangle carries a quarter of the cell.  Corners 00, 10,
// 11, 01 are numbered 0..3; a corner shared by two vertices of a ragged mesh
// keeps the coordinates of its first match, as the reference's lookup.
RTB_CUT inline void cell_triangle(size_t ix, size_t iy, size_t nx, size_t ny, int t, size_t vs[3], double c[3][2]) {
  constexpr int kTri[4][3] = {{0, 1, 3}, {2, 1, 3}, {0, 1, 2}, {0, 2, 3}};
  const double inv_nx = 1.0 / static_cast<double>(nx), inv_ny = 1.0 / static_cast<double>(ny);
  const size_t ixp = (ix + 1) % nx, iyp = (iy + 1) % ny;
  const double fx = static_cast<double>(ix), fy = static_cast<double>(iy);
  const double x0 = fx * inv_nx, x1 = (fx + 1.0) * inv_nx, y0 = fy * inv_ny, y1 = (fy + 1.0) * inv_ny;
  auto vid = [&](int k) { return (k == 1 || k == 2 ? ixp : ix) * ny + (k >= 2 ? iyp : iy); };
  for (int q = 0; q < 3; ++q) {
    const size_t v = vid(kTri[t][q]);
    int k = 0;
    while (vid(k) != v) ++k;  // first corner with this vertex
    vs[q] = v;
    c[q][0] = k == 1 || k == 2 ? x1 : x0;
    c[q][1] = k >= 2 ? y1 : y0;
  }
}

RTB_CUT inline void cell_triangles(size_t ix, size_t iy, size_t nx, size_t ny, size_t vs[4][3],
                                   double c[4][3][2]) {
  for (int t = 0; t < 4; ++t) cell_triangle(ix, iy, nx, ny, t, vs[t], c[t]);
}

// Tetrahedron t of 3D cell (ix, iy, iz) (build_tetrahedra_3d,
// tracking.rs:344-412): vertices, fractional coordinates, volume.
RTB_CUT inline void cell_tetrahedron(size_t ix, size_t iy, size_t iz, size_t nx, size_t ny, size_t nz, int t,
                                     size_t vs[4], double c[4][3], double* vol) {
  const double inv_nx = 1.0 / static_cast<double>(nx), inv_ny = 1.0 / static_cast<double>(ny),
               inv_nz = 1.0 / static_cast<double>(nz);
  const double cube_vol = inv_nx * inv_ny * inv_nz;
  const size_t ixp = (ix + 1) % nx, iyp = (iy + 1) % ny, izp = (iz + 1) % nz;
  for (int q = 0; q < 4; ++q) {
    const int k = kCubeTets[t][q];
    const size_t x = k & 1 ? ixp : ix, y = k & 2 ? iyp : iy, z = k & 4 ? izp : iz;
    vs[q] = (x * ny + y) * nz + z;
    c[q][0] = static_cast<double>(ix + (k & 1 ? 1 : 0)) * inv_nx;
    c[q][1] = static_cast<double>(iy + (k & 2 ? 1 : 0)) * inv_ny;
    c[q][2] = static_cast<double>(iz + (k & 4 ? 1 : 0)) * inv_nz;
  }
  *vol = cube_vol * kTetVolFactor[t];
}

RTB_CUT inline void cell_tetrahedra(size_t ix, size_t iy, size_t iz, size_t nx, size_t ny, size_t nz, size_t vs[5][4],
                                    double c[5][4][3], double vol[5]) {
  for (int t = 0; t < 5; ++t) cell_tetrahedron(ix, iy, iz, nx, ny, nz, t, vs[t], c[t], &vol[t]);
}

}  // namespace cut
}  // namespace rtb
