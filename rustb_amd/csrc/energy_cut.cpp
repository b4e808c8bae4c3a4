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
#include <atomic>
#include <cmath>
#include <deque>
#include <functional>
#include <thread>
#include <vector>

#include "energy_cut_core.hpp"
#include "rtb_internal.hpp"

#pragma clang fp contract(off)

namespace rtb {
namespace {

constexpr double kXCut = 18.0;  // FERMI_X_CUT (energy_cut.rs:48)
constexpr int kXSteps = 72;     // FERMI_X_STEPS (energy_cut.rs:49)

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

struct Mesh {
  size_t nx, ny, nz;
  cut::View V;
  const double* om;  // [nk][n] vertex Omega in tracked band order
};

// The zero-width cut of one simplex's band b over the sorted mu (the
// partition points of integrate_fermi_cut_2d_t0 / _3d_t0: empty below, the
// hybrid rule inside, a range add above), into acc.
template <int NV, int D>
void simplex_band(const Mesh& M, const size_t vs[NV], const double c[NV][D], double scale, double full_measure,
                  int b, const double* mu, size_t n_mu, double eta, double* acc) {
  const int n = M.V.n;
  double e_v[NV], om_v[NV];
  for (int q = 0; q < NV; ++q) {
    e_v[q] = cut::band(M.V, vs[q], b);
    om_v[q] = M.om[vs[q] * n + b];
  }
  double e_min = e_v[0], e_max = e_v[0];
  for (int q = 1; q < NV; ++q) {
    e_min = cut::dmin(e_min, e_v[q]);
    e_max = cut::dmax(e_max, e_v[q]);
  }
  double osum = om_v[0];
  for (int q = 1; q < NV; ++q) osum += om_v[q];
  const double full_val = full_measure * osum / static_cast<double>(NV);
  const size_t i_partial =
      std::partition_point(mu, mu + n_mu, [&](double x) { return x <= e_min + cut::kEps; }) - mu;
  const size_t i_full = std::partition_point(mu, mu + n_mu, [&](double x) { return x < e_max - cut::kEps; }) - mu;
  for (size_t im = i_partial; im < i_full; ++im) {
    if constexpr (NV == 3)
      acc[im] += scale * cut::hybrid_tri(M.V, c, vs, e_v, om_v, mu[im], eta, b);
    else
      acc[im] += scale * cut::hybrid_tet(M.V, c, vs, e_v, om_v, mu[im], eta, b);
  }
  if (i_full < n_mu) {
    const double add = scale * full_val;
    for (size_t im = i_full; im < n_mu; ++im) acc[im] += add;
  }
}

// integrate_fermi_cut_2d_t0 (energy_cut.rs:1231-1345) / _3d_t0 (:1650-1797)
// for the cells idx in [i0, i1) (first axis outermost), into acc [n_mu].
void cut_t0_cells(const Mesh& M, int dim, size_t i0, size_t i1, const double* mu, size_t n_mu, double eta,
                  double* acc) {
  const int n = M.V.n;
  for (size_t idx = i0; idx < i1; ++idx) {
    if (dim == 2) {
      size_t vs[4][3];
      double c[4][3][2];
      cut::cell_triangles(idx / M.ny, idx % M.ny, M.nx, M.ny, vs, c);
      const double half_vol = (1.0 / static_cast<double>(M.nx)) * (1.0 / static_cast<double>(M.ny)) / 4.0;
      for (int t = 0; t < 4; ++t) {
        const double area = cut::tri_area(c[t]);
        if (area < cut::kEps) continue;
        const double scale = half_vol / area;
        for (int b = 0; b < n; ++b) simplex_band<3, 2>(M, vs[t], c[t], scale, area, b, mu, n_mu, eta, acc);
      }
    } else {
      size_t vs[5][4];
      double c[5][4][3], vol[5];
      cut::cell_tetrahedra(idx / (M.ny * M.nz), (idx / M.nz) % M.ny, idx % M.nz, M.nx, M.ny, M.nz, vs, c, vol);
      for (int t = 0; t < 5; ++t) {
        const double v = cut::tet_vol(c[t][0], c[t][1], c[t][2], c[t][3]);
        if (v < cut::kEps) continue;
        const double scale = vol[t] / v;
        for (int b = 0; b < n; ++b) simplex_band<4, 3>(M, vs[t], c[t], scale, v, b, mu, n_mu, eta, acc);
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
    th.emplace_back(
        [&, t] { cut_t0_cells(M, dim, cells * t / nt, cells * (t + 1) / nt, mu, n_mu, eta, part[t].data()); });
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
// (summed in neighbour order, then divided by their count).
//
// Which neighbours are visited when a vertex is assigned depends only on the
// mesh, not on the data, and a vertex's assignment reads only the final
// permutations of those neighbours.  So the walk is replayed once without
// data to record, per vertex, those neighbours and a level (one more than the
// highest level among them); then the levels run in order, the vertices of
// one level on host threads at once.  Every vertex computes exactly what the
// sequential walk computes.
void energy_cut_schedule(size_t nx, size_t ny, size_t nz, TrackSchedule& ts) {
  const size_t nk = nx * ny * nz;
  ts.ndep.assign(nk, 0);
  ts.dep.assign(nk * 6, 0);
  std::vector<int> level(nk, 0);
  std::vector<size_t> order;
  order.reserve(nk);
  std::vector<char> visited(nk, 0);
  std::deque<size_t> queue;
  visited[0] = 1;
  queue.push_back(0);
  order.push_back(0);
  size_t nbs[6], vns[6];
  int max_level = 0;
  while (!queue.empty()) {
    const size_t cur = queue.front();
    queue.pop_front();
    const int cn = neighbours(cur, nx, ny, nz, nbs);
    for (int q = 0; q < cn; ++q) {
      const size_t nb = nbs[q];
      if (visited[nb]) continue;
      const int vc = neighbours(nb, nx, ny, nz, vns);
      int lv = 0;
      for (int r = 0; r < vc; ++r)
        if (visited[vns[r]]) {
          ts.dep[nb * 6 + ts.ndep[nb]++] = vns[r];
          lv = std::max(lv, level[vns[r]] + 1);
        }
      level[nb] = lv;
      max_level = std::max(max_level, lv);
      visited[nb] = 1;
      queue.push_back(nb);
      order.push_back(nb);
    }
  }
  ts.level_start.assign(static_cast<size_t>(max_level) + 2, 0);
  for (size_t v : order) ++ts.level_start[static_cast<size_t>(level[v]) + 1];
  for (size_t l = 1; l < ts.level_start.size(); ++l) ts.level_start[l] += ts.level_start[l - 1];
  ts.by_level.assign(order.size(), 0);
  std::vector<size_t> fill(ts.level_start.begin(), ts.level_start.end() - 1);
  for (size_t v : order) ts.by_level[fill[level[v]]++] = v;
}

namespace {

// A reusable barrier for a fixed set of host threads (generation counting).
class SpinBarrier {
 public:
  explicit SpinBarrier(int n) : n_(n) {}
  void wait() {
    const int gen = gen_.load(std::memory_order_acquire);
    if (count_.fetch_add(1, std::memory_order_acq_rel) + 1 == n_) {
      count_.store(0, std::memory_order_relaxed);
      gen_.fetch_add(1, std::memory_order_release);
    } else {
      while (gen_.load(std::memory_order_acquire) == gen) std::this_thread::yield();
    }
  }

 private:
  const int n_;
  std::atomic<int> count_{0};
  std::atomic<int> gen_{0};
};

}  // namespace

void energy_cut_track(size_t nx, size_t ny, size_t nz, int n, const double* ovx, const double* ovy,
                      const double* ovz, bool squared, std::vector<int>& perm) {
  const size_t zs = squared ? 1 : 2;  // doubles per overlap element
  const size_t nk = nx * ny * nz;
  const size_t n2 = static_cast<size_t>(n) * n;
  perm.resize(nk * n);
  for (size_t i = 0; i < nk; ++i)
    for (int b = 0; b < n; ++b) perm[i * n + b] = b;
  if (nk <= 1) return;
  TrackSchedule ts;
  energy_cut_schedule(nx, ny, nz, ts);
  // the assignment of vertex nb from its scheduled neighbours
  auto assign = [&](size_t nb, std::vector<double>& ov, std::vector<int>& p, std::vector<char>& assigned,
                    std::vector<std::pair<int, double>>& rows) {
    const int cnt = ts.ndep[nb];
    if (cnt == 0) return;
    for (int r = 0; r < cnt; ++r) {
      const size_t vn = ts.dep[nb * 6 + r];
      // ov[a][b] = |<tracked band a of vn | raw band b of nb>|^2: element
      // (raw row r, column b) of U_vn^H U_nb sits at base + 2 (r rs + b cs);
      // the pair's axis from its index step (x first: with ny = 1 the x and
      // y steps coincide and only x neighbours exist)
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
          o[b] = r == 0 ? x : o[b] + x;
        }
      }
    }
    for (double& o : ov) o /= static_cast<double>(cnt);
    greedy_assign(n, ov, p, assigned, rows);
    for (int b = 0; b < n; ++b) perm[nb * n + b] = p[b];
  };
  const size_t n_levels = ts.level_start.size() - 1;
  const size_t widest = [&] {
    size_t w = 0;
    for (size_t l = 0; l < n_levels; ++l) w = std::max(w, ts.level_start[l + 1] - ts.level_start[l]);
    return w;
  }();
  // threads only where the levels are wide enough to pay for the barriers
  const int nt = static_cast<int>(std::min<size_t>(host_threads(nk), std::max<size_t>(1, widest / 4)));
  SpinBarrier bar(nt);
  auto worker = [&](int t) {
    std::vector<double> ov(n2);
    std::vector<int> p(n);
    std::vector<char> assigned(n);
    std::vector<std::pair<int, double>> rows;
    for (size_t l = 0; l < n_levels; ++l) {
      const size_t a = ts.level_start[l], b = ts.level_start[l + 1];
      for (size_t i = a + t; i < b; i += nt) assign(ts.by_level[i], ov, p, assigned, rows);
      if (nt > 1) bar.wait();
    }
  };
  if (nt == 1) {
    worker(0);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(worker, t);
  worker(0);
  for (auto& x : th) x.join();
}

// integrate_fermi_cut_2d / _3d (energy_cut.rs:1355-1409, 1798-1850): the
// zero-width cut t0 at mu, or for width > 0 its convolution with the Fermi
// window: t0 on a grid of width / 10 steps reaching 12 widths past the mu
// range, interpolated linearly under the 72-point midpoint rule on |x| <= 18.
int energy_cut_convolve(const double* mu, size_t n_mu, double width,
                        const std::function<int(const double*, size_t, double*)>& t0, double* out) {
  if (width == 0.0) return t0(mu, n_mu, out);
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
  RTB_TRY(t0(mu_ext.data(), n_ext, s0.data()));
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
  return RTB_OK;
}

// The host form of the cut: vertex Omega of every vertex in its tracked band
// order once, then the cells on host threads.  Fractional measure; a 2D mesh
// is nz = 1.
void energy_cut_sigma(size_t nx, size_t ny, size_t nz, int dim, int n, const double* E, const double* Ki, size_t ks,
                      const int* perm, const double* mu, size_t n_mu, double width, double eta, double* out) {
  const size_t nk = nx * ny * nz;
  const cut::View V{n, E, Ki, ks, perm};
  std::vector<double> om(nk * n);
  {
    const size_t nt = host_threads(nk);
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t)
      th.emplace_back([&, t] {
        for (size_t v = nk * t / nt; v < nk * (t + 1) / nt; ++v)
          for (int i = 0; i < n; ++i) om[v * n + i] = cut::vertex_omega(V, v, i, eta);
      });
    for (auto& x : th) x.join();
  }
  const Mesh M{nx, ny, nz, V, om.data()};
  energy_cut_convolve(mu, n_mu, width, [&](const double* m, size_t nm, double* s0) {
    cut_t0(M, dim, m, nm, eta, s0);
    return RTB_OK;
  }, out);
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
