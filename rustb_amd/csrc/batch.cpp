// Many models in one call: the Direct anomalous Hall conductivity of every
// model of a batch, each on its own uniform mesh (the reference runs
// `hall_conductivity` once per model, response/linear/mod.rs:160-225).
//
// Plan (per precision, built on first use):
//   * models are grouped by matrix size n (the eigensolver and the Kubo
//     products are specialised per n); in each group the models follow each
//     other in decreasing cost, and the groups follow each other, so every
//     model's k-points form one range of a global k order;
//   * the global order is cut into passes whose device buffers fit the
//     workspace budget; a pass holds one portion per n (its models'
//     consecutive k-points, split into per-model segments);
//   * a pass runs the Fourier sums of all its models as work-queue launches
//     (one per k-points-per-thread variant; tiles of any n, longest first),
//     plus, in the mixed form, the separable mesh passes of the models they
//     suit; then per portion the eigensolve (FP64, or FP32 + FP64
//     refinement), the Kubo band sum and the per-k band sums f(E) Omega;
//   * a last launch sums each model's range of per-k terms.
// The launch sequence is captured into a HIP graph per (precision, number of
// chemical potentials, T, eta) and replayed; the chemical potentials are
// copied to the device before each replay.  Counters are cleared by fill
// kernels inside the graph (no memset nodes).  k-points the mixed check
// flags are re-solved in FP64 after the replay and their per-k terms
// replaced before the sums are formed again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "pipeline.hpp"

namespace rtb {
namespace {

struct Variant {
  int tk = 2;  // k-points per thread of the work-queue Fourier sum
  int nb = 0;  // register eigensolver block rows (0: by n)
};

struct Segment {  // the k-points of one model inside a portion
  int model;
  size_t off;  // first k-point in the portion
  size_t c;    // k-points
  size_t q0;   // first model-local index (gen_kmesh flat index; k1-innermost on the mesh route)
  bool mesh;
};

struct Portion {  // the k-points of one pass with one matrix size
  int n = 0;
  int group = 0;
  size_t gk0 = 0;  // first global k index
  size_t c = 0;
  std::vector<Segment> segs;
  // device buffers
  double *H = nullptr, *V = nullptr, *E = nullptr, *Om = nullptr, *Z = nullptr, *kub = nullptr, *dsc = nullptr;
  EigBufs eb{};
};

struct Queue {  // one work-queue launch
  int tk = 0;
  size_t t0 = 0, t1 = 0;  // tile range in the table
  int max_groups = 1;
  int workgroups = 1;
  int counter = 0;
};

struct Pass {
  std::vector<Portion> portions;
  std::vector<Queue> queues;
};

struct Plan {
  bool built = false;
  bool f32 = false;
  std::vector<int> order;        // model indices in global k order
  std::vector<size_t> gstart;    // per model: first global k index
  std::vector<int> group_n;      // matrix size per group
  std::vector<int> model_group;  // per model
  std::vector<bool> exact;       // per group: Kubo from A and B apart (V0, V1 kept apart)
  std::vector<bool> mesh;        // per model: separable mesh route
  std::vector<MeshPlan> mp;      // per model (mesh route)
  std::vector<FourierArgs> fa;   // per model
  std::vector<Pass> passes;
  std::vector<QueueTile> tiles;
  size_t nk_total = 0;
  int n_portions = 0, n_counters = 0;
  // device: the plan's own arena (graphs of the plan point into it)
  void* arena = nullptr;
  size_t arena_bytes = 0;
  QueueModel* d_models = nullptr;
  QueueTile* d_tiles = nullptr;
  unsigned* d_counters = nullptr;
  int* d_flags = nullptr;  // per portion: flag count; then [nk_total] flag lists
  int* d_nonfinite = nullptr;
  unsigned long long* d_maxres = nullptr;
  long long* d_start = nullptr;  // [n_models + 1] range starts in global order
  std::vector<double*> G;        // per model (mesh route)
  std::vector<int*> bidx;        // per model (mesh route)
  double* boxwork = nullptr;
};

struct GraphEntry {
  int precision = 0;
  size_t n_mu = 0;
  double T = 0.0, eta = 0.0;
  hipGraphExec_t exec = nullptr;
  size_t nodes = 0;
  double* dmu = nullptr;
  double* term = nullptr;
  double* dsig = nullptr;
};

// fermi_from_width (thermodynamics.rs:100-118), as the device's band sums
double fermi_host(double e, double mu, double width) {
  if (width == 0.0) return e < mu ? 1.0 : (e > mu ? 0.0 : 0.5);
  const double x = (e - mu) / width;
  if (x > 50.0) return 0.0;
  if (x < -50.0) return 1.0;
  return 1.0 / (1.0 + std::exp(x));
}

}  // namespace
}  // namespace rtb

using namespace rtb;

struct rtb_batch {
  int device = 0;
  std::vector<const rtb_model*> models;
  std::vector<std::array<size_t, 3>> kmesh;
  std::vector<size_t> nk;
  std::vector<std::vector<double>> dirs;  // per model: 2 rows of dim entries
  std::vector<Variant> var;
  std::mutex mu;
  Plan plans[2];                 // FP64, mixed
  std::list<GraphEntry> graphs;  // most recently used first
  int captures = 0;
  bool use_graph = true;
  int last_precision = RTB_PREC_FP64;
  size_t last_nodes = 0;
  double last_graph_ms = 0.0;
};

namespace rtb {
namespace {

constexpr size_t kMaxGraphs = 4;

// Buffers of a portion beyond H and V: eigenvalues, Omega, eigensolver,
// Kubo products (T1, A, B), the FP32-route vectors and column scales.
int shared_alloc(Work& w, int n, size_t c, bool f32, Portion& p) {
  const size_t n2 = static_cast<size_t>(n) * n;
  RTB_TRY(w.alloc(&p.E, c * n));
  RTB_TRY(w.alloc(&p.Om, c * n));
  RTB_TRY(alloc_eig(w, n, c, true, p.eb));
  RTB_TRY(w.alloc(&p.kub, c * n2 * 6));
  if (f32) {
    RTB_TRY(w.alloc(&p.Z, c * n2 * 2));
    RTB_TRY(w.alloc(&p.dsc, c * n));
  } else {
    p.Z = nullptr;
    p.dsc = nullptr;
  }
  return RTB_OK;
}

size_t shared_bytes(int n, size_t c, bool f32) {
  Work probe(nullptr, SIZE_MAX);
  Portion p;
  shared_alloc(probe, n, c, f32, p);
  return probe.used();
}

size_t hv_per_k(int n, bool exact) { return static_cast<size_t>(n) * n * 16 * (exact ? 3 : 2); }

void free_graph(GraphEntry& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.dmu) (void)hipFree(g.dmu);
  if (g.term) (void)hipFree(g.term);
  if (g.dsig) (void)hipFree(g.dsig);
  g.exec = nullptr;
  g.dmu = g.term = g.dsig = nullptr;
}

// precision < 0: every graph
void drop_graphs(rtb_batch* b, int precision) {
  for (auto it = b->graphs.begin(); it != b->graphs.end();) {
    if (precision < 0 || it->precision == precision) {
      free_graph(*it);
      it = b->graphs.erase(it);
    } else {
      ++it;
    }
  }
}

// The passes of the global k order under a budget of `budget` bytes of pass
// buffers (H and V of every portion, plus the other buffers of the largest
// portion, which the portions of a pass share one after another).  The
// budget is a target: a pass holds at least one (k2, k3) pair.
void form_passes(const rtb_batch* b, Plan& pl, size_t budget) {
  pl.passes.clear();
  const bool f32 = pl.f32;
  std::map<int, size_t> sh_cache;
  auto sh_of = [&](int n) {
    auto it = sh_cache.find(n);
    if (it != sh_cache.end()) return it->second;
    const size_t v = shared_bytes(n, 4096, f32) / 4096 + 64;
    sh_cache[n] = v;
    return v;
  };
  Pass cur;
  size_t sum_hv = 0, max_sh_other = 0;
  auto close_pass = [&]() {
    if (!cur.portions.empty()) pl.passes.push_back(std::move(cur));
    cur = Pass{};
    sum_hv = 0;
    max_sh_other = 0;
  };
  size_t gk = 0;
  for (int mi : pl.order) {
    const int n = b->models[mi]->nsta;
    const int grp = pl.model_group[mi];
    const size_t total = b->nk[mi];
    const size_t align = pl.mesh[mi] ? b->kmesh[mi][0] : 1;  // whole (k2, k3) pairs on the mesh route
    const size_t hv = hv_per_k(n, pl.exact[grp]), sh = sh_of(n);
    size_t done = 0;
    while (done < total) {
      const bool same = !cur.portions.empty() && cur.portions.back().n == n;
      const size_t c_cur = same ? cur.portions.back().c : 0;
      const size_t base_sh =
          same || cur.portions.empty()
              ? max_sh_other
              : std::max(max_sh_other, cur.portions.back().c * sh_of(cur.portions.back().n));
      const size_t slack = (cur.portions.size() + 2) * (size_t(64) << 10);
      // x more k-points fit when sum_hv + x hv + max(base_sh, (c_cur + x) sh) + slack <= budget
      size_t x = 0;
      if (budget > sum_hv + base_sh + slack && budget > sum_hv + c_cur * sh + slack) {
        const size_t x1 = (budget - sum_hv - base_sh - slack) / hv;
        const size_t x2 = (budget - sum_hv - c_cur * sh - slack) / (hv + sh);
        x = std::min(x1, x2);
      }
      x = std::min(x, total - done);
      if (x < total - done) x = x / align * align;
      if (x == 0) {
        if (!cur.portions.empty()) {
          close_pass();
          continue;
        }
        x = std::min(total - done, align);
      }
      if (!same) {
        max_sh_other = base_sh;
        Portion p;
        p.n = n;
        p.group = grp;
        p.gk0 = gk;
        cur.portions.push_back(std::move(p));
      }
      Portion& p = cur.portions.back();
      p.segs.push_back({mi, p.c, x, done, static_cast<bool>(pl.mesh[mi])});
      p.c += x;
      sum_hv += x * hv;
      done += x;
      gk += x;
      if (done < total) close_pass();
    }
  }
  close_pass();
}

// Tiles of the work-queue Fourier launches of every pass (dense-route
// segments), sorted longest first within each launch.
void form_tiles(const rtb_batch* b, Plan& pl, int cus) {
  pl.tiles.clear();
  pl.n_counters = 0;
  for (Pass& ps : pl.passes) {
    ps.queues.clear();
    for (int tk : {8, 4, 2, 1}) {
      Queue q;
      q.tk = tk;
      q.t0 = pl.tiles.size();
      for (Portion& p : ps.portions) {
        const int n = p.n;
        const size_t n2 = static_cast<size_t>(n) * n;
        const size_t vper = pl.exact[p.group] ? 2 : 1;
        for (const Segment& s : p.segs) {
          if (s.mesh || b->var[s.model].tk != tk) continue;
          const int G = fourier_queue_groups(n, tk);
          const size_t cap = static_cast<size_t>(G) * tk;
          const int eblocks = n2 <= 256 ? 1 : static_cast<int>((n2 + 255) / 256);
          q.max_groups = std::max(q.max_groups, G);
          for (size_t i = 0; i < s.c; i += cap) {
            const size_t cnt = std::min(cap, s.c - i);
            for (int e = 0; e < eblocks; ++e) {
              QueueTile t;
              t.model = s.model;
              t.e0 = e * 256;
              t.f0 = static_cast<long long>(s.q0 + i);
              t.cnt = static_cast<int>(cnt);
              // addresses only once the buffers are laid out
              t.H = p.H ? p.H + (s.off + i) * n2 * 2 : nullptr;
              t.V = p.V ? p.V + (s.off + i) * n2 * 2 * vper : nullptr;
              pl.tiles.push_back(t);
            }
          }
        }
      }
      q.t1 = pl.tiles.size();
      if (q.t1 == q.t0) continue;
      // longest first: a tile's duration follows its model's R count
      std::stable_sort(pl.tiles.begin() + q.t0, pl.tiles.begin() + q.t1,
                       [&](const QueueTile& x, const QueueTile& y) {
                         return b->models[x.model]->nR > b->models[y.model]->nR;
                       });
      int per_cu = 1;
      if (fourier_queue_occupancy(tk, q.max_groups, &per_cu) != RTB_OK) per_cu = 1;
      q.workgroups = static_cast<int>(std::min<size_t>(q.t1 - q.t0, static_cast<size_t>(cus) * per_cu));
      q.counter = pl.n_counters++;
      ps.queues.push_back(q);
    }
  }
}

// Bytes of a pass's buffers; with a base, their addresses: H and V of every
// portion, then the other buffers of each portion from one shared offset.
size_t pass_layout(Plan& pl, Pass& ps, char* base) {
  Work wp(base, SIZE_MAX / 2);
  for (Portion& p : ps.portions) {
    const size_t n2 = static_cast<size_t>(p.n) * p.n;
    wp.alloc(&p.H, p.c * n2 * 2);
    wp.alloc(&p.V, p.c * n2 * 2 * (pl.exact[p.group] ? 2 : 1));
  }
  const size_t sh0 = wp.used();
  size_t shmax = 0;
  for (Portion& p : ps.portions) {
    Work ws(base ? base + sh0 : nullptr, SIZE_MAX / 2);
    shared_alloc(ws, p.n, p.c, pl.f32, p);
    shmax = std::max(shmax, ws.used());
  }
  return sh0 + shmax;
}

// The plan's device buffers in its own arena (allocated here; the pass
// budget is halved while the allocation fails): the model and tile tables,
// counters, check flags, the mesh-route intermediates and the direction
// tables, then one region that every pass reuses.
int layout(rtb_batch* b, Plan& pl) {
  const int nm = static_cast<int>(b->models.size());
  int cus = 0;
  RTB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  cus = std::max(cus, 1);
  size_t budget = workspace_budget();
  for (;;) {
    form_passes(b, pl, budget);
    pl.n_portions = 0;
    size_t region = 0;
    for (Pass& ps : pl.passes) {
      pl.n_portions += static_cast<int>(ps.portions.size());
      for (Portion& p : ps.portions) p.H = p.V = nullptr;
      region = std::max(region, pass_layout(pl, ps, nullptr));
      for (Portion& p : ps.portions) p.H = p.V = nullptr;
    }
    form_tiles(b, pl, cus);  // counts only (no buffer addresses yet)
    auto carve = [&](Work& w, bool fill) -> int {
      RTB_TRY(w.alloc(&pl.d_models, static_cast<size_t>(nm)));
      RTB_TRY(w.alloc(&pl.d_tiles, std::max<size_t>(pl.tiles.size(), 1)));
      RTB_TRY(w.alloc(&pl.d_counters, static_cast<size_t>(std::max(pl.n_counters, 1))));
      RTB_TRY(w.alloc(&pl.d_start, static_cast<size_t>(nm) + 1));
      RTB_TRY(w.alloc(&pl.d_flags, static_cast<size_t>(pl.n_portions) + pl.nk_total));
      RTB_TRY(w.alloc(&pl.d_nonfinite, 1));
      RTB_TRY(w.alloc(&pl.d_maxres, 1));
      pl.G.assign(nm, nullptr);
      pl.bidx.assign(nm, nullptr);
      size_t maxwork = 0;
      for (int i = 0; i < nm; ++i) {
        if (!pl.mesh[i]) continue;
        RTB_TRY(w.alloc(&pl.G[i], 2 * pl.mp[i].gb_elems));
        RTB_TRY(w.alloc(&pl.bidx[i], static_cast<size_t>(b->models[i]->nR)));
        maxwork = std::max(maxwork, pl.mp[i].box_elems + pl.mp[i].ga_elems);
      }
      pl.boxwork = nullptr;
      if (maxwork) RTB_TRY(w.alloc(&pl.boxwork, 2 * maxwork));
      for (int i = 0; i < nm; ++i) {  // direction tables (w.R, w.tau) of every model
        if (fill) {
          FourierArgs a;
          RTB_TRY(upload_dirs(nullptr, w, b->models[i], b->dirs[i].data(), 2, a));
          a.gauge = RTB_GAUGE_ATOM;  // berry_curvature_at_impl: gen_v_projected(k, Gauge::Atom, ..)
          a.combine = pl.exact[pl.model_group[i]] ? 0 : 1;
          pl.fa[i] = a;
        } else {
          double* t;
          RTB_TRY(w.alloc(&t, 2 * static_cast<size_t>(b->models[i]->nR)));
          RTB_TRY(w.alloc(&t, 2 * static_cast<size_t>(b->models[i]->nsta)));
        }
      }
      char* reg = nullptr;
      RTB_TRY(w.alloc(&reg, region));
      if (fill)
        for (Pass& ps : pl.passes) pass_layout(pl, ps, reg);
      return RTB_OK;
    };
    Work probe(nullptr, SIZE_MAX);
    RTB_TRY(carve(probe, false));
    const size_t need = probe.used() + 4096;
    if (need > pl.arena_bytes) {
      drop_graphs(b, pl.f32 ? RTB_PREC_MIXED : RTB_PREC_FP64);
      if (pl.arena) RTB_HIP(hipFree(pl.arena));
      pl.arena = nullptr;
      pl.arena_bytes = 0;
      const hipError_t e = hipMalloc(&pl.arena, need);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        pl.arena = nullptr;
        if (budget > (size_t(256) << 20)) {
          budget /= 2;  // smaller passes
          continue;
        }
        set_error("batch: device allocation of %zu bytes failed: %s", need, hipGetErrorString(e));
        return RTB_ERR_DEVICE;
      }
      pl.arena_bytes = need;
    }
    Work w(pl.arena, pl.arena_bytes);
    RTB_TRY(carve(w, true));
    form_tiles(b, pl, cus);  // with the buffer addresses
    std::vector<QueueModel> qm(nm);
    for (int i = 0; i < nm; ++i) {
      qm[i].a = pl.fa[i];
      for (int d = 0; d < 3; ++d) qm[i].N[d] = static_cast<int>(b->kmesh[i][d]);
    }
    RTB_HIP(hipMemcpy(pl.d_models, qm.data(), sizeof(QueueModel) * nm, hipMemcpyHostToDevice));
    if (!pl.tiles.empty())
      RTB_HIP(hipMemcpy(pl.d_tiles, pl.tiles.data(), sizeof(QueueTile) * pl.tiles.size(), hipMemcpyHostToDevice));
    std::vector<long long> start(static_cast<size_t>(nm) + 1);
    for (int j = 0; j < nm; ++j) start[j] = static_cast<long long>(pl.gstart[pl.order[j]]);
    start[nm] = static_cast<long long>(pl.nk_total);
    RTB_HIP(hipMemcpy(pl.d_start, start.data(), sizeof(long long) * start.size(), hipMemcpyHostToDevice));
    for (int i = 0; i < nm; ++i) {
      if (!pl.mesh[i]) continue;
      std::vector<int> idx;
      mesh_box_index(b->models[i], pl.mp[i], idx);
      RTB_HIP(hipMemcpy(pl.bidx[i], idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
    }
    RTB_HIP(hipDeviceSynchronize());
    return RTB_OK;
  }
}

int build_plan(rtb_batch* b, int precision) {
  Plan& pl = b->plans[precision == RTB_PREC_MIXED ? 1 : 0];
  if (pl.built) return RTB_OK;
  const int nm = static_cast<int>(b->models.size());
  pl.f32 = precision == RTB_PREC_MIXED;
  // groups by n; models by decreasing cost inside a group
  std::vector<int> ns;
  for (const rtb_model* m : b->models) ns.push_back(m->nsta);
  std::sort(ns.begin(), ns.end());
  ns.erase(std::unique(ns.begin(), ns.end()), ns.end());
  pl.group_n = ns;
  pl.model_group.assign(nm, 0);
  pl.exact.assign(ns.size(), !pl.f32);
  for (int i = 0; i < nm; ++i) {
    const int g = static_cast<int>(std::lower_bound(ns.begin(), ns.end(), b->models[i]->nsta) - ns.begin());
    pl.model_group[i] = g;
    if (!b->models[i]->herm) pl.exact[g] = true;
  }
  pl.order.resize(nm);
  for (int i = 0; i < nm; ++i) pl.order[i] = i;
  std::stable_sort(pl.order.begin(), pl.order.end(), [&](int x, int y) {
    const rtb_model *a = b->models[x], *c = b->models[y];
    if (a->nsta != c->nsta) return a->nsta < c->nsta;
    return static_cast<double>(a->nR) * b->nk[x] > static_cast<double>(c->nR) * b->nk[y];
  });
  pl.gstart.assign(nm, 0);
  pl.nk_total = 0;
  for (int j = 0; j < nm; ++j) {
    pl.gstart[pl.order[j]] = pl.nk_total;
    pl.nk_total += b->nk[pl.order[j]];
  }
  pl.mesh.assign(nm, false);
  pl.mp.assign(nm, MeshPlan{});
  pl.fa.assign(nm, FourierArgs{});
  if (pl.f32) {
    for (int i = 0; i < nm; ++i) {
      RTB_TRY(mesh_plan(b->models[i], b->kmesh[i].data(), pl.mp[i]));
      pl.mesh[i] = pl.mp[i].ok;
    }
  }
  RTB_TRY(layout(b, pl));
  pl.built = true;
  return RTB_OK;
}

// The launch sequence of one call (captured into the graph).
// tmr: stage timing of an eager run (a capture takes a Timer without events:
// event-record nodes of a graph give no elapsed times on this runtime).
int record(const rtb_batch* b, Plan& pl, hipStream_t st, const GraphEntry& g, Timer& tmr) {
  const MixedOpts opt = mixed_opts();
  RTB_TRY(zero_ints(st, reinterpret_cast<int*>(pl.d_counters), static_cast<size_t>(std::max(pl.n_counters, 1))));
  RTB_TRY(zero_ints(st, pl.d_flags, static_cast<size_t>(pl.n_portions)));
  RTB_TRY(zero_ints(st, pl.d_nonfinite, 1));
  RTB_TRY(zero_ints(st, reinterpret_cast<int*>(pl.d_maxres), 2));
  const int nm = static_cast<int>(b->models.size());
  for (int i = 0; i < nm; ++i)
    if (pl.mesh[i]) {
      RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
      RTB_TRY(mesh_prepare_dev(st, b->models[i], pl.mp[i], pl.bidx[i], pl.boxwork, pl.G[i], true));
      RTB_TRY(tmr.end(b->models[i]->dim + 1));
    }
  int portion_no = 0;
  for (Pass& ps : pl.passes) {
    for (const Queue& q : ps.queues) {
      // 16 flops per (R, element, k): the phase product, H, two velocities
      double flop = 0.0;
      for (size_t t = q.t0; t < q.t1; ++t) {
        const QueueTile& qt = pl.tiles[t];
        const rtb_model* m = b->models[qt.model];
        const double n2 = static_cast<double>(m->nsta) * m->nsta;
        flop += 16.0 * m->nR * std::min(256.0, n2 - qt.e0) * qt.cnt;
      }
      RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
      RTB_TRY(fourier_queue(st, q.tk, pl.d_models, pl.d_tiles + q.t0, static_cast<int>(q.t1 - q.t0), q.max_groups,
                            pl.d_counters + q.counter, q.workgroups));
      RTB_TRY(tmr.end(1, flop));
    }
    for (Portion& p : ps.portions) {
      const size_t n2 = static_cast<size_t>(p.n) * p.n;
      const size_t vper = pl.exact[p.group] ? 2 : 1;
      for (const Segment& s : p.segs) {
        if (!s.mesh) continue;
        const size_t N1 = b->kmesh[s.model][0];
        RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
        RTB_TRY(mesh_final(st, b->models[s.model], pl.mp[s.model], pl.G[s.model], pl.fa[s.model],
                           b->dirs[s.model].data(), s.q0 / N1, s.c / N1, p.H + s.off * n2 * 2,
                           p.V + s.off * n2 * 2 * vper));
        RTB_TRY(tmr.end(1));
      }
    }
    for (Portion& p : ps.portions) {
      const int n = p.n;
      const size_t n2 = static_cast<size_t>(n) * n;
      const rtb_model* rep = b->models[p.segs.front().model];
      const int nb = b->var[p.segs.front().model].nb;
      for (const Segment& s : p.segs)
        if (!b->models[s.model]->herm) {
          RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
          RTB_TRY(hermitize_upper(st, n, s.c, p.H + s.off * n2 * 2));
          RTB_TRY(tmr.end(1));
        }
      double* T1 = p.kub;
      double* A = p.kub + p.c * n2 * 2;
      double* B = p.kub + p.c * n2 * 4;
      double* Zc;
      double* Ab;
      const double* zscale = nullptr;
      if (pl.f32) {
        RefineCheck chk;
        chk.res_tol = opt.res_tol;
        chk.fail_count = pl.d_flags + portion_no;
        chk.fail_list = pl.d_flags + pl.n_portions + p.gk0;
        chk.max_res = pl.d_maxres;
        RTB_TRY(eig_mixed_chunk(st, tmr, false, n, p.c, p.H, p.E, p.eb, p.Z, T1, A, B, p.dsc, chk, nb));
        Zc = A;  // refined vectors
        Ab = p.Z;
        zscale = p.dsc;
      } else {
        RTB_TRY(eigh_launch_fp64(st, &tmr, n, p.c, p.H, p.E, true, p.eb.Vg, p.eb.tau, p.eb.Y, p.eb.S, p.eb.DE,
                                 false, p.eb.band, nb));
        RTB_TRY(count_nonfinite(st, n, p.c, p.E, pl.d_nonfinite));
        RTB_TRY(tmr.begin(RTB_STAGE_BACKTRANS));
        RTB_TRY(eigh_backtransform_fp64(st, n, p.c, p.eb.Y, p.eb.DE, p.eb.S, p.eb.Vg, p.eb.tau, p.H, false,
                                        p.eb.band, nb));
        RTB_TRY(tmr.end(1, flop_backtrans(n, p.c)));
        Zc = p.H;
        Ab = A;
      }
      RTB_TRY(kubo_chunk(st, tmr, rep, p.c, p.V, Zc, zscale, p.E, RTB_SPIN_NONE, g.eta, T1, Ab, nullptr, p.Om,
                         pl.exact[p.group] ? B : nullptr, nullptr));
      RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
      RTB_TRY(kterms(st, n, p.c, p.Om, p.E, g.dmu, static_cast<int>(g.n_mu), g.T, g.term + p.gk0 * g.n_mu));
      RTB_TRY(tmr.end(1));
      ++portion_no;
    }
  }
  RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
  RTB_TRY(segment_sums(st, g.term, pl.d_start, nm, static_cast<int>(g.n_mu), g.dsig));
  RTB_TRY(tmr.end(1));
  return RTB_OK;
}

// The buffers of a parameter set: mu, the per-k terms, the model sums.
int entry_alloc(const rtb_batch* b, const Plan& pl, GraphEntry& g) {
  RTB_HIP(hipMalloc(&g.dmu, sizeof(double) * g.n_mu));
  RTB_HIP(hipMalloc(&g.term, sizeof(double) * std::max<size_t>(pl.nk_total * g.n_mu, 1)));
  RTB_HIP(hipMalloc(&g.dsig, sizeof(double) * b->models.size() * g.n_mu));
  return RTB_OK;
}

int capture(rtb_batch* b, Plan& pl, hipStream_t st, GraphEntry& g) {
  Timer none;
  RTB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = record(b, pl, st, g, none);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc != RTB_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess || !graph) {
    set_error("batch: graph capture failed: %s", hipGetErrorString(e));
    return RTB_ERR_DEVICE;
  }
  size_t nodes = 0;
  (void)hipGraphGetNodes(graph, nullptr, &nodes);
  const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    g.exec = nullptr;
    set_error("batch: graph instantiation failed: %s", hipGetErrorString(ei));
    return RTB_ERR_DEVICE;
  }
  g.nodes = nodes;
  ++b->captures;
  return RTB_OK;
}

// Model-local index q of a segment's k-point -> fractional k (model dim
// entries): gen_kmesh order (first axis outermost), or k1 innermost within
// (k2, k3) pairs on the mesh route.
void k_of(const rtb_batch* b, int mi, bool mesh, size_t q, double* k) {
  const int dim = b->models[mi]->dim;
  const auto& N = b->kmesh[mi];
  size_t x[3] = {0, 0, 0};
  if (mesh) {
    x[0] = q % N[0];
    const size_t pair = q / N[0];
    x[1] = dim == 3 ? pair / N[2] : pair;
    x[2] = dim == 3 ? pair % N[2] : 0;
  } else {
    size_t f = q;
    for (int d = dim - 1; d >= 0; --d) {
      x[d] = f % N[d];
      f /= N[d];
    }
  }
  for (int d = 0; d < dim; ++d) k[d] = static_cast<double>(x[d]) / static_cast<double>(N[d]);
}

// The k-points the mixed check flagged, re-solved in FP64: their per-k terms
// replaced, the model sums formed again.  The count goes to *resolved.
int resolve(rtb_batch* b, Plan& pl, GraphEntry& g, hipStream_t st, Lease& lease, const double* mu, double T,
            size_t* resolved) {
  *resolved = 0;
  std::vector<int> counts(static_cast<size_t>(pl.n_portions));
  RTB_HIP(hipMemcpyAsync(counts.data(), pl.d_flags, sizeof(int) * counts.size(), hipMemcpyDeviceToHost, st));
  RTB_HIP(hipStreamSynchronize(st));
  size_t total = 0;
  for (int c : counts) total += static_cast<size_t>(c);
  if (total == 0) return RTB_OK;
  if (!mixed_opts().fallback) {
    set_error("EigenvalueComputationFailed: the refined eigenpairs of %zu k-point(s) fail the residual check", total);
    return RTB_ERR_EIGEN_FAILED;
  }
  // per model: (global index, fractional k)
  std::map<int, std::vector<std::pair<size_t, std::array<double, 3>>>> per_model;
  int pno = 0;
  for (Pass& ps : pl.passes)
    for (Portion& p : ps.portions) {
      const int cnt = counts[pno++];
      if (cnt == 0) continue;
      std::vector<int> list(static_cast<size_t>(cnt));
      RTB_HIP(hipMemcpyAsync(list.data(), pl.d_flags + pl.n_portions + p.gk0, sizeof(int) * list.size(),
                             hipMemcpyDeviceToHost, st));
      RTB_HIP(hipStreamSynchronize(st));
      for (int li : list) {
        const size_t i = static_cast<size_t>(li);
        for (const Segment& s : p.segs)
          if (i >= s.off && i < s.off + s.c) {
            std::array<double, 3> k{0.0, 0.0, 0.0};
            k_of(b, s.model, s.mesh, s.q0 + (i - s.off), k.data());
            per_model[s.model].push_back({p.gk0 + i, k});
            break;
          }
      }
    }
  const double width = T <= 0.0 ? 0.0 : T * 8.617333262145e-5;
  for (auto& kv : per_model) {
    const rtb_model* m = b->models[kv.first];
    const int n = m->nsta, dim = m->dim;
    const size_t cnt = kv.second.size();
    std::vector<double> ks(cnt * dim), om(cnt * n), E(cnt * n);
    for (size_t i = 0; i < cnt; ++i)
      for (int d = 0; d < dim; ++d) ks[i * dim + d] = kv.second[i].second[d];
    Work w(lease);
    RTB_TRY(berry_points_fp64(m, ks.data(), cnt, b->dirs[kv.first].data(), g.eta, om.data(), E.data(), st, w));
    std::vector<double> t(g.n_mu);
    for (size_t i = 0; i < cnt; ++i) {
      for (size_t j = 0; j < g.n_mu; ++j) {
        double s = 0.0;
        for (int q = 0; q < n; ++q) s += om[i * n + q] * fermi_host(E[i * n + q], mu[j], width);
        t[j] = s;
      }
      RTB_HIP(hipMemcpyAsync(g.term + kv.second[i].first * g.n_mu, t.data(), sizeof(double) * g.n_mu,
                             hipMemcpyHostToDevice, st));
      RTB_HIP(hipStreamSynchronize(st));
    }
  }
  RTB_TRY(segment_sums(st, g.term, pl.d_start, static_cast<int>(b->models.size()), static_cast<int>(g.n_mu), g.dsig));
  *resolved = total;
  return RTB_OK;
}

}  // namespace
}  // namespace rtb

extern "C" {

int rtb_batch_create(const rtb_model* const* models, int n_models, const size_t* kmesh, const double* dirs,
                     const rtb_batch_variant* table, int n_table, rtb_batch** out) {
  if (!out) return fail(RTB_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  if (!models || n_models < 1 || !kmesh || !dirs)
    return fail(RTB_ERR_INVALID_ARGUMENT, "batch: need models, n_models >= 1, kmesh and dirs");
  if (n_table < 0 || (n_table > 0 && !table)) return fail(RTB_ERR_INVALID_ARGUMENT, "batch: bad variant table");
  for (int t = 0; t < n_table; ++t) {
    const rtb_batch_variant& v = table[t];
    if (v.nsta < 1 || (v.fourier_tk != 1 && v.fourier_tk != 2 && v.fourier_tk != 4 && v.fourier_tk != 8) ||
        !eigh_block_valid(v.nsta, v.eigh_nb)) {
      set_error("batch: variant table entry %d (nsta %d, fourier_tk %d, eigh_nb %d) is not valid", t, v.nsta,
                v.fourier_tk, v.eigh_nb);
      return RTB_ERR_INVALID_ARGUMENT;
    }
  }
  auto b = std::make_unique<rtb_batch>();
  for (int i = 0; i < n_models; ++i) {
    const rtb_model* m = models[i];
    if (!m) return fail(RTB_ERR_INVALID_ARGUMENT, "batch: null model");
    if (i == 0) b->device = m->device;
    if (m->device != b->device) return fail(RTB_ERR_INVALID_ARGUMENT, "batch: all models must be on one device");
    if (m->nsta > kMaxN) return fail(RTB_ERR_UNSUPPORTED, "nsta above 256 is not supported by the eigensolver");
    if (m->d_rmat)
      return fail(RTB_ERR_UNSUPPORTED, "batch: models with position matrices run one call each (rtb_hall_direct)");
    const size_t* km_in = kmesh + 3 * static_cast<size_t>(i);
    size_t nk = 0;
    RTB_TRY(check_kmesh(m, km_in, &nk));
    std::array<size_t, 3> km{1, 1, 1};
    for (int d = 0; d < m->dim; ++d) km[d] = km_in[d];
    std::vector<double> dv(2 * static_cast<size_t>(m->dim));
    for (int r = 0; r < 2; ++r)
      for (int d = 0; d < m->dim; ++d) dv[r * m->dim + d] = dirs[6 * static_cast<size_t>(i) + 3 * r + d];
    RTB_TRY(check_direction(m, dv.data(), 2));
    Variant v;
    for (int t = 0; t < n_table; ++t)
      if (table[t].nsta == m->nsta) {
        v.tk = table[t].fourier_tk;
        v.nb = table[t].eigh_nb;
      }
    b->models.push_back(m);
    b->kmesh.push_back(km);
    b->nk.push_back(nk);
    b->dirs.push_back(std::move(dv));
    b->var.push_back(v);
  }
  *out = b.release();
  return RTB_OK;
}

int rtb_batch_hall_direct(rtb_batch* b, const double* mu, size_t n_mu, double T, double eta, int precision,
                          double* sigma) {
  if (!b) return fail(RTB_ERR_INVALID_ARGUMENT, "null batch");
  // Switched off by default: its first device run stopped making progress
  // (cause not found); the host logic is checked by emulation only.
  const char* en = std::getenv("RTB_BATCH_ENABLE");
  if (!en || std::atoi(en) != 1)
    return fail(RTB_ERR_UNSUPPORTED,
                "the multi-model batch is switched off (its device run is not validated); RTB_BATCH_ENABLE=1 "
                "enables it");
  RTB_TRY(check_precision(precision));
  RTB_TRY(check_mu(mu, n_mu));
  RTB_TRY(check_eta(eta));
  RTB_TRY(check_T(T));
  if (!sigma) return fail(RTB_ERR_INVALID_ARGUMENT, "null sigma");
  std::lock_guard<std::mutex> lock(b->mu);
  RTB_TRY(prepare(b->models[0]));
  RTB_TRY(build_plan(b, precision));
  Plan& pl = b->plans[precision == RTB_PREC_MIXED ? 1 : 0];
  Lease lease;
  RTB_TRY(lease.status());
  hipStream_t st = lease.stream();
  // the buffers (and graph) of this parameter set, made on first use
  auto it = std::find_if(b->graphs.begin(), b->graphs.end(), [&](const GraphEntry& g) {
    return g.precision == precision && g.n_mu == n_mu && std::memcmp(&g.T, &T, sizeof(T)) == 0 &&
           std::memcmp(&g.eta, &eta, sizeof(eta)) == 0;
  });
  if (it == b->graphs.end()) {
    GraphEntry g;
    g.precision = precision;
    g.n_mu = n_mu;
    g.T = T;
    g.eta = eta;
    const int rc = entry_alloc(b, pl, g);
    if (rc != RTB_OK) {
      free_graph(g);
      return rc;
    }
    b->graphs.push_front(g);
    while (b->graphs.size() > kMaxGraphs) {
      free_graph(b->graphs.back());
      b->graphs.pop_back();
    }
  } else if (it != b->graphs.begin()) {
    b->graphs.splice(b->graphs.begin(), b->graphs, it);
  }
  GraphEntry& g = b->graphs.front();
  if (b->use_graph && !g.exec) RTB_TRY(capture(b, pl, st, g));
  RTB_HIP(hipMemcpyAsync(g.dmu, mu, sizeof(double) * n_mu, hipMemcpyHostToDevice, st));
  int nmax = 0;
  for (const rtb_model* m : b->models) nmax = std::max(nmax, m->nsta);
  rtb_stats s{};
  float ms = 0.f;
  if (b->use_graph) {
    hipEvent_t e0, e1;
    RTB_TRY(lease.event(0, &e0));
    RTB_TRY(lease.event(1, &e1));
    RTB_HIP(hipEventRecord(e0, st));
    RTB_HIP(hipGraphLaunch(g.exec, st));
    RTB_HIP(hipEventRecord(e1, st));
    RTB_HIP(hipStreamSynchronize(st));
    RTB_HIP(hipEventElapsedTime(&ms, e0, e1));
    s.nk = pl.nk_total;
    s.nsta = nmax;
    s.ms[RTB_STAGE_TOTAL] = ms;
    s.launches[RTB_STAGE_TOTAL] = static_cast<int>(g.nodes);
  } else {
    Timer tmr(lease);
    RTB_TRY(record(b, pl, st, g, tmr));
    RTB_HIP(hipStreamSynchronize(st));
    RTB_TRY(tmr.gather(pl.nk_total, nmax, &s));
    ms = static_cast<float>(s.ms[RTB_STAGE_TOTAL]);
  }
  size_t resolved = 0;
  if (pl.f32) RTB_TRY(resolve(b, pl, g, st, lease, mu, T, &resolved));
  const size_t nm = b->models.size();
  std::vector<double> sums(nm * n_mu);
  int nf = 0;
  unsigned long long mr = 0;
  RTB_HIP(hipMemcpyAsync(sums.data(), g.dsig, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, st));
  RTB_HIP(hipMemcpyAsync(&nf, pl.d_nonfinite, sizeof(int), hipMemcpyDeviceToHost, st));
  RTB_HIP(hipMemcpyAsync(&mr, pl.d_maxres, sizeof(mr), hipMemcpyDeviceToHost, st));
  RTB_HIP(hipStreamSynchronize(st));
  if (nf > 0) {
    set_error("EigenvalueComputationFailed: the eigenvalues of %d k-point(s) are not finite", nf);
    return RTB_ERR_EIGEN_FAILED;
  }
  // sum / nk / det per model (response/linear/mod.rs:221)
  for (size_t j = 0; j < nm; ++j) {
    const int mi = pl.order[j];
    for (size_t q = 0; q < n_mu; ++q)
      sigma[static_cast<size_t>(mi) * n_mu + q] =
          sums[j * n_mu + q] / static_cast<double>(b->nk[mi]) / b->models[mi]->det_lat;
  }
  double r;
  std::memcpy(&r, &mr, sizeof(r));
  s.max_residual = r;
  s.resolved = resolved;
  publish_stats(s);
  b->last_precision = precision;
  b->last_nodes = b->use_graph ? g.nodes : 0;
  b->last_graph_ms = ms;
  return RTB_OK;
}

int rtb_batch_set_graph(rtb_batch* b, int use_graph) {
  if (!b) return fail(RTB_ERR_INVALID_ARGUMENT, "null batch");
  std::lock_guard<std::mutex> lock(b->mu);
  b->use_graph = use_graph != 0;
  return RTB_OK;
}

int rtb_batch_info_get(rtb_batch* b, rtb_batch_info* out) {
  if (!b || !out) return fail(RTB_ERR_INVALID_ARGUMENT, "null batch or output");
  std::lock_guard<std::mutex> lock(b->mu);
  std::memset(out, 0, sizeof(*out));
  const Plan& pl = b->plans[b->last_precision == RTB_PREC_MIXED ? 1 : 0];
  out->n_models = static_cast<int>(b->models.size());
  out->n_passes = static_cast<int>(pl.passes.size());
  out->n_portions = pl.n_portions;
  out->n_tiles = static_cast<int>(pl.tiles.size());
  for (const Pass& ps : pl.passes) out->n_queues += static_cast<int>(ps.queues.size());
  for (size_t i = 0; i < pl.mesh.size(); ++i) out->n_mesh_models += pl.mesh[i] ? 1 : 0;
  out->graphs_cached = static_cast<int>(b->graphs.size());
  out->captures = b->captures;
  out->graph_nodes = static_cast<int>(b->last_nodes);
  size_t nk = 0;
  for (size_t v : b->nk) nk += v;
  out->nk_total = nk;
  out->arena_bytes = b->plans[0].arena_bytes + b->plans[1].arena_bytes;
  out->graph_ms = b->last_graph_ms;
  return RTB_OK;
}

void rtb_batch_destroy(rtb_batch* b) {
  if (!b) return;
  {
    std::lock_guard<std::mutex> lock(b->mu);
    if (hipSetDevice(b->device) == hipSuccess) {
      drop_graphs(b, -1);
      for (Plan& pl : b->plans)
        if (pl.arena) (void)hipFree(pl.arena);
    }
  }
  delete b;
}

}  // extern "C"
