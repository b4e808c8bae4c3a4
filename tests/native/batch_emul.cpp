// Host emulation of the multi-model batch (rustb_amd/csrc/batch.cpp).
//
// batch.cpp and workspace.cpp are compiled unchanged with the host compiler
// and linked against this file, which stands in for the HIP runtime and for
// the device stages: device memory is host memory (every allocation is
// registered, and every buffer a stage receives is checked to lie inside one),
// graph capture records the stages and a graph launch replays them.  The
// work-queue Fourier launch is emulated tile by tile with the kernel's own
// thread -> (element, k-point) mapping; every element it writes must be
// written once.  The other stages compute stand-ins that depend on k:
// E_j(k) = Re H_jj(k), Omega_j(k) = Re V_jj(k) (first velocity matrix, or
// V0 + i V1 combined).  Consumed buffers are poisoned with NaN, so a
// k-point the plan fails to produce in a later pass shows up as a non-finite
// eigenvalue.  The conductivities of every model are compared with the same
// stand-ins summed over the model's gen_kmesh points; the mixed form flags
// some k-points so that the FP64 re-solve path (k-point recovery on both
// Fourier routes) runs too.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "../../rustb_amd/csrc/pipeline.hpp"

namespace {

int g_errors = 0;
void report(const char* fmt, const char* what) {
  if (g_errors < 20) std::fprintf(stderr, fmt, what);
  ++g_errors;
}

std::map<uintptr_t, size_t> g_allocs;  // base -> bytes
bool g_capturing = false;
std::vector<std::function<void()>> g_recorded;

bool inside(const void* p, size_t bytes) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  auto it = g_allocs.upper_bound(a);
  if (it == g_allocs.begin()) return false;
  --it;
  return a >= it->first && a + bytes <= it->first + it->second;
}

void need(const void* p, size_t bytes, const char* what) {
  if (bytes == 0) return;
  if (!p || !inside(p, bytes)) report("buffer outside every allocation: %s\n", what);
}

void submit(std::function<void()> f) {
  if (g_capturing)
    g_recorded.push_back(std::move(f));
  else
    f();
}

void poison(double* p, size_t count) {
  for (size_t i = 0; i < count; ++i) p[i] = std::nan("");
}

struct FakeGraph {
  std::vector<std::function<void()>> ops;
};

}  // namespace

// ---------------------------------------------------------------- HIP runtime
extern "C" {
hipError_t hipMalloc(void** p, size_t n) {
  const size_t sz = (n + 255) / 256 * 256 + 256;
  void* q = std::aligned_alloc(256, sz);
  if (!q) return hipErrorOutOfMemory;
  poison(static_cast<double*>(q), sz / sizeof(double));
  g_allocs[reinterpret_cast<uintptr_t>(q)] = n;
  *p = q;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  if (!g_allocs.erase(reinterpret_cast<uintptr_t>(p))) report("hipFree of %s\n", "an unknown pointer");
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  if (g_capturing) report("%s during capture\n", "hipMemcpy");
  std::memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  if (g_capturing) report("%s during capture\n", "hipMemcpyAsync");
  std::memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) {
  if (g_capturing) report("%s during capture\n", "hipStreamSynchronize");
  return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) {
  if (g_capturing) report("%s during capture\n", "hipDeviceSynchronize");
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
  *v = 8;  // compute units
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emulated"; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  static int dummy[16];
  static int next = 0;
  *s = reinterpret_cast<hipStream_t>(&dummy[next++ % 16]);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  static int dummy[256];
  static int next = 0;
  *e = reinterpret_cast<hipEvent_t>(&dummy[next++ % 256]);
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) {
  if (g_capturing) report("%s during capture\n", "hipEventRecord");
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 1.0f;
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  if (g_capturing) report("%s\n", "nested capture");
  g_capturing = true;
  g_recorded.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  g_capturing = false;
  auto* fg = new FakeGraph{std::move(g_recorded)};
  g_recorded.clear();
  *g = reinterpret_cast<hipGraph_t>(fg);
  return hipSuccess;
}
hipError_t hipGraphGetNodes(hipGraph_t g, hipGraphNode_t*, size_t* n) {
  *n = reinterpret_cast<FakeGraph*>(g)->ops.size();
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *e = reinterpret_cast<hipGraphExec_t>(new FakeGraph(*reinterpret_cast<FakeGraph*>(g)));
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete reinterpret_cast<FakeGraph*>(g);
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete reinterpret_cast<FakeGraph*>(e);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& f : reinterpret_cast<FakeGraph*>(e)->ops) f();
  return hipSuccess;
}
}  // extern "C"

// ------------------------------------------------------- the model stand-in
namespace {

// H_mc, V0_mc, V1_mc at fractional k as the dense Fourier kernel forms them
// (stored R order, phases of k.R reduced mod 1, Atom gauge, tau term).
void bloch(const rtb::FourierArgs& a, const double* k, int m, int c, double out[6]) {
  const int n2 = a.n * a.n;
  const double2* ham = reinterpret_cast<const double2*>(a.ham);
  auto phase = [](double arg, double& co, double& si) {
    const double t = arg - std::rint(arg);
    co = std::cos(2.0 * M_PI * t);
    si = std::sin(2.0 * M_PI * t);
  };
  double hr = 0, hi = 0, vr[2] = {0, 0}, vi[2] = {0, 0};
  for (int r = 0; r < a.nR; ++r) {
    const double* Rv = a.R + 3 * r;
    double ur, ui;
    phase(Rv[0] * k[0] + Rv[1] * k[1] + Rv[2] * k[2], ur, ui);
    const double2 h = ham[static_cast<size_t>(r) * n2 + m * a.n + c];
    const double tr = ur * h.x - ui * h.y, ti = ur * h.y + ui * h.x;
    hr += tr;
    hi += ti;
    for (int p = 0; p < 2; ++p) {
      vr[p] += a.wR[p * a.nR + r] * tr;
      vi[p] += a.wR[p * a.nR + r] * ti;
    }
  }
  const double* tm = a.tau + 3 * m;
  const double* tc = a.tau + 3 * c;
  double cm, sm, cc, sc;
  phase(tm[0] * k[0] + tm[1] * k[1] + tm[2] * k[2], cm, sm);
  phase(tc[0] * k[0] + tc[1] * k[1] + tc[2] * k[2], cc, sc);
  const double gr = cm * cc + sm * sc, gi = cm * sc - sm * cc;
  out[0] = hr * gr - hi * gi;
  out[1] = hr * gi + hi * gr;
  for (int p = 0; p < 2; ++p) {
    double xr = -vi[p], xi = vr[p];
    const double diff = a.wtau[p * a.n + c] - a.wtau[p * a.n + m];
    xr += -diff * hi;
    xi += diff * hr;
    out[2 + 2 * p] = xr * gr - xi * gi;
    out[3 + 2 * p] = xr * gi + xi * gr;
  }
}

// writes element (m, c) of H(k) and V(k) (combined or two matrices)
void store(const rtb::FourierArgs& a, const double* k, int m, int c, double* H, double* V) {
  double o[6];
  bloch(a, k, m, c, o);
  const size_t e = static_cast<size_t>(m) * a.n + c, n2 = static_cast<size_t>(a.n) * a.n;
  H[2 * e] = o[0];
  H[2 * e + 1] = o[1];
  if (a.combine) {
    V[2 * e] = o[2] - o[5];
    V[2 * e + 1] = o[3] + o[4];
  } else {
    V[2 * e] = o[2];
    V[2 * e + 1] = o[3];
    V[2 * (n2 + e)] = o[4];
    V[2 * (n2 + e) + 1] = o[5];
  }
}

// stand-ins: E_j = Re H_jj, Omega_j = Re V_jj (first matrix)
void standins(int n, const double* H, const double* V, double* E, double* Om) {
  for (int j = 0; j < n; ++j) {
    E[j] = H[2 * (static_cast<size_t>(j) * n + j)];
    Om[j] = V[2 * (static_cast<size_t>(j) * n + j)];
  }
}

double fermi(double e, double mu, double width) {
  if (width == 0.0) return e < mu ? 1.0 : (e > mu ? 0.0 : 0.5);
  const double x = (e - mu) / width;
  if (x > 50.0) return 0.0;
  if (x < -50.0) return 1.0;
  return 1.0 / (1.0 + std::exp(x));
}

std::vector<double> g_mixed_flags_seen;  // global indices flagged by the stand-in check
int g_flag_every = 0;

}  // namespace

// --------------------------------------------------------------- rtb stages
namespace rtb {

thread_local std::string g_err;
rtb_stats g_last{};
void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
void publish_stats(const rtb_stats& s) { g_last = s; }

int prepare(const rtb_model*) { return RTB_OK; }
int check_kmesh(const rtb_model* m, const size_t* kmesh, size_t* nk) {
  size_t t = 1;
  for (int d = 0; d < m->dim; ++d) {
    if (kmesh[d] == 0) return fail(RTB_ERR_INVALID_KMESH, "kmesh");
    t *= kmesh[d];
  }
  *nk = t;
  return RTB_OK;
}
int check_direction(const rtb_model*, const double*, int) { return RTB_OK; }
int check_eta(double eta) { return eta >= 0.0 ? RTB_OK : fail(RTB_ERR_INVALID_RESPONSE, "eta"); }
int check_mu(const double* mu, size_t n) { return mu && n ? RTB_OK : fail(RTB_ERR_INVALID_RESPONSE, "mu"); }
int check_T(double T) { return T >= 0.0 ? RTB_OK : fail(RTB_ERR_INVALID_RESPONSE, "T"); }
int check_precision(int p) { return p == 0 || p == 1 ? RTB_OK : fail(RTB_ERR_INVALID_ARGUMENT, "precision"); }

int upload_dirs(hipStream_t, Work& w, const rtb_model* m, const double* dirs, int np, FourierArgs& a) {
  a.dim = m->dim;
  a.n = m->nsta;
  a.nR = m->nR;
  a.np = np;
  a.ham = m->d_ham;
  a.R = m->d_R;
  a.tau = m->d_tau;
  double *wR, *wt;
  RTB_TRY(w.alloc(&wR, static_cast<size_t>(np) * m->nR));
  RTB_TRY(w.alloc(&wt, static_cast<size_t>(np) * m->nsta));
  need(wR, sizeof(double) * np * m->nR, "w.R table");
  for (int p = 0; p < np; ++p) {
    for (int r = 0; r < m->nR; ++r) {
      double s = 0.0;
      for (int d = 0; d < m->dim; ++d) {
        double rc = 0.0;
        for (int i = 0; i < m->dim; ++i) rc += static_cast<double>(m->hamR[r * m->dim + i]) * m->lat[i * m->dim + d];
        s += dirs[p * m->dim + d] * rc;
      }
      wR[p * m->nR + r] = s;
    }
    for (int o = 0; o < m->nsta; ++o) {
      double s = 0.0;
      for (int d = 0; d < m->dim; ++d) {
        double tc = 0.0;
        for (int i = 0; i < m->dim; ++i) tc += m->orb[(o % m->norb) * m->dim + i] * m->lat[i * m->dim + d];
        s += dirs[p * m->dim + d] * tc;
      }
      wt[p * m->nsta + o] = s;
    }
  }
  a.wR = wR;
  a.wtau = wt;
  return RTB_OK;
}

size_t workspace_budget() {
  if (const char* env = std::getenv("RTB_WORKSPACE_MB")) return static_cast<size_t>(std::atoll(env)) << 20;
  if (const char* env = std::getenv("EMUL_BUDGET_KB")) return static_cast<size_t>(std::atoll(env)) << 10;
  return size_t(1) << 30;
}

size_t eigh_band_elems(int n) { return n > kMaxEighN ? static_cast<size_t>(n) * 40 : 0; }
size_t eig_extra_bytes(int n, bool vectors) {
  if (n <= kMaxEighN) return 0;
  return eigh_band_elems(n) * 8 + (vectors ? 0 : static_cast<size_t>(n) * n * 16);
}
int alloc_eig(Work& w, int n, size_t K, bool vectors, EigBufs& b) {
  const size_t n2 = static_cast<size_t>(n) * n;
  b.Vg = b.tau = b.band = nullptr;
  RTB_TRY(w.alloc(&b.DE, 3 * K * n));
  if (vectors || n > kMaxEighN) RTB_TRY(w.alloc(&b.Vg, K * n2 * 2));
  if (n > kMaxEighN) RTB_TRY(w.alloc(&b.band, K * eigh_band_elems(n)));
  if (vectors) {
    RTB_TRY(w.alloc(&b.tau, K * n * 2));
    RTB_TRY(w.alloc(&b.Y, K * n2));
    RTB_TRY(w.alloc(&b.S, K * n2 * 3));
  } else {
    b.Y = b.S = nullptr;
  }
  return RTB_OK;
}
MixedOpts mixed_opts() { return MixedOpts{}; }

bool eigh_block_valid(int n, int nb) {
  if (nb == 0) return true;
  return n <= kMaxEighN && (nb == 32 || nb == 64 || nb == 128) && nb >= n;
}

void check_eig(int n, size_t c, double* H, double* E, const EigBufs& eb) {
  const size_t n2 = static_cast<size_t>(n) * n;
  need(H, c * n2 * 16, "H");
  need(E, c * n * 8, "E");
  need(eb.DE, 3 * c * n * 8, "DE");
  need(eb.Vg, c * n2 * 16, "Vg");
  need(eb.tau, c * n * 16, "tau");
  need(eb.Y, c * n2 * 8, "Y");
  need(eb.S, c * n2 * 24, "S");
  if (n > kMaxEighN) need(eb.band, c * eigh_band_elems(n) * 8, "band");
}

// the stand-in eigenvalues from H (which the stage consumes)
void eig_standin(int n, size_t c, double* H, double* E) {
  const size_t n2 = static_cast<size_t>(n) * n;
  for (size_t k = 0; k < c; ++k)
    for (int j = 0; j < n; ++j) E[k * n + j] = H[2 * (k * n2 + static_cast<size_t>(j) * n + j)];
  poison(H, c * n2 * 2);
}

int eigh_launch_fp64(hipStream_t, Timer*, int n, size_t c, double* H, double* E, bool vectors, double* Vg,
                     double* tau, double* Y, double* S, double* DE, bool, double* band, int nb) {
  if (!vectors || !eigh_block_valid(n, nb)) report("%s\n", "eigh_launch_fp64 arguments");
  EigBufs eb{Vg, tau, Y, S, DE, band};
  submit([=] {
    check_eig(n, c, H, E, eb);
    eig_standin(n, c, H, E);
  });
  return RTB_OK;
}
int eigh_backtransform_fp64(hipStream_t, int n, size_t c, const double*, const double*, const double*,
                            const double*, const double*, double* Z, bool, const double*, int) {
  submit([=] { need(Z, c * n * n * 16, "Z"); });
  return RTB_OK;
}

int eig_mixed_chunk(hipStream_t, Timer&, bool, int n, size_t c, double* H, double* E, EigBufs& eb, double* Z,
                    double* W1, double* W2, double* W3, double* scale, const RefineCheck& chk, int nb) {
  if (!eigh_block_valid(n, nb)) report("%s\n", "eig_mixed_chunk block");
  const EigBufs b = eb;
  const RefineCheck ck = chk;
  submit([=] {
    const size_t n2 = static_cast<size_t>(n) * n;
    check_eig(n, c, H, E, b);
    need(Z, c * n2 * 16, "Z");
    need(W1, c * n2 * 16, "W1");
    need(W2, c * n2 * 16, "W2");
    need(W3, c * n2 * 16, "W3");
    need(scale, c * n * 8, "scale");
    need(ck.fail_count, 4, "fail count");
    need(ck.fail_list, c * 4, "fail list");
    need(ck.max_res, 8, "max residual");
    eig_standin(n, c, H, E);
    for (size_t i = 0; i < c; ++i) scale[i * n] = 1.0;
    // flag every g_flag_every-th k-point of the chunk (by local index)
    if (g_flag_every > 0)
      for (size_t i = 0; i < c; ++i)
        if (i % g_flag_every == 1) ck.fail_list[(*ck.fail_count)++] = static_cast<int>(i);
    const double r = 1e-12;
    std::memcpy(ck.max_res, &r, 8);
  });
  return RTB_OK;
}

int kubo_chunk(hipStream_t, Timer&, const rtb_model* m, size_t K, const double* V, const double* Z,
               const double* zscale, const double* E, int spin_dir, double, double* T1, double* C, double* J,
               double* omega, double* B, double* metric) {
  const int n = m->nsta;
  if (spin_dir != RTB_SPIN_NONE || J || metric) report("%s\n", "kubo_chunk arguments");
  submit([=] {
    const size_t n2 = static_cast<size_t>(n) * n;
    const size_t vper = B ? 2 : 1;
    need(V, K * n2 * 16 * vper, "V");
    need(Z, K * n2 * 16, "Z (Kubo)");
    if (zscale) need(zscale, K * n * 8, "zscale");
    need(E, K * n * 8, "E (Kubo)");
    need(T1, K * n2 * 16, "T1");
    need(C, K * n2 * 16, "C");
    need(omega, K * n * 8, "omega");
    if (B) need(B, K * n2 * 16, "B");
    for (size_t k = 0; k < K; ++k)
      for (int j = 0; j < n; ++j) omega[k * n + j] = V[2 * (k * vper * n2 + static_cast<size_t>(j) * n + j)];
    poison(const_cast<double*>(V), K * n2 * 2 * vper);
  });
  return RTB_OK;
}

int berry_points_fp64(const rtb_model* m, const double* k, size_t nk, const double* dir, double, double* omega,
                      double* E, hipStream_t, Work& w) {
  FourierArgs a;
  RTB_TRY(upload_dirs(nullptr, w, m, dir, 2, a));
  a.gauge = RTB_GAUGE_ATOM;
  a.combine = 0;
  const int n = m->nsta;
  std::vector<double> H(static_cast<size_t>(n) * n * 2), V(static_cast<size_t>(n) * n * 4);
  for (size_t i = 0; i < nk; ++i) {
    double kv[3] = {0, 0, 0};
    for (int d = 0; d < m->dim; ++d) kv[d] = k[i * m->dim + d];
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) store(a, kv, r, c, H.data(), V.data());
    standins(n, H.data(), V.data(), E + i * n, omega + i * n);
  }
  return RTB_OK;
}

int hermitize_upper(hipStream_t, int, size_t, double*) { return RTB_OK; }
int count_nonfinite(hipStream_t, int n, size_t c, const double* E, int* count) {
  submit([=] {
    for (size_t i = 0; i < c; ++i) {
      bool bad = false;
      for (int j = 0; j < n; ++j) bad = bad || !std::isfinite(E[i * n + j]);
      *count += bad ? 1 : 0;
    }
  });
  return RTB_OK;
}
int zero_ints(hipStream_t, int* p, size_t count) {
  submit([=] {
    need(p, count * 4, "zero_ints");
    std::memset(p, 0, count * 4);
  });
  return RTB_OK;
}
int kterms(hipStream_t, int n, size_t nk, const double* omega, const double* E, const double* mu, int n_mu,
           double T, double* out) {
  submit([=] {
    need(out, nk * n_mu * 8, "per-k terms");
    const double width = T <= 0.0 ? 0.0 : T * 8.617333262145e-5;
    for (size_t k = 0; k < nk; ++k)
      for (int q = 0; q < n_mu; ++q) {
        double s = 0.0;
        for (int b = 0; b < n; ++b) s += omega[k * n + b] * fermi(E[k * n + b], mu[q], width);
        out[k * n_mu + q] = s;
      }
    poison(const_cast<double*>(E), nk * n);
    poison(const_cast<double*>(omega), nk * n);
  });
  return RTB_OK;
}
int segment_sums(hipStream_t, const double* term, const long long* start, int n_seg, int n_mu, double* out) {
  submit([=] {
    for (int s = 0; s < n_seg; ++s)
      for (int q = 0; q < n_mu; ++q) {
        double acc = 0.0;
        for (long long k = start[s]; k < start[s + 1]; ++k) acc += term[k * n_mu + q];
        out[s * n_mu + q] = acc;
      }
  });
  return RTB_OK;
}

// the queue kernel's thread mapping, one workgroup at a time
int fourier_queue_groups(int n, int tk) {
  const int n2 = n * n;
  return n2 <= 256 ? std::min(256 / n2, kQueueSlots / tk) : 1;
}
int fourier_queue_occupancy(int, int, int* per_cu) {
  *per_cu = 2;
  return RTB_OK;
}
int fourier_queue(hipStream_t, int tk, const QueueModel* models, const QueueTile* tiles, int ntiles, int max_groups,
                  unsigned* next, int workgroups) {
  if (tk < 1 || max_groups < 1 || max_groups * tk > kQueueSlots || workgroups < 1) report("%s\n", "queue shape");
  submit([=] {
    need(tiles, sizeof(QueueTile) * ntiles, "tile table");
    need(next, 4, "queue counter");
    if (*next != 0) report("%s\n", "queue counter not cleared");
    std::set<const double*> written;
    for (unsigned t = (*next)++; static_cast<int>(t) < ntiles; t = (*next)++) {
      const QueueTile tile = tiles[t];
      const QueueModel& md = models[tile.model];
      const FourierArgs& a = md.a;
      const int n2 = a.n * a.n;
      const int EB = n2 <= 256 ? n2 : 256;
      const int G = n2 <= 256 ? std::min(256 / n2, kQueueSlots / tk) : 1;
      if (G > max_groups) report("%s\n", "tile groups above the launch's max_groups");
      if (tile.cnt < 1 || tile.cnt > G * tk) report("%s\n", "tile k-point count");
      const size_t vper = a.combine ? 1 : 2;
      need(tile.H, sizeof(double) * 2 * n2 * tile.cnt, "tile H");
      need(tile.V, sizeof(double) * 2 * n2 * vper * tile.cnt, "tile V");
      for (int tid = 0; tid < 256; ++tid) {
        const int grp = tid / EB;
        const int e = tile.e0 + tid % EB;
        if (!(grp < G && e < n2)) continue;
        for (int kk = 0; kk < tk; ++kk) {
          const int kq = grp * tk + kk;
          if (kq >= tile.cnt) break;
          long long f = tile.f0 + kq;
          long long x[3] = {0, 0, 0};
          for (int d = a.dim - 1; d >= 0; --d) {
            x[d] = f % md.N[d];
            f /= md.N[d];
          }
          double kv[3] = {0, 0, 0};
          for (int d = 0; d < a.dim; ++d) kv[d] = static_cast<double>(x[d]) / static_cast<double>(md.N[d]);
          double* H = tile.H + static_cast<size_t>(kq) * n2 * 2;
          double* V = tile.V + static_cast<size_t>(kq) * n2 * 2 * vper;
          const double* addr = H + 2 * e;
          if (!written.insert(addr).second) report("%s\n", "an element written twice by one launch");
          store(a, kv, e / a.n, e % a.n, H, V);
        }
      }
    }
  });
  return RTB_OK;
}

// the separable route: its plan (as mesh_plan) and the final pass's output
// order (k1 innermost within (k2, k3) pairs)
int mesh_plan(const rtb_model* m, const size_t* kmesh, MeshPlan& p) {
  p.dim = m->dim;
  for (int d = 0; d < 3; ++d) {
    p.lo[d] = 0;
    p.L[d] = 1;
    p.N[d] = d < m->dim ? static_cast<int>(kmesh[d]) : 1;
  }
  for (int d = 0; d < m->dim; ++d) {
    long long lo = 0, hi = 0;
    for (int r = 0; r < m->nR; ++r) {
      lo = std::min<long long>(lo, m->hamR[r * m->dim + d]);
      hi = std::max<long long>(hi, m->hamR[r * m->dim + d]);
    }
    p.lo[d] = static_cast<int>(lo);
    p.L[d] = static_cast<int>(hi - lo + 1);
  }
  const size_t n2 = static_cast<size_t>(m->nsta) * m->nsta;
  p.box = static_cast<size_t>(p.L[0]) * p.L[1] * p.L[2];
  p.P = static_cast<size_t>(p.N[1]) * p.N[2];
  p.box_elems = p.box * n2;
  p.ga_elems = p.dim == 3 ? 2 * static_cast<size_t>(p.L[0]) * p.L[1] * p.N[2] * n2 : 0;
  p.gb_elems = static_cast<size_t>(p.dim) * p.L[0] * p.P * n2;
  p.ok = p.box <= 4 * static_cast<size_t>(m->nR) + 64;
  return RTB_OK;
}
void mesh_box_index(const rtb_model* m, const MeshPlan&, std::vector<int>& idx) { idx.assign(m->nR, 0); }
int mesh_prepare_dev(hipStream_t, const rtb_model* m, const MeshPlan& p, const int* bidx, double* work, double* G,
                     bool) {
  submit([=] {
    need(bidx, sizeof(int) * m->nR, "box index");
    need(work, sizeof(double) * 2 * (p.box_elems + p.ga_elems), "box work");
    need(G, sizeof(double) * 2 * p.gb_elems, "G");
  });
  return RTB_OK;
}
int mesh_final(hipStream_t, const rtb_model* m, const MeshPlan& p, const double* G, const FourierArgs& a,
               const double*, size_t p0, size_t np_pairs, double* H, double* V) {
  const FourierArgs fa = a;
  submit([=] {
    need(G, sizeof(double) * 2 * p.gb_elems, "G (final)");
    const int n = m->nsta;
    const size_t n2 = static_cast<size_t>(n) * n, vper = fa.combine ? 1 : 2;
    need(H, sizeof(double) * 2 * n2 * np_pairs * p.N[0], "H (final)");
    need(V, sizeof(double) * 2 * n2 * vper * np_pairs * p.N[0], "V (final)");
    if (p0 + np_pairs > p.P) report("%s\n", "pairs beyond the mesh");
    for (size_t q = 0; q < np_pairs; ++q)
      for (int i1 = 0; i1 < p.N[0]; ++i1) {
        const size_t pg = p0 + q, i2 = pg / p.N[2], i3 = pg % p.N[2];
        double kv[3] = {static_cast<double>(i1) / p.N[0], m->dim >= 2 ? static_cast<double>(i2) / p.N[1] : 0.0,
                        m->dim >= 3 ? static_cast<double>(i3) / p.N[2] : 0.0};
        const size_t kl = q * p.N[0] + i1;
        for (int r = 0; r < n; ++r)
          for (int c = 0; c < n; ++c) store(fa, kv, r, c, H + kl * n2 * 2, V + kl * n2 * 2 * vper);
      }
  });
  return RTB_OK;
}

}  // namespace rtb

// ----------------------------------------------------------------- the test
namespace {

struct Spec {
  int dim, norb, spin, nRhalf, rmax;
  size_t N[3];
};

rtb_model* make_model(const Spec& s, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> g(0.0, 1.0);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  auto* m = new rtb_model();
  m->dim = s.dim;
  m->spin = s.spin;
  m->norb = s.norb;
  m->nsta = s.spin ? 2 * s.norb : s.norb;
  const int n = m->nsta;
  // R = 0 and nRhalf pairs +-R inside the box |R_i| <= rmax (distinct)
  std::vector<std::vector<long long>> Rs = {std::vector<long long>(s.dim, 0)};
  std::set<std::vector<long long>> seen = {Rs[0]};
  while (static_cast<int>(Rs.size()) < 1 + 2 * s.nRhalf) {
    std::vector<long long> R(s.dim), Rm(s.dim);
    for (int d = 0; d < s.dim; ++d) {
      R[d] = static_cast<long long>(u(rng) * (2 * s.rmax + 1)) - s.rmax;
      Rm[d] = -R[d];
    }
    if (seen.count(R)) continue;
    seen.insert(R);
    seen.insert(Rm);
    Rs.push_back(R);
    Rs.push_back(Rm);
  }
  m->nR = static_cast<int>(Rs.size());
  for (auto& R : Rs)
    for (long long v : R) m->hamR.push_back(v);
  m->lat.assign(s.dim * s.dim, 0.0);
  for (int i = 0; i < s.dim; ++i)
    for (int j = 0; j < s.dim; ++j) m->lat[i * s.dim + j] = (i == j ? 1.0 : 0.0) + 0.1 * g(rng);
  m->det_lat = s.dim == 1 ? m->lat[0]
               : s.dim == 2 ? m->lat[0] * m->lat[3] - m->lat[1] * m->lat[2]
                            : m->lat[0] * (m->lat[4] * m->lat[8] - m->lat[5] * m->lat[7]) -
                                  m->lat[1] * (m->lat[3] * m->lat[8] - m->lat[5] * m->lat[6]) +
                                  m->lat[2] * (m->lat[3] * m->lat[7] - m->lat[4] * m->lat[6]);
  for (int i = 0; i < s.norb * s.dim; ++i) m->orb.push_back(u(rng));
  const size_t n2 = static_cast<size_t>(n) * n;
  (void)hipMalloc(reinterpret_cast<void**>(&m->d_ham), sizeof(double) * 2 * n2 * m->nR);
  (void)hipMalloc(reinterpret_cast<void**>(&m->d_R), sizeof(double) * 3 * m->nR);
  (void)hipMalloc(reinterpret_cast<void**>(&m->d_tau), sizeof(double) * 3 * n);
  double2* h = reinterpret_cast<double2*>(m->d_ham);
  for (int r = 0; r < m->nR; ++r)
    for (int d = 0; d < 3; ++d) m->d_R[3 * r + d] = d < s.dim ? static_cast<double>(Rs[r][d]) : 0.0;
  for (int o = 0; o < n; ++o)
    for (int d = 0; d < 3; ++d) m->d_tau[3 * o + d] = d < s.dim ? m->orb[(o % s.norb) * s.dim + d] : 0.0;
  // R = 0 Hermitian; pairs h_{-R} = h_R^H
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      const double re = g(rng), im = i == j ? 0.0 : g(rng);
      h[static_cast<size_t>(i) * n + j] = make_double2(re, im);
      h[static_cast<size_t>(j) * n + i] = make_double2(re, -im);
    }
  for (int r = 1; r < m->nR; r += 2)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double re = 0.5 * g(rng), im = 0.5 * g(rng);
        h[(r * n2) + static_cast<size_t>(i) * n + j] = make_double2(re, im);
        h[((r + 1) * n2) + static_cast<size_t>(j) * n + i] = make_double2(re, -im);
      }
  m->herm = true;
  return m;
}

// the stand-in conductivity of one model summed in gen_kmesh order
std::vector<double> reference(const rtb_model* m, const size_t* N, const double* dirs, bool combine,
                              const std::vector<double>& mu) {
  rtb::FourierArgs a;
  void* buf;
  (void)hipMalloc(&buf, size_t(16) << 20);
  rtb::Work w(buf, size_t(16) << 20);
  rtb::upload_dirs(nullptr, w, m, dirs, 2, a);
  a.gauge = RTB_GAUGE_ATOM;
  a.combine = combine ? 1 : 0;
  const int n = m->nsta;
  size_t nk = 1;
  for (int d = 0; d < m->dim; ++d) nk *= N[d];
  std::vector<double> H(static_cast<size_t>(n) * n * 2), V(static_cast<size_t>(n) * n * 4), E(n), Om(n);
  std::vector<double> out(mu.size(), 0.0);
  for (size_t f = 0; f < nk; ++f) {
    size_t x[3] = {0, 0, 0}, r = f;
    for (int d = m->dim - 1; d >= 0; --d) {
      x[d] = r % N[d];
      r /= N[d];
    }
    double kv[3] = {0, 0, 0};
    for (int d = 0; d < m->dim; ++d) kv[d] = static_cast<double>(x[d]) / static_cast<double>(N[d]);
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < n; ++c) store(a, kv, i, c, H.data(), V.data());
    standins(n, H.data(), V.data(), E.data(), Om.data());
    for (size_t q = 0; q < mu.size(); ++q)
      for (int b = 0; b < n; ++b) out[q] += Om[b] * fermi(E[b], mu[q], 0.0);
  }
  for (double& v : out) v = v / static_cast<double>(nk) / m->det_lat;
  (void)hipFree(buf);
  return out;
}

int run_case(const char* name, const std::vector<Spec>& specs, unsigned seed, int precision, const char* budget_kb,
             int flag_every, int use_graph, const int* variants /* per model: tk, nb */) {
  const int errors0 = g_errors;
  if (budget_kb)
    setenv("EMUL_BUDGET_KB", budget_kb, 1);
  else
    unsetenv("EMUL_BUDGET_KB");
  g_flag_every = flag_every;
  const int nm = static_cast<int>(specs.size());
  std::vector<rtb_model*> models;
  std::vector<size_t> kmesh(3 * nm, 1);
  std::vector<double> dirs(6 * nm, 0.0);
  std::vector<rtb_batch_variant> table;
  for (int i = 0; i < nm; ++i) {
    models.push_back(make_model(specs[i], seed + 17 * i));
    for (int d = 0; d < 3; ++d) kmesh[3 * i + d] = specs[i].N[d];
    const int dim = specs[i].dim;
    for (int r = 0; r < 2; ++r)
      for (int d = 0; d < dim; ++d) dirs[6 * i + 3 * r + d] = (d == r ? 1.0 : 0.0) + (dim == 1 ? 0.7 * r : 0.0);
    if (variants) {
      bool have = false;
      for (auto& t : table) have = have || t.nsta == models[i]->nsta;
      if (!have) table.push_back({models[i]->nsta, variants[2 * i], variants[2 * i + 1]});
    }
  }
  rtb_batch* b = nullptr;
  int rc = rtb_batch_create(const_cast<const rtb_model* const*>(models.data()), nm, kmesh.data(), dirs.data(),
                            table.empty() ? nullptr : table.data(), static_cast<int>(table.size()), &b);
  if (rc != RTB_OK) {
    std::fprintf(stderr, "%s: create failed %d %s\n", name, rc, rtb::g_err.c_str());
    return 1;
  }
  rtb_batch_set_graph(b, use_graph);
  const std::vector<double> mu = {-0.5, 0.0, 0.7};
  std::vector<double> sig(nm * mu.size()), sig2(nm * mu.size());
  for (int rep = 0; rep < 2; ++rep) {
    rc = rtb_batch_hall_direct(b, mu.data(), mu.size(), 0.0, 1e-3, precision, rep ? sig2.data() : sig.data());
    if (rc != RTB_OK) {
      std::fprintf(stderr, "%s: call failed %d %s\n", name, rc, rtb::g_err.c_str());
      return 1;
    }
  }
  if (std::memcmp(sig.data(), sig2.data(), sizeof(double) * sig.size()) != 0) report("%s: repeat differs\n", name);
  rtb_batch_info info;
  rtb_batch_info_get(b, &info);
  double worst = 0.0;
  for (int i = 0; i < nm; ++i) {
    std::vector<double> dv(2 * specs[i].dim);
    for (int r = 0; r < 2; ++r)
      for (int d = 0; d < specs[i].dim; ++d) dv[r * specs[i].dim + d] = dirs[6 * i + 3 * r + d];
    const auto ref = reference(models[i], specs[i].N, dv.data(), precision == RTB_PREC_MIXED, mu);
    double scale = 1e-300;
    for (double v : ref) scale = std::max(scale, std::fabs(v));
    for (size_t q = 0; q < mu.size(); ++q) worst = std::max(worst, std::fabs(sig[i * mu.size() + q] - ref[q]) / scale);
  }
  if (!(worst <= 1e-12)) report("%s: conductivities differ from the stand-in reference\n", name);
  const bool flagged = precision == RTB_PREC_MIXED && flag_every > 0;
  if (flagged && rtb::g_last.resolved == 0) report("%s: no k-point re-solved\n", name);
  std::printf("%-28s models %2d passes %4d portions %4d tiles %6d queues %3d mesh %2d nodes %5d nk %7zu  "
              "worst %.2e resolved %zu %s\n",
              name, info.n_models, info.n_passes, info.n_portions, info.n_tiles, info.n_queues, info.n_mesh_models,
              info.graph_nodes, info.nk_total, worst, rtb::g_last.resolved, g_errors == errors0 ? "ok" : "FAIL");
  rtb_batch_destroy(b);
  for (auto* m : models) {
    (void)hipFree(m->d_ham);
    (void)hipFree(m->d_R);
    (void)hipFree(m->d_tau);
    delete m;
  }
  return g_errors == errors0 ? 0 : 1;
}

}  // namespace

int main() {
  // sizes across the queue's shapes (n^2 <= 256 with several k-groups, above
  // with element blocks, n > 128), 1D/2D/3D, meshes smaller than the R extent
  const std::vector<Spec> small = {
      {2, 2, 0, 6, 2, {9, 7, 1}},   {2, 4, 0, 12, 3, {8, 8, 1}}, {2, 12, 0, 10, 2, {6, 5, 1}},
      {2, 20, 0, 15, 3, {7, 6, 1}}, {2, 4, 0, 20, 3, {3, 2, 1}}, {3, 8, 1, 13, 1, {3, 4, 2}},
      {1, 3, 1, 2, 2, {11, 1, 1}},  {3, 65, 1, 13, 1, {2, 2, 2}}, {2, 1, 0, 3, 1, {5, 4, 1}},
  };
  // 64 models of the workload's sizes (norb 4..96, R extent |R_i| <= 11) on
  // small meshes and with fewer R, so that the emulated sums stay short
  std::vector<Spec> wide;
  std::mt19937 r(4000);
  for (int i = 0; i < 64; ++i) {
    const int norb = 4 * (1 + static_cast<int>(r() % 24));
    const int half = 8 + static_cast<int>(r() % 24);
    const size_t nk = 2 + r() % 2;
    wide.push_back({2, norb, 0, half, 11, {nk, nk + 1, 1}});
  }
  int bad = 0;
  {  // switched off unless enabled
    rtb_batch* b = nullptr;
    const Spec s = small[0];
    rtb_model* m = make_model(s, 1);
    const double dirs[6] = {1, 0, 0, 0, 1, 0};
    const double mu = 0.0;
    double sig = 0.0;
    unsetenv("RTB_BATCH_ENABLE");
    if (rtb_batch_create(const_cast<const rtb_model* const*>(&m), 1, s.N, dirs, nullptr, 0, &b) != RTB_OK ||
        rtb_batch_hall_direct(b, &mu, 1, 0.0, 1e-3, RTB_PREC_FP64, &sig) != RTB_ERR_UNSUPPORTED)
      report("%s\n", "the batch ran without RTB_BATCH_ENABLE=1");
    rtb_batch_destroy(b);
    (void)hipFree(m->d_ham);
    (void)hipFree(m->d_R);
    (void)hipFree(m->d_tau);
    delete m;
    setenv("RTB_BATCH_ENABLE", "1", 1);
  }
  for (int precision : {RTB_PREC_FP64, RTB_PREC_MIXED}) {
    const char* pn = precision ? "mixed" : "fp64";
    char name[64];
    for (int graph : {1, 0}) {
      std::snprintf(name, sizeof(name), "small %s graph=%d", pn, graph);
      bad += run_case(name, small, 11, precision, nullptr, 0, graph, nullptr);
    }
    std::snprintf(name, sizeof(name), "small %s 64 KiB passes", pn);
    bad += run_case(name, small, 11, precision, "64", 0, 1, nullptr);
    std::snprintf(name, sizeof(name), "small %s 2 MiB passes", pn);
    bad += run_case(name, small, 11, precision, "2048", 3, 1, nullptr);
    const int var[] = {8, 32, 1, 64, 4, 0, 2, 128, 1, 0, 4, 32, 8, 64, 8, 0, 1, 32};
    std::snprintf(name, sizeof(name), "small %s variants", pn);
    bad += run_case(name, small, 11, precision, nullptr, precision ? 5 : 0, 1, var);
  }
  bad += run_case("wide64 fp64", wide, 4000, RTB_PREC_FP64, nullptr, 0, 1, nullptr);
  bad += run_case("wide64 mixed 8 MiB", wide, 4000, RTB_PREC_MIXED, "8192", 7, 1, nullptr);
  rtb_workspace_release();  // the slot the calls leased
  std::printf("%s (%d problems)\n", bad ? "FAILED" : "all cases ok", g_errors);
  return bad ? 1 : 0;
}
