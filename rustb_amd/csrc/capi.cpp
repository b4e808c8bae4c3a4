// C ABI entry points: validation (response/config.rs, model_physics.rs),
// k-chunking, device work buffers and the kernel pipeline of each call.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>

#include "eigh.hpp"
#include "pipeline.hpp"

namespace rtb {

namespace {

thread_local std::string g_error;
thread_local rtb_stats g_stats;

std::once_flag g_smem_once;
int g_smem_rc = RTB_OK;

}  // namespace

// The helpers below are shared with the multi-model batch (batch.cpp)
// through pipeline.hpp.
int prepare(const rtb_model* m) {
  if (!m) return fail(RTB_ERR_INVALID_ARGUMENT, "null model");
  RTB_HIP(hipSetDevice(m->device));
  std::call_once(g_smem_once, [] {
    g_smem_rc = mesh_set_smem_limits();
    if (g_smem_rc == RTB_OK) g_smem_rc = eigh_set_smem_limits();
    if (g_smem_rc == RTB_OK) g_smem_rc = energy_cut_track_smem_limits();
  });
  return g_smem_rc;
}

bool finite(double x) { return std::isfinite(x); }

// validate_k_mesh (response/config.rs:246-259), validate_kmesh (kpoints.rs:39-47)
int check_kmesh(const rtb_model* m, const size_t* kmesh, size_t* nk) {
  if (!kmesh) return fail(RTB_ERR_INVALID_ARGUMENT, "null kmesh");
  size_t t = 1;
  for (int d = 0; d < m->dim; ++d) {
    if (kmesh[d] == 0) return fail(RTB_ERR_INVALID_KMESH, "k-mesh entries must be at least 1");
    t *= kmesh[d];
  }
  *nk = t;
  return RTB_OK;
}

// validate_direction_matrix (response/config.rs:202-236), rank 2
int check_direction(const rtb_model* m, const double* dir, int rows = 2) {
  if (!dir) return fail(RTB_ERR_INVALID_ARGUMENT, "null direction");
  for (int p = 0; p < rows; ++p) {
    bool allzero = true;
    for (int d = 0; d < m->dim; ++d) {
      double x = dir[p * m->dim + d];
      if (!finite(x)) return fail(RTB_ERR_INVALID_RESPONSE, "direction: all components must be finite");
      allzero = allzero && x == 0.0;
    }
    if (allzero) return fail(RTB_ERR_INVALID_RESPONSE, "direction: each direction row must not be the zero vector");
  }
  return RTB_OK;
}

int check_eta(double eta) {  // validate_broadening (config.rs:265-273)
  if (!finite(eta) || eta < 0.0) return fail(RTB_ERR_INVALID_RESPONSE, "broadening: must be finite and non-negative");
  return RTB_OK;
}

int check_mu(const double* mu, size_t n_mu) {  // validate_chemical_potentials (config.rs:275-289)
  if (n_mu == 0 || !mu) return fail(RTB_ERR_INVALID_RESPONSE, "chemical_potentials: must contain at least one value");
  for (size_t i = 0; i < n_mu; ++i)
    if (!finite(mu[i])) return fail(RTB_ERR_INVALID_RESPONSE, "chemical_potentials: all values must be finite");
  return RTB_OK;
}

int check_T(double T) {  // validate_temperature (config.rs:183-200)
  if (!finite(T) || T < 0.0) return fail(RTB_ERR_INVALID_RESPONSE, "T: all values must be finite and non-negative");
  return RTB_OK;
}

int check_spin(const rtb_model* m, int spin_dir) {
  if (spin_dir < RTB_SPIN_NONE || spin_dir > RTB_SPIN_Z) return fail(RTB_ERR_INVALID_ARGUMENT, "spin_dir must be 0..3");
  if (!m->spin && spin_dir != RTB_SPIN_NONE)
    return fail(RTB_ERR_SPIN_NOT_ALLOWED, "spin current requested on a spinless model");
  return RTB_OK;
}

int check_precision(int precision) {
  if (precision != RTB_PREC_FP64 && precision != RTB_PREC_MIXED)
    return fail(RTB_ERR_INVALID_ARGUMENT, "precision must be RTB_PREC_FP64 or RTB_PREC_MIXED");
  return RTB_OK;
}

// Per-call constant tables: w_p . R_cart and w_p . tau_cart (velocity.rs:530-579)
int upload_dirs(hipStream_t st, Work& w, const rtb_model* m, const double* dirs, int np, FourierArgs& a) {
  a.dim = m->dim;
  a.n = m->nsta;
  a.nR = m->nR;
  a.np = np;
  a.ham = m->d_ham;
  a.R = m->d_R;
  a.tau = m->d_tau;
  a.wR = nullptr;
  a.wtau = nullptr;
  if (np == 0) return RTB_OK;
  std::vector<double> wR(static_cast<size_t>(np) * m->nR), wt(static_cast<size_t>(np) * m->nsta);
  const int dim = m->dim;
  for (int p = 0; p < np; ++p) {
    const double* w = dirs + p * dim;
    for (int r = 0; r < m->nR; ++r) {
      double s = 0.0;
      for (int d = 0; d < dim; ++d) {
        double rc = 0.0;  // R_cart[d] = sum_i R_i lat[i][d]
        for (int i = 0; i < dim; ++i) rc += static_cast<double>(m->hamR[r * dim + i]) * m->lat[i * dim + d];
        s += w[d] * rc;
      }
      wR[static_cast<size_t>(p) * m->nR + r] = s;
    }
    for (int o = 0; o < m->nsta; ++o) {
      const int orb = o % m->norb;
      double s = 0.0;
      for (int d = 0; d < dim; ++d) {
        double tc = 0.0;
        for (int i = 0; i < dim; ++i) tc += m->orb[orb * dim + i] * m->lat[i * dim + d];
        s += w[d] * tc;
      }
      wt[static_cast<size_t>(p) * m->nsta + o] = s;
    }
  }
  double *dwR, *dwt;
  RTB_TRY(w.alloc(&dwR, wR.size()));
  RTB_TRY(w.alloc(&dwt, wt.size()));
  RTB_HIP(hipMemcpyAsync(dwR, wR.data(), wR.size() * sizeof(double), hipMemcpyHostToDevice, st));
  RTB_HIP(hipMemcpyAsync(dwt, wt.data(), wt.size() * sizeof(double), hipMemcpyHostToDevice, st));
  a.wR = dwR;
  a.wtau = dwt;
  return RTB_OK;
}

// k rows of `dim` doubles -> device [nk][3] (zero padded)
int upload_k(hipStream_t st, Work& w, int dim, const double* k, size_t nk, double** dk) {
  std::vector<double> kp(nk * 3, 0.0);
  for (size_t i = 0; i < nk; ++i)
    for (int d = 0; d < dim; ++d) kp[i * 3 + d] = k[i * dim + d];
  RTB_TRY(w.alloc(dk, kp.size()));
  RTB_HIP(hipMemcpyAsync(*dk, kp.data(), kp.size() * sizeof(double), hipMemcpyHostToDevice, st));
  return RTB_OK;
}

// Device bytes one call may hold for its k-point chunk: RTB_WORKSPACE_MB
// when set, else an eighth of the device memory (36 GB on a 288 GB part),
// at most 32 GiB.  Larger chunks mean fewer, fuller launches.
size_t workspace_budget() {
  if (const char* env = std::getenv("RTB_WORKSPACE_MB")) {
    const long long mb = std::atoll(env);
    if (mb > 0) return static_cast<size_t>(mb) << 20;
  }
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || total_b == 0) return size_t(4) << 30;
  return std::min(total_b / 8, size_t(32) << 30);
}

size_t chunk_for(size_t nk, size_t per_k_bytes) {
  const size_t budget = workspace_budget();
  size_t c = std::max<size_t>(1, budget / std::max<size_t>(per_k_bytes, 1));
  return std::min(nk, c);
}

// Bytes per k-point of alloc_eig's buffers beyond those of n <= 128 (the
// two-stage route's reflector store, and its matrix when no vectors are
// asked for).
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

// ---- eigenpair checks ----
// Options of the mixed form's check, read from the environment per call:
//   RTB_MIXED_RESIDUAL  the accepted residual |H u - lambda u| of a unit
//                       refined eigenvector, in eV per max(1, max |lambda|)
//                       (default 1e-9);
//   RTB_MIXED_FALLBACK  0: a flagged k-point fails the call with
//                       RTB_ERR_EIGEN_FAILED instead of being re-solved in FP64.
MixedOpts mixed_opts() {
  MixedOpts o;
  if (const char* e = std::getenv("RTB_MIXED_RESIDUAL")) {
    char* end = nullptr;
    const double v = std::strtod(e, &end);
    if (end != e && v >= 0.0) o.res_tol = v;
  }
  if (const char* e = std::getenv("RTB_MIXED_FALLBACK")) o.fallback = std::atoi(e) != 0;
  return o;
}

// Device counters of a call: the k-points the mixed check flagged in the
// current chunk (count, local indices), the k-points whose FP64 eigenvalues
// are not finite, the largest refined residual (bits of a double); and the
// buffers of the FP64 re-solve, taken from the workspace on first use.
struct Checks {
  int* flag_count = nullptr;
  int* flag_list = nullptr;
  int* nonfinite = nullptr;
  unsigned long long* max_res = nullptr;
  size_t resolved = 0;
  MixedOpts opt;
  size_t cap = 0;
  double *dk = nullptr, *H = nullptr, *E = nullptr, *Z = nullptr;
  EigBufs eb{};
  RefineCheck refine() const {
    RefineCheck r;
    r.res_tol = opt.res_tol;
    r.fail_count = flag_count;
    r.fail_list = flag_list;
    r.max_res = max_res;
    return r;
  }
};

// The counters (and, with reserve > 0, the FP64 re-solve buffers for that
// many k-points at a time) from the workspace; checks_zero clears the counters.
int checks_alloc(Work& w, int n, size_t K, size_t reserve, Checks& ck) {
  ck.opt = mixed_opts();
  int* ints;
  RTB_TRY(w.alloc(&ints, 4 + K));  // flag count, non-finite count, max residual (8 bytes), list
  ck.flag_count = ints;
  ck.nonfinite = ints + 1;
  ck.max_res = reinterpret_cast<unsigned long long*>(ints + 2);
  ck.flag_list = ints + 4;
  if (reserve > 0) {
    const size_t n2 = static_cast<size_t>(n) * n;
    ck.cap = reserve;
    RTB_TRY(w.alloc(&ck.dk, ck.cap * 3));
    RTB_TRY(w.alloc(&ck.H, ck.cap * n2 * 2));
    RTB_TRY(w.alloc(&ck.E, ck.cap * n));
    RTB_TRY(w.alloc(&ck.Z, ck.cap * n2 * 2));
    RTB_TRY(alloc_eig(w, n, ck.cap, true, ck.eb));
  }
  return RTB_OK;
}

int checks_zero(hipStream_t st, Checks& ck) { return zero_ints(st, ck.flag_count, 4); }

int checks_init(hipStream_t st, Work& w, size_t K, Checks& ck) {
  RTB_TRY(checks_alloc(w, 0, K, 0, ck));
  return checks_zero(st, ck);
}

// The k-points of a chunk the mixed check flagged, re-solved in FP64: H by
// the dense Fourier sum at kof(local index), the FP64 eigensolver, and the
// results written over the refined ones (E, unit columns into Zc, scale 1).
// With the fallback switched off a flagged k-point fails the call.
template <typename KOf>
int resolve_flagged(hipStream_t st, Work& w, Timer& tmr, const rtb_model* m, const FourierArgs& a, bool from_upper,
                    size_t c, KOf kof, Checks& ck, double* E, double* Zc, double* scale) {
  int cnt = 0;
  RTB_HIP(hipMemcpyAsync(&cnt, ck.flag_count, sizeof(int), hipMemcpyDeviceToHost, st));
  RTB_HIP(hipStreamSynchronize(st));
  if (cnt == 0) return RTB_OK;
  if (!ck.opt.fallback) {
    set_error("EigenvalueComputationFailed: the refined eigenpairs of %d k-point(s) fail the residual check", cnt);
    return RTB_ERR_EIGEN_FAILED;
  }
  std::vector<int> list(static_cast<size_t>(cnt));
  RTB_HIP(hipMemcpyAsync(list.data(), ck.flag_list, list.size() * sizeof(int), hipMemcpyDeviceToHost, st));
  RTB_HIP(hipStreamSynchronize(st));
  const int n = m->nsta;
  const size_t n2 = static_cast<size_t>(n) * n;
  if (ck.cap == 0) {
    ck.cap = std::min<size_t>(c, 1024);
    RTB_TRY(w.alloc(&ck.dk, ck.cap * 3));
    RTB_TRY(w.alloc(&ck.H, ck.cap * n2 * 2));
    RTB_TRY(w.alloc(&ck.E, ck.cap * n));
    RTB_TRY(w.alloc(&ck.Z, ck.cap * n2 * 2));
    RTB_TRY(alloc_eig(w, n, ck.cap, true, ck.eb));
  }
  FourierArgs a0 = a;
  a0.np = 0;
  a0.combine = 0;
  std::vector<double> kp;
  for (size_t b0 = 0; b0 < list.size(); b0 += ck.cap) {
    const size_t bc = std::min(ck.cap, list.size() - b0);
    kp.assign(bc * 3, 0.0);
    for (size_t i = 0; i < bc; ++i) kof(static_cast<size_t>(list[b0 + i]), &kp[3 * i]);
    RTB_HIP(hipMemcpyAsync(ck.dk, kp.data(), kp.size() * sizeof(double), hipMemcpyHostToDevice, st));
    RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
    RTB_TRY(fourier_dense(st, a0, ck.dk, bc, ck.H, nullptr));
    RTB_TRY(tmr.end(1));
    RTB_TRY(eigh_launch_fp64(st, &tmr, n, bc, ck.H, ck.E, true, ck.eb.Vg, ck.eb.tau, ck.eb.Y, ck.eb.S, ck.eb.DE,
                             from_upper, ck.eb.band));
    RTB_TRY(tmr.begin(RTB_STAGE_BACKTRANS));
    RTB_TRY(eigh_backtransform_fp64(st, n, bc, ck.eb.Y, ck.eb.DE, ck.eb.S, ck.eb.Vg, ck.eb.tau, ck.Z, false,
                                    ck.eb.band));
    RTB_TRY(tmr.end(1, flop_backtrans(n, bc)));
    RTB_TRY(tmr.begin(RTB_STAGE_REFINE));
    RTB_TRY(count_nonfinite(st, n, bc, ck.E, ck.nonfinite));
    RTB_TRY(scatter_resolved(st, n, static_cast<int>(bc), ck.flag_list + b0, ck.E, ck.Z, E, Zc, scale));
    RTB_TRY(tmr.end(2));
    RTB_HIP(hipStreamSynchronize(st));  // kp is refilled next
  }
  RTB_TRY(zero_ints(st, ck.flag_count, 1));
  ck.resolved += list.size();
  return RTB_OK;
}

// End of a call: the largest refined residual and the re-solved count into
// the stats; RTB_ERR_EIGEN_FAILED when an eigenvalue is not finite.
int checks_finish(hipStream_t st, Checks& ck, Timer& tmr) {
  int nf = 0;
  unsigned long long mr = 0;
  RTB_HIP(hipMemcpyAsync(&nf, ck.nonfinite, sizeof(int), hipMemcpyDeviceToHost, st));
  RTB_HIP(hipMemcpyAsync(&mr, ck.max_res, sizeof(mr), hipMemcpyDeviceToHost, st));
  RTB_HIP(hipStreamSynchronize(st));
  double r;
  std::memcpy(&r, &mr, sizeof(r));
  tmr.note_check(r, ck.resolved);
  if (nf > 0) {
    set_error("EigenvalueComputationFailed: the eigenvalues of %d k-point(s) are not finite", nf);
    return RTB_ERR_EIGEN_FAILED;
  }
  return RTB_OK;
}

// Mixed-form eigenpairs of a chunk: FP32 Householder / bisection / inverse
// iteration / back-transformation (the FP64 buffers of `eb` reused as FP32
// storage) into Z, then FP64 refinement against the FP64 H.  The refined
// eigenvectors end in W2 as columns that are unit after multiplication by
// scale[k][j]; Z, W1 and W3 are scratch afterwards.  from_upper: H is first
// rebuilt from its upper triangle (models whose storage is not exactly
// Hermitian), so the solve and the refinement see the reference's matrix.
int eig_mixed_chunk(hipStream_t st, Timer& tmr, bool from_upper, int n, size_t c, double* H, double* E, EigBufs& eb,
                    double* Z, double* W1, double* W2, double* W3, double* scale, const RefineCheck& chk, int nb) {
  float* Vg = reinterpret_cast<float*>(eb.Vg);
  float* tau = reinterpret_cast<float*>(eb.tau);
  float* Y = reinterpret_cast<float*>(eb.Y);
  float* S = reinterpret_cast<float*>(eb.S);
  float* DE = reinterpret_cast<float*>(eb.DE);
  float* band = reinterpret_cast<float*>(eb.band);
  RTB_TRY(eigh_launch_fp32(st, &tmr, n, c, H, E, true, Vg, tau, Y, S, DE, from_upper, band, nb));
  RTB_TRY(tmr.begin(RTB_STAGE_BACKTRANS));
  float* Zf = reinterpret_cast<float*>(Z);  // exact: the route's values are FP32
  RTB_TRY(eigh_backtransform_fp32(st, n, c, Y, DE, S, Vg, tau, Zf, false, scale, band, nb));  // |z_j|^2 into scale
  RTB_TRY(tmr.end(1, flop_backtrans(n, c)));
  // one Ogita-Aishima step in FP64 (vector error squared), Rayleigh quotients,
  // and Rayleigh-Ritz inside clusters closer than 1e-3 of the spectral radius
  RTB_TRY(refine_eigenpairs(st, &tmr, n, c, H, Zf, W1, W2, W3, E, scale, 1e-3, chk));
  return RTB_OK;
}

// Kubo for a chunk: V = [K][n][n] (V0 + i V1) or [K][2][n][n] (spin current);
// Z = [K][n][n] (columns); omega [K][n].  The reference forms A = Z^H J Z and
// B = Z^H v_b Z (response/traits.rs:177-180); here C = Z^H (J + i v_b) Z = A + i B,
// two products instead of four, and the band sum reads A, B out of C.
int kubo_chunk(hipStream_t st, Timer& tmr, const rtb_model* m, size_t K, const double* V, const double* Z,
               const double* zscale, const double* E, int spin_dir, double eta, double* T1, double* C, double* J,
               double* omega, double* B, double* metric) {
  const int n = m->nsta;
  const size_t n2 = static_cast<size_t>(n) * n;
  if (B) {
    // A = U^H J_a U and B = U^H v_b U as the reference forms them (V holds
    // V0, V1 apart), then its band sum term by term; U = Z diag(zscale)
    const double* Ma = V;
    size_t sMa = 2 * n2;
    if (m->spin && spin_dir != RTB_SPIN_NONE) {
      RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
      RTB_TRY(spin_current(st, m->norb, K, spin_dir, V, 2 * n2 * 2, nullptr, J));
      RTB_TRY(tmr.end(1));
      Ma = J;
      sMa = n2;
    }
    RTB_TRY(tmr.begin(RTB_STAGE_GEMM64));
    RTB_TRY(zgemm_batched(st, n, K, Ma, sMa, false, Z, n2, T1, n2));
    RTB_TRY(zgemm_batched_scaled(st, n, K, Z, n2, true, T1, n2, C, n2, false, zscale));
    RTB_TRY(zgemm_batched(st, n, K, V + n2 * 2, 2 * n2, false, Z, n2, T1, n2));
    RTB_TRY(zgemm_batched_scaled(st, n, K, Z, n2, true, T1, n2, B, n2, false, zscale));
    RTB_TRY(tmr.end(4, 4 * flop_zgemm(n, K), 4 * mflop_zgemm(n, K, false)));
    RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
    RTB_TRY(kubo_omega_exact(st, n, K, C, B, E, eta, omega, metric));
    RTB_TRY(tmr.end(1));
    return RTB_OK;
  }
  // the operator pair as one matrix M = A_op + i B_op: V holds V0 + i V1
  // (combined by the Fourier stage), or V0, V1 apart when the first
  // operator is the spin current J(V0)
  const double* M = V;
  if (m->spin && spin_dir != RTB_SPIN_NONE) {
    RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
    RTB_TRY(spin_current(st, m->norb, K, spin_dir, V, 2 * n2 * 2, V + n2 * 2, J));
    RTB_TRY(tmr.end(1));
    M = J;
  }
  RTB_TRY(tmr.begin(RTB_STAGE_GEMM64));
  // C = U^H M U for the unit eigenvectors U = Z diag(zscale) (zscale null: Z
  // is unit), consumed by the product's epilogue: per-tile sums of the band-
  // sum weights into C's buffer, then Omega from them
  RTB_TRY(zgemm_batched(st, n, K, M, n2, false, Z, n2, T1, n2));
  RTB_TRY(zgemm_kubo(st, n, K, Z, T1, zscale, E, eta, C));
  RTB_TRY(tmr.end(2, 2 * flop_zgemm(n, K), 2 * mflop_zgemm(n, K, false)));
  RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
  RTB_TRY(kubo_omega_from_parts(st, n, K, C, omega));
  RTB_TRY(tmr.end(1));
  return RTB_OK;
}

// Berry curvature dipole kernel of a chunk (berry_curvature_dipole_n_onek,
// response/nonlinear/mod.rs:89-163): V holds v_cur, v_f1, v_f2 per k (stride
// 3 n^2); A1 = U^H J U (J = v_cur or its spin current), A2 = U^H v_f1 U,
// A3 = U^H v_f2 U with U = Z diag(zscale); Omega^{cur,f1} (and with sym
// Omega^{cur,f2}) by the term-by-term band sum, then the per-band kernel.
int dipole_chunk(hipStream_t st, Timer& tmr, const rtb_model* m, size_t K, const double* V, const double* Z,
                 const double* zscale, const double* E, int spin_dir, double eta, bool sym, double* T1, double* A1,
                 double* A2, double* A3, double* J, double* om1, double* om2, double* kern) {
  const int n = m->nsta;
  const size_t n2 = static_cast<size_t>(n) * n;
  const double* Ma = V;
  size_t sMa = 3 * n2;
  if (m->spin && spin_dir != RTB_SPIN_NONE) {
    RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
    RTB_TRY(spin_current(st, m->norb, K, spin_dir, V, 3 * n2 * 2, nullptr, J));
    RTB_TRY(tmr.end(1));
    Ma = J;
    sMa = n2;
  }
  RTB_TRY(tmr.begin(RTB_STAGE_GEMM64));
  RTB_TRY(zgemm_batched(st, n, K, Ma, sMa, false, Z, n2, T1, n2));
  RTB_TRY(zgemm_batched_scaled(st, n, K, Z, n2, true, T1, n2, A1, n2, false, zscale));
  RTB_TRY(zgemm_batched(st, n, K, V + n2 * 2, 3 * n2, false, Z, n2, T1, n2));
  RTB_TRY(zgemm_batched_scaled(st, n, K, Z, n2, true, T1, n2, A2, n2, false, zscale));
  RTB_TRY(zgemm_batched(st, n, K, V + n2 * 4, 3 * n2, false, Z, n2, T1, n2));
  RTB_TRY(zgemm_batched_scaled(st, n, K, Z, n2, true, T1, n2, A3, n2, false, zscale));
  RTB_TRY(tmr.end(6, 6 * flop_zgemm(n, K), 6 * mflop_zgemm(n, K, false)));
  RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
  RTB_TRY(kubo_omega_exact(st, n, K, A1, A2, E, eta, om1, nullptr));
  if (sym) RTB_TRY(kubo_omega_exact(st, n, K, A1, A3, E, eta, om2, nullptr));
  RTB_TRY(dipole_combine(st, n, K, A2, A3, om1, om2, sym, kern));
  RTB_TRY(tmr.end(sym ? 3 : 2));
  return RTB_OK;
}

void publish_stats(const rtb_stats& s) { g_stats = s; }

// Host-to-device copy of a large pageable buffer (a model's hoppings): HIP
// moves pageable memory through its own small staging at a few GB/s, so the
// copy goes through two pinned 32 MB buffers instead, host threads filling
// one while the DMA engine drains the other.  Small copies go direct.
int upload_pageable(void* dst, const void* src, size_t bytes) {
  constexpr size_t kStage = size_t(32) << 20;
  if (bytes < 4 * kStage) {
    RTB_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return RTB_OK;
  }
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipStream_t st = nullptr;
  int rc = RTB_OK;
  auto check = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == RTB_OK) rc = fail(RTB_ERR_DEVICE, what);
    return e == hipSuccess;
  };
  if (check(hipStreamCreate(&st), "upload stream") && check(hipHostMalloc(&stage[0], kStage), "pinned staging") &&
      check(hipHostMalloc(&stage[1], kStage), "pinned staging") && check(hipEventCreate(&ev[0]), "event") &&
      check(hipEventCreate(&ev[1]), "event")) {
    const char* s8 = static_cast<const char*>(src);
    char* d8 = static_cast<char*>(dst);
    size_t i = 0;
    for (size_t off = 0; off < bytes && rc == RTB_OK; off += kStage, ++i) {
      const int b = static_cast<int>(i & 1);
      if (i >= 2 && !check(hipEventSynchronize(ev[b]), "staging wait")) break;
      const size_t len = std::min(kStage, bytes - off);
      constexpr int kT = 4;
      std::vector<std::thread> th;
      for (int t = 0; t < kT; ++t)
        th.emplace_back([&, t] {
          const size_t a = len / kT * t, e = t == kT - 1 ? len : len / kT * (t + 1);
          std::memcpy(static_cast<char*>(stage[b]) + a, s8 + off + a, e - a);
        });
      for (auto& x : th) x.join();
      if (!check(hipMemcpyAsync(d8 + off, stage[b], len, hipMemcpyHostToDevice, st), "staged copy")) break;
      check(hipEventRecord(ev[b], st), "event record");
    }
    check(hipStreamSynchronize(st), "upload sync");
  }
  for (int b = 0; b < 2; ++b) {
    if (ev[b]) (void)hipEventDestroy(ev[b]);
    if (stage[b]) (void)hipHostFree(stage[b]);
  }
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

// ham[-R] == ham[R]^H exactly for every stored R (the builder's storage, by
// construction); models read from files are Hermitian only to their printed
// precision.  Decides whether the velocity operators may be combined into
// one Kubo product (exact for Hermitian operators only).
bool storage_is_hermitian(int dim, int n, int n_R, const int64_t* hamR, const double* ham) {
  std::map<std::vector<int64_t>, int> index;
  for (int r = 0; r < n_R; ++r) index.emplace(std::vector<int64_t>(hamR + r * dim, hamR + (r + 1) * dim), r);
  std::vector<int> partner(n_R);
  for (int r = 0; r < n_R; ++r) {
    std::vector<int64_t> neg(dim);
    for (int d = 0; d < dim; ++d) neg[d] = -hamR[r * dim + d];
    auto it = index.find(neg);
    if (it == index.end()) return false;
    partner[r] = it->second;
  }
  // each pair once (r <= partner), blocks spread over host threads: a full
  // pass over the storage (349 MB for 1331 R at n = 128)
  const size_t n2 = static_cast<size_t>(n) * n;
  const int T = std::max(1, std::min(16, std::min(n_R, static_cast<int>(std::thread::hardware_concurrency()))));
  std::atomic<bool> ok{true};
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      for (int r = t; r < n_R && ok.load(std::memory_order_relaxed); r += T) {
        if (partner[r] < r) continue;
        const double* a = ham + r * n2 * 2;
        const double* b = ham + partner[r] * n2 * 2;
        bool good = true;
        for (int i = 0; i < n && good; ++i)
          for (int j = 0; j < n; ++j) {
            const double* x = a + (static_cast<size_t>(i) * n + j) * 2;
            const double* y = b + (static_cast<size_t>(j) * n + i) * 2;
            good = good && x[0] == y[0] && x[1] == -y[1];
          }
        if (!good) ok.store(false, std::memory_order_relaxed);
      }
    });
  for (auto& x : th) x.join();
  return ok.load();
}

// Every value of a large array finite (a model's hoppings), spread over
// host threads.
bool all_finite(const double* x, size_t count) {
  const int T = static_cast<int>(std::max<size_t>(1, std::min<size_t>({16, count / (size_t(1) << 20) + 1,
                                                                          std::thread::hardware_concurrency()})));
  std::atomic<bool> ok{true};
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      const size_t a = count / T * t, e = t == T - 1 ? count : count / T * (t + 1);
      bool good = true;
      for (size_t i = a; i < e; ++i) good = good && std::isfinite(x[i]);
      if (!good) ok.store(false, std::memory_order_relaxed);
    });
  for (auto& x : th) x.join();
  return ok.load();
}

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
}

}  // namespace rtb

using namespace rtb;

extern "C" {

const char* rtb_last_error(void) { return g_error.c_str(); }

int rtb_last_stats(rtb_stats* out) {
  if (!out) return fail(RTB_ERR_INVALID_ARGUMENT, "null stats");
  *out = g_stats;
  return RTB_OK;
}

int rtb_model_create(int dim, int spin, int norb, int n_R, const int64_t* hamR, const double* ham,
                     const double* orb, const double* lat, rtb_model** out) {
  if (!out || !hamR || !ham || !orb || !lat) return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (dim < 1 || dim > 3) return fail(RTB_ERR_INVALID_MODEL, "dim must be 1, 2 or 3");
  if (norb < 1) return fail(RTB_ERR_INVALID_MODEL, "the model has no orbitals");
  if (n_R < 1) return fail(RTB_ERR_INVALID_MODEL, "hamR must hold at least R = 0");
  for (int d = 0; d < dim; ++d)
    if (hamR[d] != 0) return fail(RTB_ERR_INVALID_MODEL, "hamR row 0 must be R = 0");
  for (int i = 0; i < dim * dim; ++i)
    if (!std::isfinite(lat[i])) return fail(RTB_ERR_INVALID_MODEL, "lattice must be finite");
  for (int i = 0; i < norb * dim; ++i)
    if (!std::isfinite(orb[i])) return fail(RTB_ERR_INVALID_MODEL, "orbital positions must be finite");
  const int n = spin ? 2 * norb : norb;
  const size_t n2 = static_cast<size_t>(n) * n;
  // a NaN or infinite hopping would come back as eigenvalues that are not
  // numbers (the reference's zheev reports such input as an error)
  if (!all_finite(ham, 2 * n2 * static_cast<size_t>(n_R)))
    return fail(RTB_ERR_INVALID_MODEL, "hoppings must be finite");
  auto* m = new rtb_model();
  m->dim = dim;
  m->spin = spin ? 1 : 0;
  m->norb = norb;
  m->nsta = n;
  m->nR = n_R;
  m->hamR.assign(hamR, hamR + static_cast<size_t>(n_R) * dim);
  m->orb.assign(orb, orb + static_cast<size_t>(norb) * dim);
  m->lat.assign(lat, lat + dim * dim);
  if (dim == 1) {
    m->det_lat = lat[0];
  } else if (dim == 2) {
    m->det_lat = lat[0] * lat[3] - lat[1] * lat[2];
  } else {
    m->det_lat = lat[0] * (lat[4] * lat[8] - lat[5] * lat[7]) - lat[1] * (lat[3] * lat[8] - lat[5] * lat[6]) +
                 lat[2] * (lat[3] * lat[7] - lat[4] * lat[6]);
  }
  auto cleanup = [&](int rc) {
    rtb_model_destroy(m);
    return rc;
  };
  if (hipGetDevice(&m->device) != hipSuccess) return cleanup(fail(RTB_ERR_DEVICE, "no HIP device"));
  m->herm = storage_is_hermitian(dim, n, n_R, hamR, ham);
  std::vector<double> R(static_cast<size_t>(n_R) * 3, 0.0), Rc(static_cast<size_t>(n_R) * 3, 0.0);
  for (int r = 0; r < n_R; ++r)
    for (int d = 0; d < dim; ++d) {
      R[r * 3 + d] = static_cast<double>(hamR[r * dim + d]);
      double s = 0.0;
      for (int i = 0; i < dim; ++i) s += static_cast<double>(hamR[r * dim + i]) * lat[i * dim + d];
      Rc[r * 3 + d] = s;
    }
  std::vector<double> tau(static_cast<size_t>(n) * 3, 0.0), tauc(static_cast<size_t>(n) * 3, 0.0);
  for (int o = 0; o < n; ++o) {
    int b = o % norb;
    for (int d = 0; d < dim; ++d) {
      tau[o * 3 + d] = orb[b * dim + d];
      double s = 0.0;
      for (int i = 0; i < dim; ++i) s += orb[b * dim + i] * lat[i * dim + d];
      tauc[o * 3 + d] = s;
    }
  }
  if (hipMalloc(&m->d_ham, sizeof(double) * 2 * n2 * n_R) != hipSuccess ||
      hipMalloc(&m->d_R, sizeof(double) * R.size()) != hipSuccess ||
      hipMalloc(&m->d_Rcart, sizeof(double) * Rc.size()) != hipSuccess ||
      hipMalloc(&m->d_tau, sizeof(double) * tau.size()) != hipSuccess ||
      hipMalloc(&m->d_taucart, sizeof(double) * tauc.size()) != hipSuccess)
    return cleanup(fail(RTB_ERR_DEVICE, "device allocation for the model failed"));
  if (upload_pageable(m->d_ham, ham, sizeof(double) * 2 * n2 * n_R) != RTB_OK) return cleanup(RTB_ERR_DEVICE);
  if (hipMemcpy(m->d_R, R.data(), sizeof(double) * R.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_Rcart, Rc.data(), sizeof(double) * Rc.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_tau, tau.data(), sizeof(double) * tau.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_taucart, tauc.data(), sizeof(double) * tauc.size(), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(RTB_ERR_DEVICE, "model upload failed"));
  *out = m;
  return RTB_OK;
}

void rtb_model_destroy(rtb_model* m) {
  if (!m) return;
  (void)hipFree(m->d_ham);
  (void)hipFree(m->d_R);
  (void)hipFree(m->d_Rcart);
  (void)hipFree(m->d_tau);
  (void)hipFree(m->d_taucart);
  (void)hipFree(m->d_rmat);
  delete m;
}

int rtb_model_set_rmatrix(rtb_model* m, const double* rmatrix) {
  if (!m || !rmatrix) return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  const size_t count = static_cast<size_t>(m->nR) * m->dim * m->nsta * m->nsta * 2;
  for (size_t i = 0; i < count; ++i)
    if (!std::isfinite(rmatrix[i])) return fail(RTB_ERR_INVALID_MODEL, "position matrices must be finite");
  if (!m->d_rmat && hipMalloc(&m->d_rmat, count * sizeof(double)) != hipSuccess) {
    m->d_rmat = nullptr;
    return fail(RTB_ERR_DEVICE, "device allocation for the position matrices failed");
  }
  return upload_pageable(m->d_rmat, rmatrix, count * sizeof(double));
}

int rtb_model_nsta(const rtb_model* m) { return m ? m->nsta : RTB_ERR_INVALID_ARGUMENT; }

int rtb_gen_ham_batch(const rtb_model* m, const double* k, size_t nk, int gauge, const double* dirs, int n_proj,
                      double* H, double* V) {
  RTB_TRY(prepare(m));
  if ((nk > 0 && (!k || !H)) || (n_proj > 0 && (!dirs || !V)))
    return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  if (gauge != RTB_GAUGE_ATOM && gauge != RTB_GAUGE_LATTICE) return fail(RTB_ERR_INVALID_ARGUMENT, "bad gauge");
  if (n_proj < 0 || n_proj > 3) return fail(RTB_ERR_UNSUPPORTED, "n_proj must be 0..3");
  if (nk == 0) return RTB_OK;
  Lease lease;
  RTB_TRY(lease.status());
  hipStream_t st = lease.stream();
  Work w(lease);
  Timer tmr(lease);
  FourierArgs a;
  RTB_TRY(upload_dirs(st, w, m, dirs, n_proj, a));
  a.gauge = gauge;
  const int n = m->nsta;
  const size_t n2 = static_cast<size_t>(n) * n;
  const bool rterm = m->d_rmat && n_proj > 0;  // the position-matrix term of v
  const size_t K = chunk_for(nk, (1 + n_proj + (rterm ? 3 : 0)) * n2 * 16 + 24);
  double *dk, *dH, *dV = nullptr, *rproj = nullptr, *Rk = nullptr, *T1 = nullptr, *T2 = nullptr;
  RTB_TRY(w.alloc(&dk, K * 3));
  RTB_TRY(w.alloc(&dH, K * n2 * 2));
  if (n_proj) RTB_TRY(w.alloc(&dV, K * n_proj * n2 * 2));
  if (rterm) {
    RTB_TRY(w.alloc(&rproj, static_cast<size_t>(m->nR) * n2 * 2));
    RTB_TRY(w.alloc(&Rk, K * n2 * 2));
    RTB_TRY(w.alloc(&T1, K * n2 * 2));
    RTB_TRY(w.alloc(&T2, K * n2 * 2));
  }
  std::vector<double> kp;
  for (size_t k0 = 0; k0 < nk; k0 += K) {
    const size_t c = std::min(K, nk - k0);
    kp.assign(c * 3, 0.0);
    for (size_t i = 0; i < c; ++i)
      for (int d = 0; d < m->dim; ++d) kp[i * 3 + d] = k[(k0 + i) * m->dim + d];
    RTB_HIP(hipMemcpyAsync(dk, kp.data(), kp.size() * sizeof(double), hipMemcpyHostToDevice, st));
    RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
    RTB_TRY(fourier_dense(st, a, dk, c, dH, dV));
    RTB_TRY(tmr.end(1));
    if (rterm) {
      RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
      RTB_TRY(rmatrix_velocity(st, a, m->d_rmat, dirs, dk, c, dH, dV, rproj, Rk, T1, T2));
      RTB_TRY(tmr.end(1));
    }
    RTB_HIP(hipMemcpyAsync(H + k0 * n2 * 2, dH, c * n2 * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    if (n_proj)
      RTB_HIP(hipMemcpyAsync(V + k0 * n_proj * n2 * 2, dV, c * n_proj * n2 * 2 * sizeof(double),
                             hipMemcpyDeviceToHost, st));
    RTB_HIP(hipStreamSynchronize(st));
  }
  RTB_TRY(tmr.collect(nk, n));
  return RTB_OK;
}

// lower_vectors: eigenvectors of the Hermitian matrix of H's lower triangle
// (the reference's zheevr path) instead of its upper one (its eigh)
static int solve_impl(const rtb_model* m, const double* k, size_t nk, int precision, double* E, double* evec,
                      bool lower_vectors = false) {
  RTB_TRY(prepare(m));
  RTB_TRY(check_precision(precision));
  if (nk > 0 && (!k || !E)) return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  if (nk == 0) return RTB_OK;
  const int n = m->nsta;
  if (n > kMaxN) return fail(RTB_ERR_UNSUPPORTED, "nsta above 256 is not supported by the eigensolver");
  const bool vec = evec != nullptr;
  Lease lease;
  RTB_TRY(lease.status());
  hipStream_t st = lease.stream();
  Work w(lease);
  Timer tmr(lease);
  FourierArgs a;
  RTB_TRY(upload_dirs(st, w, m, nullptr, 0, a));
  a.gauge = RTB_GAUGE_ATOM;  // solve_band_onek / solve_onek use Gauge::Atom
  const size_t n2 = static_cast<size_t>(n) * n;
  // the mixed form runs the FP32 route (with FP64 refinement) when
  // eigenvectors are requested; eigenvalue-only solves stay FP64
  const bool f32 = vec && precision == RTB_PREC_MIXED;
  const size_t per_k = 24 + n2 * 16 * 2 + n * 32 + (vec ? n2 * 32 + n2 * 16 : 0) + (f32 ? n2 * 32 + n * 8 : 0) +
                       eig_extra_bytes(n, vec);
  const size_t K = chunk_for(nk, per_k);
  double *dk, *dH, *dE, *dZ = nullptr, *dZc = nullptr, *dHZ = nullptr, *dsc = nullptr;
  EigBufs eb;
  RTB_TRY(w.alloc(&dk, K * 3));
  RTB_TRY(w.alloc(&dH, K * n2 * 2));
  RTB_TRY(w.alloc(&dE, K * n));
  RTB_TRY(alloc_eig(w, n, K, vec, eb));
  if (vec) RTB_TRY(w.alloc(&dZ, K * n2 * 2));
  if (f32) {
    RTB_TRY(w.alloc(&dZc, K * n2 * 2));
    RTB_TRY(w.alloc(&dHZ, K * n2 * 2));
    RTB_TRY(w.alloc(&dsc, K * n));
  }
  Checks ck;
  RTB_TRY(checks_init(st, w, K, ck));
  std::vector<double> kp;
  for (size_t k0 = 0; k0 < nk; k0 += K) {
    const size_t c = std::min(K, nk - k0);
    kp.assign(c * 3, 0.0);
    for (size_t i = 0; i < c; ++i)
      for (int d = 0; d < m->dim; ++d) kp[i * 3 + d] = k[(k0 + i) * m->dim + d];
    RTB_HIP(hipMemcpyAsync(dk, kp.data(), kp.size() * sizeof(double), hipMemcpyHostToDevice, st));
    RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
    RTB_TRY(fourier_dense(st, a, dk, c, dH, nullptr));
    RTB_TRY(tmr.end(1));
    if (f32) {
      // W2, W3: the reflector and inverse-iteration scratch, dead by then
      const bool up = !m->herm && !lower_vectors;
      RTB_TRY(eig_mixed_chunk(st, tmr, up, n, c, dH, dE, eb, dZc, dHZ, eb.Vg, eb.S, dsc, ck.refine()));
      auto kof = [&](size_t i, double* k3) {
        for (int d = 0; d < m->dim; ++d) k3[d] = k[(k0 + i) * m->dim + d];
      };
      RTB_TRY(resolve_flagged(st, w, tmr, m, a, up, c, kof, ck, dE, eb.Vg, dsc));
      RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
      RTB_TRY(transpose_batched(st, n, c, eb.Vg, dsc, dZ));
      RTB_TRY(tmr.end(1));
    } else {
      RTB_TRY(eigh_launch_fp64(st, &tmr, n, c, dH, dE, vec, eb.Vg, eb.tau, eb.Y, eb.S, eb.DE,
                               vec && !m->herm && !lower_vectors, eb.band));
      RTB_TRY(count_nonfinite(st, n, c, dE, ck.nonfinite));
      if (vec) {
        RTB_TRY(tmr.begin(RTB_STAGE_BACKTRANS));
        RTB_TRY(eigh_backtransform_fp64(st, n, c, eb.Y, eb.DE, eb.S, eb.Vg, eb.tau, dZ, true, eb.band));
        RTB_TRY(tmr.end(1, flop_backtrans(n, c)));
      }
    }
    RTB_HIP(hipMemcpyAsync(E + k0 * n, dE, c * n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (vec)
      RTB_HIP(hipMemcpyAsync(evec + k0 * n2 * 2, dZ, c * n2 * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    RTB_HIP(hipStreamSynchronize(st));
  }
  RTB_TRY(checks_finish(st, ck, tmr));
  RTB_TRY(tmr.collect(nk, n));
  return RTB_OK;
}

int rtb_solve_band_all(const rtb_model* m, const double* k, size_t nk, int precision, double* E) {
  return solve_impl(m, k, nk, precision, E, nullptr);
}

int rtb_solve_all(const rtb_model* m, const double* k, size_t nk, int precision, double* E, double* evec) {
  if (nk > 0 && !evec) return fail(RTB_ERR_INVALID_ARGUMENT, "null evec");
  return solve_impl(m, k, nk, precision, E, evec);
}

// The eigenpairs with E in (vl, vu] (zheevr RANGE = 'V'), packed to the
// front of each k's row, ascending; count[k] of them.  The whole spectrum
// comes from the device solve, the selection is a host pass over its output.
static int solve_range_impl(const rtb_model* m, const double* k, size_t nk, double vl, double vu, double epsilon,
                            int precision, double* E, double* evec, int* count) {
  RTB_TRY(prepare(m));
  if (!(vl < vu) || !std::isfinite(vl) || !std::isfinite(vu))
    return fail(RTB_ERR_INVALID_ARGUMENT, "range: need finite vl < vu (zheevr RANGE = 'V')");
  if (!std::isfinite(epsilon)) return fail(RTB_ERR_INVALID_ARGUMENT, "epsilon must be finite");
  if (nk > 0 && !count) return fail(RTB_ERR_INVALID_ARGUMENT, "null count");
  RTB_TRY(solve_impl(m, k, nk, precision, E, evec, true));
  const int n = m->nsta;
  const size_t n2 = static_cast<size_t>(n) * n;
  for (size_t i = 0; i < nk; ++i) {
    double* e = E + i * n;
    int c = 0;
    for (int j = 0; j < n; ++j) {
      if (!(e[j] > vl && e[j] <= vu)) continue;
      if (c != j) {
        e[c] = e[j];
        if (evec) std::memmove(evec + (i * n2 + static_cast<size_t>(c) * n) * 2,
                               evec + (i * n2 + static_cast<size_t>(j) * n) * 2, n * 2 * sizeof(double));
      }
      ++c;
    }
    count[i] = c;
  }
  return RTB_OK;
}

int rtb_solve_band_range(const rtb_model* m, const double* k, size_t nk, double vl, double vu, double epsilon,
                         int precision, double* E, int* count) {
  return solve_range_impl(m, k, nk, vl, vu, epsilon, precision, E, nullptr, count);
}

int rtb_solve_range(const rtb_model* m, const double* k, size_t nk, double vl, double vu, double epsilon,
                    int precision, double* E, double* evec, int* count) {
  if (nk > 0 && !evec) return fail(RTB_ERR_INVALID_ARGUMENT, "null evec");
  return solve_range_impl(m, k, nk, vl, vu, epsilon, precision, E, evec, count);
}

// Shared Berry-curvature pipeline over either a k list (k != null) or a mesh.
//   mode 0: omega + E to host; mode 1: occupied sum per k to host;
//   mode 2: Hall partial sums over mu
struct BerryJob {
  const double* k = nullptr;        // host k list, or null for the mesh
  const size_t* kmesh = nullptr;
  size_t nk = 0;
  const double* dir = nullptr;
  int spin_dir = RTB_SPIN_NONE;
  double eta = 1e-3;
  int mode = 0;
  double* omega = nullptr;
  double* E = nullptr;
  double mu0 = 0.0;
  double T = 0.0;
  double* occ = nullptr;
  const double* mu = nullptr;
  size_t n_mu = 0;
  double* sigma = nullptr;
  int precision = RTB_PREC_FP64;
  // quantum geometry (quantum_geometry.rs): the metric beside Omega; mode 0
  // writes it per k into `metric`, mode 2 integrates it into `msum`
  bool geometry = false;
  double* metric = nullptr;
  double* msum = nullptr;
  // Berry curvature dipole (response/nonlinear/mod.rs): dir has 3 rows
  // (current, field_1, field_2); mode 2 only, weights -df/dE
  bool dipole = false;
  bool sym = false;
  // layer-resolved AHC (mode 2): rows of H by layer (perm, bounds: host,
  // nsta and n_layers + 1 entries), output [n_mu][n_layers]
  const int* layer_perm = nullptr;
  const int* layer_bounds = nullptr;
  int n_layers = 0;
  double* sigma_layers = nullptr;
};

// A chunk of consecutive pipeline k-points; for the mesh path it covers the
// pairs [p0, p0 + np) with every i1 (k1-innermost order).
struct Chunk {
  size_t k0, c, p0, np;
};

// Mesh intermediates for the separable Bloch sum (mixed form on mesh calls).
static int mesh_setup(hipStream_t st, Work& w, Timer& tmr, const rtb_model* m, const size_t* kmesh, int precision,
                      size_t nk, size_t K, bool want_v, MeshPlan& mp, double** G, std::vector<Chunk>& chunks) {
  bool use_mesh = false;
  *G = nullptr;
  if (kmesh && precision == RTB_PREC_MIXED) {
    RTB_TRY(mesh_plan(m, kmesh, mp));
    use_mesh = mp.ok && static_cast<size_t>(mp.N[0]) <= K;
  }
  chunks.clear();
  // equal chunks (a short last chunk would run its launches under-filled)
  if (!use_mesh) {
    const size_t nc = (nk + K - 1) / K, Kb = (nk + nc - 1) / nc;
    for (size_t k0 = 0; k0 < nk; k0 += Kb) chunks.push_back({k0, std::min(Kb, nk - k0), 0, 0});
    return RTB_OK;
  }
  const size_t KP0 = K / mp.N[0], ncp = (mp.P + KP0 - 1) / KP0, KP = (mp.P + ncp - 1) / ncp;
  for (size_t p0 = 0; p0 < mp.P; p0 += KP) {
    const size_t np = std::min(KP, mp.P - p0);
    chunks.push_back({p0 * mp.N[0], np * mp.N[0], p0, np});
  }
  double* work;
  int* bidx;
  RTB_TRY(w.alloc(&work, 2 * (mp.box_elems + mp.ga_elems)));
  RTB_TRY(w.alloc(G, 2 * mp.gb_elems));
  RTB_TRY(w.alloc(&bidx, static_cast<size_t>(m->nR)));
  RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
  RTB_TRY(mesh_prepare(st, m, mp, bidx, work, *G, want_v));
  RTB_TRY(tmr.end(m->dim == 3 ? (want_v ? 4 : 3) : 2));
  return RTB_OK;
}

// Device buffers of a Berry-curvature call (besides the direction tables
// and the mesh intermediates), for chunks of K k-points.
struct BerryBufs {
  double *dk = nullptr, *dH = nullptr, *dV = nullptr, *dE = nullptr, *dOm = nullptr, *dJ = nullptr, *dZ = nullptr;
  double *dMet = nullptr, *dA3 = nullptr, *dOm2 = nullptr, *dKern = nullptr, *kub = nullptr, *dsc = nullptr;
  double *rproj = nullptr, *dmu = nullptr, *dpart = nullptr, *dsig = nullptr, *docc = nullptr;
  double *dmpart = nullptr, *dmsig = nullptr;
  int *dperm = nullptr, *dbounds = nullptr;
  double *dlpart = nullptr, *dlsum = nullptr, *dlay = nullptr;
  EigBufs eb{};
  int total_blk = 0;
};

static int berry_alloc(Work& w, const rtb_model* m, const BerryJob& j, size_t K, const std::vector<Chunk>& chunks,
                       bool f32, bool spinj, bool rterm, BerryBufs& b) {
  const int n = m->nsta;
  const int nv = j.dipole ? 3 : 2;
  const size_t n2 = static_cast<size_t>(n) * n;
  RTB_TRY(w.alloc(&b.dk, K * 3));
  RTB_TRY(w.alloc(&b.dH, K * n2 * 2));
  RTB_TRY(w.alloc(&b.dV, K * n2 * 2 * nv));
  RTB_TRY(w.alloc(&b.dE, K * n));
  RTB_TRY(w.alloc(&b.dOm, K * n));
  if (j.geometry) RTB_TRY(w.alloc(&b.dMet, K * n));
  if (j.dipole) {
    RTB_TRY(w.alloc(&b.dA3, K * n2 * 2));
    RTB_TRY(w.alloc(&b.dOm2, K * n));
    RTB_TRY(w.alloc(&b.dKern, K * n));
  }
  RTB_TRY(alloc_eig(w, n, K, true, b.eb));
  RTB_TRY(w.alloc(&b.kub, K * n2 * 6));  // T1, A, B (complex)
  if (spinj) RTB_TRY(w.alloc(&b.dJ, K * n2 * 2));
  if (f32) {
    RTB_TRY(w.alloc(&b.dZ, K * n2 * 2));  // the refinement needs H after the solve
    RTB_TRY(w.alloc(&b.dsc, K * n));      // column scales of the refined eigenvectors
  } else {
    b.dZ = b.dH;  // H is consumed by the Householder kernel before Z is written
  }
  if (rterm) RTB_TRY(w.alloc(&b.rproj, static_cast<size_t>(m->nR) * n2 * 2));
  b.total_blk = 0;
  if (j.mode == 2) {
    for (const Chunk& ch : chunks) b.total_blk += hall_blocks(ch.c);
    RTB_TRY(w.alloc(&b.dmu, j.n_mu));
    RTB_TRY(w.alloc(&b.dpart, static_cast<size_t>(b.total_blk) * j.n_mu));
    RTB_TRY(w.alloc(&b.dsig, j.n_mu));
    if (j.geometry) {
      RTB_TRY(w.alloc(&b.dmpart, static_cast<size_t>(b.total_blk) * j.n_mu));
      RTB_TRY(w.alloc(&b.dmsig, j.n_mu));
    }
  }
  if (j.mode == 1) RTB_TRY(w.alloc(&b.docc, K));
  if (j.n_layers > 0) {
    const size_t ml = j.n_mu * static_cast<size_t>(j.n_layers);
    RTB_TRY(w.alloc(&b.dperm, static_cast<size_t>(n)));
    RTB_TRY(w.alloc(&b.dbounds, static_cast<size_t>(j.n_layers) + 1));
    RTB_TRY(w.alloc(&b.dlpart, K * ml));
    RTB_TRY(w.alloc(&b.dlsum, chunks.size() * ml));
    RTB_TRY(w.alloc(&b.dlay, ml));
  }
  return RTB_OK;
}

// Equal chunks of at most K consecutive k-points (a short last chunk would
// run its launches under-filled).
static void equal_chunks(size_t nk, size_t K, std::vector<Chunk>& chunks) {
  chunks.clear();
  const size_t nc = (nk + K - 1) / K, Kb = (nk + nc - 1) / nc;
  for (size_t k0 = 0; k0 < nk; k0 += Kb) chunks.push_back({k0, std::min(Kb, nk - k0), 0, 0});
}

// Where a Berry-curvature call runs: a leased workspace slot (grown as
// needed), or a fixed device buffer (the allocation-free entry point: the
// dense Fourier sum, chunks sized so that every buffer of the call fits).
struct Exec {
  hipStream_t st = nullptr;
  Work* w = nullptr;
  Timer* tmr = nullptr;
  bool fixed = false;
  size_t cap = 0;  // fixed: bytes of the buffer
};

// k-points at a time of the FP64 re-solve buffers reserved up front in a
// fixed buffer
constexpr size_t kFixedResolve = 16;

static int berry_run(const rtb_model* m, BerryJob& j, Exec& ex) {
  const int n = m->nsta;
  if (n > kMaxN) return fail(RTB_ERR_UNSUPPORTED, "nsta above 256 is not supported by the eigensolver");
  hipStream_t st = ex.st;
  Work& w = *ex.w;
  Timer& tmr = *ex.tmr;
  FourierArgs a;
  RTB_TRY(upload_dirs(st, w, m, j.dir, j.dipole ? 3 : 2, a));
  a.gauge = RTB_GAUGE_ATOM;  // berry_curvature_at_impl: gen_v_projected(k, Gauge::Atom, ..)
  const int nv = j.dipole ? 3 : 2;  // projected velocities per k
  const size_t n2 = static_cast<size_t>(n) * n;
  const bool spinj = m->spin && j.spin_dir != RTB_SPIN_NONE;
  const bool f32 = j.precision == RTB_PREC_MIXED;  // FP32 eigensolve + FP64 refinement
  // the plain form, and any model whose storage is not exactly Hermitian,
  // forms the reference's A and B apart (V0, V1 kept apart); the mixed form
  // of a Hermitian model takes V0 + i V1 straight from the Fourier stage
  const bool rterm = m->d_rmat != nullptr;  // the position-matrix term of v
  const bool exact_kubo = !f32 || !m->herm || j.geometry || j.dipole || rterm;
  a.combine = spinj || exact_kubo ? 0 : 1;
  // H/Z, V(2), Vg, Y+S aliased with T1/A/B, J
  const size_t per_k = 24 + n2 * 16 + n2 * 16 * nv + n2 * 16 + n2 * 48 + (spinj ? n2 * 16 : 0) + n * 56 +
                       (f32 ? n2 * 16 + n * 8 : 0) + (j.dipole ? n2 * 16 + n * 24 : 0) + eig_extra_bytes(n, true);
  MeshPlan mp;
  double* G = nullptr;
  std::vector<Chunk> chunks;
  BerryBufs bb;
  Checks ck;
  size_t K = 0;
  if (!ex.fixed) {
    K = chunk_for(j.nk, per_k);
    // the separable mesh passes form two projected velocities; the dipole's
    // three take the dense sum over mesh points
    RTB_TRY(mesh_setup(st, w, tmr, m, j.k || j.dipole || rterm ? nullptr : j.kmesh, j.precision, j.nk, K, true, mp,
                       &G, chunks));
    RTB_TRY(berry_alloc(w, m, j, K, chunks, f32, spinj, rterm, bb));
    RTB_TRY(checks_alloc(w, n, K, 0, ck));
  } else {
    // the largest chunk for which every buffer of the call fits what is left
    const size_t left = ex.cap > w.used() ? ex.cap - w.used() : 0;
    K = std::max<size_t>(1, std::min(j.nk, left / per_k));
    for (;;) {
      equal_chunks(j.nk, K, chunks);
      Work probe(nullptr, SIZE_MAX);
      BerryBufs pb;
      Checks pc;
      RTB_TRY(berry_alloc(probe, m, j, K, chunks, f32, spinj, rterm, pb));
      RTB_TRY(checks_alloc(probe, n, K, f32 ? std::min(K, kFixedResolve) : 0, pc));
      if (probe.used() <= left) break;
      if (K == 1) {
        set_error("the allocation-free call needs %zu bytes of device buffers for this model, above the %zu "
                  "bytes left of its fixed %zu-byte scratch",
                  probe.used(), left, ex.cap);
        return RTB_ERR_UNSUPPORTED;
      }
      K = std::max<size_t>(1, static_cast<size_t>(static_cast<double>(K) * static_cast<double>(left) /
                                                  static_cast<double>(probe.used()) * 0.98));
    }
    RTB_TRY(berry_alloc(w, m, j, K, chunks, f32, spinj, rterm, bb));
    RTB_TRY(checks_alloc(w, n, K, f32 ? std::min(K, kFixedResolve) : 0, ck));
  }
  RTB_TRY(checks_zero(st, ck));
  double *dk = bb.dk, *dH = bb.dH, *dV = bb.dV, *dE = bb.dE, *dOm = bb.dOm, *dJ = bb.dJ, *dZ = bb.dZ;
  double *dMet = bb.dMet, *dA3 = bb.dA3, *dOm2 = bb.dOm2, *dKern = bb.dKern, *dsc = bb.dsc, *rproj = bb.rproj;
  double *dmu = bb.dmu, *dpart = bb.dpart, *dsig = bb.dsig, *docc = bb.docc, *dmpart = bb.dmpart;
  double* dmsig = bb.dmsig;
  EigBufs& eb = bb.eb;
  const int total_blk = bb.total_blk;
  double* T1 = bb.kub;
  double* A = bb.kub + K * n2 * 2;
  double* B = bb.kub + K * n2 * 4;
  if (j.mode == 2) RTB_HIP(hipMemcpyAsync(dmu, j.mu, j.n_mu * sizeof(double), hipMemcpyHostToDevice, st));
  const size_t ml = j.n_mu * static_cast<size_t>(j.n_layers);
  if (j.n_layers > 0) {
    RTB_HIP(hipMemcpyAsync(bb.dperm, j.layer_perm, sizeof(int) * n, hipMemcpyHostToDevice, st));
    RTB_HIP(hipMemcpyAsync(bb.dbounds, j.layer_bounds, sizeof(int) * (j.n_layers + 1), hipMemcpyHostToDevice, st));
  }
  size_t chunk_no = 0;
  std::vector<double> kp;
  int blk_off = 0;
  for (const Chunk& ch : chunks) {
    const size_t k0 = ch.k0, c = ch.c;
    if (j.k) {
      kp.assign(c * 3, 0.0);
      for (size_t i = 0; i < c; ++i)
        for (int d = 0; d < m->dim; ++d) kp[i * 3 + d] = j.k[(k0 + i) * m->dim + d];
      RTB_HIP(hipMemcpyAsync(dk, kp.data(), kp.size() * sizeof(double), hipMemcpyHostToDevice, st));
    }
    if (G) {
      RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
      RTB_TRY(mesh_final(st, m, mp, G, a, j.dir, ch.p0, ch.np, dH, dV));
      RTB_TRY(tmr.end(1));
    } else {
      if (!j.k) {
        RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
        RTB_TRY(mesh_points(st, m->dim, j.kmesh, k0, c, dk));
        RTB_TRY(tmr.end(1));
      }
      RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
      RTB_TRY(fourier_dense(st, a, dk, c, dH, dV));
      RTB_TRY(tmr.end(1));
    }
    if (rterm) {  // the Kubo buffers are free until the eigensolve
      RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
      RTB_TRY(rmatrix_velocity(st, a, m->d_rmat, j.dir, dk, c, dH, dV, rproj, T1, A, B));
      RTB_TRY(tmr.end(1));
    }
    double* Zc = dZ;  // eigenvectors (columns) for the Kubo products
    double* Ab = A;   // scratch of the Kubo products
    if (f32) {
      RTB_TRY(eig_mixed_chunk(st, tmr, !m->herm, n, c, dH, dE, eb, dZ, T1, A, B, dsc, ck.refine()));
      Zc = A;  // refined vectors end in the A buffer; dZ is free
      Ab = dZ;
      // the k-point of local index i, in the order the Fourier stage used
      auto kof = [&](size_t i, double* k3) {
        size_t x[3] = {0, 0, 0};
        if (G) {  // k1 innermost within the chunk's (k2, k3) pairs
          const size_t pair = ch.p0 + i / mp.N[0];
          x[0] = i % mp.N[0];
          x[1] = pair / mp.N[2];
          x[2] = pair % mp.N[2];
        } else if (j.k) {
          for (int d = 0; d < m->dim; ++d) k3[d] = j.k[(k0 + i) * m->dim + d];
          return;
        } else {  // gen_kmesh order: first axis outermost
          size_t f = k0 + i;
          for (int d = m->dim - 1; d >= 0; --d) {
            x[d] = f % j.kmesh[d];
            f /= j.kmesh[d];
          }
        }
        for (int d = 0; d < m->dim; ++d)
          k3[d] = static_cast<double>(x[d]) / static_cast<double>(G ? mp.N[d] : static_cast<int>(j.kmesh[d]));
      };
      RTB_TRY(resolve_flagged(st, w, tmr, m, a, !m->herm, c, kof, ck, dE, Zc, dsc));
    } else {
      RTB_TRY(eigh_launch_fp64(st, &tmr, n, c, dH, dE, true, eb.Vg, eb.tau, eb.Y, eb.S, eb.DE, !m->herm, eb.band));
      RTB_TRY(count_nonfinite(st, n, c, dE, ck.nonfinite));
      RTB_TRY(tmr.begin(RTB_STAGE_BACKTRANS));
      RTB_TRY(eigh_backtransform_fp64(st, n, c, eb.Y, eb.DE, eb.S, eb.Vg, eb.tau, dZ, false, eb.band));
      RTB_TRY(tmr.end(1, flop_backtrans(n, c)));
    }
    if (j.dipole)
      RTB_TRY(dipole_chunk(st, tmr, m, c, dV, Zc, dsc, dE, j.spin_dir, j.eta, j.sym, T1, Ab, B, dA3, dJ, dOm, dOm2,
                           dKern));
    else
      RTB_TRY(kubo_chunk(st, tmr, m, c, dV, Zc, dsc, dE, j.spin_dir, j.eta, T1, Ab, dJ, dOm, exact_kubo ? B : nullptr,
                         dMet));
    int l2 = 0;
    RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
    if (j.mode == 0) {
      RTB_HIP(hipMemcpyAsync(j.omega + k0 * n, dOm, c * n * sizeof(double), hipMemcpyDeviceToHost, st));
      RTB_HIP(hipMemcpyAsync(j.E + k0 * n, dE, c * n * sizeof(double), hipMemcpyDeviceToHost, st));
      if (j.geometry)
        RTB_HIP(hipMemcpyAsync(j.metric + k0 * n, dMet, c * n * sizeof(double), hipMemcpyDeviceToHost, st));
    } else if (j.mode == 1) {
      RTB_TRY(occupied_sum(st, n, c, dOm, dE, j.mu0, j.T, docc));
      ++l2;
      RTB_HIP(hipMemcpyAsync(j.occ + k0, docc, c * sizeof(double), hipMemcpyDeviceToHost, st));
    } else {
      const int nb = hall_blocks(c);
      RTB_TRY(hall_partials(st, n, c, j.dipole ? dKern : dOm, dE, dmu, static_cast<int>(j.n_mu), j.T,
                            dpart + static_cast<size_t>(blk_off) * j.n_mu, nb, j.dipole));
      ++l2;
      if (j.geometry) {
        RTB_TRY(hall_partials(st, n, c, dMet, dE, dmu, static_cast<int>(j.n_mu), j.T,
                              dmpart + static_cast<size_t>(blk_off) * j.n_mu, nb));
        ++l2;
      }
      if (j.n_layers > 0) {  // Zc still holds the eigenvector columns
        RTB_TRY(layer_partials(st, n, c, Zc, f32 ? dsc : nullptr, dE, dOm, dmu, static_cast<int>(j.n_mu), j.T,
                               bb.dperm, bb.dbounds, j.n_layers, bb.dlpart));
        RTB_TRY(reduce_partials(st, bb.dlpart, static_cast<int>(c), static_cast<int>(ml), 1.0,
                                bb.dlsum + chunk_no * ml));
        l2 += 2;
      }
      blk_off += nb;
    }
    RTB_TRY(tmr.end(l2));
    if (j.mode != 2) RTB_HIP(hipStreamSynchronize(st));
    ++chunk_no;
  }
  if (j.mode == 2) {
    RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
    RTB_TRY(reduce_partials(st, dpart, total_blk, static_cast<int>(j.n_mu), 1.0, dsig));
    if (j.geometry) RTB_TRY(reduce_partials(st, dmpart, total_blk, static_cast<int>(j.n_mu), 1.0, dmsig));
    RTB_TRY(tmr.end(j.geometry ? 2 : 1));
    std::vector<double> sums(j.n_mu), msums(j.geometry ? j.n_mu : 0);
    RTB_HIP(hipMemcpyAsync(sums.data(), dsig, j.n_mu * sizeof(double), hipMemcpyDeviceToHost, st));
    if (j.geometry)
      RTB_HIP(hipMemcpyAsync(msums.data(), dmsig, j.n_mu * sizeof(double), hipMemcpyDeviceToHost, st));
    RTB_HIP(hipStreamSynchronize(st));
    // sum / nk / det (response/linear/mod.rs:221; quantum_geometry.rs:219,234)
    for (size_t i = 0; i < j.n_mu; ++i) j.sigma[i] = sums[i] / static_cast<double>(j.nk) / m->det_lat;
    for (size_t i = 0; i < msums.size(); ++i) j.msum[i] = msums[i] / static_cast<double>(j.nk) / m->det_lat;
    if (j.n_layers > 0) {
      RTB_TRY(reduce_partials(st, bb.dlsum, static_cast<int>(chunks.size()), static_cast<int>(ml), 1.0, bb.dlay));
      RTB_HIP(hipMemcpyAsync(j.sigma_layers, bb.dlay, ml * sizeof(double), hipMemcpyDeviceToHost, st));
      RTB_HIP(hipStreamSynchronize(st));
      for (size_t i = 0; i < ml; ++i) j.sigma_layers[i] /= static_cast<double>(j.nk) * m->det_lat;
    }
  }
  RTB_HIP(hipStreamSynchronize(st));
  RTB_TRY(checks_finish(st, ck, tmr));
  RTB_TRY(tmr.collect(j.nk, n));
  return RTB_OK;
}

static int berry_impl(const rtb_model* m, BerryJob& j) {
  Lease lease;
  RTB_TRY(lease.status());
  Work w(lease);
  Timer tmr(lease);
  Exec ex;
  ex.st = lease.stream();
  ex.w = &w;
  ex.tmr = &tmr;
  return berry_run(m, j, ex);
}

int rtb_berry_curvature(const rtb_model* m, const double* k, size_t nk, const double* dir, int spin_dir, double eta,
                        int precision, double* omega, double* E) {
  RTB_TRY(prepare(m));
  RTB_TRY(check_precision(precision));
  // berry_curvature_at (traits.rs:65-77): direction, broadening, spin
  RTB_TRY(check_direction(m, dir));
  RTB_TRY(check_eta(eta));
  RTB_TRY(check_spin(m, spin_dir));
  if (nk > 0 && (!k || !omega || !E)) return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  if (nk == 0) return RTB_OK;
  BerryJob j;
  j.k = k;
  j.nk = nk;
  j.dir = dir;
  j.spin_dir = spin_dir;
  j.eta = eta;
  j.mode = 0;
  j.omega = omega;
  j.E = E;
  j.precision = precision;
  return berry_impl(m, j);
}

int rtb_occupied_berry_curvature(const rtb_model* m, const double* k, size_t nk, const double* dir, int spin_dir,
                                 double eta, double mu, double T, int precision, double* out) {
  RTB_TRY(prepare(m));
  RTB_TRY(check_precision(precision));
  // occupied_berry_curvature_on (traits.rs:109-129)
  RTB_TRY(check_mu(&mu, 1));
  RTB_TRY(check_T(T));
  RTB_TRY(check_direction(m, dir));
  RTB_TRY(check_eta(eta));
  RTB_TRY(check_spin(m, spin_dir));
  if (nk > 0 && (!k || !out)) return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  if (nk == 0) return RTB_OK;
  BerryJob j;
  j.k = k;
  j.nk = nk;
  j.dir = dir;
  j.spin_dir = spin_dir;
  j.eta = eta;
  j.mode = 1;
  j.mu0 = mu;
  j.T = T;
  j.occ = out;
  j.precision = precision;
  return berry_impl(m, j);
}

}  // extern "C"

namespace rtb {

int berry_points_fp64(const rtb_model* m, const double* k, size_t nk, const double* dir, double eta, double* omega,
                      double* E, hipStream_t st, Work& w) {
  if (nk == 0) return RTB_OK;
  BerryJob j;
  j.k = k;
  j.nk = nk;
  j.dir = dir;
  j.eta = eta;
  j.mode = 0;
  j.omega = omega;
  j.E = E;
  j.precision = RTB_PREC_FP64;
  Timer tmr;  // no timing events
  Exec ex;
  ex.st = st;
  ex.w = &w;
  ex.tmr = &tmr;
  return berry_run(m, j, ex);
}

}  // namespace rtb


// The job of hall_conductivity, Integration::Direct, after its validation:
// validate_rank2 (config.rs:141-147) then the spin check (linear/mod.rs:162-165).
static int hall_direct_job(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu, double T,
                           double eta, const double* dir, int spin_dir, int precision, double* sigma, BerryJob& j) {
  RTB_TRY(check_precision(precision));
  size_t nk = 0;
  RTB_TRY(check_kmesh(m, kmesh, &nk));
  RTB_TRY(check_direction(m, dir));
  RTB_TRY(check_mu(mu, n_mu));
  RTB_TRY(check_eta(eta));
  RTB_TRY(check_T(T));
  RTB_TRY(check_spin(m, spin_dir));
  if (!sigma) return fail(RTB_ERR_INVALID_ARGUMENT, "null sigma");
  j.kmesh = kmesh;
  j.nk = nk;
  j.dir = dir;
  j.spin_dir = spin_dir;
  j.eta = eta;
  j.mode = 2;
  j.T = T;
  j.mu = mu;
  j.n_mu = n_mu;
  j.sigma = sigma;
  j.precision = precision;
  return RTB_OK;
}

// Device buffers of the allocation-free entry point: a static device array,
// one copy per GPU the library runs on, held until the process exits.  Host
// code never uses the array's name as an address: each device's copy has its
// own address, taken once per device with hipGetSymbolAddress.
constexpr size_t kNoAllocBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

// the device address of `scratch` on device `dev` (the current device)
int scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= 64) return fail(RTB_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress of the allocation-free scratch failed: %s", hipGetErrorString(e));
      return RTB_ERR_DEVICE;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return RTB_OK;
}
}  // namespace

extern "C" {

int rtb_hall_direct(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu, double T, double eta,
                    const double* dir, int spin_dir, int precision, double* sigma) {
  RTB_TRY(prepare(m));
  BerryJob j;
  RTB_TRY(hall_direct_job(m, kmesh, mu, n_mu, T, eta, dir, spin_dir, precision, sigma, j));
  return berry_impl(m, j);
}

int rtb_hall_direct_layers(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu, double T,
                           double eta, const double* dir, int spin_dir, int precision, int n_layers,
                           const int* orbital_layer, double* sigma, double* sigma_layers) {
  RTB_TRY(prepare(m));
  BerryJob j;
  RTB_TRY(hall_direct_job(m, kmesh, mu, n_mu, T, eta, dir, spin_dir, precision, sigma, j));
  if (n_layers < 1 || !orbital_layer || !sigma_layers)
    return fail(RTB_ERR_INVALID_ARGUMENT, "layers: need n_layers >= 1, orbital_layer and sigma_layers");
  if (static_cast<long long>(n_mu) * n_layers > 1024)
    return fail(RTB_ERR_UNSUPPORTED, "layers: n_mu * n_layers above 1024");
  for (int o = 0; o < m->norb; ++o)
    if (orbital_layer[o] < 0 || orbital_layer[o] >= n_layers)
      return fail(RTB_ERR_INVALID_ARGUMENT, "layers: orbital_layer entries must be in [0, n_layers)");
  // rows of H grouped by layer (both spin copies of an orbital in its layer)
  std::vector<int> perm, bounds(static_cast<size_t>(n_layers) + 1, 0);
  for (int l = 0; l < n_layers; ++l) {
    bounds[l] = static_cast<int>(perm.size());
    for (int r = 0; r < m->nsta; ++r)
      if (orbital_layer[r % m->norb] == l) perm.push_back(r);
  }
  bounds[n_layers] = static_cast<int>(perm.size());
  j.layer_perm = perm.data();
  j.layer_bounds = bounds.data();
  j.n_layers = n_layers;
  j.sigma_layers = sigma_layers;
  return berry_impl(m, j);
}

int rtb_hall_direct_noalloc(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu, double T,
                            double eta, const double* dir, int spin_dir, int precision, double* sigma) {
  RTB_TRY(prepare(m));
  BerryJob j;
  RTB_TRY(hall_direct_job(m, kmesh, mu, n_mu, T, eta, dir, spin_dir, precision, sigma, j));
  void* base = nullptr;
  RTB_TRY(scratch_address(m->device, &base));
  Work w(base, kNoAllocBytes);
  Timer tmr;  // no timing events
  Exec ex;
  ex.st = nullptr;  // the device's null stream
  ex.w = &w;
  ex.tmr = &tmr;
  ex.fixed = true;
  ex.cap = kNoAllocBytes;
  return berry_run(m, j, ex);
}

int rtb_hall_energy_cut(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu, double T, double eta,
                        const double* dir, int spin_dir, double* sigma) {
  RTB_TRY(prepare(m));
  // validate_rank2 and the spin check as the Direct form, then EnergyCut's
  // own checks (linear/mod.rs:177-185)
  size_t nk = 0;
  RTB_TRY(check_kmesh(m, kmesh, &nk));
  RTB_TRY(check_direction(m, dir));
  RTB_TRY(check_mu(mu, n_mu));
  RTB_TRY(check_eta(eta));
  RTB_TRY(check_T(T));
  RTB_TRY(check_spin(m, spin_dir));
  for (size_t i = 0; i + 1 < n_mu; ++i)
    if (mu[i] > mu[i + 1]) return fail(RTB_ERR_INVALID_RESPONSE, "mu: must be sorted in ascending order");
  if (m->dim != 2 && m->dim != 3) {
    set_error("Invalid dimension for operation: %d. Supported dimensions: [2, 3]", m->dim);
    return RTB_ERR_INVALID_DIMENSION;
  }
  if (!sigma) return fail(RTB_ERR_INVALID_ARGUMENT, "null sigma");
  const int n = m->nsta;
  if (n > kMaxN) return fail(RTB_ERR_UNSUPPORTED, "nsta above 256 is not supported by the eigensolver");
  // vertex (ix ny + iy) nz + iz of gen_kmesh's order; a 2D mesh is nz = 1
  const size_t nx = kmesh[0], ny = kmesh[1], nz = m->dim == 3 ? kmesh[2] : 1;
  const size_t n2 = static_cast<size_t>(n) * n;
  Lease lease;
  RTB_TRY(lease.status());
  hipStream_t st = lease.stream();
  Work w(lease);
  Timer tmr(lease);
  FourierArgs a;
  RTB_TRY(upload_dirs(st, w, m, dir, 2, a));
  a.gauge = RTB_GAUGE_ATOM;  // compute_velocity_kernel(.., Gauge::Atom, ..) (linear/mod.rs:235-242)
  a.combine = 0;
  const bool spinj = m->spin && spin_dir != RTB_SPIN_NONE;
  const bool rterm = m->d_rmat != nullptr;
  // resident for the whole mesh: eigenvectors U, Im K and E per vertex, and
  // the squared neighbour overlaps (what the walk reads); per chunk: H, V (2),
  // eigensolver scratch, T1/A/B, J
  // the neighbour pairs of each axis, indexed by the lower vertex: x (i, i +
  // ny nz), y (i, i + nz), z (i, i + 1), every i whose partner is on the mesh
  // (the pairs across a row end are formed and never read)
  const size_t step[3] = {ny * nz, nz, 1};
  const size_t cnt_ax[3] = {nx > 1 ? nk - step[0] : 0, ny > 1 ? nk - step[1] : 0, nz > 1 ? nk - step[2] : 0};
  double *dU, *dKv, *dEv, *dOv[3] = {nullptr, nullptr, nullptr};
  RTB_TRY(w.alloc(&dU, nk * n2 * 2));
  RTB_TRY(w.alloc(&dKv, nk * n2));
  RTB_TRY(w.alloc(&dEv, nk * n));
  for (int ax = 0; ax < 3; ++ax)
    if (cnt_ax[ax]) RTB_TRY(w.alloc(&dOv[ax], cnt_ax[ax] * n2));
  const size_t per_k = 24 + n2 * 16 * 3 + n2 * 48 + n2 * 48 + (spinj ? n2 * 16 : 0) + n * 56 +
                       eig_extra_bytes(n, true);
  const size_t K = chunk_for(nk, per_k);
  double *dk, *dH, *dV, *kub, *dJ = nullptr, *rproj = nullptr;
  EigBufs eb;
  RTB_TRY(w.alloc(&dk, K * 3));
  RTB_TRY(w.alloc(&dH, K * n2 * 2));
  RTB_TRY(w.alloc(&dV, K * n2 * 4));
  RTB_TRY(alloc_eig(w, n, K, true, eb));
  RTB_TRY(w.alloc(&kub, K * n2 * 6));
  if (spinj) RTB_TRY(w.alloc(&dJ, K * n2 * 2));
  if (rterm) RTB_TRY(w.alloc(&rproj, static_cast<size_t>(m->nR) * n2 * 2));
  double* T1 = kub;
  double* A = kub + K * n2 * 2;
  double* B = kub + K * n2 * 4;
  // the tracking walk's schedule (mesh only) for the device walk, uploaded
  // before the vertex stage; RTB_TRACK_HOST=1 takes the host walk instead
  const char* th = std::getenv("RTB_TRACK_HOST");
  const bool dev_track = n <= kTrackDevMaxN && nk <= 0xffffffffull && !(th && std::atoi(th) == 1);
  TrackSchedule ts;
  TrackEntry* sched = nullptr;
  if (dev_track) {
    energy_cut_schedule(nx, ny, nz, ts);
    std::vector<TrackEntry> he(nk);  // level order: vertex, count, visited neighbours
    for (size_t i = 0; i < nk; ++i) {
      const size_t v = ts.by_level[i];
      TrackEntry& e = he[i];
      e.v = static_cast<unsigned>(v);
      e.cnt = ts.ndep[v];
      for (int r = 0; r < 6; ++r) e.dep[r] = r < static_cast<int>(e.cnt) ? static_cast<unsigned>(ts.dep[v * 6 + r]) : 0u;
    }
    RTB_TRY(w.alloc(&sched, nk));
    RTB_HIP(hipMemcpyAsync(sched, he.data(), nk * sizeof(TrackEntry), hipMemcpyHostToDevice, st));
    RTB_HIP(hipStreamSynchronize(st));  // the host vector goes out of scope here
  }
  Checks ck;
  RTB_TRY(checks_init(st, w, 0, ck));
  for (size_t k0 = 0; k0 < nk; k0 += K) {
    const size_t c = std::min(K, nk - k0);
    double* Zc = dU + k0 * n2 * 2;
    RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
    RTB_TRY(mesh_points(st, m->dim, kmesh, k0, c, dk));
    RTB_TRY(fourier_dense(st, a, dk, c, dH, dV));
    if (rterm) RTB_TRY(rmatrix_velocity(st, a, m->d_rmat, dir, dk, c, dH, dV, rproj, T1, A, B));
    RTB_TRY(tmr.end(rterm ? 3 : 2));
    // hamk.eigh(UPLO::Lower) (primitives.rs:87): the upper triangle of a
    // storage that is not exactly Hermitian
    RTB_TRY(eigh_launch_fp64(st, &tmr, n, c, dH, dEv + k0 * n, true, eb.Vg, eb.tau, eb.Y, eb.S, eb.DE, !m->herm,
                             eb.band));
    RTB_TRY(count_nonfinite(st, n, c, dEv + k0 * n, ck.nonfinite));
    RTB_TRY(tmr.begin(RTB_STAGE_BACKTRANS));
    RTB_TRY(eigh_backtransform_fp64(st, n, c, eb.Y, eb.DE, eb.S, eb.Vg, eb.tau, Zc, false, eb.band));
    RTB_TRY(tmr.end(1, flop_backtrans(n, c)));
    // v^a = U^H J_a U, v^b = U^H v_b U (the reference's evec^T v evec^* with
    // evec = conj(U)), the spin current on the first direction only
    const double* Ma = dV;
    size_t sMa = 2 * n2;
    if (spinj) {
      RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
      RTB_TRY(spin_current(st, m->norb, c, spin_dir, dV, 2 * n2 * 2, nullptr, dJ));
      RTB_TRY(tmr.end(1));
      Ma = dJ;
      sMa = n2;
    }
    RTB_TRY(tmr.begin(RTB_STAGE_GEMM64));
    RTB_TRY(zgemm_batched(st, n, c, Ma, sMa, false, Zc, n2, T1, n2));
    RTB_TRY(zgemm_batched(st, n, c, Zc, n2, true, T1, n2, A, n2));
    RTB_TRY(zgemm_batched(st, n, c, dV + n2 * 2, 2 * n2, false, Zc, n2, T1, n2));
    RTB_TRY(zgemm_batched(st, n, c, Zc, n2, true, T1, n2, B, n2));
    RTB_TRY(tmr.end(4, 4 * flop_zgemm(n, c), 4 * mflop_zgemm(n, c, false)));
    RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
    RTB_TRY(energy_cut_vertex(st, n, c, A, B, dKv + k0 * n2));
    RTB_TRY(tmr.end(1));
  }
  // overlaps of mesh neighbours U_i^H U_j in pieces of at most K products
  // through T1, then their squared moduli
  const size_t piece = std::min<size_t>(K, 32768);
  for (int ax = 0; ax < 3; ++ax) {
    for (size_t i0 = 0; i0 < cnt_ax[ax]; i0 += piece) {
      const size_t c = std::min(piece, cnt_ax[ax] - i0);
      RTB_TRY(tmr.begin(RTB_STAGE_GEMM64));
      RTB_TRY(zgemm_batched(st, n, c, dU + i0 * n2 * 2, n2, true, dU + (i0 + step[ax]) * n2 * 2, n2, T1, n2));
      RTB_TRY(tmr.end(1, flop_zgemm(n, c), mflop_zgemm(n, c, false)));
      RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
      RTB_TRY(abs2(st, c * n2, T1, dOv[ax] + i0 * n2));
      RTB_TRY(tmr.end(1));
    }
  }
  // band tracking: on the device, level by level over the walk's schedule
  // (which depends on the mesh only); with RTB_TRACK_HOST=1 the squared
  // overlaps go to the host walk.  Either way the permutations are the same.
  int* dperm;
  RTB_TRY(w.alloc(&dperm, nk * n));
  std::vector<int> perm;
  if (dev_track) {
    // above kTrackLdsMaxN bands each vertex's overlap matrix is in a
    // workspace slice: at most 1024 vertices of a level per launch
    // (RTB_TRACK_SLICES=s sets another bound, s >= 1)
    double* mwork = nullptr;
    size_t mblocks = 0;
    if (n > kTrackLdsMaxN) {
      for (size_t l = 0; l + 1 < ts.level_start.size(); ++l)
        mblocks = std::max(mblocks, ts.level_start[l + 1] - ts.level_start[l]);
      const char* sl = std::getenv("RTB_TRACK_SLICES");
      const long cap = sl ? std::atol(sl) : 1024;
      if (cap < 1) return fail(RTB_ERR_INVALID_ARGUMENT, "RTB_TRACK_SLICES must be at least 1");
      mblocks = std::min<size_t>(mblocks, static_cast<size_t>(cap));
      RTB_TRY(w.alloc(&mwork, mblocks * n2));
    }
    RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
    RTB_TRY(energy_cut_track_dev(st, nx, ny, nz, n, dOv[0], dOv[1], dOv[2], sched, ts.level_start, mwork, mblocks,
                                 dperm));
    size_t launches = 1;  // the identity permutations, then each level's launches
    for (size_t l = 1; l + 1 < ts.level_start.size(); ++l) {
      const size_t lw = ts.level_start[l + 1] - ts.level_start[l];
      launches += mblocks ? (lw + mblocks - 1) / mblocks : (lw ? 1 : 0);
    }
    RTB_TRY(tmr.end(static_cast<int>(launches)));
  } else {
    std::vector<double> hOv[3];
    for (int ax = 0; ax < 3; ++ax) {
      hOv[ax].resize(cnt_ax[ax] * n2);
      if (!hOv[ax].empty())
        RTB_HIP(hipMemcpyAsync(hOv[ax].data(), dOv[ax], hOv[ax].size() * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    RTB_HIP(hipStreamSynchronize(st));
    energy_cut_track(nx, ny, nz, n, hOv[0].data(), hOv[1].data(), hOv[2].data(), true, perm);
    RTB_HIP(hipMemcpyAsync(dperm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, st));
  }
  // vertex Omega in tracked band order, then the cut on the device (E and
  // Im K stayed on the device)
  double* dom;
  RTB_TRY(w.alloc(&dom, nk * n));
  RTB_TRY(tmr.begin(RTB_STAGE_KUBO));
  RTB_TRY(energy_cut_omega(st, n, nk, dEv, dKv, dperm, eta, dom));
  RTB_TRY(tmr.end(1));
  const size_t waves = energy_cut_waves(m->dim, nx, ny, nz, n);
  const size_t chunk = std::max<size_t>(1, std::min<size_t>(256, (size_t(512) << 20) / (waves * sizeof(double))));
  double* dpart;
  RTB_TRY(w.alloc(&dpart, waves * chunk));
  const double width = T <= 0.0 ? 0.0 : T * 8.617333262145e-5;  // thermodynamics.rs:10, 50-59
  RTB_TRY(energy_cut_convolve(mu, n_mu, width, [&](const double* mv, size_t nm, double* s0) -> int {
    double *dmu, *ds0;
    RTB_TRY(w.alloc(&dmu, nm));
    RTB_TRY(w.alloc(&ds0, nm));
    RTB_HIP(hipMemcpyAsync(dmu, mv, nm * sizeof(double), hipMemcpyHostToDevice, st));
    RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
    RTB_TRY(energy_cut_t0(st, m->dim, nx, ny, nz, n, dEv, dKv, dperm, dom, dmu, nm, eta, dpart, chunk, ds0));
    RTB_TRY(tmr.end(static_cast<int>(2 * ((nm + chunk - 1) / chunk))));
    RTB_HIP(hipMemcpyAsync(s0, ds0, nm * sizeof(double), hipMemcpyDeviceToHost, st));
    RTB_HIP(hipStreamSynchronize(st));
    return RTB_OK;
  }, sigma));
  RTB_TRY(checks_finish(st, ck, tmr));
  RTB_TRY(tmr.collect(nk, n));
  for (size_t i = 0; i < n_mu; ++i) sigma[i] /= m->det_lat;  // linear/mod.rs:267
  return RTB_OK;
}

int rtb_quantum_geometry(const rtb_model* m, const double* k, size_t nk, const double* dir, double eta,
                         int precision, double* metric, double* omega, double* E) {
  RTB_TRY(prepare(m));
  RTB_TRY(check_precision(precision));
  // quantum_geometry_on (quantum_geometry.rs:117-118): direction, broadening
  RTB_TRY(check_direction(m, dir));
  RTB_TRY(check_eta(eta));
  if (nk > 0 && (!k || !metric || !omega || !E)) return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  if (nk == 0) return RTB_OK;
  BerryJob j;
  j.k = k;
  j.nk = nk;
  j.dir = dir;
  j.eta = eta;
  j.mode = 0;
  j.omega = omega;
  j.E = E;
  j.geometry = true;
  j.metric = metric;
  j.precision = precision;
  return berry_impl(m, j);
}

int rtb_quantum_geometry_direct(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu, double T,
                                double eta, const double* dir, int precision, double* metric, double* berry) {
  RTB_TRY(prepare(m));
  RTB_TRY(check_precision(precision));
  // validate_rank2 (config.rs:141-147), as Model::quantum_geometry (quantum_geometry.rs:196)
  size_t nk = 0;
  RTB_TRY(check_kmesh(m, kmesh, &nk));
  RTB_TRY(check_direction(m, dir));
  RTB_TRY(check_mu(mu, n_mu));
  RTB_TRY(check_eta(eta));
  RTB_TRY(check_T(T));
  if (!metric || !berry) return fail(RTB_ERR_INVALID_ARGUMENT, "null output");
  BerryJob j;
  j.kmesh = kmesh;
  j.nk = nk;
  j.dir = dir;
  j.eta = eta;
  j.mode = 2;
  j.T = T;
  j.mu = mu;
  j.n_mu = n_mu;
  j.sigma = berry;
  j.geometry = true;
  j.msum = metric;
  j.precision = precision;
  return berry_impl(m, j);
}

int rtb_nonlinear_hall_direct(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu, double T,
                              double eta, const double* dir, int spin_dir, int symmetrized, int precision,
                              double* sigma) {
  RTB_TRY(prepare(m));
  RTB_TRY(check_precision(precision));
  // extrinsic_nonlinear_hall (response/nonlinear/mod.rs:259-266): validate_rank3
  // (config.rs:153-158), broadening, spin; the Direct sum needs T > 0 (:318-326)
  size_t nk = 0;
  RTB_TRY(check_kmesh(m, kmesh, &nk));
  RTB_TRY(check_direction(m, dir, 3));
  RTB_TRY(check_mu(mu, n_mu));
  RTB_TRY(check_T(T));
  RTB_TRY(check_eta(eta));
  RTB_TRY(check_spin(m, spin_dir));
  if (!(T > 0.0))
    return fail(RTB_ERR_INVALID_THERMO, "T: direct nonlinear Hall integration requires a finite temperature");
  if (!sigma) return fail(RTB_ERR_INVALID_ARGUMENT, "null sigma");
  // symmetrized only when the two field rows differ (mod.rs:284-285)
  bool differ = false;
  for (int d = 0; d < m->dim; ++d) differ = differ || dir[m->dim + d] != dir[2 * m->dim + d];
  BerryJob j;
  j.kmesh = kmesh;
  j.nk = nk;
  j.dir = dir;
  j.spin_dir = spin_dir;
  j.eta = eta;
  j.mode = 2;
  j.T = T;
  j.mu = mu;
  j.n_mu = n_mu;
  j.sigma = sigma;
  j.dipole = true;
  j.sym = symmetrized != 0 && differ;
  j.precision = precision;
  return berry_impl(m, j);
}

int rtb_dos(const rtb_model* m, const size_t* kmesh, double Emin, double Emax, size_t En, double sigma,
            int precision, double* E, double* rho) {
  RTB_TRY(prepare(m));
  RTB_TRY(check_precision(precision));
  // model_physics.rs:187-205
  if (Emin >= Emax) return fail(RTB_ERR_INVALID_ENERGY_RANGE, "E_min must be below E_max");
  if (En == 0) return fail(RTB_ERR_INVALID_DOS_PARAMETER, "E_n: number of energy bins must be at least 1");
  if (!std::isfinite(sigma) || sigma <= 0.0)
    return fail(RTB_ERR_INVALID_DOS_PARAMETER, "sigma: Gaussian smearing width must be finite and positive");
  size_t nk = 0;
  RTB_TRY(check_kmesh(m, kmesh, &nk));
  if (!E || !rho) return fail(RTB_ERR_INVALID_ARGUMENT, "null output");
  const int n = m->nsta;
  if (n > kMaxN) return fail(RTB_ERR_UNSUPPORTED, "nsta above 256 is not supported by the eigensolver");
  const double step = En > 1 ? (Emax - Emin) / static_cast<double>(En - 1) : 0.0;
  for (size_t i = 0; i < En; ++i) E[i] = Emin + step * static_cast<double>(i);
  Lease lease;
  RTB_TRY(lease.status());
  hipStream_t st = lease.stream();
  Work w(lease);
  Timer tmr(lease);
  FourierArgs a;
  RTB_TRY(upload_dirs(st, w, m, nullptr, 0, a));
  a.gauge = RTB_GAUGE_ATOM;
  const size_t n2 = static_cast<size_t>(n) * n;
  const size_t K = chunk_for(nk, 24 + n2 * 32 + n * 32 + eig_extra_bytes(n, false));
  MeshPlan mp;
  double* G = nullptr;
  std::vector<Chunk> chunks;
  RTB_TRY(mesh_setup(st, w, tmr, m, kmesh, precision, nk, K, false, mp, &G, chunks));
  double *dk, *dH, *dE;
  EigBufs eb;
  RTB_TRY(w.alloc(&dk, K * 3));
  RTB_TRY(w.alloc(&dH, K * n2 * 2));
  RTB_TRY(w.alloc(&dE, K * n));
  RTB_TRY(alloc_eig(w, n, K, false, eb));
  int total_blk = 0;
  for (const Chunk& ch : chunks) total_blk += dos_blocks(ch.c * n);
  double *dpart, *drho;
  RTB_TRY(w.alloc(&dpart, static_cast<size_t>(total_blk) * En));
  RTB_TRY(w.alloc(&drho, En));
  Checks ck;
  RTB_TRY(checks_init(st, w, 0, ck));
  int blk_off = 0;
  for (const Chunk& ch : chunks) {
    const size_t k0 = ch.k0, c = ch.c;
    if (G) {
      RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
      RTB_TRY(mesh_final(st, m, mp, G, a, nullptr, ch.p0, ch.np, dH, nullptr));
      RTB_TRY(tmr.end(1));
    } else {
      RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
      RTB_TRY(mesh_points(st, m->dim, kmesh, k0, c, dk));
      RTB_TRY(tmr.end(1));
      RTB_TRY(tmr.begin(RTB_STAGE_FOURIER));
      RTB_TRY(fourier_dense(st, a, dk, c, dH, nullptr));
      RTB_TRY(tmr.end(1));
    }
    RTB_TRY(eigh_launch_fp64(st, &tmr, n, c, dH, dE, false, eb.Vg, nullptr, nullptr, nullptr, eb.DE, false, eb.band));
    RTB_TRY(count_nonfinite(st, n, c, dE, ck.nonfinite));
    RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
    const int nb = dos_blocks(c * n);
    RTB_TRY(dos_partials(st, c * n, dE, Emin, step, static_cast<int>(En), sigma,
                         dpart + static_cast<size_t>(blk_off) * En, nb));
    RTB_TRY(tmr.end(1));
    blk_off += nb;
  }
  RTB_TRY(tmr.begin(RTB_STAGE_REDUCE));
  RTB_TRY(reduce_partials(st, dpart, total_blk, static_cast<int>(En), 1.0, drho));
  RTB_TRY(tmr.end(1));
  RTB_HIP(hipMemcpyAsync(rho, drho, En * sizeof(double), hipMemcpyDeviceToHost, st));
  RTB_HIP(hipStreamSynchronize(st));
  for (size_t i = 0; i < En; ++i) rho[i] /= static_cast<double>(nk);  // model_physics.rs:224
  RTB_TRY(checks_finish(st, ck, tmr));
  RTB_TRY(tmr.collect(nk, n));
  return RTB_OK;
}

}  // extern "C"
