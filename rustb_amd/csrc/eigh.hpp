// Launch wrappers for the hot-path kernels (one header for all units).
#pragma once

#include "rtb_internal.hpp"

namespace rtb {

// ---- eigensolver (eigh.hip) ----
// Householder + bisection (+ the tridiagonal eigenvector data when
// `vectors`: twisted factorisations, inverse iteration inside clusters).
// The matrix is the Hermitian one of H's lower triangle (with Re of the
// diagonal), as the reference's eigvalsh_v reads it; with from_upper H is
// first rebuilt in place from its upper triangle, as the reference's eigh
// reads it (the two agree to rounding when the model storage is exactly
// Hermitian, so callers ask for it only for the other models).
//   H: [nmat][n][n] complex; E: [nmat][n]; Vg: [nmat][n][n] complex;
//   taug: [nmat][n] complex; Y: [nmat][n][n]; S: [nmat][3][n][n];
//   DE: [3][nmat][n] the tridiagonal (diagonal, off-diagonal) and the twist
//   indices; Vg, taug, Y, S, DE are the inputs of the back-transformation.
//   n > 128 (two-stage route): Vg also holds the matrix being reduced (also
//   without vectors), and `band` ([nmat][eigh_band_elems(n)] in T) the
//   reflectors of both stages; null for n <= 128.
//   nb (n <= 128): rows of the register route's block, 32, 64 or 128 and at
//   least n, or 0 for the smallest that holds n (eigh_block_valid).
int eigh_launch_fp64(hipStream_t st, Timer* tmr, int n, size_t nmat, double* H, double* E, bool vectors,
                     double* Vg, double* taug, double* Y, double* S, double* DE, bool from_upper, double* band,
                     int nb = 0);
bool eigh_block_valid(int n, int nb);
// FP32 route (mixed form): H read as FP64 and rounded; E written as FP64.
int eigh_launch_fp32(hipStream_t st, Timer* tmr, int n, size_t nmat, double* H, double* E, bool vectors,
                     float* Vg, float* taug, float* Y, float* S, float* DE, bool from_upper, float* band,
                     int nb = 0);
// Z: [nmat][n][n] complex FP32 (the values of the FP32 route, exact);
// gdiag (optional, [nmat][n]): the squared column norms of Z in FP64.
int eigh_backtransform_fp32(hipStream_t st, int n, size_t nmat, const float* Y, const float* DE, const float* S,
                            const float* Vg, const float* taug, float* Z, bool rows_out, double* gdiag,
                            const float* band, int nb = 0);
// H[k] (complex [nmat][n][n]) rebuilt in place from its upper triangle as a
// Hermitian matrix (what from_upper does before a solve).
int hermitize_upper(hipStream_t st, int n, size_t nmat, double* H);
// Real elements of T per matrix the two-stage route keeps in `band` (0 for
// n <= 128).
size_t eigh_band_elems(int n);
// Raises the dynamic LDS limits of the two-stage route's kernels (once).
int eigh_set_smem_limits();
// FP64 refinement of FP32 eigenpairs (refine.hip): one Ogita-Aishima step,
// Rayleigh quotients, unit columns, Rayleigh-Ritz in clusters with gaps
// below gap_rel * max(1, spectral radius).  Z: [nk][n][n] FP32 columns (in);
// H is overwritten; the refined columns are written to W2 (FP64), unit
// after multiplication by scale[k][j] ([nk][n]: the squared column norms of
// Z on entry, as eigh_backtransform_fp32 writes them); W1, W3 scratch.
//
// Check: per eigenpair the residual |H u - lambda u| of the unit refined
// vector u (from the correction product's epilogue, recomputed after the
// Rayleigh-Ritz rotation inside clusters); a k-point with an eigenvalue that
// is not finite, a cluster the Rayleigh-Ritz step cannot handle, or a
// residual above res_tol * max(1, max |lambda|) eV gets its local index
// appended to fail_list (fail_count: an int the caller zeroes); max_res
// (the bits of a non-negative double, zeroed by the caller) receives the
// largest finite residual.
struct RefineCheck {
  double res_tol = 1e-9;
  int* fail_count = nullptr;
  int* fail_list = nullptr;
  unsigned long long* max_res = nullptr;
};
int refine_eigenpairs(hipStream_t st, Timer* tmr, int n, size_t nk, double* H, const float* Z, double* W1,
                      double* W2, double* W3, double* E, double* scale, double gap_rel, const RefineCheck& chk);
// Count of k-points ([nk][n] eigenvalues) with an eigenvalue that is not
// finite, added to *count.
int count_nonfinite(hipStream_t st, int n, size_t nk, const double* E, int* count);
// For i < cnt: E[list[i]] = Ef[i], Z[list[i]] = Zf[i] ([n][n] complex
// columns), scale[list[i]][:] = 1 (the FP64 re-solve of flagged k-points).
int scatter_resolved(hipStream_t st, int n, int cnt, const int* list, const double* Ef, const double* Zf,
                     double* E, double* Z, double* scale);
// p[0..count) = 0 (a fill kernel: usable where hipMemsetAsync is not).
int zero_ints(hipStream_t st, int* p, size_t count);
// out[k][r][c] = Z[k][c][r] * scale[k][r] (eigenvector columns to rows;
// scale may be null)
int transpose_batched(hipStream_t st, int n, size_t nk, const double* Z, const double* scale, double* out);
// Z = Q Y (complex, [nmat][n][n]); rows_out stores eigenvectors as rows.
int eigh_backtransform_fp64(hipStream_t st, int n, size_t nmat, const double* Y, const double* DE, const double* S,
                            const double* Vg, const double* taug, double* Z, bool rows_out, const double* band,
                            int nb = 0);

// ---- Fourier sum (fourier.hip) ----
struct FourierArgs {
  int dim, n, nR, np;     // np: number of projected velocity directions (0..3)
  int gauge;              // RTB_GAUGE_*
  int combine = 0;        // np == 2: write V0 + i V1 as one matrix per k
  const double* ham;      // [nR][n][n] complex
  const double* R;        // [nR][3]
  const double* wR;       // [np][nR]  w_p . R_cart
  const double* tau;      // [n][3]    fractional
  const double* wtau;     // [np][n]   w_p . tau_cart
};
// k: [nk][3] (zero padded); H: [nk][n][n]; V: [nk][np][n][n] (or null when
// np == 0; [nk][n][n] with a.combine)
int fourier_dense(hipStream_t st, const FourierArgs& a, const double* k, size_t nk, double* H,
                  double* V);
// Many models' H and two projected velocities on their uniform meshes in one
// persistent launch over a table of tiles (fourier.hip).  QueueModel: the
// Fourier arguments of a model (np == 2) and its mesh; QueueTile: cnt
// k-points from flat mesh index f0 (gen_kmesh order) of model `model`,
// elements [e0, e0 + 256) (all n^2 when n^2 <= 256), outputs H [cnt][n][n]
// and V [cnt][2][n][n] ([cnt][n][n] with combine) from the given pointers.
// Tiles hold at most fourier_queue_groups(n, tk) * tk k-points; workgroups
// take tiles from *next (zeroed beforehand) in table order; max_groups: the
// largest group count among the launch's tiles.
constexpr int kQueueSlots = 64;  // k-points per tile at most
struct QueueModel {
  FourierArgs a;
  int N[3];
};
struct QueueTile {
  int model;
  int e0;
  long long f0;
  int cnt;
  double* H;
  double* V;
};
int fourier_queue_groups(int n, int tk);
int fourier_queue(hipStream_t st, int tk, const QueueModel* models, const QueueTile* tiles, int ntiles,
                  int max_groups, unsigned* next, int workgroups);
// Workgroups of the queue kernel that fit one CU at once (dynamic LDS of
// max_groups * tk k-point slots).
int fourier_queue_occupancy(int tk, int max_groups, int* per_cu);
// The position-matrix term of the projected velocities (velocity.rs:377-402,
// 627-668): V[k][p] += i [H(k), r_p(k)] with r_p(k) the Bloch sum (same
// phases and gauge as H) of sum_d dirs[p][d] rmat[R][d] and, in the Atom
// gauge, its diagonal zeroed.  rmat: [nR][dim][n][n]; V: [nk][np][n][n] (not
// combined); rproj: nR n^2 complex scratch; Rk, T1, T2: nk n^2 complex.
int rmatrix_velocity(hipStream_t st, const FourierArgs& a, const double* rmat, const double* dirs, const double* k,
                     size_t nk, const double* H, double* V, double* rproj, double* Rk, double* T1, double* T2);
// Axis-separable Bloch sum on a uniform mesh (fourier_mesh.hip).
struct MeshPlan {
  int dim = 0;
  int lo[3], L[3], N[3];
  size_t box = 0, P = 0;               // box cells, pairs (N2*N3)
  size_t box_elems = 0, ga_elems = 0, gb_elems = 0;  // complex element counts
  bool ok = false;
};
int mesh_plan(const rtb_model* m, const size_t* kmesh, MeshPlan& p);
// passes A/B: work >= box_elems + ga_elems complex; G = gb_elems complex
int mesh_prepare(hipStream_t st, const rtb_model* m, const MeshPlan& p, int* d_boxidx, double* work, double* G,
                 bool want_v);
// The same from a device copy of the flat box index of every stored R
// (mesh_box_index); kernels only (no copies from the host, no memset).
void mesh_box_index(const rtb_model* m, const MeshPlan& p, std::vector<int>& idx);
int mesh_prepare_dev(hipStream_t st, const rtb_model* m, const MeshPlan& p, const int* d_boxidx, double* work,
                     double* G, bool want_v);
// pass C for pairs [p0, p0 + np_pairs): H [np_pairs*N1][n][n], V [..][2][n][n]
// ([..][n][n] holding V0 + i V1 with a.combine) in k1-innermost order: local
// k = (pair - p0) * N1 + i1
int mesh_final(hipStream_t st, const rtb_model* m, const MeshPlan& p, const double* G, const FourierArgs& a,
               const double* dirs, size_t p0, size_t np_pairs, double* H, double* V);
int mesh_set_smem_limits();
// Uniform mesh points k = (i/N1, j/N2, l/N3) for flat indices [k0, k0 + nk).
int mesh_points(hipStream_t st, int dim, const size_t* kmesh, size_t k0, size_t nk, double* k);

// ---- batched complex products (zgemm.hip) ----
// A batched complex operand, (re, im) pairs in FP64 or in FP32.
struct ZOperand {
  const void* p = nullptr;
  size_t stride = 0;  // complex elements between consecutive matrices
  bool f32 = false;
};
inline ZOperand zop(const double* p, size_t stride) { return {p, stride, false}; }
inline ZOperand zop(const float* p, size_t stride) { return {p, stride, true}; }
// C = op(A) B (FP64 arithmetic), diag(s) op(A) B diag(s) when s ([nb][n]) is
// given; herm: the product is Hermitian (lower tiles computed, mirrored).
int zgemm(hipStream_t st, int n, size_t nb, ZOperand A, bool conj_a, ZOperand B, double* C, size_t sC, bool herm,
          const double* s);
// The Gram product G = Z^H Z (FP64, Z stored in FP32, stride n^2) of the
// Ogita-Aishima step with G consumed in the epilogue: F (FP32, [nb][n][n])
// receives the correction F' - I from G, S = Z^H H Z ([nb][n][n]) and the
// exact diagonal gd ([nb][n]) of G (n <= 128); lt ([nb][n]) the Rayleigh
// quotients S_ii / G_ii.
int zgemm_gram_oa(hipStream_t st, int n, size_t nb, const float* Z, const double* S, const double* gd, float* F,
                  double gap_rel, double* lt);
// The Kubo product C = diag(s) U^H T1 diag(s) (U, T1: [nb][n][n] FP64; s
// may be null) consumed by its epilogue: with Q_ij = |C_ij|^2 /
// ((E_i - E_j)^2 + eta^2), i != j, part ([nb][2][ceil(n/64)][n]) receives the
// per-tile column sums ([0][tile row]) and row sums ([1][tile column]) of Q.
int zgemm_kubo(hipStream_t st, int n, size_t nb, const double* U, const double* T1, const double* s,
               const double* E, double eta, double* part);
// Both refinement corrections in one pass: Zp = Z + Z F (FP64 out) and the
// Rayleigh sums of the Zp columns against HZ + HZ F (not stored) with the
// squared residuals |HZ + HZ F - lt_j Zp_j|^2, per 64-row tile into part
// ([nb][ceil(n/64)][3][n]); products on the FP32 matrix cores.
int zgemm_corrections_f32(hipStream_t st, int n, size_t nb, const float* Z, const double* HZ, const float* F,
                          const double* lt, double* Zp, double* part);
// C[b] = op(A[b]) * B[b], n x n complex; conj_a: op = conjugate transpose.
// sA, sB, sC: batch strides in complex elements.  herm: the product is known
// to be Hermitian; the upper triangle is filled from the lower one.
int zgemm_batched(hipStream_t st, int n, size_t nb, const double* A, size_t sA, bool conj_a, const double* B,
                  size_t sB, double* C, size_t sC, bool herm = false);
// The same product with C[b][i][j] multiplied by s[b][i] s[b][j] (s: [nb][n],
// stride n): diag(s) op(A) B diag(s), e.g. U^H J U for unit columns U diag(s).
int zgemm_batched_scaled(hipStream_t st, int n, size_t nb, const double* A, size_t sA, bool conj_a, const double* B,
                         size_t sB, double* C, size_t sC, bool herm, const double* s);

// ---- Kubo band sum, reductions, DOS (kubo.hip) ----
// J = (S v + v S) / 2 for the spin matrix S = sigma/2 (x) I (helpers.rs:11-39);
// with w (stride v_stride like v) the output is J + i w.
int spin_current(hipStream_t st, int norb, size_t nb, int spin_dir, const double* v, size_t v_stride,
                 const double* w, double* J);
// Omega_n = sum_{m != n} -2 Im(A_nm B_mn) / ((E_n - E_m)^2 + eta^2) for
// Hermitian A, B, from C = A + i B (Im(A_nm B_mn) = (|C_nm|^2 - |C_mn|^2) / 4):
// (sum over tiles of the column sums - sum of the row sums) / 2 from the
// partial sums zgemm_kubo writes (tiles = ceil(n/64)).
int kubo_omega_from_parts(hipStream_t st, int n, size_t nk, const double* part, double* omega);
// The same band sum from A and B themselves ([nk][n][n] each, not assumed
// Hermitian): Omega_n = sum_{m != n} -2 Im(A_nm B_mn) / ((E_n - E_m)^2 + eta^2);
// metric (optional): g_n = sum_{m != n} Re(A_nm B_mn) / ((E_n - E_m)^2 + eta^2).
int kubo_omega_exact(hipStream_t st, int n, size_t nk, const double* A, const double* B, const double* E, double eta,
                     double* omega, double* metric);
// out[k] = sum_n omega[k][n] f(E[k][n], mu, T)
int occupied_sum(hipStream_t st, int n, size_t nk, const double* omega, const double* E, double mu,
                 double T, double* out);
// partial[blk][imu] += sum_{k in blk} sum_n omega f(E, mu_imu)   (fixed blocking)
// deriv: the weights are -df/dE = f (1 - f) / (k_B T) instead of f (T > 0).
int hall_partials(hipStream_t st, int n, size_t nk, const double* omega, const double* E,
                  const double* mu, int n_mu, double T, double* partial, int n_blk, bool deriv = false);
// Berry curvature dipole kernel per band: out = om1 * Re(A3_nn), or with sym
// (om1 * Re(A3_nn) + om2 * Re(A2_nn)) / 2 (A2, A3: [nk][n][n], om: [nk][n]).
int dipole_combine(hipStream_t st, int n, size_t nk, const double* A2, const double* A3, const double* om1,
                   const double* om2, bool sym, double* out);
int hall_blocks(size_t nk);
// out[k][m] = sum_n omega[k][n] f(E[k][n], mu_m, T) per k-point ([nk][n_mu]).
int kterms(hipStream_t st, int n, size_t nk, const double* omega, const double* E, const double* mu, int n_mu,
           double T, double* out);
// out[s][m] = sum over k in [start[s], start[s + 1]) of term[k][m] (start:
// device, n_seg + 1 entries), in a fixed order.
int segment_sums(hipStream_t st, const double* term, const long long* start, int n_seg, int n_mu, double* out);
// Layer-resolved occupied curvature: partial[k][m][l] = sum_j f(E_j, mu_m)
// Omega_j sum_{i in layer l} |U_ij|^2 (U = Z diag(scale), Z: [nk][n][n]
// columns; the rows of layer l are perm[bounds[l] .. bounds[l+1])).
int layer_partials(hipStream_t st, int n, size_t nk, const double* Z, const double* scale, const double* E,
                   const double* omega, const double* mu, int n_mu, double T, const int* perm, const int* bounds,
                   int n_layers, double* partial);
// EnergyCut Fermi cut on the device (energy_cut_dev.hip): vertex Omega in tracked
// band order (om [nk][n]; E [nk][n], Ki = Im K [nk][n][n], perm [nk][n]), and
// the zero-width cut of every (cell, simplex, band) for the sorted mu (device,
// n_mu) into out (device, n_mu), in launches of `chunk` chemical potentials
// with partial [energy_cut_waves(...)][chunk] of scratch.
int energy_cut_omega(hipStream_t st, int n, size_t nk, const double* E, const double* Ki, const int* perm, double eta,
                     double* om);
size_t energy_cut_waves(int dim, size_t nx, size_t ny, size_t nz, int n);
// Band tracking on the device for n <= kTrackDevMaxN: the walk of
// energy_cut_track over the schedule of energy_cut_schedule, given as one
// entry per vertex in level order (the vertex, its count of visited
// neighbours and those neighbours, 32-bit indices) on the device and the
// level bounds on the host; from the squared overlaps; perm [nk][n] on the
// device.  Up to kTrackLdsMaxN bands a vertex's averaged overlap matrix is
// in LDS; above, mwork (device, [mwork_blocks][n][n]) holds it for up to
// mwork_blocks vertices of a level at a time (null below).
constexpr int kTrackDevMaxN = 256;
constexpr int kTrackLdsMaxN = 128;
struct TrackEntry {
  unsigned v, cnt, dep[6];
};
int energy_cut_track_dev(hipStream_t st, size_t nx, size_t ny, size_t nz, int n, const double* ovx,
                         const double* ovy, const double* ovz, const TrackEntry* sched,
                         const std::vector<size_t>& level_start, double* mwork, size_t mwork_blocks, int* perm);
int energy_cut_track_smem_limits();

int energy_cut_t0(hipStream_t st, int dim, size_t nx, size_t ny, size_t nz, int n, const double* E, const double* Ki,
                  const int* perm, const double* om, const double* mu, size_t n_mu, double eta, double* partial,
                  size_t chunk, double* out);
// EnergyCut vertex kernel Im K[b][i][j] = Im(A[b][i][j] B[b][j][i]) ([nk][n][n] real).
int energy_cut_vertex(hipStream_t st, int n, size_t nk, const double* A, const double* B, double* Ki);
// x = |z|^2 elementwise (total complex z, real x).
int abs2(hipStream_t st, size_t total, const double* z, double* x);
// DOS partial sums: partial[blk][j] += sum_{eps in blk} g(E_j - eps)
int dos_partials(hipStream_t st, size_t ne, const double* evals, double Emin, double dE, int En,
                 double sigma, double* partial, int n_blk);
int dos_blocks(size_t ne);
// out[j] = scale * sum_blk partial[blk][j]
int reduce_partials(hipStream_t st, const double* partial, int n_blk, int m, double scale, double* out);
// out[j] = sum over n_rows rows of partial[r][j] (one workgroup per column, a
// fixed order): for many rows and few columns.
int column_sums(hipStream_t st, const double* partial, size_t n_rows, int m, double* out);

}  // namespace rtb
