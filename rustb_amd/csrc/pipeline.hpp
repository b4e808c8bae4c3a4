// Pieces of the per-call pipeline (capi.cpp) that the multi-model batch
// (batch.cpp) reuses: argument checks, direction tables, the eigen buffers,
// the mixed eigensolve of a chunk, the Kubo products and the FP64 Berry
// curvature of a k list.
#pragma once

#include "eigh.hpp"

namespace rtb {

// Device set-up before a model's first call (current device, LDS limits).
int prepare(const rtb_model* m);
int check_kmesh(const rtb_model* m, const size_t* kmesh, size_t* nk);
int check_direction(const rtb_model* m, const double* dir, int rows);
int check_eta(double eta);
int check_mu(const double* mu, size_t n_mu);
int check_T(double T);
int check_precision(int precision);

// The Fourier arguments of a model with np projected directions (rows of
// `dirs`, model dim entries each); the w.R and w.tau tables from `w`.
int upload_dirs(hipStream_t st, Work& w, const rtb_model* m, const double* dirs, int np, FourierArgs& a);

// Device bytes one call may hold for its k-point buffers (RTB_WORKSPACE_MB,
// else an eighth of the device memory, at most 32 GiB).
size_t workspace_budget();

// Eigensolver buffers of a chunk of K k-points (alloc_eig).
struct EigBufs {
  double *Vg, *tau, *Y, *S, *DE, *band;
};
size_t eig_extra_bytes(int n, bool vectors);
int alloc_eig(Work& w, int n, size_t K, bool vectors, EigBufs& b);

// Options of the mixed form's eigenpair check (RTB_MIXED_RESIDUAL,
// RTB_MIXED_FALLBACK), read from the environment per call.
struct MixedOpts {
  double res_tol = 1e-9;
  bool fallback = true;
};
MixedOpts mixed_opts();

// FP32 eigensolve + FP64 refinement of a chunk (refined unit-after-scale
// columns into W2); nb: the register route's block rows (0: by n).
int eig_mixed_chunk(hipStream_t st, Timer& tmr, bool from_upper, int n, size_t c, double* H, double* E, EigBufs& eb,
                    double* Z, double* W1, double* W2, double* W3, double* scale, const RefineCheck& chk, int nb = 0);

// Omega per band of a chunk from V ([K][n][n] holding V0 + i V1, or [K][2][n][n]
// with B given: the exact band sum of A = U^H V0 U and B = U^H V1 U) and the
// eigenvector columns Z (unit after zscale when given).
int kubo_chunk(hipStream_t st, Timer& tmr, const rtb_model* m, size_t K, const double* V, const double* Z,
               const double* zscale, const double* E, int spin_dir, double eta, double* T1, double* C, double* J,
               double* omega, double* B = nullptr, double* metric = nullptr);

// Omega_n(k) and E_n(k) in FP64 for nk k-points (host rows of model dim
// entries; outputs host [nk][n]) on stream st with buffers from w: the plain
// Berry-curvature path (Atom gauge, directions `dir`, broadening eta).
int berry_points_fp64(const rtb_model* m, const double* k, size_t nk, const double* dir, double eta, double* omega,
                      double* E, hipStream_t st, Work& w);

}  // namespace rtb
