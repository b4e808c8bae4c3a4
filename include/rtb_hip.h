/*
 * rtb_hip.h — C ABI of librtb_hip.so, the MI355X (gfx950) k-mesh hot path of
 * the Rustb tight-binding crate (LiuyichenYanwushang/Rustb v0.7.2).
 *
 * Conventions (all entry points):
 *   - Synchronous; return RTB_OK (0) or a negative RTB_ERR_* code; the message
 *     of the last failure on the calling thread is rtb_last_error().
 *   - Every host buffer is owned by the caller; outputs are caller-allocated.
 *   - Complex numbers are interleaved (re, im) doubles; matrices are C-order.
 *   - k-points are fractional reciprocal coordinates, one row of `dim` doubles.
 *   - n = nsta = norb (spinless) or 2*norb (spinful, [up block, down block]).
 *   - Thread safety: concurrent calls on one rtb_model are allowed.  Each
 *     call leases a workspace slot of its device (a HIP stream and one device
 *     block); at most RTB_MAX_CONCURRENT_CALLS (environment, default 4) calls
 *     run on one device at once and further callers wait for a slot.
 *   - Device memory: a slot's block grows to the largest workspace a call on
 *     it has needed (for a k-mesh call, a chunk budget of min(device/8,
 *     32 GiB) or RTB_WORKSPACE_MB, plus the mesh intermediates) and is kept
 *     for later calls until rtb_workspace_release() or process exit, so that
 *     repeated calls allocate nothing.  Concurrent large calls each hold
 *     their own block: an allocation the device cannot satisfy is
 *     RTB_ERR_DEVICE.  Host allocation is per call and proportional to the
 *     outputs it stages (EnergyCut: the tracking walk's schedule, 32 bytes
 *     per k-point; its vertex data only under RTB_TRACK_HOST=1).
 */
#ifndef RTB_HIP_H
#define RTB_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: one per reference TbError variant raised on this path ---- */
#define RTB_OK 0
#define RTB_ERR_KVECTOR_LENGTH -1        /* TbError::KVectorLengthMismatch (error.rs:291) */
#define RTB_ERR_DIMENSION_MISMATCH -2    /* TbError::DimensionMismatch (error.rs:99) */
#define RTB_ERR_INVALID_KMESH -3         /* TbError::InvalidKmeshDimensions (error.rs:135) */
#define RTB_ERR_INVALID_ENERGY_RANGE -4  /* TbError::InvalidEnergyRange (error.rs:138) */
#define RTB_ERR_INVALID_DOS_PARAMETER -5 /* TbError::InvalidDosParameter (error.rs:185) */
#define RTB_ERR_INVALID_RESPONSE -6      /* TbError::InvalidResponseParameter (error.rs:179) */
#define RTB_ERR_SPIN_NOT_ALLOWED -7      /* TbError::SpinNotAllowed (error.rs:132) */
#define RTB_ERR_EIGEN_FAILED -8          /* TbError::EigenvalueComputationFailed (error.rs:92) */
#define RTB_ERR_ONSITE_NOT_REAL -9       /* TbError::OnsiteHoppingMustBeReal */
#define RTB_ERR_R_LENGTH -10             /* TbError::RVectorLengthError */
#define RTB_ERR_INVALID_MODEL -11        /* TbError::InvalidModelInvariant / LatticeDimensionError */
#define RTB_ERR_INVALID_ARGUMENT -12     /* null pointer, zero size where one is required */
#define RTB_ERR_DEVICE -13               /* a HIP runtime call failed */
#define RTB_ERR_UNSUPPORTED -14          /* a size outside what the kernels implement */
#define RTB_ERR_FILE_OPEN -15            /* TbError::FileCreation (error.rs:79), an input file missing */
#define RTB_ERR_FILE_PARSE -16           /* TbError::FileParse (error.rs:64) */
#define RTB_ERR_ORBITAL_PROJECTION -17   /* TbError::InvalidOrbitalProjection (error.rs:67) */
#define RTB_ERR_SPIN_MISMATCH -18        /* TbError::Other "Spin mismatch" (wannier90.rs:559-564) */
#define RTB_ERR_IO -19                   /* TbError::Io (error.rs:61), e.g. text that is not UTF-8 */
#define RTB_ERR_NO_ORBITALS -20          /* TbError::NoOrbitals (error.rs:279) */
#define RTB_ERR_LINALG -21               /* TbError::Linalg (error.rs:83): a singular lattice */
#define RTB_ERR_INVALID_THERMO -22       /* TbError::InvalidThermodynamicParameter (thermodynamics.rs) */
#define RTB_ERR_MISSING_CONJUGATE -23    /* TbError::MissingHermitianConjugate (error.rs:285) */
#define RTB_ERR_INVALID_DIMENSION -24    /* TbError::InvalidDimension (error.rs:239) */

/* ---- enums ---- */
#define RTB_GAUGE_LATTICE 0 /* Gauge::Lattice (model.rs:449-453) */
#define RTB_GAUGE_ATOM 1    /* Gauge::Atom */

#define RTB_SPIN_NONE 0 /* Option<SpinDirection>::None (model.rs:470-475) */
#define RTB_SPIN_X 1
#define RTB_SPIN_Y 2
#define RTB_SPIN_Z 3

#define RTB_PREC_FP64 0  /* plain form: the reference's formulas in FP64 */
/* mixed form: FP32 eigensolve + FP64 refinement.  Every refined eigenpair is
 * checked: a k-point with an eigenvalue that is not finite or a residual
 * |H u - lambda u| of a unit eigenvector above RTB_MIXED_RESIDUAL (environment,
 * eV per max(1, max |lambda|), default 1e-9) is re-solved by the FP64
 * eigensolver; with RTB_MIXED_FALLBACK=0 it fails the call with
 * RTB_ERR_EIGEN_FAILED instead.  Any call whose final eigenvalues are not all
 * finite returns RTB_ERR_EIGEN_FAILED (the reference's zheev info != 0,
 * ndarray_lapack.rs:443-447, and TbError from eigh, response/traits.rs:168). */
#define RTB_PREC_MIXED 1

/* =====================================================================
 * Model builder (host only, no GPU).  Mirrors the reference's hopping API
 * so C/C++ hosts can build `Model` storage with the same semantics.
 * ===================================================================== */
typedef struct rtb_builder rtb_builder;

/* Model::tb_model (model_build.rs:212-285): ham = one zero block at R = 0. */
int rtb_builder_new(int dim, int spin, int norb, const double* lat /*dim*dim rows*/,
                    const double* orb /*norb*dim*/, rtb_builder** out);
void rtb_builder_destroy(rtb_builder* b);
/* set_hop (model_build.rs:354-470) / add_hop (model_build.rs:477-558);
 * pauli = RTB_SPIN_* decoration, ignored for spinless models. */
int rtb_builder_set_hop(rtb_builder* b, double re, double im, int i, int j, const int64_t* R,
                        int pauli);
int rtb_builder_add_hop(rtb_builder* b, double re, double im, int i, int j, const int64_t* R,
                        int pauli);
/* add_element (model_build.rs:613-687): spin-doubled indices, no decoration. */
int rtb_builder_add_element(rtb_builder* b, double re, double im, int i, int j, const int64_t* R);
/* set_onsite (model_build.rs:689-700) / add_onsite (model_build.rs:737-750). */
int rtb_builder_set_onsite(rtb_builder* b, const double* values /*norb*/, int pauli);
int rtb_builder_add_onsite(rtb_builder* b, const double* values /*norb*/, int pauli);
/* Storage queries and export (model.rs:150-158 layout). */
int rtb_builder_n_R(const rtb_builder* b);
int rtb_builder_nsta(const rtb_builder* b);
int rtb_builder_export(const rtb_builder* b, int64_t* hamR /*n_R*dim*/, double* ham /*n_R*n*n*2*/);

/* Seeded synthetic workloads (SplitMix64-seeded xoshiro256**, Box-Muller),
 * identical on every host.  `kind` selects the recipe:
 *   1: random spinful 3D model, |R_i|<=rmax, amplitude 0.5*exp(-|R|) (scale=0.5, decay=1)
 *   2: Wannier-like 3D model, amplitude exp(-|R|/decay)
 *   3: 2D model with n_R drawn R vectors within |R_i|<=rmax, amplitude exp(-|R|)
 *   4: 2D slab of `layers` layers x norb_layer orbitals (see the synthetic-model notes)
 * The sizes are returned first with ham == NULL, then the arrays are filled. */
typedef struct rtb_synth_spec {
  int kind;
  int dim;
  int spin;
  int norb;        /* orbitals (per layer for kind 4) */
  int rmax;        /* |R_i| bound */
  int n_R_target;  /* kind 3: number of stored R (odd, includes R = 0) */
  int layers;      /* kind 4 */
  double scale;    /* amplitude prefactor */
  double decay;    /* exp(-|R|/decay) */
  uint64_t seed;
} rtb_synth_spec;
int rtb_synth_model(const rtb_synth_spec* spec, int* n_R, int* nsta, int64_t* hamR, double* ham,
                    double* orb, double* lat);

/* =====================================================================
 * Wannier90 input (host only).  Wannier90::from_hr (wannier90.rs:49-1228)
 * for Model<SPIN, 3, NoRMatrix>: reads {path}{seedname}_hr.dat and .win,
 * and _centres.xyz and _wsvec.dat when present.  Storage follows model.rs:
 * hamR row 0 is R = 0 (Fermi level `zero_energy` subtracted from its
 * diagonal), then the other R in file order (or in order of first appearance
 * after the Wigner-Seitz shifts of _wsvec.dat).
 * ===================================================================== */
typedef struct rtb_w90 rtb_w90;
/* spin: 0 or 1 is the caller's SPIN (a different `spinors` setting in .win is
 * RTB_ERR_SPIN_MISMATCH), -1 takes the file's.  n_threads: parse threads for
 * the _hr.dat body (0: up to 16). */
int rtb_w90_read(const char* path, const char* seedname, double zero_energy, int spin, int n_threads,
                 rtb_w90** out);
/* Model::from_hr_with_rmatrix (wannier90.rs:1208-1211), R = HasRMatrix: also
 * reads {path}{seedname}_r.dat (required) into the position matrices, moved
 * with the Wigner-Seitz shifts and made Hermitian across +-R. */
int rtb_w90_read_with_rmatrix(const char* path, const char* seedname, double zero_energy, int spin,
                              int n_threads, rtb_w90** out);
void rtb_w90_destroy(rtb_w90* m);
/* Model::rmatrix (model.rs): n_R x 3 x nsta x nsta complex (RTB_ERR_INVALID_ARGUMENT
 * for a model read without _r.dat). */
int rtb_w90_export_rmatrix(const rtb_w90* m, double* rmatrix);
/* Model::norb / nsta / hamR.nrows() / natom (model.rs:697-709). */
int rtb_w90_info(const rtb_w90* m, int* spin, int* norb, int* nsta, int* n_R, int* natom);
/* Model fields lat (3x3 rows, Angstrom), orb (norb x 3 fractional),
 * orb_projection (norb OrbProj codes in the enum's order, atom_struct.rs:15-77),
 * hamR (n_R x 3), ham (n_R x nsta x nsta complex).  Null outputs are skipped. */
int rtb_w90_export(const rtb_w90* m, int64_t* hamR, double* ham, double* orb, double* lat, int* orb_proj);
/* Model::atoms[i] (atom_struct.rs Atom): species as its atomic number
 * (AtomType, H = 1 .. Ra = 88), fractional position, orbitals
 * [first_orb, first_orb + n_orb). */
int rtb_w90_atom(const rtb_w90* m, int i, int* species, double* pos, int* first_orb, int* n_orb);
const char* rtb_element_symbol(int z);
const char* rtb_orbital_projection_name(int code);

/* =====================================================================
 * Device-resident model and the hot path
 * ===================================================================== */
typedef struct rtb_model rtb_model;

/* Copies a1 storage (model.rs:150-158) to HBM once.  hamR row 0 must be R = 0.
 * Replaces the per-k reads of `ham` by BLAS zaxpy (ndarray_lapack.rs:35-39). */
int rtb_model_create(int dim, int spin, int norb, int n_R, const int64_t* hamR,
                     const double* ham /*n_R*n*n complex*/, const double* orb /*norb*dim*/,
                     const double* lat /*dim*dim rows*/, rtb_model** out);
void rtb_model_destroy(rtb_model* m);
/* The device model of a loaded Wannier90 model (as rtb_model_create, with
 * the position matrices when it was read with _r.dat). */
int rtb_w90_model(const rtb_w90* w, rtb_model** out);
/* Model<_, _, HasRMatrix>: position matrices rmatrix (n_R x dim x nsta x nsta
 * complex, rows as hamR) for the velocity term i[H(k), r_p(k)]
 * (velocity.rs:377-402 Atom gauge, diagonal of r_p zeroed; :645-668 and the
 * Lattice branch).  Uploaded once; every velocity path then includes it. */
int rtb_model_set_rmatrix(rtb_model* m, const double* rmatrix);
int rtb_model_nsta(const rtb_model* m);

/* Model::gen_ham (model_physics.rs:43-148) and Model::gen_v_projected
 * (velocity.rs:470-731) for nk k-points.  H: nk*n*n complex.  If n_proj > 0,
 * dirs is n_proj*dim and V receives nk*n_proj*n*n complex. */
int rtb_gen_ham_batch(const rtb_model* m, const double* k, size_t nk, int gauge,
                      const double* dirs, int n_proj, double* H, double* V);

/* Solve::solve_band_all(_parallel) (solve_ham.rs:82-107): E nk*n ascending. */
int rtb_solve_band_all(const rtb_model* m, const double* k, size_t nk, int precision, double* E);

/* Solve::solve_all(_parallel) (solve_ham.rs:150-184): E nk*n and evec
 * nk*n*n complex with the eigenvectors as ROWS (solve_ham.rs:127). */
int rtb_solve_all(const rtb_model* m, const double* k, size_t nk, int precision, double* E,
                  double* evec);

/* Solve::solve_band_range_onek (solve_ham.rs:64-80) and solve_range_onek
 * (solve_ham.rs:130-147), LAPACK zheevr RANGE = 'V' (ndarray_lapack.rs:217-395):
 * per k the eigenvalues in (vl, vu], ascending, packed to the front of its
 * row of E (nk*n), their number in count[k]; solve_range also the matching
 * eigenvectors as the first count[k] rows of evec (nk*n*n complex).  These
 * read H's lower triangle, as the reference's raw-buffer zheevr call does.
 * epsilon (zheevr's ABSTOL) only bounds the reference's accuracy; the
 * eigenvalues here are FP64-accurate whatever it is. */
int rtb_solve_band_range(const rtb_model* m, const double* k, size_t nk, double vl, double vu, double epsilon,
                         int precision, double* E, int* count);
int rtb_solve_range(const rtb_model* m, const double* k, size_t nk, double vl, double vu, double epsilon,
                    int precision, double* E, double* evec, int* count);

/* BerryCurvature::berry_curvature_at (response/traits.rs:60-78,160-200) for
 * nk k-points: omega, E nk*n.  dir is 2*dim (rows a, b). */
int rtb_berry_curvature(const rtb_model* m, const double* k, size_t nk, const double* dir,
                        int spin_dir, double eta, int precision, double* omega, double* E);

/* BerryCurvature::occupied_berry_curvature_on (response/traits.rs:104-150):
 * out[k] = sum_n Omega_n f(E_n, mu) at temperature T (kelvin). */
int rtb_occupied_berry_curvature(const rtb_model* m, const double* k, size_t nk,
                                 const double* dir, int spin_dir, double eta, double mu,
                                 double T, int precision, double* out);

/* Model::hall_conductivity, Integration::Direct (response/linear/mod.rs:160-225):
 * sigma[n_mu] = sum_k sum_n Omega_n f(E_n, mu) / N_k / det(lat). kmesh has dim entries. */
int rtb_hall_direct(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu,
                    double T, double eta, const double* dir, int spin_dir, int precision,
                    double* sigma);

/* The AHC of rtb_hall_direct resolved by layer, for layered (slab) models.
 * No counterpart in the reference, which defines no layer-resolved
 * curvature: here sigma_layers[m][l] = sum_k sum_n f(E_n, mu_m) Omega_n(k)
 * sum_{i in layer l} |u_n(k)_i|^2 / N_k / det(lat), the band curvature
 * weighted by the band's weight on the layer's orbitals (both spin copies of
 * orbital o belong to layer orbital_layer[o], o < norb); summed over l it is
 * sigma (also returned).  For degenerate bands the split between layers
 * depends on the basis chosen inside the degenerate subspace.
 * n_mu * n_layers <= 1024. */
int rtb_hall_direct_layers(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu,
                           double T, double eta, const double* dir, int spin_dir, int precision,
                           int n_layers, const int* orbital_layer, double* sigma, double* sigma_layers);

/* rtb_hall_direct without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * device buffers are a fixed static device array of 256 MiB per GPU the
 * library runs on (held until the process exits); the k-mesh is processed
 * in chunks sized to fit it, with the dense Fourier sum in both precision
 * forms.  A call whose buffers cannot fit (even one k-point per chunk, plus
 * the direction tables, position matrices and the n_mu-long sums) returns
 * RTB_ERR_UNSUPPORTED.  Runs on the device's null stream; every call on one
 * GPU uses the same array, so calls that overlap on one GPU race. */
int rtb_hall_direct_noalloc(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu,
                            double T, double eta, const double* dir, int spin_dir, int precision,
                            double* sigma);

/* ---------------------------------------------------------------------
 * Many models in one call.  The reference evaluates Model::hall_conductivity
 * (Integration::Direct) one model at a time (response/linear/mod.rs:160-225);
 * a batch runs that loop for every model of the batch, each on its own
 * uniform mesh, with the Atom gauge and no spin current, as device work
 * planned once and replayed: the models' k-points in passes that fit the
 * workspace budget, the Fourier sums of all models of a pass in persistent
 * work-queue launches (tiles longest first), the eigensolve, Kubo band sum
 * and per-k band sums per matrix size, captured in a HIP graph per
 * (precision, n_mu, T, eta).  Results equal rtb_hall_direct per model up to
 * the order of the k sum (FP64), or within the mixed form's bars.
 * Switched off unless the environment sets RTB_BATCH_ENABLE=1 (otherwise
 * rtb_batch_hall_direct returns RTB_ERR_UNSUPPORTED): its first device run
 * stopped making progress and the cause is not known; its host logic is
 * checked by emulation.
 * --------------------------------------------------------------------- */
typedef struct rtb_batch rtb_batch;
/* A per-size choice of kernel variants (the autotuned table): fourier_tk,
 * k-points per thread of the work-queue Fourier sum (1, 2, 4 or 8); eigh_nb,
 * rows of the register eigensolver's block (0: the smallest of 32, 64, 128
 * that holds nsta; else one of them, at least nsta, nsta <= 128).  Sizes not
 * in the table take fourier_tk 2, eigh_nb 0. */
typedef struct rtb_batch_variant {
  int nsta;
  int fourier_tk;
  int eigh_nb;
} rtb_batch_variant;
/* models: n_models handles on one device (no position matrices), which must
 * outlive the batch; kmesh: [n_models][3] mesh sizes (entries beyond a
 * model's dim ignored); dirs: [n_models][2][3] the two directions of each
 * model's conductivity (first dim entries of each row used); table may be
 * null (n_table 0).  Checks as rtb_hall_direct's. */
int rtb_batch_create(const rtb_model* const* models, int n_models, const size_t* kmesh, const double* dirs,
                     const rtb_batch_variant* table, int n_table, rtb_batch** out);
/* sigma: [n_models][n_mu] in the order of `models`; sum / nk / det(lat) per
 * model.  The first call of a parameter set plans (per precision) and
 * captures its graph; later ones replay it.  Calls on one batch are
 * serialised.  Mixed form: k-points whose refined eigenpairs fail the check
 * are re-solved in FP64 after the replay (RTB_MIXED_FALLBACK=0: the call
 * fails with RTB_ERR_EIGEN_FAILED); rtb_last_stats gives nk (all models),
 * ms[RTB_STAGE_TOTAL] (the graph's device time), launches[RTB_STAGE_TOTAL]
 * (graph nodes), max_residual and resolved. */
int rtb_batch_hall_direct(rtb_batch* b, const double* mu, size_t n_mu, double T, double eta, int precision,
                          double* sigma);
/* use_graph 1 (the default): capture once per parameter set, replay;
 * 0: launch the same sequence directly at every call, with per-stage times
 * in rtb_last_stats (event nodes of a graph give no elapsed times). */
int rtb_batch_set_graph(rtb_batch* b, int use_graph);
typedef struct rtb_batch_info {
  int n_models;
  int n_passes;      /* of the last call's precision */
  int n_portions;    /* (pass, matrix size) pairs */
  int n_tiles;       /* work-queue Fourier tiles */
  int n_queues;      /* work-queue launches */
  int n_mesh_models; /* models on the separable mesh route (mixed form) */
  int graphs_cached;
  int captures;      /* graphs captured over the batch's life */
  int graph_nodes;   /* of the last call's graph */
  size_t nk_total;
  size_t arena_bytes; /* device memory the batch holds */
  double graph_ms;    /* the last replay */
} rtb_batch_info;
int rtb_batch_info_get(rtb_batch* b, rtb_batch_info* out);
void rtb_batch_destroy(rtb_batch* b);

/* Model::hall_conductivity, Integration::EnergyCut (response/linear/mod.rs:177-185,
 * 226-268) for a 2D or 3D model: per-vertex E and K^{ab}_nm = v^a_nm v^b_mn
 * (compute_velocity_kernel, response/primitives.rs:33-137, FP64), band tracking
 * over the mesh (global_band_track, response/tracking.rs:123-224), the Fermi cut
 * of each cell's four triangles (integrate_fermi_cut_2d, response/energy_cut.rs:
 * 1231-1409) or five tetrahedra (integrate_fermi_cut_3d, response/energy_cut.rs:
 * 1650-1850), / det(lat).  mu must be sorted ascending (RTB_ERR_INVALID_RESPONSE);
 * a 1D model gives RTB_ERR_INVALID_DIMENSION.  Everything runs on the device
 * (the tracking walk one launch per level of its breadth-first schedule;
 * above nsta = 128 each vertex's overlap matrix is in a workspace slice of
 * nsta^2 doubles, at most 1024 vertices of a level per launch, or
 * RTB_TRACK_SLICES=s in the environment).  RTB_TRACK_HOST=1 selects the host
 * walk instead (the squared overlaps copied to the host; the same
 * permutations: for checking). */
int rtb_hall_energy_cut(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu,
                        double T, double eta, const double* dir, int spin_dir, double* sigma);

/* The host part of rtb_hall_energy_cut from given vertex data (no device
 * work): E [nx ny][nsta], K [nx ny][nsta][nsta] complex, the raw eigenvector
 * overlaps ovx [(nx-1) ny][nsta][nsta] = U_i^H U_{i+ny} and ovy [nx ny - 1]
 * [nsta][nsta] = U_i^H U_{i+1} (complex, U: eigenvectors as columns), sorted
 * mu; sigma[n_mu] = integrate_fermi_cut_2d after global_band_track, / det. */
int rtb_energy_cut_host(size_t nx, size_t ny, int nsta, const double* E, const double* K,
                        const double* ovx, const double* ovy, const double* mu, size_t n_mu,
                        double T, double eta, double det, double* sigma);
/* The same for a 3D mesh (vertex (ix ny + iy) nz + iz): ovx [nk - ny nz] =
 * U_i^H U_{i+ny nz}, ovy [nk - nz] = U_i^H U_{i+nz}, ovz [nk - 1] = U_i^H U_{i+1}
 * ([nsta][nsta] complex each; ovx, ovy, ovz may be null when nx, ny, nz is 1);
 * sigma = integrate_fermi_cut_3d after global_band_track, / det. */
int rtb_energy_cut_host_3d(size_t nx, size_t ny, size_t nz, int nsta, const double* E, const double* K,
                           const double* ovx, const double* ovy, const double* ovz, const double* mu,
                           size_t n_mu, double T, double eta, double det, double* sigma);

/* QuantumGeometry::quantum_geometry_on (quantum_geometry.rs:104-140, kernel
 * quantum_geometry_at_impl :149-190) for nk k-points: with A = U^H v_a U,
 * B = U^H v_b U, G_n = sum_{m != n} A_nm B_mn / ((E_n - E_m)^2 + eta^2);
 * metric = Re G, omega = -2 Im G, E: nk*n each.  dir is 2*dim (rows a, b).
 * Both precisions form A and B apart (FP64 products). */
int rtb_quantum_geometry(const rtb_model* m, const double* k, size_t nk, const double* dir, double eta,
                         int precision, double* metric, double* omega, double* E);

/* Model::quantum_geometry, Integration::Direct (quantum_geometry.rs:196-240):
 * metric[n_mu], berry[n_mu] = sum_k sum_n {g_n, Omega_n} f(E_n, mu) / N_k / det(lat). */
int rtb_quantum_geometry_direct(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu, double T,
                                double eta, const double* dir, int precision, double* metric, double* berry);

/* Model::extrinsic_nonlinear_hall, Integration::Direct (response/nonlinear/mod.rs:256-347):
 * the Berry curvature dipole sigma[n_mu] = sum_k sum_n d_{f2}E_n Omega_n^{cur,f1}
 * (-df/dE)(E_n) / N_k / det(lat), per band from berry_curvature_dipole_n_onek
 * (mod.rs:89-163).  dir is 3*dim (rows current, field_1, field_2); T > 0;
 * symmetrized != 0 (FieldSymmetry::Symmetrized, the reference's default)
 * averages the two field orderings when the field rows differ. */
int rtb_nonlinear_hall_direct(const rtb_model* m, const size_t* kmesh, const double* mu, size_t n_mu, double T,
                              double eta, const double* dir, int spin_dir, int symmetrized, int precision,
                              double* sigma);

/* Model::dos (model_physics.rs:179-226): E[En] = linspace(Emin, Emax, En), rho[En]. */
int rtb_dos(const rtb_model* m, const size_t* kmesh, double Emin, double Emax, size_t En,
            double sigma, int precision, double* E, double* rho);

/* Frees the workspace blocks and streams of the idle slots of the current
 * device (slots in use by running calls are kept).  No counterpart in the
 * reference (whose calls hold only host memory): for hosts that want the
 * device memory back between bursts of calls. */
int rtb_workspace_release(void);
/* Device bytes held by the workspace slots of the current device and their
 * number (slots may be null).  No counterpart in the reference. */
int rtb_workspace_bytes(size_t* bytes, int* slots);

/* Thread-local message of the last failure on this thread ("" if none). */
const char* rtb_last_error(void);

/* Per-kernel device time of the last hot-path call on this thread, measured
 * with HIP events on the stream the kernels ran on (one event pair around
 * each launch or group of launches of one kernel kind).  Stages:
 *   RTB_STAGE_FOURIER    Fourier sum kernels (H and projected velocities)
 *   RTB_STAGE_TRIDIAG    Householder tridiagonalisation kernel
 *   RTB_STAGE_TRIEIG     tridiagonal eigenvalues (+ eigenvectors) kernel
 *   RTB_STAGE_BACKTRANS  eigenvector back-transformation kernel
 *   RTB_STAGE_GEMM64     FP64 matrix-core products (Kubo transforms and the
 *                        refinement products)
 *   RTB_STAGE_GEMM32     FP32 matrix-core correction products (refinement)
 *   RTB_STAGE_REFINE     other refinement kernels (mixed form)
 *   RTB_STAGE_KUBO       spin current and the Omega band sum
 *   RTB_STAGE_REDUCE     mesh points, occupation/DOS reductions, copies
 *   RTB_STAGE_TOTAL      first to last event of the call
 * ms[] in milliseconds; launches[] kernel launches per stage; flop[] the
 * algorithmic floating-point operations of those launches where the stage
 * has a closed form (complex product 8 n^3, Hermitian product 4 n^2 (n+1),
 * tridiagonalisation 16/3 n^3, back-transformation 8 n^3; 0 otherwise);
 * mflop[] the matrix-core operations the product kernels executed (three
 * real products per complex one, 64 x 64 tiles, the k loop padded to 16). */
#define RTB_STAGE_FOURIER 0
#define RTB_STAGE_TRIDIAG 1
#define RTB_STAGE_TRIEIG 2
#define RTB_STAGE_BACKTRANS 3
#define RTB_STAGE_GEMM64 4
#define RTB_STAGE_GEMM32 5
#define RTB_STAGE_REFINE 6
#define RTB_STAGE_KUBO 7
#define RTB_STAGE_REDUCE 8
#define RTB_STAGE_TOTAL 9
#define RTB_N_STAGES 10
typedef struct rtb_stats {
  double ms[RTB_N_STAGES];
  int launches[RTB_N_STAGES];
  double flop[RTB_N_STAGES];
  size_t nk;
  int nsta;
  double mflop[RTB_N_STAGES];
  /* mixed form: the largest residual |H u - lambda u| (eV) of a unit refined
   * eigenvector over the call, and the number of k-points whose refined
   * eigenpairs failed the check and were re-solved in FP64 */
  double max_residual;
  size_t resolved;
} rtb_stats;
int rtb_last_stats(rtb_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* RTB_HIP_H */
