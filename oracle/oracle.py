"""CPU oracle for the Rustb k-mesh hot path (FP64, numpy + LAPACK zheev).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product path (``rustb_amd``) never imports, calls or links it.

This module restates, in plain numpy, the per-k algorithm of the reference crate
(LiuyichenYanwushang/Rustb v0.7.2).  Every function cites the reference file and
lines it follows.  The eigensolve is LAPACK ``zheev`` (``scipy.linalg.lapack``),
the same routine the reference reaches through ``ndarray_lapack.rs:412-448`` and
ndarray-linalg's ``Eigh``.

Pinning status (what checks each function, see ``tests/test_oracle.py``):

* ``Model`` builder (``set_hop``/``add_hop``/``add_element``/``set_onsite``/
  ``add_onsite``): storage layout pinned by the reference's own test models
  (``lib.rs:671-755``) reproducing its tabulated values, by the Hermitian
  invariant ``ham[-R] = ham[R]^H`` and by the spinful graphene band fixture
  ``examples/graphene/BAND.dat`` (Pauli-Z decoration path).
* ``gen_ham``: exact Gamma eigenvalues ``[-3, 3]`` and the ``(1/3, 2/3)`` Dirac
  point of ``lib.rs:701-706``; the 2001-point ``BAND.dat`` fixture (5e-7 eV);
  the graphene closed form ``+-|t||1+e^{i2pi k1}+e^{i2pi k2}|``; gauge
  invariance of the spectrum; brute-force triple loop on tiny models.
* ``gen_v``/``gen_v_projected``: the tabulated Atom- and Lattice-gauge values of
  ``lib.rs:707-742`` (1e-8); central finite difference of ``gen_ham`` in the
  Lattice gauge; linearity in the direction rows.
* ``eigh``/``solve_onek``: the ``U^T H U^*`` convention of ``lib.rs:1231-1277``
  and the ``U^H H U`` diagonal check of ``lib.rs:743-755``.
* ``berry_curvature_at``: sum rule ``sum_n Omega_n = 0``; Haldane Chern number
  ``2 pi sigma = 1 +- 1e-3`` at 31x31 (``lib.rs:902-1021``); Wilson-loop
  plaquette at ``(1/3, 2/3)`` (``lib.rs:757-823``, 1e-4); time-reversal
  antisymmetry ``sigma(t2) + sigma(-t2) <= 1e-10`` (``lib.rs:2665-2740``).
* ``hall_conductivity_direct``: as above plus single- vs multi-mu agreement
  (``lib.rs:825-845``, ``lib.rs:962-998``).
* ``quantum_geometry_at``: the Berry part equals ``berry_curvature_at``
  (pinned above); the metric equals the Fubini-Study metric
  (1/2) Tr(d_a P_n d_b P_n) of each isolated band from central finite
  differences of the band projector P_n (gauge invariant, independent of the
  sum-over-states formula), and for a two-band model g_0 = g_1 and
  Omega_0 = -Omega_1.  ``quantum_geometry_direct`` sums it like
  ``hall_conductivity_direct`` (its Berry part equals that function's result).
* ``berry_curvature_dipole_onek`` / ``extrinsic_nonlinear_hall_direct``: the
  band velocity Re(U^H v_c U)_nn equals the finite-difference derivative of
  the band energy (Hellmann-Feynman); the Omega factor equals
  ``berry_curvature_at``; the dipole response is even under time reversal
  (Haldane t2 -> -t2, the reference's ``dipole_energy_cut_tr``,
  lib.rs:2745-2790, here for the Direct sum on the k -> -k symmetric mesh);
  Ordered with field_1 = field_2 equals Symmetrized.
* ``dos``: normalisation (integral of rho over a window that contains the whole
  spectrum equals nsta) and the closed-form single-level Gaussian.
* ``fermi``: the step/half/zero cases of ``thermodynamics.rs:152-158`` and the
  300 K equivalence of ``thermodynamics.rs:160-172``.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg.lapack as _lapack

TWO_PI = 2.0 * math.pi
# thermodynamics.rs:10
BOLTZMANN_CONSTANT_EV_PER_K = 8.617_333_262_145e-5

PAULI = {None: 0, "X": 1, "Y": 2, "Z": 3}


class OracleError(ValueError):
    """Mirror of the reference ``TbError`` variants raised on this path."""

    def __init__(self, kind: str, message: str = ""):
        super().__init__(f"{kind}: {message}")
        self.kind = kind


# --------------------------------------------------------------------------
# Model storage and builder (model.rs:150-158, model_build.rs:71-145,212-558)
# --------------------------------------------------------------------------
class Model:
    """``Model<SPIN, DIM, NoRMatrix>`` storage.

    ``ham[iR, i, j] = <i,0|H|j,R>`` with ``hamR[iR] = R``; row 0 is R = 0
    (``model_build.rs:244-245``).  A new R appends R then -R
    (``model_build.rs:549-555``).
    """

    def __init__(self, lat, orb, spin: bool):
        lat = np.asarray(lat, dtype=np.float64)
        orb = np.asarray(orb, dtype=np.float64)
        dim = lat.shape[1]
        if lat.shape[0] != dim:
            raise OracleError("LatticeDimensionError")
        if orb.ndim != 2 or orb.shape[1] != dim:
            raise OracleError("InvalidModelInvariant", "orbital_position_shape")
        self.lat = lat
        self.orb = orb
        self.spin = bool(spin)
        self.dim = dim
        self.norb = orb.shape[0]
        self.nsta = 2 * self.norb if spin else self.norb  # model.rs:705-707
        # model_build.rs:244-245: ham starts as one zero block at R = 0
        self.ham = np.zeros((1, self.nsta, self.nsta), dtype=np.complex128)
        self.hamR = np.zeros((1, dim), dtype=np.int64)
        self.rmatrix = None  # [N_R, dim, n, n] for Model<_, _, HasRMatrix>

    # model_utils.rs:18-34 (linear scan)
    def find_R(self, R):
        R = np.asarray(R, dtype=np.int64)
        hits = np.nonzero(np.all(self.hamR == R[None, :], axis=1))[0]
        return int(hits[0]) if hits.size else None

    def _decorate(self, block, tmp, i, j, pauli, accumulate):
        """update_hamiltonian!/add_hamiltonian! (model_build.rs:71-145)."""
        n = self.norb
        op = (lambda a, b: a + b) if accumulate else (lambda a, b: b)
        if not self.spin:
            block[i, j] = op(block[i, j], tmp)
            return
        if pauli is None:
            block[i, j] = op(block[i, j], tmp)
            block[i + n, j + n] = op(block[i + n, j + n], tmp)
        elif pauli == "X":
            block[i + n, j] = op(block[i + n, j], tmp)
            block[i, j + n] = op(block[i, j + n], tmp)
        elif pauli == "Y":
            block[i + n, j] = op(block[i + n, j], tmp * 1j)
            block[i, j + n] = op(block[i, j + n], -tmp * 1j)
        elif pauli == "Z":
            block[i, j] = op(block[i, j], tmp)
            block[i + n, j + n] = op(block[i + n, j + n], -tmp)
        else:
            raise OracleError("InvalidPauli", str(pauli))

    def _hop(self, tmp, i, j, R, pauli, accumulate):
        """set_hop (model_build.rs:354-470) / add_hop (model_build.rs:477-558)."""
        tmp = complex(tmp)
        R = np.asarray(R, dtype=np.int64)
        if R.shape != (self.dim,):
            raise OracleError("RVectorLengthError")
        if not (0 <= i < self.norb and 0 <= j < self.norb):
            raise OracleError("DimensionMismatch", "orbital indices")
        onsite = i == j and not np.any(R)
        if onsite and tmp.imag != 0.0:  # model_build.rs:505
            raise OracleError("OnsiteHoppingMustBeReal")
        if not self.spin:
            pauli = None
        idx = self.find_R(R)
        if idx is not None:
            inv = self.find_R(-R)
            self._decorate(self.ham[idx], tmp, i, j, pauli, accumulate)
            if idx != 0 or i != j:  # model_build.rs:532
                self._decorate(self.ham[inv], tmp.conjugate(), j, i, pauli, accumulate)
        else:
            # a new R uses set semantics for both blocks (model_build.rs:545-556)
            b1 = np.zeros((self.nsta, self.nsta), dtype=np.complex128)
            self._decorate(b1, tmp, i, j, pauli, False)
            b2 = np.zeros((self.nsta, self.nsta), dtype=np.complex128)
            self._decorate(b2, tmp.conjugate(), j, i, pauli, False)
            self.ham = np.concatenate([self.ham, b1[None], b2[None]], axis=0)
            self.hamR = np.concatenate([self.hamR, R[None], -R[None]], axis=0)

    def set_hop(self, tmp, i, j, R, pauli=None):
        self._hop(tmp, i, j, R, pauli, accumulate=False)

    def add_hop(self, tmp, i, j, R, pauli=None):
        self._hop(tmp, i, j, R, pauli, accumulate=True)

    def add_element(self, tmp, i, j, R):
        """add_element (model_build.rs:613-687): full spin-doubled indices."""
        tmp = complex(tmp)
        R = np.asarray(R, dtype=np.int64)
        if R.shape != (self.dim,):
            raise OracleError("RVectorLengthError")
        if not (0 <= i < self.nsta and 0 <= j < self.nsta):
            raise OracleError("DimensionMismatch", "orbital indices")
        if i == j and not np.any(R) and tmp.imag != 0.0:
            raise OracleError("OnsiteHoppingMustBeReal")
        idx = self.find_R(R)
        if idx is not None:
            inv = self.find_R(-R)
            self.ham[idx, i, j] = tmp
            if idx != 0 or i != j:
                self.ham[inv, j, i] = tmp.conjugate()
        else:
            b1 = np.zeros((self.nsta, self.nsta), dtype=np.complex128)
            b1[i, j] = tmp
            b2 = np.zeros((self.nsta, self.nsta), dtype=np.complex128)
            b2[j, i] = tmp.conjugate()
            self.ham = np.concatenate([self.ham, b1[None], b2[None]], axis=0)
            self.hamR = np.concatenate([self.hamR, R[None], -R[None]], axis=0)

    def set_onsite(self, values, pauli=None):
        """set_onsite -> set_onsite_one -> set_hop at R = 0 (model_build.rs:689-793)."""
        values = np.asarray(values, dtype=np.float64)
        if values.shape != (self.norb,):
            raise OracleError("DimensionMismatch", "onsite length")
        for i, v in enumerate(values):
            self.set_hop(complex(v, 0.0), i, i, np.zeros(self.dim, np.int64), pauli)

    def add_onsite(self, values, pauli=None):
        """add_onsite -> add_hop at R = 0 (model_build.rs:737-750)."""
        values = np.asarray(values, dtype=np.float64)
        if values.shape != (self.norb,):
            raise OracleError("DimensionMismatch", "onsite length")
        for i, v in enumerate(values):
            self.add_hop(complex(v, 0.0), i, i, np.zeros(self.dim, np.int64), pauli)

    def orb_sta(self):
        """tau duplicated for spin (model_physics.rs:131-136, velocity.rs:545-550)."""
        return np.concatenate([self.orb, self.orb], axis=0) if self.spin else self.orb


def model_from_arrays(lat, orb, spin, hamR, ham, rmatrix=None) -> Model:
    """rmatrix: [N_R, dim, n, n] position matrices (Model<_, _, HasRMatrix>)."""
    m = Model(lat, orb, spin)
    m.hamR = np.ascontiguousarray(hamR, dtype=np.int64)
    m.ham = np.ascontiguousarray(ham, dtype=np.complex128)
    m.rmatrix = None if rmatrix is None else np.ascontiguousarray(rmatrix, dtype=np.complex128)
    return m


def _rmatrix_term(model: Model, us, gauge_mat, hamk, directions):
    """The HasRMatrix term of gen_v_projected (velocity.rs:645-668; Lattice
    branch without the gauge factor and the diagonal zeroing): rk[d] =
    sum_R U_R rmatrix[R][d]; for each direction rk_p = sum_d w_d rk[d]
    (w_d != 0), Atom gauge applied and its diagonal zeroed; returns the
    i [hamk, rk_p] to add to v_p."""
    rm = model.rmatrix
    rk = np.einsum("r,rdij->dij", us[:rm.shape[0]], rm)
    out = []
    for w in directions:
        rp = np.zeros(rk.shape[1:], complex)
        for d in range(model.dim):
            if w[d] != 0.0:
                rp += w[d] * rk[d]
        if gauge_mat is not None:
            rp = rp * gauge_mat
            np.fill_diagonal(rp, 0.0)
        out.append(1j * (hamk @ rp - rp @ hamk))
    return out


# --------------------------------------------------------------------------
# k-space sampling (kpoints.rs:39-87, kpath.rs:17-70)
# --------------------------------------------------------------------------
def gen_kmesh(k_mesh):
    """gen_kmesh (kpoints.rs:49-87): k = (i/N1, j/N2, ...), first axis outermost."""
    k_mesh = [int(x) for x in k_mesh]
    if len(k_mesh) == 0 or any(n == 0 for n in k_mesh):  # kpoints.rs:39-47
        raise OracleError("InvalidKmeshDimensions", str(k_mesh))
    axes = [np.arange(n, dtype=np.float64) / float(n) for n in k_mesh]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.reshape(-1) for g in grids], axis=1)


def k_path(lat, path, nk):
    """Kpath::k_path (kpath.rs:17-70), including its node-index truncation."""
    lat = np.asarray(lat, dtype=np.float64)
    path = np.asarray(path, dtype=np.float64)
    n_node = path.shape[0]
    k_metric = np.linalg.inv(lat @ lat.T)
    k_node = np.zeros(n_node)
    for n in range(1, n_node):
        dk = path[n] - path[n - 1]
        k_node[n] = k_node[n - 1] + math.sqrt(dk @ (k_metric @ dk))
    node_index = [0]
    for n in range(1, n_node - 1):
        frac = k_node[n] / k_node[n_node - 1]
        node_index.append(int(frac * float(round(nk - 1))))  # kpath.rs:46 (truncating cast)
    node_index.append(nk - 1)
    k_dist = np.zeros(nk)
    k_vec = np.zeros((nk, path.shape[1]))
    k_vec[0] = path[0]
    for n in range(1, n_node):
        n_i, n_f = node_index[n - 1], node_index[n]
        for j in range(n_i, n_f + 1):
            frac = (j - n_i) / (n_f - n_i)
            k_dist[j] = k_node[n - 1] + frac * (k_node[n] - k_node[n - 1])
            k_vec[j] = (1.0 - frac) * path[n - 1] + frac * path[n]
    return k_vec, k_dist, k_node


# --------------------------------------------------------------------------
# Per-k kernels (model_physics.rs:43-148, velocity.rs:226-441,470-731)
# --------------------------------------------------------------------------
def _phases_R(model: Model, k):
    # model_physics.rs:58-90: exp(i 2 pi R.k), argument not reduced mod 1
    return np.exp(1j * (TWO_PI * (model.hamR.astype(np.float64) @ np.asarray(k, np.float64))))


def _phases_orb(model: Model, k):
    # model_physics.rs:102-136 (tau duplicated for spin)
    return np.exp(1j * (TWO_PI * (model.orb_sta() @ np.asarray(k, np.float64))))


def gen_ham(model: Model, k, gauge: str = "atom"):
    """gen_ham (model_physics.rs:43-148): zaxpy in stored R order, then the
    Atom-gauge similarity H[m,n] *= conj(phi_m) phi_n."""
    k = np.asarray(k, np.float64)
    if k.shape != (model.dim,):
        raise OracleError("KVectorLengthMismatch")
    us = _phases_R(model, k)
    hamk = np.zeros((model.nsta, model.nsta), dtype=np.complex128)
    for iR in range(model.ham.shape[0]):  # model_physics.rs:93-96
        hamk += us[iR] * model.ham[iR]
    if gauge == "lattice":
        return hamk
    phi = _phases_orb(model, k)
    return hamk * (np.conj(phi)[:, None] * phi[None, :])


def gen_v_projected(model: Model, k, gauge: str, directions):
    """gen_v_projected (velocity.rs:470-731), NoRMatrix branch, plus the
    HasRMatrix term when the model carries ``rmatrix`` (velocity.rs:645-668).

    v_p = D^* [ sum_R i (w_p . R_cart) U_R h_R + i (w_p.tau_n - w_p.tau_m) H_lat ] D
    with R_cart = R . lat (velocity.rs:530-531) and tau_cart = tau . lat
    (velocity.rs:575-579); rows with w.R == 0 are skipped (velocity.rs:594).
    Returns (v[n_proj, n, n], H) with H in the requested gauge.
    """
    k = np.asarray(k, np.float64)
    directions = np.atleast_2d(np.asarray(directions, np.float64))
    if k.shape != (model.dim,) or directions.shape[1] != model.dim:
        raise OracleError("KVectorLengthMismatch")
    n = model.nsta
    us = _phases_R(model, k)
    R0 = model.hamR.astype(np.float64) @ model.lat
    hamk = np.zeros((n, n), dtype=np.complex128)
    for iR in range(model.ham.shape[0]):
        hamk += us[iR] * model.ham[iR]
    n_proj = directions.shape[0]
    v = np.zeros((n_proj, n, n), dtype=np.complex128)
    for p in range(n_proj):
        w = directions[p]
        vv = np.zeros((n, n), dtype=np.complex128)
        for iR in range(model.ham.shape[0]):  # velocity.rs:587-599
            r_dot_w = float(R0[iR] @ w)
            if r_dot_w != 0.0:
                vv += (us[iR] * 1j * r_dot_w) * model.ham[iR]
        v[p] = vv
    rmat = getattr(model, "rmatrix", None) is not None
    if gauge == "lattice":
        if rmat:
            v = v + np.array(_rmatrix_term(model, us, None, hamk, directions))
        return v, hamk
    phi = _phases_orb(model, k)
    gauge_mat = np.conj(phi)[:, None] * phi[None, :]
    tau_proj = directions @ (model.orb_sta() @ model.lat).T  # velocity.rs:579
    for p in range(n_proj):  # velocity.rs:602-610
        diff = tau_proj[p][None, :] - tau_proj[p][:, None]  # [m, n] = tau_n - tau_m
        v[p] = (v[p] + 1j * diff * hamk) * gauge_mat
    hamk = hamk * gauge_mat  # velocity.rs:616-622
    if rmat:
        v = v + np.array(_rmatrix_term(model, us, gauge_mat, hamk, directions))
    return v, hamk


def gen_v(model: Model, k, gauge: str = "atom"):
    """Velocity::gen_v (velocity.rs:236-441), NoRMatrix: one Cartesian
    direction per axis.  Equal to gen_v_projected with identity rows."""
    return gen_v_projected(model, k, gauge, np.eye(model.dim))


# --------------------------------------------------------------------------
# Eigensolve (ndarray_lapack.rs:412-448, solve_ham.rs:52-63,110-129)
# --------------------------------------------------------------------------
def eigvalsh(H):
    """eigvalsh_v(&hamk, UPLO::Upper) -> LAPACK zheev('N', 'U')
    (solve_ham.rs:61, ndarray_lapack.rs:412-448).  The row-major buffer is
    handed to LAPACK as column-major, i.e. H^T, so 'U' reads the LOWER
    triangle of H (H[i][j], i >= j): for an H that is Hermitian only to
    rounding this is the matrix whose eigenvalues the reference returns."""
    w, _, info = _lapack.zheev(np.asarray(H, np.complex128), compute_v=0, lower=1)
    if info != 0:
        raise OracleError("EigenvalueComputationFailed", f"info={info}")
    return w


def eigh(H):
    """LAPACK zheev('V'): ascending eigenvalues E and the true eigenvectors
    u_j as the columns of U (H U = U diag(E)).

    The reference's ``hamk.eigh(UPLO::Lower)`` hands back ``conj(U)`` because
    the row-major buffer is read column-major (``evec_transform_sanity``,
    lib.rs:1231-1277: ``E_vec^T H E_vec^* = diag``); for the same reason its
    'L' reads the UPPER triangle of H (H[i][j], i <= j), as zheev('U') on H
    does here."""
    w, v, info = _lapack.zheev(np.asarray(H, np.complex128), compute_v=1, lower=0)
    if info != 0:
        raise OracleError("EigenvalueComputationFailed", f"info={info}")
    return w, v


def solve_band_range_onek(model: Model, k, range_, epsilon):
    """solve_band_range_onek (solve_ham.rs:64-80) -> eigvalsh_r, zheevr with
    RANGE = 'V' (ndarray_lapack.rs:301-395): the eigenvalues in the half-open
    interval (vl, vu] of the matrix eigvalsh reads (H's lower triangle).
    epsilon is zheevr's ABSTOL (accuracy only; not used here)."""
    e = eigvalsh(gen_ham(model, k, "atom"))
    return e[(e > range_[0]) & (e <= range_[1])]


def solve_range_onek(model: Model, k, range_, epsilon):
    """solve_range_onek (solve_ham.rs:130-147) -> eigh_r (ndarray_lapack.rs:
    217-286) on the raw buffer ('U' of H^T, i.e. H's lower triangle), then
    conjugated: the eigenpairs in (vl, vu] with the eigenvectors as rows."""
    w, v, info = _lapack.zheev(np.asarray(gen_ham(model, k, "atom"), np.complex128), compute_v=1, lower=1)
    if info != 0:
        raise OracleError("EigenvalueComputationFailed", f"info={info}")
    keep = (w > range_[0]) & (w <= range_[1])
    return w[keep], v[:, keep].T.copy()


def solve_band_onek(model: Model, k):
    """solve_band_onek (solve_ham.rs:52-63): eigenvalues of the Atom-gauge H."""
    return eigvalsh(gen_ham(model, k, "atom"))


def solve_onek(model: Model, k):
    """solve_onek (solve_ham.rs:110-129): (E, vectors) with the eigenvectors as
    ROWS (``conjugate(E_vec)`` = Hermitian transpose of ``conj(U)`` = U^T)."""
    E, U = eigh(gen_ham(model, k, "atom"))
    return E, U.T.copy()


def solve_band_all(model: Model, kvec):
    """solve_band_all(_parallel) (solve_ham.rs:82-107)."""
    kvec = np.atleast_2d(np.asarray(kvec, np.float64))
    return np.stack([solve_band_onek(model, k) for k in kvec]) if len(kvec) else np.zeros((0, model.nsta))


def solve_all(model: Model, kvec):
    """solve_all(_parallel) (solve_ham.rs:150-184)."""
    kvec = np.atleast_2d(np.asarray(kvec, np.float64))
    Es, Vs = zip(*[solve_onek(model, k) for k in kvec])
    return np.stack(Es), np.stack(Vs)


# --------------------------------------------------------------------------
# Occupation (thermodynamics.rs:86-118)
# --------------------------------------------------------------------------
def fermi(energy, mu, T_kelvin):
    """Occupation::value_unchecked -> fermi_from_width (thermodynamics.rs:86-118)."""
    width = 0.0 if T_kelvin <= 0.0 else T_kelvin * BOLTZMANN_CONSTANT_EV_PER_K
    e = np.asarray(energy, np.float64)
    if width == 0.0:
        return np.where(e < mu, 1.0, np.where(e > mu, 0.0, 0.5))
    x = (e - mu) / width
    with np.errstate(over="ignore"):
        mid = 1.0 / (1.0 + np.exp(np.clip(x, -50.0, 50.0)))
    return np.where(x > 50.0, 0.0, np.where(x < -50.0, 1.0, mid))


def spin_matrix(norb, spin):
    """build_spin_matrix (response/helpers.rs:11-39)."""
    n = 2 * norb
    m = np.zeros((n, n), dtype=np.complex128)
    idx = np.arange(norb)
    if spin is None:
        return np.eye(n, dtype=np.complex128)
    if spin == "X":
        m[idx, idx + norb] = 0.5
        m[idx + norb, idx] = 0.5
    elif spin == "Y":
        m[idx, idx + norb] = -0.5j
        m[idx + norb, idx] = 0.5j
    elif spin == "Z":
        m[idx, idx] = 0.5
        m[idx + norb, idx + norb] = -0.5
    else:
        raise OracleError("InvalidSpinDirection", repr(spin))
    return m


# --------------------------------------------------------------------------
# Berry curvature and Hall conductivity (response/traits.rs:160-200,
# response/linear/mod.rs:160-225, response/config.rs:141-303)
# --------------------------------------------------------------------------
def validate_direction(direction, rank, dim):
    """validate_direction_matrix (response/config.rs:202-236)."""
    d = np.atleast_2d(np.asarray(direction, np.float64))
    if d.shape[0] != rank or d.shape[1] != dim:
        raise OracleError("DimensionMismatch", "direction")
    if not np.all(np.isfinite(d)):
        raise OracleError("InvalidResponseParameter", "direction")
    if np.any(np.all(d == 0.0, axis=1)):
        raise OracleError("InvalidResponseParameter", "direction")
    return d


def berry_curvature_at(model: Model, k, direction, eta, spin=None):
    """berry_curvature_at_impl (response/traits.rs:160-200).

    A = U^H J U, B = U^H v_b U (traits.rs:177-180), K_nm = A_nm B_mn (:181),
    Omega_n = sum_{m != n} -2 Im K_nm / ((E_n - E_m)^2 + eta^2) (:185-195).
    Returns (Omega[n], E[n])."""
    direction = validate_direction(direction, 2, model.dim)
    if not (math.isfinite(eta) and eta >= 0.0):  # config.rs:265-273
        raise OracleError("InvalidResponseParameter", "broadening")
    if spin is not None and not model.spin:
        raise OracleError("SpinNotAllowed", str(spin))
    v, H = gen_v_projected(model, k, "atom", direction)
    E, U = eigh(H)
    if model.spin and spin is not None:  # traits.rs:170-175
        S = spin_matrix(model.norb, spin)
        J = 0.5 * (S @ v[0] + v[0] @ S)
    else:
        J = v[0]
    Uh = U.conj().T
    A = Uh @ (J @ U)
    B = Uh @ (v[1] @ U)
    K = A * B.T
    dE = E[:, None] - E[None, :]
    denom = dE * dE + eta * eta
    with np.errstate(divide="ignore", invalid="ignore"):
        terms = -2.0 * K.imag / denom
    np.fill_diagonal(terms, 0.0)
    return terms.sum(axis=1), E


def occupied_berry_curvature_on(model: Model, kvec, direction, mu, T, eta, spin=None):
    """occupied_berry_curvature_on (response/traits.rs:104-150)."""
    out = []
    for k in np.atleast_2d(kvec):
        om, E = berry_curvature_at(model, k, direction, eta, spin)
        out.append(float(np.sum(om * fermi(E, mu, T))))
    return np.array(out)


def hall_conductivity_direct(model: Model, kmesh, direction, mu, T=0.0, eta=1e-3, spin=None):
    """hall_conductivity, Integration::Direct (response/linear/mod.rs:160-225):
    sigma(mu) = sum_k sum_n Omega_n f(E_n, mu) / N_k / det(lat)."""
    mu = np.atleast_1d(np.asarray(mu, np.float64))
    if mu.size == 0 or not np.all(np.isfinite(mu)):
        raise OracleError("InvalidResponseParameter", "chemical_potentials")
    if not (math.isfinite(T) and T >= 0.0):
        raise OracleError("InvalidResponseParameter", "T")
    kvec = gen_kmesh(kmesh)
    det = float(np.linalg.det(model.lat))
    om_all, E_all = [], []
    for k in kvec:
        om, E = berry_curvature_at(model, k, direction, eta, spin)
        om_all.append(om)
        E_all.append(E)
    om_all = np.array(om_all)
    E_all = np.array(E_all)
    nk = kvec.shape[0]
    return np.array([np.sum(om_all * fermi(E_all, m, T)) / nk / det for m in mu])


def dos(model: Model, kmesh, E_min, E_max, E_n, sigma):
    """Model::dos (model_physics.rs:179-226): Gaussian-smeared DOS."""
    if E_min >= E_max:
        raise OracleError("InvalidEnergyRange")
    if E_n == 0:
        raise OracleError("InvalidDosParameter", "E_n")
    if not (math.isfinite(sigma) and sigma > 0.0):
        raise OracleError("InvalidDosParameter", "sigma")
    kvec = gen_kmesh(kmesh)
    evals = solve_band_all(model, kvec).reshape(-1)
    E = np.linspace(E_min, E_max, E_n)
    return E, dos_from_eigenvalues(evals, E, sigma) / kvec.shape[0]


def dos_from_eigenvalues(evals, E, sigma):
    """The fold of model_physics.rs:210-223 (sum over every eigenvalue)."""
    sigma0 = 1.0 / sigma
    pi0 = 1.0 / math.sqrt(2.0 * math.pi)
    acc = np.zeros_like(E)
    for chunk in np.array_split(evals, max(1, evals.size // 4096)):
        A = (E[None, :] - chunk[:, None]) * sigma0
        acc += (np.exp(-A * A / 2.0) * sigma0 * pi0).sum(axis=0)
    return acc


# --------------------------------------------------------------------------
# Batched helpers for larger fixtures (same formulas; summation by matmul)
# --------------------------------------------------------------------------
def quantum_geometry_at(model: Model, k, direction, eta):
    """quantum_geometry_at_impl (quantum_geometry.rs:149-190): with A = U^H v_a U,
    B = U^H v_b U (Atom gauge), G_n = sum_{m != n} A_nm B_mn / ((E_n - E_m)^2 +
    eta^2); returns (metric = Re G, berry = -2 Im G, E)."""
    direction = validate_direction(direction, 2, model.dim)
    v, H = gen_v_projected(model, k, "atom", direction)
    e, U = eigh(H)
    Uh = U.conj().T
    A = Uh @ (v[0] @ U)
    B = Uh @ (v[1] @ U)
    K = A * B.T
    dE = e[:, None] - e[None, :]
    t = K / (dE * dE + eta * eta)
    np.fill_diagonal(t, 0.0)
    G = t.sum(axis=1)
    return G.real, -2.0 * G.imag, e


def quantum_geometry_direct(model: Model, kmesh, direction, mu, T=0.0, eta=1e-3):
    """Model::quantum_geometry, Integration::Direct (quantum_geometry.rs:196-240):
    (metric[mu], berry[mu]) = sum over the mesh and bands of (g, Omega) f(E, mu)
    / N_k / det(lat)."""
    ks = gen_kmesh(kmesh)
    mu = np.atleast_1d(np.asarray(mu, float))
    gm = np.zeros(mu.size)
    bm = np.zeros(mu.size)
    for k in ks:
        g, b, e = quantum_geometry_at(model, k, direction, eta)
        for i, m in enumerate(mu):
            f = fermi(e, m, T)
            gm[i] += (g * f).sum()
            bm[i] += (b * f).sum()
    norm = 1.0 / len(ks) / np.linalg.det(model.lat)
    return gm * norm, bm * norm


def berry_curvature_dipole_onek(model: Model, k, current, dir_2, dir_3, eta, spin=None):
    """berry_curvature_dipole_n_onek (response/nonlinear/mod.rs:89-163):
    per band d_{dir_3}E_n * Omega_n^{current, dir_2}, with
    d E_n = Re (U^H v_3 U)_nn and Omega from A1 = U^H J U, A2 = U^H v_2 U
    (J = v_current, or its spin current for a spinful model)."""
    d = np.array([current, dir_2, dir_3], float)
    v, H = gen_v_projected(model, k, "atom", d)
    e, U = eigh(H)
    S = spin_matrix(model.norb, spin) if (model.spin and spin is not None) else None
    J = v[0] if S is None else 0.5 * (S @ v[0] + v[0] @ S)
    Uh = U.conj().T
    A1 = Uh @ (J @ U)
    A2 = Uh @ (v[1] @ U)
    dE = np.real(np.diag(Uh @ (v[2] @ U)))
    U0 = 1.0 / ((e[:, None] - e[None, :]) ** 2 + eta * eta)
    np.fill_diagonal(U0, 0.0)
    om = -2.0 * np.einsum("ij,ji->i", A1 * U0, A2).imag
    return om * dE, e


def fermi_derivative(energy, mu, T_kelvin):
    """fermi_derivative_from_width (thermodynamics.rs:122-126): f (1 - f) / width."""
    width = T_kelvin * 8.617333262145e-5
    f = fermi(energy, mu, T_kelvin)
    return f * (1.0 - f) / width


def extrinsic_nonlinear_hall_direct(model: Model, kmesh, direction, mu, T, eta, spin=None, symmetrized=True):
    """Model::extrinsic_nonlinear_hall, Integration::Direct
    (response/nonlinear/mod.rs:256-347): rows (current, field_1, field_2);
    Symmetrized averages the (field_1, field_2) and (field_2, field_1) kernels
    when the field rows differ."""
    direction = np.asarray(direction, float)
    cur, f1, f2 = direction
    mu = np.atleast_1d(np.asarray(mu, float))
    ks = gen_kmesh(kmesh)
    norm = 1.0 / len(ks) / np.linalg.det(model.lat)

    def component(a, b, c):
        out = np.zeros(mu.size)
        for k in ks:
            kern, e = berry_curvature_dipole_onek(model, k, a, b, c, eta, spin)
            for i, m in enumerate(mu):
                out[i] += (kern * fermi_derivative(e, m, T)).sum()
        return out * norm

    first = component(cur, f1, f2)
    if symmetrized and not np.array_equal(f1, f2):
        return 0.5 * (first + component(cur, f2, f1))
    return first


def gen_ham_batch(model: Model, kvec, gauge="atom"):
    """gen_ham for many k at once: phases [N_k, N_R] times ham [N_R, n*n].
    Same formula as gen_ham; the R sum is done by BLAS instead of zaxpy."""
    kvec = np.atleast_2d(np.asarray(kvec, np.float64))
    n = model.nsta
    U = np.exp(1j * (TWO_PI * (kvec @ model.hamR.astype(np.float64).T)))
    H = (U @ model.ham.reshape(model.ham.shape[0], n * n)).reshape(-1, n, n)
    if gauge == "lattice":
        return H
    phi = np.exp(1j * (TWO_PI * (kvec @ model.orb_sta().T)))
    return H * (np.conj(phi)[:, :, None] * phi[:, None, :])


def gen_v_projected_batch(model: Model, kvec, directions):
    """Atom-gauge gen_v_projected for many k (same formula, BLAS R-sum)."""
    kvec = np.atleast_2d(np.asarray(kvec, np.float64))
    directions = np.atleast_2d(np.asarray(directions, np.float64))
    n = model.nsta
    nR = model.ham.shape[0]
    hflat = model.ham.reshape(nR, n * n)
    U = np.exp(1j * (TWO_PI * (kvec @ model.hamR.astype(np.float64).T)))
    Hl = (U @ hflat).reshape(-1, n, n)
    wR = directions @ (model.hamR.astype(np.float64) @ model.lat).T  # [p, R]
    phi = np.exp(1j * (TWO_PI * (kvec @ model.orb_sta().T)))
    gm = np.conj(phi)[:, :, None] * phi[:, None, :]
    tau_proj = directions @ (model.orb_sta() @ model.lat).T
    vs = []
    for p in range(directions.shape[0]):
        vR = ((U * (1j * wR[p])[None, :]) @ hflat).reshape(-1, n, n)
        diff = tau_proj[p][None, :] - tau_proj[p][:, None]
        vs.append((vR + 1j * diff[None] * Hl) * gm)
    v, H = np.stack(vs, axis=1), Hl * gm
    if getattr(model, "rmatrix", None) is not None:
        for ik in range(kvec.shape[0]):
            v[ik] += np.array(_rmatrix_term(model, U[ik], gm[ik], H[ik], directions))
    return v, H


def berry_curvature_batch(model: Model, kvec, direction, eta, spin=None):
    """berry_curvature_at_impl for many k (same formula)."""
    direction = validate_direction(direction, 2, model.dim)
    v, H = gen_v_projected_batch(model, kvec, direction)
    nk = H.shape[0]
    n = model.nsta
    om = np.zeros((nk, n))
    E = np.zeros((nk, n))
    S = spin_matrix(model.norb, spin) if (model.spin and spin is not None) else None
    for ik in range(nk):
        e, U = eigh(H[ik])
        J = v[ik, 0] if S is None else 0.5 * (S @ v[ik, 0] + v[ik, 0] @ S)
        Uh = U.conj().T
        A = Uh @ (J @ U)
        B = Uh @ (v[ik, 1] @ U)
        K = A * B.T
        dE = e[:, None] - e[None, :]
        t = -2.0 * K.imag / (dE * dE + eta * eta)
        np.fill_diagonal(t, 0.0)
        om[ik] = t.sum(axis=1)
        E[ik] = e
    return om, E
