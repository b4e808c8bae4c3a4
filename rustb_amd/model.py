"""Host-side mirror of the reference `Model<SPIN, DIM, R>` surface for the
k-mesh hot path.  Construction goes through the C++ builder of librtb_hip.so
(model_build.rs semantics); every physics call goes to the HIP kernels through
the C ABI.  Method names, argument meaning and error behaviour follow
src/model_build.rs, src/model_physics.rs, src/velocity.rs, src/solve_ham.rs
and src/response/{traits.rs,linear/mod.rs,config.rs} of the reference.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from ._lib import TbError, check, load, ptr

GAUGE = {"lattice": 0, "atom": 1, "Lattice": 0, "Atom": 1}
SPIN_DIR = {None: 0, "X": 1, "Y": 2, "Z": 3, 0: 0, 1: 1, 2: 2, 3: 3}
PRECISION = {"fp64": 0, "mixed": 1, 0: 0, 1: 1}


@dataclass
class Parameters:
    """response/config.rs:41-87 (Integration::Direct only on this path).
    Defaults: T=[0], eta=1e-3."""

    kmesh: tuple
    direction: np.ndarray
    mu: np.ndarray
    T: np.ndarray = field(default_factory=lambda: np.array([0.0]))
    eta: float = 1e-3
    spin: Optional[str] = None
    integration: str = "Direct"
    precision: str = "fp64"
    field_symmetry: str = "Symmetrized"  # FieldSymmetry (config.rs): or "Ordered"

    @classmethod
    def rank2(cls, kmesh, direction_a, direction_b, mu, **kw):
        """Parameters::rank2 (config.rs:95-102)."""
        return cls(tuple(kmesh), np.array([direction_a, direction_b], float), np.atleast_1d(np.asarray(mu, float)), **kw)

    @classmethod
    def rank3(cls, kmesh, current, field_1, field_2, mu, **kw):
        """Parameters::rank3 (config.rs): direction rows (current, field_1, field_2)."""
        return cls(tuple(kmesh), np.array([current, field_1, field_2], float),
                   np.atleast_1d(np.asarray(mu, float)), **kw)

    @classmethod
    def at_mu(cls, kmesh, direction, mu: float, **kw):
        return cls(tuple(kmesh), np.asarray(direction, float), np.array([float(mu)]), **kw)

    def with_temperature(self, kelvin: float) -> "Parameters":
        self.T = np.array([float(kelvin)])
        return self


@dataclass
class BandBerryCurvature:
    """response/traits.rs:22-27."""

    berry_curvature: np.ndarray
    energies: np.ndarray


@dataclass
class HallConductivityResult:
    """response/linear/mod.rs:51-63."""

    chemical_potentials: np.ndarray
    conductivity: np.ndarray

    def single(self):
        return float(self.conductivity[0]) if self.conductivity.size == 1 else None


@dataclass
class NonlinearHallResult:
    """response/nonlinear/mod.rs:52-61."""

    chemical_potentials: np.ndarray
    conductivity: np.ndarray
    diagnostics: Optional[dict] = None


@dataclass
class BandQuantumGeometry:
    """quantum_geometry.rs:30-40: per band at one k."""

    metric: np.ndarray
    berry_curvature: np.ndarray
    energies: np.ndarray


@dataclass
class QuantumGeometryMap:
    """quantum_geometry.rs:42-51: [nk, nsta] each."""

    metric: np.ndarray
    berry_curvature: np.ndarray
    energies: np.ndarray


@dataclass
class QuantumGeometryResult:
    """quantum_geometry.rs:53-66: occupation-weighted, per chemical potential."""

    chemical_potentials: np.ndarray
    metric: np.ndarray
    berry_curvature: np.ndarray
    diagnostics: Optional[dict] = None


@dataclass
class Atom:
    """An atomic site (atom_struct.rs Atom): species, fractional position and
    the indices of its orbitals."""

    species: str
    position: np.ndarray
    orbitals: list

    def norb(self):
        return len(self.orbitals)


class Model:
    """Tight-binding model (model.rs:150-158) with device-resident storage."""

    def __init__(self, lat, orb, spin: bool, _builder=True):
        lib = load()
        self.lat = np.ascontiguousarray(lat, dtype=np.float64)
        self.orb = np.ascontiguousarray(orb, dtype=np.float64)
        if self.lat.ndim != 2 or self.lat.shape[0] != self.lat.shape[1]:
            raise TbError(-11, "LatticeDimensionError: lattice must be square")
        self.dim = self.lat.shape[0]
        if self.orb.ndim != 2 or self.orb.shape[1] != self.dim:
            raise TbError(-11, "orbital_position_shape")
        self.spin = bool(spin)
        self.norb = self.orb.shape[0]
        self.nsta = 2 * self.norb if self.spin else self.norb
        self._b = None
        self._dev = None
        self._arrays = None
        self.rmatrix = None  # [N_R, dim, n, n] position matrices (HasRMatrix)
        if _builder:
            h = ctypes.c_void_p()
            check(lib.rtb_builder_new(self.dim, int(self.spin), self.norb, ptr(self.lat), ptr(self.orb), ctypes.byref(h)))
            self._b = h

    # ---- construction (model_build.rs) ----
    @classmethod
    def tb_model(cls, lat, orb, spin: bool = False) -> "Model":
        """Model::tb_model (model_build.rs:212-285)."""
        return cls(lat, orb, spin)

    @classmethod
    def from_arrays(cls, lat, orb, spin, hamR, ham, rmatrix=None) -> "Model":
        """A model from a1 storage arrays (hamR [N_R, dim], ham [N_R, n, n];
        rmatrix [N_R, dim, n, n] for a model with position matrices)."""
        m = cls(lat, orb, spin, _builder=False)
        hamR = np.ascontiguousarray(hamR, dtype=np.int64)
        ham = np.ascontiguousarray(ham, dtype=np.complex128)
        if ham.shape != (hamR.shape[0], m.nsta, m.nsta):
            raise TbError(-11, "hamiltonian_shape")
        m._arrays = (hamR, ham)
        if rmatrix is not None:
            rmatrix = np.ascontiguousarray(rmatrix, dtype=np.complex128)
            if rmatrix.shape != (hamR.shape[0], m.dim, m.nsta, m.nsta):
                raise TbError(-11, "position_matrix_shape")
            m.rmatrix = rmatrix
        return m

    @classmethod
    def from_hr_with_rmatrix(cls, path: str, file_name: str, zero_energy: float, spin=None,
                             n_threads: int = 0) -> "Model":
        """Model::from_hr_with_rmatrix (wannier90.rs:1208-1211): from_hr plus
        {path}{file_name}_r.dat (required), kept as ``rmatrix``; every
        velocity then carries the term i[H, r] (velocity.rs:377-402, 627-668)."""
        return cls.from_hr(path, file_name, zero_energy, spin, n_threads, _rmatrix=True)

    @classmethod
    def from_hr(cls, path: str, file_name: str, zero_energy: float, spin=None, n_threads: int = 0,
                _rmatrix: bool = False) -> "Model":
        """Wannier90::from_hr (wannier90.rs:49-1228), R = NoRMatrix: reads
        {path}{file_name}_hr.dat, .win and, when present, _centres.xyz and
        _wsvec.dat.  spin: the Model's SPIN (True/False; a different `spinors`
        setting in the .win file raises), or None to take the file's.
        Adds ``atoms`` (species, fractional position, orbital indices) and
        ``orb_projection`` (orbital names) to the model."""
        lib = load()
        h = ctypes.c_void_p()
        sp = -1 if spin is None else int(bool(spin))
        reader = lib.rtb_w90_read_with_rmatrix if _rmatrix else lib.rtb_w90_read
        check(reader(path.encode(), file_name.encode(), float(zero_energy), sp, int(n_threads), ctypes.byref(h)))
        try:
            v = [ctypes.c_int() for _ in range(5)]
            check(lib.rtb_w90_info(h, *[ctypes.byref(x) for x in v]))
            file_spin, norb, nsta, nR, natom = (x.value for x in v)
            hamR = np.zeros((nR, 3), np.int64)
            ham = np.empty((nR, nsta, nsta), np.complex128)
            orb = np.zeros((norb, 3))
            lat = np.zeros((3, 3))
            proj = np.zeros(norb, np.int32)
            check(lib.rtb_w90_export(h, ptr(hamR), ptr(ham), ptr(orb), ptr(lat), ptr(proj)))
            rmat = None
            if _rmatrix:
                rmat = np.empty((nR, 3, nsta, nsta), np.complex128)
                check(lib.rtb_w90_export_rmatrix(h, ptr(rmat)))
            atoms = []
            for i in range(natom):
                z, first, cnt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
                pos = np.zeros(3)
                check(lib.rtb_w90_atom(h, i, ctypes.byref(z), ptr(pos), ctypes.byref(first), ctypes.byref(cnt)))
                atoms.append(Atom(lib.rtb_element_symbol(z.value).decode(), pos,
                                  list(range(first.value, first.value + cnt.value))))
        finally:
            lib.rtb_w90_destroy(h)
        m = cls(lat, orb, bool(file_spin), _builder=False)
        m._arrays = (hamR, ham)
        m.rmatrix = rmat
        m.atoms = atoms
        m.orb_projection = [lib.rtb_orbital_projection_name(int(c)).decode() for c in proj]
        return m

    def natom(self):
        return len(getattr(self, "atoms", []))

    def _R(self, R):
        R = np.ascontiguousarray(R, dtype=np.int64)
        if R.shape != (self.dim,):
            raise TbError(-10, "the R length should equal to dim_r")
        return R

    def _mutate(self):
        if self._b is None:
            raise TbError(-12, "model built from arrays is read-only")
        self._dev_free()
        self._arrays = None

    def set_hop(self, tmp, ind_i, ind_j, R, pauli=None):
        """set_hop (model_build.rs:354-470)."""
        self._mutate()
        t = complex(tmp)
        check(load().rtb_builder_set_hop(self._b, t.real, t.imag, ind_i, ind_j, ptr(self._R(R)), SPIN_DIR[pauli]))

    def add_hop(self, tmp, ind_i, ind_j, R, pauli=None):
        """add_hop (model_build.rs:477-558)."""
        self._mutate()
        t = complex(tmp)
        check(load().rtb_builder_add_hop(self._b, t.real, t.imag, ind_i, ind_j, ptr(self._R(R)), SPIN_DIR[pauli]))

    def add_element(self, tmp, ind_i, ind_j, R):
        """add_element (model_build.rs:613-687)."""
        self._mutate()
        t = complex(tmp)
        check(load().rtb_builder_add_element(self._b, t.real, t.imag, ind_i, ind_j, ptr(self._R(R))))

    def set_onsite(self, values, pauli=None):
        """set_onsite (model_build.rs:689-700)."""
        self._mutate()
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != (self.norb,):
            raise TbError(-2, "onsite input length must equal norb")
        check(load().rtb_builder_set_onsite(self._b, ptr(v), SPIN_DIR[pauli]))

    def add_onsite(self, values, pauli=None):
        """add_onsite (model_build.rs:737-750)."""
        self._mutate()
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != (self.norb,):
            raise TbError(-2, "onsite input length must equal norb")
        check(load().rtb_builder_add_onsite(self._b, ptr(v), SPIN_DIR[pauli]))

    # ---- storage ----
    def _storage(self):
        if self._arrays is None:
            lib = load()
            nR = lib.rtb_builder_n_R(self._b)
            hamR = np.zeros((nR, self.dim), np.int64)
            ham = np.zeros((nR, self.nsta, self.nsta), np.complex128)
            check(lib.rtb_builder_export(self._b, ptr(hamR), ptr(ham)))
            self._arrays = (hamR, ham)
        return self._arrays

    @property
    def hamR(self):
        return self._storage()[0]

    @property
    def ham(self):
        return self._storage()[1]

    def dim_r(self):
        return self.dim

    def det_lat(self):
        return float(np.linalg.det(self.lat))

    # ---- device model ----
    def _device(self):
        if self._dev is None:
            hamR, ham = self._storage()
            h = ctypes.c_void_p()
            check(load().rtb_model_create(self.dim, int(self.spin), self.norb, hamR.shape[0], ptr(hamR), ptr(ham),
                                          ptr(self.orb), ptr(self.lat), ctypes.byref(h)))
            self._dev = h
            if self.rmatrix is not None:
                check(load().rtb_model_set_rmatrix(h, ptr(self.rmatrix)))
        return self._dev

    def _dev_free(self):
        if self._dev is not None:
            load().rtb_model_destroy(self._dev)
            self._dev = None

    def __del__(self):
        try:
            self._dev_free()
            if self._b is not None:
                load().rtb_builder_destroy(self._b)
                self._b = None
        except Exception:
            pass

    def _kvec(self, kvec):
        k = np.ascontiguousarray(np.atleast_2d(np.asarray(kvec, dtype=np.float64)))
        if k.shape[1] != self.dim:
            raise TbError(-1, f"the k-vector's length must equal the dimension of the model ({self.dim})")
        return k

    # ---- hot path ----
    def gen_ham_all(self, kvec, gauge="atom"):
        """gen_ham (model_physics.rs:43-148) for every row of kvec."""
        k = self._kvec(kvec)
        H = np.zeros((k.shape[0], self.nsta, self.nsta), np.complex128)
        check(load().rtb_gen_ham_batch(self._device(), ptr(k), k.shape[0], GAUGE[gauge], None, 0, ptr(H), None))
        return H

    def gen_ham(self, k, gauge="atom"):
        return self.gen_ham_all(np.asarray(k, float)[None, :], gauge)[0]

    def gen_v_projected_all(self, kvec, gauge, directions):
        """gen_v_projected (velocity.rs:470-731) for every row of kvec:
        returns (v [nk, n_proj, n, n], H [nk, n, n])."""
        k = self._kvec(kvec)
        d = np.ascontiguousarray(np.atleast_2d(np.asarray(directions, np.float64)))
        if d.shape[1] != self.dim:
            raise TbError(-2, "directions must have dim_r columns")
        np_ = d.shape[0]
        H = np.zeros((k.shape[0], self.nsta, self.nsta), np.complex128)
        V = np.zeros((k.shape[0], np_, self.nsta, self.nsta), np.complex128)
        check(load().rtb_gen_ham_batch(self._device(), ptr(k), k.shape[0], GAUGE[gauge], ptr(d), np_, ptr(H), ptr(V)))
        return V, H

    def gen_v_projected(self, k, gauge, directions):
        V, H = self.gen_v_projected_all(np.asarray(k, float)[None, :], gauge, directions)
        return V[0], H[0]

    def gen_v(self, k, gauge="atom"):
        """Velocity::gen_v (velocity.rs:236-441): one Cartesian axis per row."""
        return self.gen_v_projected(k, gauge, np.eye(self.dim))

    def solve_band_all(self, kvec, precision="fp64"):
        """solve_band_all(_parallel) (solve_ham.rs:82-107)."""
        k = self._kvec(kvec)
        E = np.zeros((k.shape[0], self.nsta))
        check(load().rtb_solve_band_all(self._device(), ptr(k), k.shape[0], PRECISION[precision], ptr(E)))
        return E

    solve_band_all_parallel = solve_band_all

    def solve_band_onek(self, k, precision="fp64"):
        return self.solve_band_all(np.asarray(k, float)[None, :], precision)[0]

    def solve_all(self, kvec, precision="fp64"):
        """solve_all(_parallel) (solve_ham.rs:150-184): eigenvectors as rows."""
        k = self._kvec(kvec)
        E = np.zeros((k.shape[0], self.nsta))
        V = np.zeros((k.shape[0], self.nsta, self.nsta), np.complex128)
        check(load().rtb_solve_all(self._device(), ptr(k), k.shape[0], PRECISION[precision], ptr(E), ptr(V)))
        return E, V

    solve_all_parallel = solve_all

    def solve_band_range_onek(self, k, range_, epsilon, precision="fp64"):
        """solve_band_range_onek (solve_ham.rs:64-80): eigenvalues in
        (range_[0], range_[1]] (zheevr RANGE = 'V')."""
        kk = self._kvec(np.asarray(k, float)[None, :])
        E = np.zeros((1, self.nsta))
        cnt = np.zeros(1, np.int32)
        check(load().rtb_solve_band_range(self._device(), ptr(kk), 1, float(range_[0]), float(range_[1]),
                                          float(epsilon), PRECISION[precision], ptr(E), ptr(cnt)))
        return E[0, :cnt[0]].copy()

    def solve_range_onek(self, k, range_, epsilon, precision="fp64"):
        """solve_range_onek (solve_ham.rs:130-147): eigenvalues in (range_[0],
        range_[1]] and their eigenvectors as rows."""
        kk = self._kvec(np.asarray(k, float)[None, :])
        E = np.zeros((1, self.nsta))
        V = np.zeros((1, self.nsta, self.nsta), np.complex128)
        cnt = np.zeros(1, np.int32)
        check(load().rtb_solve_range(self._device(), ptr(kk), 1, float(range_[0]), float(range_[1]), float(epsilon),
                                     PRECISION[precision], ptr(E), ptr(V), ptr(cnt)))
        return E[0, :cnt[0]].copy(), V[0, :cnt[0]].copy()

    def solve_onek(self, k, precision="fp64"):
        E, V = self.solve_all(np.asarray(k, float)[None, :], precision)
        return E[0], V[0]

    def _dir(self, params: Parameters):
        d = np.ascontiguousarray(np.asarray(params.direction, np.float64))
        if d.shape != (2, self.dim):
            raise TbError(-2, f"direction must have shape (2, {self.dim})")
        return d

    def berry_curvature_on(self, kvec, params: Parameters):
        """berry_curvature_at (response/traits.rs:60-78) for every row of kvec."""
        k = self._kvec(kvec)
        om = np.zeros((k.shape[0], self.nsta))
        E = np.zeros((k.shape[0], self.nsta))
        check(load().rtb_berry_curvature(self._device(), ptr(k), k.shape[0], ptr(self._dir(params)),
                                         SPIN_DIR[params.spin], float(params.eta), PRECISION[params.precision],
                                         ptr(om), ptr(E)))
        return om, E

    def berry_curvature_at(self, k, params: Parameters) -> BandBerryCurvature:
        om, E = self.berry_curvature_on(np.asarray(k, float)[None, :], params)
        return BandBerryCurvature(om[0], E[0])

    def occupied_berry_curvature_on(self, kvec, params: Parameters):
        """occupied_berry_curvature_on (response/traits.rs:104-150)."""
        k = self._kvec(kvec)
        mu = np.atleast_1d(np.asarray(params.mu, float))
        if mu.size != 1:
            raise TbError(-6, "mu: occupied_berry_curvature_on expects a single chemical potential")
        T = np.atleast_1d(np.asarray(params.T, float))
        if T.size == 0:
            raise TbError(-6, "T: must contain at least one value")
        out = np.zeros(k.shape[0])
        check(load().rtb_occupied_berry_curvature(self._device(), ptr(k), k.shape[0], ptr(self._dir(params)),
                                                  SPIN_DIR[params.spin], float(params.eta), float(mu[0]),
                                                  float(T[0]), PRECISION[params.precision], ptr(out)))
        return out

    def occupied_berry_curvature_at(self, k, params: Parameters) -> float:
        return float(self.occupied_berry_curvature_on(np.asarray(k, float)[None, :], params)[0])

    def hall_conductivity(self, params: Parameters) -> HallConductivityResult:
        """hall_conductivity (response/linear/mod.rs:160-276): Integration::Direct
        (the k-sum of occupied Omega), or EnergyCut for a 2D or 3D model (band
        tracking and the per-triangle / per-tetrahedron Fermi cut, always FP64);
        Simplex is rejected as the
        reference does (:166-176)."""
        if params.integration not in ("Direct", "EnergyCut"):
            raise TbError(-6, "integration: hall_conductivity supports Integration::Direct or EnergyCut, not Simplex")
        kmesh = np.ascontiguousarray(np.asarray(params.kmesh, dtype=np.uint64))
        if kmesh.shape != (self.dim,):
            raise TbError(-3, "kmesh must have dim_r entries")
        mu = np.ascontiguousarray(np.atleast_1d(np.asarray(params.mu, np.float64)))
        T = np.atleast_1d(np.asarray(params.T, float))
        if T.size == 0:
            raise TbError(-6, "T: must contain at least one value")
        sig = np.zeros(mu.size)
        if params.integration == "EnergyCut":
            check(load().rtb_hall_energy_cut(self._device(), ptr(kmesh), ptr(mu), mu.size, float(T[0]),
                                             float(params.eta), ptr(self._dir(params)), SPIN_DIR[params.spin],
                                             ptr(sig)))
            return HallConductivityResult(mu.copy(), sig)
        check(load().rtb_hall_direct(self._device(), ptr(kmesh), ptr(mu), mu.size, float(T[0]), float(params.eta),
                                     ptr(self._dir(params)), SPIN_DIR[params.spin], PRECISION[params.precision],
                                     ptr(sig)))
        return HallConductivityResult(mu.copy(), sig)

    def hall_conductivity_layers(self, params: Parameters, orbital_layer, n_layers=None):
        """The Direct AHC resolved by layer (no reference counterpart): returns
        (sigma [n_mu], sigma_layers [n_mu, n_layers]) with sigma_layers[m, l] =
        sum_k sum_n f(E_n, mu_m) Omega_n sum_{i in layer l} |u_n(i)|^2 / N_k / det,
        orbital o (both spin copies) in layer orbital_layer[o]; the layers sum
        to sigma."""
        lay = np.ascontiguousarray(np.asarray(orbital_layer, dtype=np.int32))
        if lay.shape != (self.norb,):
            raise TbError(-2, "orbital_layer must have norb entries")
        nl = int(lay.max()) + 1 if n_layers is None else int(n_layers)
        kmesh = np.ascontiguousarray(np.asarray(params.kmesh, dtype=np.uint64))
        mu = np.ascontiguousarray(np.atleast_1d(np.asarray(params.mu, np.float64)))
        T = float(np.atleast_1d(np.asarray(params.T, float))[0])
        sig = np.zeros(mu.size)
        sl = np.zeros((mu.size, nl))
        check(load().rtb_hall_direct_layers(self._device(), ptr(kmesh), ptr(mu), mu.size, T, float(params.eta),
                                            ptr(self._dir(params)), SPIN_DIR[params.spin],
                                            PRECISION[params.precision], nl, ptr(lay), ptr(sig), ptr(sl)))
        return sig, sl

    # ---- quantum geometry (quantum_geometry.rs) ----
    def quantum_geometry_on(self, kvec, params: Parameters) -> "QuantumGeometryMap":
        """quantum_geometry_on (quantum_geometry.rs:104-140): metric, Berry
        curvature and energies [nk, nsta] (Atom gauge, rows of params.direction)."""
        k = self._kvec(kvec)
        n = self.nsta
        g = np.zeros((k.shape[0], n))
        om = np.zeros((k.shape[0], n))
        E = np.zeros((k.shape[0], n))
        check(load().rtb_quantum_geometry(self._device(), ptr(k), k.shape[0], ptr(self._dir(params)),
                                          float(params.eta), PRECISION[params.precision], ptr(g), ptr(om), ptr(E)))
        return QuantumGeometryMap(g, om, E)

    def quantum_geometry_at(self, k, params: Parameters) -> "BandQuantumGeometry":
        """quantum_geometry_at (quantum_geometry.rs:88-102)."""
        r = self.quantum_geometry_on(np.asarray(k, float)[None, :], params)
        return BandQuantumGeometry(r.metric[0], r.berry_curvature[0], r.energies[0])

    def quantum_geometry(self, params: Parameters) -> "QuantumGeometryResult":
        """Model::quantum_geometry, Integration::Direct (quantum_geometry.rs:196-240)."""
        if params.integration == "EnergyCut":
            raise TbError(-6, "integration: quantum_geometry supports Integration::Direct or Simplex, not EnergyCut")
        if params.integration != "Direct":
            raise TbError(-14, "integration: this path implements Integration::Direct")
        kmesh = np.ascontiguousarray(np.asarray(params.kmesh, dtype=np.uint64))
        if kmesh.shape != (self.dim,):
            raise TbError(-3, "kmesh must have dim_r entries")
        mu = np.ascontiguousarray(np.atleast_1d(np.asarray(params.mu, np.float64)))
        T = np.atleast_1d(np.asarray(params.T, float))
        if T.size == 0:
            raise TbError(-6, "T: must contain at least one value")
        g = np.zeros(mu.size)
        b = np.zeros(mu.size)
        check(load().rtb_quantum_geometry_direct(self._device(), ptr(kmesh), ptr(mu), mu.size, float(T[0]),
                                                 float(params.eta), ptr(self._dir(params)),
                                                 PRECISION[params.precision], ptr(g), ptr(b)))
        return QuantumGeometryResult(mu.copy(), g, b, None)

    def extrinsic_nonlinear_hall(self, params: Parameters) -> NonlinearHallResult:
        """Model::extrinsic_nonlinear_hall, Integration::Direct
        (response/nonlinear/mod.rs:256-347): the Berry curvature dipole
        response; params.direction rows (current, field_1, field_2)."""
        if params.integration == "Simplex":
            raise TbError(-6, "integration: extrinsic_nonlinear_hall supports Integration::Direct or EnergyCut, "
                              "not Simplex")
        if params.integration != "Direct":
            raise TbError(-14, "integration: this path implements Integration::Direct")
        if params.field_symmetry not in ("Symmetrized", "Ordered"):
            raise TbError(-12, "field_symmetry must be 'Symmetrized' or 'Ordered'")
        kmesh = np.ascontiguousarray(np.asarray(params.kmesh, dtype=np.uint64))
        if kmesh.shape != (self.dim,):
            raise TbError(-3, "kmesh must have dim_r entries")
        d = np.ascontiguousarray(np.asarray(params.direction, np.float64))
        if d.shape != (3, self.dim):
            raise TbError(-2, f"direction must have shape (3, {self.dim})")
        mu = np.ascontiguousarray(np.atleast_1d(np.asarray(params.mu, np.float64)))
        T = np.atleast_1d(np.asarray(params.T, float))
        if T.size == 0:
            raise TbError(-6, "T: must contain at least one value")
        sig = np.zeros(mu.size)
        check(load().rtb_nonlinear_hall_direct(self._device(), ptr(kmesh), ptr(mu), mu.size, float(T[0]),
                                               float(params.eta), ptr(d), SPIN_DIR[params.spin],
                                               int(params.field_symmetry == "Symmetrized"),
                                               PRECISION[params.precision], ptr(sig)))
        return NonlinearHallResult(mu.copy(), sig)

    def dos(self, k_mesh, E_min, E_max, E_n, sigma, precision="fp64"):
        """Model::dos (model_physics.rs:179-226) -> (E, rho)."""
        kmesh = np.ascontiguousarray(np.asarray(k_mesh, dtype=np.uint64))
        if kmesh.shape != (self.dim,):
            raise TbError(-3, "kmesh must have dim_r entries")
        E_n = int(E_n)
        E = np.zeros(max(E_n, 0))
        rho = np.zeros(max(E_n, 0))
        check(load().rtb_dos(self._device(), ptr(kmesh), float(E_min), float(E_max), E_n, float(sigma),
                             PRECISION[precision], ptr(E), ptr(rho)))
        return E, rho


def synth_model(kind, dim, spin, norb, rmax, seed, scale=1.0, decay=1.0, n_R_target=0, layers=0) -> Model:
    """Seeded synthetic model (rtb_synth_model): kind 1/2 random 3D, 3 sparse 2D,
    4 layered slab."""
    lib = load()
    spec = _lib.SynthSpec(kind, dim, int(spin), norb, rmax, n_R_target, layers, scale, decay, seed)
    nR, n = ctypes.c_int(), ctypes.c_int()
    check(lib.rtb_synth_model(ctypes.byref(spec), ctypes.byref(nR), ctypes.byref(n), None, None, None, None))
    norb_tot = norb * layers if kind == 4 else norb
    hamR = np.zeros((nR.value, dim), np.int64)
    ham = np.zeros((nR.value, n.value, n.value), np.complex128)
    orb = np.zeros((norb_tot, dim))
    lat = np.zeros((dim, dim))
    check(lib.rtb_synth_model(ctypes.byref(spec), ctypes.byref(nR), ctypes.byref(n), ptr(hamR), ptr(ham), ptr(orb),
                              ptr(lat)))
    return Model.from_arrays(lat, orb, spin, hamR, ham)
