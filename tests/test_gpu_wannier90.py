"""GPU parity for models read from Wannier90 files: the loader's storage
through the device hot path (Fourier sum, eigensolve, Berry curvature, AHC)
against the oracle's reading of the same files through the CPU oracle.

Tolerances as in test_gpu_parity.py (FP64): H(k) 1e-12 relative to
max|h| N_R, eigenvalues 1e-12 relative to the spectral radius, Omega 1e-8
relative to max|Omega|; the mixed form 1e-9 eV on eigenvalues."""
import numpy as np
import pytest

import rustb_amd as ra
from oracle import oracle as O
from oracle import wannier90_ref as OW
from rustb_amd import workloads as W

pytestmark = pytest.mark.gpu


def written_model(tmp_path, norb=6, rmax=2, seed=41, wsvec=False):
    m0 = ra.synth_model(2, 3, True, norb, rmax, seed, scale=1.0, decay=1.5)
    ndegen = np.random.default_rng(seed).integers(1, 5, m0.hamR.shape[0])
    prefix = str(tmp_path / "wan")
    W.write_wannier90(m0, prefix, ndegen=ndegen, fermi=0.75, centres=True)
    if wsvec:
        rng = np.random.default_rng(seed + 1)
        lines = ["## written by test_gpu_wannier90"]
        n = m0.nsta
        for R in m0.hamR:
            for i in range(1, n + 1):
                for j in range(1, n + 1):
                    k = int(rng.integers(1, 3))
                    lines += [f"{R[0]} {R[1]} {R[2]} {i} {j}", str(k)]
                    lines += ["0 0 0"] + [f"{int(rng.integers(-1, 2)) * 4} 0 0" for _ in range(k - 1)]
        with open(prefix + "_wsvec.dat", "w") as f:
            f.write("\n".join(lines) + "\n")
    d = str(tmp_path) + "/"
    m = ra.Model.from_hr(d, "wan", 0.75)
    w = OW.from_hr(d, "wan", 0.75)
    o = O.model_from_arrays(w.lat, w.orb, w.spin, w.hamR, w.ham)
    return m, o, w


@pytest.mark.parametrize("wsvec", [False, True])
def test_hamiltonian_and_bands_of_loaded_model(wsvec, tmp_path):
    m, o, w = written_model(tmp_path, wsvec=wsvec)
    print(f"[w90] nsta={m.nsta} n_R={m.hamR.shape[0]} wsvec={wsvec}", flush=True)
    assert np.array_equal(m.hamR, w.hamR)
    assert np.array_equal(m.ham.view(np.float64), w.ham.view(np.float64))
    ks = np.random.default_rng(7).random((24, 3))
    H = m.gen_ham_all(ks, "atom")
    Hr = O.gen_ham_batch(o, ks, "atom")
    scale = np.abs(w.ham).max() * w.hamR.shape[0]
    assert np.abs(H - Hr).max() <= 1e-12 * scale
    E = m.solve_band_all(ks)
    Er = O.solve_band_all(o, ks)
    assert np.abs(E - Er).max() <= 1e-12 * max(1.0, np.abs(Er).max())
    Em = m.solve_band_all(ks, precision="mixed")
    assert np.abs(Em - Er).max() <= 1e-9


def test_berry_curvature_and_hall_of_loaded_model(tmp_path):
    m, o, _ = written_model(tmp_path, norb=5, rmax=1, seed=43)
    print(f"[w90] berry nsta={m.nsta}", flush=True)
    ks = np.random.default_rng(8).random((12, 3))
    dirs = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    p = ra.Parameters(kmesh=(1, 1, 1), direction=dirs, mu=np.array([0.0]), eta=1e-3)
    om, E = m.berry_curvature_on(ks, p)
    ob, Eb = O.berry_curvature_batch(o, ks, dirs, 1e-3)
    assert np.abs(E - Eb).max() < 1e-12 * max(1.0, np.abs(Eb).max())
    assert np.abs(om - ob).max() <= 1e-8 * max(1.0, np.abs(ob).max())
    mu = np.array([-0.5, 0.0, 0.5])
    s = m.hall_conductivity(ra.Parameters.rank2([4, 4, 3], [1, 0, 0], [0, 1, 0], mu, eta=0.01)).conductivity
    r = O.hall_conductivity_direct(o, [4, 4, 3], dirs, mu, 0.0, 0.01)
    assert np.abs(s - r).max() <= 1e-9 * max(1.0, np.abs(r).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_triangle_conventions_on_non_hermitian_storage(precision):
    """ham[-R] != ham[R]^H at the 1e-5 level (as in files printed to a few
    decimals): eigenvalue-only solves follow the lower triangle of H(k)
    (the reference's eigvalsh_v), eigenvector and Berry paths its upper
    triangle (the reference's eigh)."""
    m0 = ra.synth_model(2, 3, False, 10, 1, 55, scale=1.0, decay=1.5)
    rng = np.random.default_rng(3)
    ham = m0.ham + 1e-5 * (rng.standard_normal(m0.ham.shape) + 1j * rng.standard_normal(m0.ham.shape))
    m = ra.Model.from_arrays(m0.lat, m0.orb, False, m0.hamR, ham)
    o = O.model_from_arrays(m0.lat, m0.orb, False, m0.hamR, ham)
    print(f"[w90] triangle conventions {precision}", flush=True)
    ks = np.random.default_rng(4).random((16, 3))
    scale = max(1.0, np.abs(O.solve_band_all(o, ks)).max())
    etol = 1e-12 * scale if precision == "fp64" else 1e-9  # eV; the mixed-form bar is 1e-9 eV
    assert np.abs(m.solve_band_all(ks, precision=precision) - O.solve_band_all(o, ks)).max() <= etol
    E, V = m.solve_all(ks, precision=precision)
    for i, k in enumerate(ks):
        Er, _ = O.eigh(O.gen_ham(o, k, "atom"))
        assert np.abs(E[i] - Er).max() <= etol
        U = V[i].T  # eigenvectors as rows
        up = np.triu(O.gen_ham(o, k, "atom"), 1)
        M = up + up.conj().T + np.diag(O.gen_ham(o, k, "atom").diagonal().real)
        assert np.abs(M @ U - U * E[i]).max() <= (1e-12 if precision == "fp64" else 1e-9) * scale
    dirs = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    p = ra.Parameters(kmesh=(1, 1, 1), direction=dirs, mu=np.array([0.0]), eta=1e-2, precision=precision)
    om, Eb = m.berry_curvature_on(ks, p)
    ob, Ebr = O.berry_curvature_batch(o, ks, dirs, 1e-2)
    assert np.abs(Eb - Ebr).max() <= etol
    assert np.abs(om - ob).max() <= (1e-8 if precision == "fp64" else 1e-6) * max(1.0, np.abs(ob).max())


def rmatrix_pair(seed=91, norb=4, spin=False):
    """A model with position matrices, Hermitian across +-R, on both sides."""
    m0 = ra.synth_model(2, 3, spin, norb, 1, seed, scale=1.0, decay=1.5)
    rng = np.random.default_rng(seed)
    n, nR = m0.nsta, m0.hamR.shape[0]
    rm = 0.3 * (rng.standard_normal((nR, 3, n, n)) + 1j * rng.standard_normal((nR, 3, n, n)))
    idx = {tuple(R): i for i, R in enumerate(m0.hamR)}
    for i, R in enumerate(m0.hamR):
        rm[idx[tuple(-R)]] = np.conj(np.swapaxes(rm[i], 1, 2))
    z = idx[(0, 0, 0)]
    rm[z] = 0.5 * (rm[z] + np.conj(np.swapaxes(rm[z], 1, 2)))
    m = ra.Model.from_arrays(m0.lat, m0.orb, spin, m0.hamR, m0.ham, rm)
    o = O.model_from_arrays(m0.lat, m0.orb, spin, m0.hamR, m0.ham, rm)
    return m, o


@pytest.mark.parametrize("gauge", ["atom", "lattice"])
def test_rmatrix_velocity_vs_oracle(gauge):
    """gen_v_projected with the HasRMatrix term i[H, r_p] (velocity.rs:377-402,
    627-668) through the C ABI."""
    m, o = rmatrix_pair()
    print(f"[w90] r-matrix velocity {gauge}", flush=True)
    ks = np.random.default_rng(12).uniform(-1, 1, (10, 3))
    dirs = np.array([[1.0, 0.0, 0.0], [0.2, 1.0, -0.3], [0.0, 0.0, 1.0]])
    V, H = m.gen_v_projected_all(ks, gauge, dirs)
    scale = (np.abs(o.ham).sum(axis=0).max() + 1.0) * (1 + np.abs(o.rmatrix).sum(axis=0).max())
    for i, k in enumerate(ks):
        ov, oh = O.gen_v_projected(o, k, gauge, dirs)
        assert np.abs(H[i] - oh).max() <= 1e-13 * scale
        assert np.abs(V[i] - ov).max() <= 1e-12 * scale


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rmatrix_berry_dipole_geometry_vs_oracle(precision):
    m, o = rmatrix_pair(seed=93, norb=3, spin=True)
    print(f"[w90] r-matrix responses {precision}", flush=True)
    ks = np.random.default_rng(13).random((8, 3))
    dirs = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    tol = 1e-8 if precision == "fp64" else 1e-6
    p = ra.Parameters(kmesh=(1, 1, 1), direction=dirs, mu=np.array([0.0]), eta=1e-2, precision=precision)
    om, E = m.berry_curvature_on(ks, p)
    ob, Eb = O.berry_curvature_batch(o, ks, dirs, 1e-2)
    assert np.abs(om - ob).max() <= tol * max(1.0, np.abs(ob).max())
    g = m.quantum_geometry_on(ks, p)
    for i, k in enumerate(ks):
        gm, gb, _ = O.quantum_geometry_at(o, k, dirs, 1e-2)
        assert np.abs(g.metric[i] - gm).max() <= tol * max(1.0, np.abs(gm).max())
    mu = np.array([-0.4, 0.3])
    s = m.hall_conductivity(ra.Parameters.rank2([3, 3, 2], [1, 0, 0], [0, 1, 0], mu, eta=0.02,
                                                precision=precision)).conductivity
    r = O.hall_conductivity_direct(o, [3, 3, 2], dirs, mu, 0.0, 0.02)
    assert np.abs(s - r).max() <= (1e-9 if precision == "fp64" else 1e-6) * max(1.0, np.abs(r).max())
    d3 = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.4, 1.0]])
    pn = ra.Parameters(kmesh=(3, 3, 2), direction=d3, mu=mu, T=np.array([900.0]), eta=0.02, precision=precision)
    got = m.extrinsic_nonlinear_hall(pn).conductivity
    ref = O.extrinsic_nonlinear_hall_direct(o, [3, 3, 2], d3, mu, 900.0, 0.02)
    assert np.abs(got - ref).max() <= (1e-9 if precision == "fp64" else 1e-6) * max(1e-3, np.abs(ref).max())


def test_rmatrix_from_files_through_device(tmp_path):
    """from_hr_with_rmatrix: files -> loader -> device velocities, against the
    oracle's reading of the same files."""
    from tests.test_wannier90 import rmatrix_dataset
    p, *_ = rmatrix_dataset(tmp_path, seed=5, n=4, wsvec=True)
    m = ra.Model.from_hr_with_rmatrix(p, "seedname", 0.0)
    w = OW.from_hr(p, "seedname", 0.0, rmatrix=True)
    o = O.model_from_arrays(w.lat, w.orb, w.spin, w.hamR, w.ham, w.rmatrix)
    ks = np.random.default_rng(14).random((6, 3))
    dirs = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    V, _ = m.gen_v_projected_all(ks, "atom", dirs)
    scale = (np.abs(w.ham).sum(axis=0).max() + 1.0) * (1 + np.abs(w.rmatrix).sum(axis=0).max())
    for i, k in enumerate(ks):
        ov, _ = O.gen_v_projected(o, k, "atom", dirs)
        assert np.abs(V[i] - ov).max() <= 1e-12 * scale
