"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
seeded inputs, the reference's own pins, and size-independent properties.

Tolerances (FP64 plain form): Fourier sums 1e-12 relative to max|h|*N_R;
eigenvalues 1e-12 relative to the spectral radius; Berry curvature 1e-8
relative to max|Omega| (eta-amplified); AHC 1e-10 absolute at O(1) values."""
import math
import os

import numpy as np
import pytest

import rustb_amd as ra
from oracle import oracle as O
from tests import refmodels as RM

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def both(recipe, *a):
    return recipe(ra.Model.tb_model, *a), recipe(O.Model, *a)


def random_pair(norb, spin, dim, nR_half, seed, rmax=2):
    """A random model built through both builders with the same add_element calls."""
    rng = np.random.default_rng(seed)
    lat = np.eye(dim) + 0.2 * rng.standard_normal((dim, dim))
    orb = rng.random((norb, dim))
    m = ra.Model.tb_model(lat, orb, spin)
    o = O.Model(lat, orb, spin)
    n = m.nsta
    z = np.zeros(dim, int)
    for i in range(n):
        v = rng.uniform(-1, 1)
        m.add_element(v, i, i, z), o.add_element(v, i, i, z)
        for j in range(i + 1, n):
            t = complex(*rng.standard_normal(2))
            m.add_element(t, i, j, z), o.add_element(t, i, j, z)
    for _ in range(nR_half):
        R = rng.integers(-rmax, rmax + 1, dim)
        if not R.any():
            continue
        amp = math.exp(-np.linalg.norm(R))
        for i in range(n):
            for j in range(n):
                t = complex(*rng.standard_normal(2)) * amp
                m.add_element(t, i, j, R), o.add_element(t, i, j, R)
    return m, o


# ------------------------------------------------------------- reference pins
def test_gen_v_pins_gpu():
    """lib.rs:671-755 through the GPU path."""
    m = RM.gen_v_model(ra.Model.tb_model)
    assert np.array_equal(m.solve_band_onek([0.0, 0.0]), np.array([-3.0, 3.0]))
    assert np.allclose(m.solve_band_onek([1 / 3, 2 / 3]), [0.0, 0.0], atol=1e-5)
    v, _ = m.gen_v([1 / 3, 1 / 3], "atom")
    vy = np.array([[0, -0.4698463103929542 - 0.17101007166283436j],
                   [-0.4698463103929542 + 0.17101007166283436j, 0]])
    vx = np.array([[0, -0.8137976813493737 - 0.2961981327260237j],
                   [-0.8137976813493737 + 0.2961981327260237j, 0]])
    assert np.abs(v[0] - vx).max() < 1e-8 and np.abs(v[1] - vy).max() < 1e-8
    v, _ = m.gen_v([1 / 3, 1 / 3], "lattice")
    vxl = np.array([[0, -3 * math.sqrt(3) / 4 + 0.75j], [-3 * math.sqrt(3) / 4 - 0.75j, 0]])
    assert np.abs(v[0] - vxl).max() < 1e-8
    band, evec = m.solve_onek([1 / 3, 1 / 3])
    H = m.gen_ham([1 / 3, 1 / 3])
    assert np.allclose(np.diag(evec.conj() @ H @ evec.T).real, band, atol=1e-5)


def test_graphene_band_fixture_gpu():
    """examples/graphene/BAND.dat through the GPU eigensolver (5e-7 eV)."""
    data = np.loadtxt(os.path.join(GOLDEN, "graphene_soc_BAND.dat"))
    m = RM.graphene_soc_example(ra.Model.tb_model)
    kvec, _, _ = ra.k_path(m.lat, [[0, 0], [1 / 3, 2 / 3], [2 / 3, 1 / 3], [0, 0]], 2001)
    E = m.solve_band_all(kvec)
    assert np.abs(E - data[:, 1:]).max() < 6e-7


def test_haldane_chern_and_mu_grid_gpu():
    """lib.rs:926-998: 2 pi sigma = 1 +- 1e-3 at 31x31; mu-grid consistency."""
    m = RM.haldane(ra.Model.tb_model)
    p = ra.Parameters.rank2([31, 31], [1, 0], [0, 1], [0.0, -2.0, 2.0])
    s = m.hall_conductivity(p).conductivity
    assert abs(s[0] * 2 * math.pi - 1.0) < 1e-3
    g = m.hall_conductivity(ra.Parameters.rank2([31, 31], [1, 0], [0, 1], np.linspace(-2, 2, 101))).conductivity
    assert abs(g[50] - s[0]) < 1e-3 and abs(g[0] - s[1]) < 1e-3 and abs(g[100] - s[2]) < 1e-3
    o = O.hall_conductivity_direct(RM.haldane(O.Model), [31, 31], np.eye(2), [0.0, -2.0, 2.0], 0.0, 1e-3)
    assert np.abs(s - o).max() < 1e-10


def test_hall_time_reversal_gpu():
    """lib.rs:2665-2740: |sigma(t2) + sigma(-t2)| <= 1e-10, 51x51, eta=0.05."""
    mu = np.linspace(-3, 3, 101)
    p = ra.Parameters.rank2([51, 51], [1, 0], [0, 1], mu, eta=0.05)
    a = RM.haldane(ra.Model.tb_model, -0.3).hall_conductivity(p).conductivity
    b = RM.haldane(ra.Model.tb_model, 0.3).hall_conductivity(p).conductivity
    assert np.abs(a + b).max() < 1e-10
    assert np.abs(a).max() > 1e-3


def test_kubo_vs_wilson_loop_gpu():
    """lib.rs:757-823 with GPU Berry curvature and GPU eigenvectors."""
    m = RM.haldane(ra.Model.tb_model)
    k = np.array([1 / 3, 2 / 3])
    p = ra.Parameters.rank2([1, 1], [1, 0], [0, 1], [0.0])
    bb = m.berry_curvature_at(k, p)
    r1 = bb.berry_curvature[bb.energies <= 0.0].sum() * 2 * math.pi
    d1, d2, dk = np.array([1.0, 0.0]), np.array([0.0, 1.0]), 1e-4
    steps = [d1, d1 + d2, d2, -d1 + d2, -d1, -d1 - d2, -d2, d1 - d2, d1]
    _, vecs = m.solve_all(np.array([k + dk * s for s in steps]))
    prod = 1.0 + 0j
    for j in range(len(steps) - 1):
        prod *= np.vdot(vecs[j, 0], vecs[j + 1, 0])
    r2 = -np.angle(prod) / dk**2 / 4 / (2 * math.pi) * math.sqrt(3) / 2
    assert abs(r2 - r1) < 1e-4


def test_qwz_3d_plateau_gpu():
    m = RM.qwz_stacked_3d(ra.Model.tb_model)
    s = m.hall_conductivity(ra.Parameters.rank2([14, 14, 4], [1, 0, 0], [0, 1, 0], [0.0], eta=0.1)).conductivity
    o = O.hall_conductivity_direct(RM.qwz_stacked_3d(O.Model), [14, 14, 4], [[1, 0, 0], [0, 1, 0]], [0.0], 0.0, 0.1)
    assert abs(abs(s[0]) - 1 / (2 * math.pi)) < 5.3e-2
    assert abs(s[0] - o[0]) < 1e-10


# ------------------------------------------------------------- oracle parity
@pytest.mark.parametrize("norb,spin,dim", [(1, False, 1), (3, False, 2), (4, True, 3), (16, True, 3)])
@pytest.mark.parametrize("gauge", ["atom", "lattice"])
def test_fourier_and_velocity_vs_oracle(norb, spin, dim, gauge):
    m, o = random_pair(norb, spin, dim, 6, seed=norb * 10 + dim)
    rng = np.random.default_rng(1)
    ks = rng.uniform(-1.5, 1.5, (37, dim))
    dirs = rng.standard_normal((3, dim))
    V, H = m.gen_v_projected_all(ks, gauge, dirs)
    scale = np.abs(o.ham).sum(axis=0).max()
    for i, k in enumerate(ks):
        ov, oh = O.gen_v_projected(o, k, gauge, dirs)
        assert np.abs(H[i] - oh).max() <= 1e-13 * scale
        assert np.abs(V[i] - ov).max() <= 1e-12 * scale * (1 + np.abs(o.hamR).max())
    Hl = m.gen_ham_all(ks, gauge)
    assert np.array_equal(Hl, H)


@pytest.mark.parametrize("n_orb,spin", [(2, False), (5, False), (16, True), (32, True), (64, True)])
def test_eigenvalues_and_vectors_vs_zheev(n_orb, spin):
    m, o = random_pair(n_orb, spin, 3, 3, seed=n_orb)
    ks = np.random.default_rng(2).random((24, 3))
    E = m.solve_band_all(ks)
    E2, V = m.solve_all(ks)
    # eigenvalue-only solves read H(k)'s lower triangle, eigenvector solves
    # its upper one (as the reference's eigvalsh_v / eigh): the two triangles'
    # Fourier sums round differently, so the spectra agree to rounding only
    assert np.abs(E - E2).max() <= 1e-14 * max(1.0, np.abs(E).max()) * m.nsta
    H = O.gen_ham_batch(o, ks)
    for i in range(len(ks)):
        w = O.eigvalsh(H[i])
        scale = max(1.0, np.abs(w).max())
        assert np.abs(E[i] - w).max() <= 1e-13 * scale * m.nsta
        U = V[i].T  # rows are eigenvectors (solve_ham.rs:127)
        assert np.abs(H[i] @ U - U * E[i]).max() <= 1e-13 * scale * m.nsta
        assert np.abs(U.conj().T @ U - np.eye(m.nsta)).max() <= 1e-13 * m.nsta


def test_eigensolver_n128_vs_zheev():
    m, o = random_pair(64, True, 3, 2, seed=128, rmax=1)
    ks = np.random.default_rng(3).random((6, 3))
    E, V = m.solve_all(ks)
    H = O.gen_ham_batch(o, ks)
    for i in range(len(ks)):
        w = O.eigvalsh(H[i])
        scale = np.abs(w).max()
        assert np.abs(E[i] - w).max() <= 2e-14 * scale * 128
        U = V[i].T
        assert np.abs(H[i] @ U - U * E[i]).max() <= 2e-14 * scale * 128
        assert np.abs(U.conj().T @ U - np.eye(128)).max() <= 1e-12


@pytest.mark.parametrize("spin_dir", [None, "X", "Y", "Z"])
def test_berry_curvature_vs_oracle(spin_dir):
    m, o = random_pair(3, True, 2, 4, seed=31)
    ks = np.random.default_rng(4).random((16, 2))
    dirs = np.array([[1.0, 0.2], [-0.3, 1.0]])
    p = ra.Parameters(kmesh=(1, 1), direction=dirs, mu=np.array([0.0]), eta=1e-3, spin=spin_dir)
    om, E = m.berry_curvature_on(ks, p)
    ob, Eb = O.berry_curvature_batch(o, ks, dirs, 1e-3, spin_dir)
    assert np.abs(E - Eb).max() < 1e-12
    assert np.abs(om - ob).max() <= 1e-8 * max(1.0, np.abs(ob).max())
    if spin_dir is None:
        assert np.abs(om.sum(axis=1)).max() <= 1e-9 * max(1.0, np.abs(om).max())
    occ = m.occupied_berry_curvature_on(ks, ra.Parameters(kmesh=(1, 1), direction=dirs, mu=np.array([0.1]),
                                                          T=np.array([300.0]), eta=1e-3, spin=spin_dir))
    ref = (ob * O.fermi(Eb, 0.1, 300.0)).sum(axis=1)
    assert np.abs(occ - ref).max() <= 1e-8 * max(1.0, np.abs(ob).max())


def test_hall_random_model_vs_oracle():
    m, o = random_pair(4, False, 2, 5, seed=77)
    mu = np.array([-1.0, 0.0, 0.5])
    s = m.hall_conductivity(ra.Parameters.rank2([9, 11], [1, 0], [0, 1], mu, eta=0.01)).conductivity
    r = O.hall_conductivity_direct(o, [9, 11], np.eye(2), mu, 0.0, 0.01)
    assert np.abs(s - r).max() <= 1e-9 * max(1.0, np.abs(r).max())


def test_dos_graphene_config1():
    """The graphene DOS workload (101x101, [-4,4], 801 points, sigma=0.02)."""
    m, o = both(RM.graphene_config1)
    E, rho = m.dos([101, 101], -4.0, 4.0, 801, 0.02)
    Er, rr = O.dos(o, [101, 101], -4.0, 4.0, 801, 0.02)
    assert np.array_equal(E, Er)
    assert np.abs(rho - rr).max() <= 1e-12 * rr.max()


# ------------------------------------------------------------- errors
def test_validation_errors_gpu():
    m = RM.haldane(ra.Model.tb_model)
    cases = [
        (lambda: m.dos([0, 3], -1, 1, 10, 0.1), "InvalidKmeshDimensions"),
        (lambda: m.dos([3, 3], 1, 1, 10, 0.1), "InvalidEnergyRange"),
        (lambda: m.dos([3, 3], -1, 1, 0, 0.1), "InvalidDosParameter"),
        (lambda: m.dos([3, 3], -1, 1, 10, 0.0), "InvalidDosParameter"),
        (lambda: m.hall_conductivity(ra.Parameters.rank2([3, 3], [0, 0], [0, 1], [0.0])), "InvalidResponseParameter"),
        (lambda: m.hall_conductivity(ra.Parameters.rank2([3, 3], [1, 0], [0, 1], [np.nan])), "InvalidResponseParameter"),
        (lambda: m.hall_conductivity(ra.Parameters.rank2([3, 3], [1, 0], [0, 1], [0.0], eta=-1.0)),
         "InvalidResponseParameter"),
        (lambda: m.hall_conductivity(ra.Parameters.rank2([3, 3], [1, 0], [0, 1], [0.0], T=np.array([-1.0]))),
         "InvalidResponseParameter"),
        (lambda: m.hall_conductivity(ra.Parameters.rank2([3, 3], [1, 0], [0, 1], [0.0], spin="Z")), "SpinNotAllowed"),
        (lambda: m.solve_band_all(np.zeros((2, 3))), "KVectorLengthMismatch"),
    ]
    for fn, kind in cases:
        with pytest.raises(ra.TbError) as e:
            fn()
        assert e.value.kind == kind, (kind, e.value)
    # empty input is not an error
    assert m.solve_band_all(np.zeros((0, 2))).shape == (0, 2)


def test_last_stats_reports_stages():
    from rustb_amd._lib import STAGES
    m = ra.workloads.random3d_n32()
    m.solve_band_all(np.random.default_rng(0).random((64, 3)))
    st = ra.last_stats()
    i = STAGES.index
    assert st["nk"] == 64 and st["nsta"] == 32 and st["ms"][i("total")] > 0
    assert st["launches"][i("tridiag")] == 1 and st["launches"][i("trieig")] == 1
    assert st["flop"][i("tridiag")] == pytest.approx(16.0 / 3.0 * 32**3 * 64, rel=1e-12)
    # the mixed Hall path reports its matrix-core products with their flops
    p = ra.Parameters(kmesh=(4, 4, 4), direction=np.eye(3)[:2], mu=np.array([0.0]), eta=1e-3, precision="mixed")
    m.hall_conductivity(p)
    st = ra.last_stats()
    # refinement: H Z, Z^H (H Z), Z^H Z; Kubo: M Z, Z^H (M Z) for M = v_a + i v_b
    assert st["launches"][i("gemm64")] == 5 and st["launches"][i("gemm32")] == 1  # Z F', HZ F' in one kernel
    assert st["launches"][i("trieig")] == 2  # bisection + twisted factors, cluster members
    herm = 4.0 * 32 * 32 * 33 * 64
    assert st["flop"][i("gemm64")] == pytest.approx(3 * 8.0 * 32**3 * 64 + 2 * herm, rel=1e-12)
    assert all(st["ms"][i(s)] > 0 for s in ("fourier", "tridiag", "trieig", "backtransform", "gemm64", "refine"))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_quantum_geometry_vs_oracle(precision):
    """quantum_geometry_on / quantum_geometry (Direct) through the C ABI."""
    m, o = random_pair(3, True, 3, 4, seed=61)
    ks = np.random.default_rng(5).random((12, 3))
    dirs = np.array([[1.0, 0.3, 0.0], [0.0, 1.0, -0.2]])
    p = ra.Parameters(kmesh=(1, 1, 1), direction=dirs, mu=np.array([0.0]), eta=1e-3, precision=precision)
    r = m.quantum_geometry_on(ks, p)
    tol = 1e-8 if precision == "fp64" else 1e-7
    for i, k in enumerate(ks):
        g, om, E = O.quantum_geometry_at(o, k, dirs, 1e-3)
        assert np.abs(r.energies[i] - E).max() <= (1e-12 if precision == "fp64" else 1e-9) * max(1.0, np.abs(E).max())
        assert np.abs(r.metric[i] - g).max() <= tol * max(1.0, np.abs(g).max())
        assert np.abs(r.berry_curvature[i] - om).max() <= tol * max(1.0, np.abs(om).max())
    m2, o2 = random_pair(4, False, 2, 5, seed=62)
    mu = np.array([-0.5, 0.0, 0.7])
    q = m2.quantum_geometry(ra.Parameters.rank2([7, 9], [1, 0], [0, 1], mu, eta=0.02, precision=precision))
    gm, bm = O.quantum_geometry_direct(o2, [7, 9], np.eye(2), mu, 0.0, 0.02)
    rel = 1e-9 if precision == "fp64" else 1e-6
    assert np.abs(q.metric - gm).max() <= rel * max(1.0, np.abs(gm).max())
    assert np.abs(q.berry_curvature - bm).max() <= rel * max(1.0, np.abs(bm).max())
    s = m2.hall_conductivity(ra.Parameters.rank2([7, 9], [1, 0], [0, 1], mu, eta=0.02, precision="fp64"))
    assert np.abs(q.berry_curvature - s.conductivity).max() <= rel * max(1.0, np.abs(bm).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nonlinear_hall_dipole_vs_oracle(precision):
    """extrinsic_nonlinear_hall (Direct) through the C ABI: a C3-broken Haldane
    model (symmetrized and ordered), and a spinful 3D model with a spin current."""
    rel = 1e-9 if precision == "fp64" else 1e-6
    m = RM.haldane(ra.Model.tb_model, t2_imag=-0.3, t=-1.0, delta=0.7)
    m.add_hop(-0.4, 0, 1, [0, 0])
    o = RM.haldane(O.Model, t2_imag=-0.3, t=-1.0, delta=0.7)
    o.add_hop(-0.4, 0, 1, np.array([0, 0]))
    mu = np.linspace(-1.5, 1.5, 7)
    for sym in ("Symmetrized", "Ordered"):
        p = ra.Parameters.rank3([9, 11], [1.0, 0.0], [0.0, 1.0], [1.0, 0.2], mu, T=np.array([600.0]), eta=0.05,
                                precision=precision, field_symmetry=sym)
        got = m.extrinsic_nonlinear_hall(p).conductivity
        ref = O.extrinsic_nonlinear_hall_direct(o, [9, 11], p.direction, mu, 600.0, 0.05,
                                                symmetrized=sym == "Symmetrized")
        assert np.abs(got - ref).max() <= rel * max(1e-3, np.abs(ref).max()), sym
    m3, o3 = random_pair(3, True, 3, 3, seed=71)
    d3 = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.3, 1.0]])
    p3 = ra.Parameters(kmesh=(3, 4, 3), direction=d3, mu=np.array([-0.3, 0.4]), T=np.array([900.0]), eta=0.02,
                       spin="Z", precision=precision)
    got = m3.extrinsic_nonlinear_hall(p3).conductivity
    ref = O.extrinsic_nonlinear_hall_direct(o3, [3, 4, 3], d3, [-0.3, 0.4], 900.0, 0.02, spin="Z")
    assert np.abs(got - ref).max() <= rel * max(1e-3, np.abs(ref).max())
    with pytest.raises(ra.TbError) as e:
        m.extrinsic_nonlinear_hall(ra.Parameters.rank3([3, 3], [1, 0], [0, 1], [1, 0], [0.0], precision=precision))
    assert e.value.kind == "InvalidThermodynamicParameter"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_range_solves_vs_oracle(precision):
    """solve_band_range_onek / solve_range_onek through the C ABI: the same
    eigenvalues in (vl, vu] as the oracle, and eigenvectors (rows) of the
    lower-triangle matrix; the half-open edges on exact eigenvalues."""
    m, o = random_pair(5, True, 3, 3, seed=81)
    tol = 1e-12 if precision == "fp64" else 1e-9
    for k in np.random.default_rng(9).random((5, 3)):
        full = O.solve_band_onek(o, k)
        vl, vu = 0.5 * (full[2] + full[3]), 0.5 * (full[7] + full[8])
        e = m.solve_band_range_onek(k, (vl, vu), 1e-10, precision=precision)
        er = O.solve_band_range_onek(o, k, (vl, vu), 1e-10)
        assert e.shape == er.shape and np.abs(e - er).max() <= tol * max(1.0, np.abs(full).max())
        w, V = m.solve_range_onek(k, (vl, vu), 1e-10, precision=precision)
        H = O.gen_ham(o, k, "atom")
        low = np.tril(H, -1) + np.tril(H, -1).conj().T + np.diag(H.diagonal().real)
        assert w.shape == er.shape and np.abs(low @ V.T - V.T * w).max() <= tol * max(1.0, np.abs(full).max())
    # edges between eigenvalues (one exactly at an edge is decided by rounding
    # in any FP64 solver; the half-open rule itself is pinned on the oracle)
    d = ra.Model.tb_model(np.eye(2), np.zeros((4, 2)), False)
    d.set_onsite([-1.0, 0.0, 1.0, 2.0])
    e = d.solve_band_range_onek([0.1, 0.2], (-0.5, 1.5), 1e-10, precision=precision)
    assert e.shape == (2,) and np.abs(e - [0.0, 1.0]).max() <= 1e-14
    with pytest.raises(ra.TbError):
        d.solve_band_range_onek([0.1, 0.2], (1.0, 0.0), 1e-10)
