"""Pins the CPU oracle against the reference's own numeric tests, its committed
band fixture, closed forms and invariants (CPU only)."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests import refmodels as RM

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def make(lat, orb, spin):
    return O.Model(lat, orb, spin)


# ---------------------------------------------------------------- test_gen_v
def test_gen_v_pins():
    """lib.rs:671-755 — exact Gamma eigenvalues, Dirac point, tabulated v."""
    m = RM.gen_v_model(make)
    assert np.array_equal(O.solve_band_onek(m, [0.0, 0.0]), np.array([-3.0, 3.0]))
    assert np.allclose(O.solve_band_onek(m, [1 / 3, 2 / 3]), [0.0, 0.0], atol=1e-5)
    v, _ = O.gen_v(m, [1 / 3, 1 / 3], "atom")
    vy = np.array([[0, -0.4698463103929542 - 0.17101007166283436j],
                   [-0.4698463103929542 + 0.17101007166283436j, 0]])
    vx = np.array([[0, -0.8137976813493737 - 0.2961981327260237j],
                   [-0.8137976813493737 + 0.2961981327260237j, 0]])
    assert np.abs(v[0] - vx).max() < 1e-8
    assert np.abs(v[1] - vy).max() < 1e-8
    v, _ = O.gen_v(m, [1 / 3, 1 / 3], "lattice")
    t = 1.0
    vx_l = np.array([[0, -3 * math.sqrt(3) / 4 * t + 0.75j * t],
                     [-3 * math.sqrt(3) / 4 * t - 0.75j * t, 0]])
    assert np.abs(v[0] - vx_l).max() < 1e-8
    # lib.rs:743-755: conj(evec) H evec^T is diagonal with the bands
    k = [1 / 3, 1 / 3]
    band, evec = O.solve_onek(m, k)
    H = O.gen_ham(m, k)
    new_band = np.diag(evec.conj() @ H @ evec.T).real
    assert np.allclose(new_band, band, atol=1e-5)


def test_evec_transform_sanity():
    """lib.rs:1231-1277 — E_vec^T H E_vec^* = diag(E) with E_vec = conj(U)."""
    m = RM.haldane(make)
    H = O.gen_ham(m, [0.3, 0.4])
    E, U = O.eigh(H)
    evec = U.conj()  # what ndarray-linalg's eigh hands back
    D = evec.T @ H @ evec.conj()
    assert np.abs(np.diag(D).real - E).max() < 1e-10
    assert np.abs(D - np.diag(np.diag(D))).max() < 1e-10


# ------------------------------------------------------ conductivity_test
def _wilson_phase(m, klist, band):
    """Single-band Wilson loop of geometry.rs:156-209 (1x1 SVD = phase)."""
    _, vecs = O.solve_all(m, klist)
    prod = 1.0 + 0j
    for j in range(len(klist) - 1):
        prod *= np.vdot(vecs[j, band], vecs[j + 1, band])
    return -np.angle(prod)


def test_kubo_vs_wilson_plaquette():
    """lib.rs:757-823: 2 pi sum_occ Omega vs a small Wilson loop, 1e-4."""
    m = RM.haldane(make)
    k = np.array([1 / 3, 2 / 3])
    d1, d2 = np.array([1.0, 0.0]), np.array([0.0, 1.0])
    om, E = O.berry_curvature_at(m, k, [d1, d2], 1e-3)
    result1 = om[E <= 0.0].sum() * 2 * math.pi
    dk = 1e-4
    steps = [d1, d1 + d2, d2, -d1 + d2, -d1, -d1 - d2, -d2, d1 - d2, d1]
    klist = np.array([k + dk * s for s in steps])
    result2 = _wilson_phase(m, klist, 0) / dk**2 / 4 / (2 * math.pi) * math.sqrt(3) / 2
    assert abs(result2 - result1) < 1e-4


def test_hall_single_vs_multi_mu():
    """lib.rs:825-845 (a 40x40 mesh keeps the CPU suite short)."""
    m = RM.haldane(make)
    a1 = O.hall_conductivity_direct(m, [40, 40], [[0, 1], [1, 0]], [-1.0], 0.0, 1e-3)
    a2 = O.hall_conductivity_direct(m, [40, 40], [[0, 1], [1, 0]], [-1.0, 0.0], 0.0, 1e-3)
    assert abs(a1[0] - a2[0]) < 1e-5


def test_haldane_chern_number():
    """lib.rs:926-998: 2 pi sigma(mu=0) = 1 +- 1e-3 at 31x31, eta=1e-3."""
    m = RM.haldane(make)
    s = O.hall_conductivity_direct(m, [31, 31], [[1, 0], [0, 1]], [0.0, -2.0, 2.0], 0.0, 1e-3)
    assert abs(s[0] * 2 * math.pi - 1.0) < 1e-3
    grid = O.hall_conductivity_direct(m, [31, 31], [[1, 0], [0, 1]], np.linspace(-2, 2, 101), 0.0, 1e-3)
    assert abs(grid[50] - s[0]) < 1e-3
    assert abs(grid[0] - s[1]) < 1e-3 and abs(grid[100] - s[2]) < 1e-3


def test_hall_time_reversal():
    """lib.rs:2665-2740: sigma(t2) + sigma(-t2) <= 1e-10 (Direct), 51x51."""
    mu = np.linspace(-3, 3, 101)
    a = O.hall_conductivity_direct(RM.haldane(make, -0.3), [51, 51], np.eye(2), mu, 0.0, 0.05)
    b = O.hall_conductivity_direct(RM.haldane(make, 0.3), [51, 51], np.eye(2), mu, 0.0, 0.05)
    assert np.abs(a + b).max() < 1e-10
    assert np.abs(a).max() > 1e-3


def test_qwz_3d_plateau():
    """lib.rs:3219-3270: stacked QWZ, eta=0.1; the EC plateau is 1/(2 pi)
    within 3e-3 and direct-vs-EC within 5e-2, so direct is within 5.3e-2."""
    m = RM.qwz_stacked_3d(make)
    for nk in (10, 14):
        s = O.hall_conductivity_direct(m, [nk, nk, 4], [[1, 0, 0], [0, 1, 0]], [0.0], 0.0, 0.1)
        assert abs(abs(s[0]) - 1 / (2 * math.pi)) < 5.3e-2


# --------------------------------------------------------- band fixture
def test_graphene_band_fixture():
    """examples/graphene/BAND.dat (2001 k, 4 bands, 6 decimals) from the
    spinful SOC graphene of examples/graphene/main.rs:11-47 along G-K-K'-G."""
    data = np.loadtxt(os.path.join(GOLDEN, "graphene_soc_BAND.dat"))
    m = RM.graphene_soc_example(make)
    path = [[0, 0], [1 / 3, 2 / 3], [2 / 3, 1 / 3], [0, 0]]
    kvec, kdist, _ = O.k_path(m.lat, path, 2001)
    E = np.stack([O.eigvalsh(H) for H in O.gen_ham_batch(m, kvec)])
    assert np.abs(kdist - data[:, 0]).max() < 6e-7
    assert np.abs(E - data[:, 1:]).max() < 6e-7


# ------------------------------------------------- closed forms, invariants
def test_graphene_closed_form_and_gauge():
    m = RM.graphene_config1(make)
    rng = np.random.default_rng(0)
    for k in rng.random((20, 2)):
        f = abs(1 + np.exp(-2j * np.pi * k[0]) + np.exp(-2j * np.pi * k[1]))
        e = O.solve_band_onek(m, k)
        assert np.allclose(e, [-2.85 * f, 2.85 * f][::-1] if -2.85 * f > 2.85 * f else [-2.85 * f, 2.85 * f], atol=1e-12)
        el = O.eigvalsh(O.gen_ham(m, k, "lattice"))
        assert np.allclose(e, el, atol=1e-12)
        # lattice gauge periodic under k -> k + G
        assert np.allclose(O.gen_ham(m, k, "lattice"), O.gen_ham(m, k + [1, -2], "lattice"), atol=1e-12)


def _random_model(n_orb, spin, dim, nR_half, seed):
    rng = np.random.default_rng(seed)
    lat = np.eye(dim) + 0.1 * rng.standard_normal((dim, dim))
    m = O.Model(lat, rng.random((n_orb, dim)), spin)
    nsta = m.nsta
    m.set_onsite(rng.uniform(-1, 1, n_orb))
    for i in range(nsta):
        for j in range(i + 1, nsta):
            m.add_element(complex(*rng.standard_normal(2)), i, j, np.zeros(dim, int))
    for _ in range(nR_half):
        R = rng.integers(-2, 3, dim)
        if not np.any(R):
            continue
        for i in range(nsta):
            for j in range(nsta):
                m.add_element(complex(*rng.standard_normal(2)) * 0.3, i, j, R)
    return m


def test_hermitian_storage_and_brute_force():
    m = _random_model(3, True, 3, 4, 1)
    for iR, R in enumerate(m.hamR):
        inv = m.find_R(-R)
        assert np.allclose(m.ham[inv], m.ham[iR].conj().T)
    k = np.array([0.1, 0.27, -0.4])
    H = O.gen_ham(m, k, "atom")
    tau = m.orb_sta()
    Hb = np.zeros_like(H)
    for a in range(m.nsta):
        for b in range(m.nsta):
            for iR, R in enumerate(m.hamR):
                Hb[a, b] += m.ham[iR, a, b] * np.exp(2j * np.pi * k @ (R - tau[a] + tau[b]))
    assert np.abs(H - Hb).max() < 1e-12
    assert np.abs(H - H.conj().T).max() < 1e-12
    assert np.allclose(O.gen_ham_batch(m, k[None])[0], H, atol=1e-12)


def test_velocity_is_finite_difference():
    """Lattice gauge: v_p = (1/2pi) sum_d (lat w_p)_d dH/dk_d."""
    m = _random_model(2, False, 2, 5, 2)
    k = np.array([0.13, 0.71])
    w = np.array([[0.6, -0.8], [1.0, 2.0]])
    v, _ = O.gen_v_projected(m, k, "lattice", w)
    h = 1e-6
    for p in range(2):
        c = m.lat @ w[p]
        fd = sum(c[d] * (O.gen_ham(m, k + h * np.eye(2)[d], "lattice") -
                         O.gen_ham(m, k - h * np.eye(2)[d], "lattice")) / (2 * h)
                 for d in range(2)) / (2 * np.pi)
        assert np.abs(v[p] - fd).max() < 1e-6
    # linearity in the direction rows
    v2, _ = O.gen_v_projected(m, k, "atom", [w[0] + 2 * w[1], w[1]])
    va, _ = O.gen_v_projected(m, k, "atom", w)
    assert np.abs(v2[0] - (va[0] + 2 * va[1])).max() < 1e-12
    vb, Hb = O.gen_v_projected_batch(m, k[None], w)
    assert np.abs(vb[0] - va).max() < 1e-12


def test_berry_sum_rule_and_batch():
    m = _random_model(3, False, 2, 4, 3)
    ks = np.random.default_rng(4).random((5, 2))
    for k in ks:
        om, _ = O.berry_curvature_at(m, k, np.eye(2), 1e-3)
        assert abs(om.sum()) < 1e-9 * max(1.0, np.abs(om).max())
    ob, Eb = O.berry_curvature_batch(m, ks, np.eye(2), 1e-3)
    oa = np.stack([O.berry_curvature_at(m, k, np.eye(2), 1e-3)[0] for k in ks])
    assert np.abs(ob - oa).max() < 1e-9 * max(1.0, np.abs(oa).max())


def test_dos_normalisation_and_single_level():
    m = RM.graphene_config1(make)
    E, rho = O.dos(m, [21, 21], -10.0, 10.0, 2001, 0.1)
    assert abs(np.trapezoid(rho, E) - 2.0) < 1e-6
    ev = np.array([0.3])
    Eg = np.linspace(-1, 1, 11)
    g = O.dos_from_eigenvalues(ev, Eg, 0.2)
    ref = np.exp(-((Eg - 0.3) / 0.2) ** 2 / 2) / (0.2 * math.sqrt(2 * math.pi))
    assert np.allclose(g, ref, rtol=1e-14, atol=0)


def test_fermi_cases():
    """thermodynamics.rs:152-172."""
    assert O.fermi(-1.0, 0.0, 0.0) == 1.0
    assert O.fermi(0.0, 0.0, 0.0) == 0.5
    assert O.fermi(1.0, 0.0, 0.0) == 0.0
    a = O.fermi(0.02, 0.0, 300.0)
    assert abs(a - 1.0 / (1.0 + math.exp(0.02 / (300 * O.BOLTZMANN_CONSTANT_EV_PER_K)))) < 1e-14
    assert O.fermi(10.0, 0.0, 1.0) == 0.0 and O.fermi(-10.0, 0.0, 1.0) == 1.0


def test_kmesh_order_and_errors():
    k = O.gen_kmesh([2, 3])
    assert k.shape == (6, 2)
    assert np.allclose(k[:3], [[0, 0], [0, 1 / 3], [0, 2 / 3]])
    with pytest.raises(O.OracleError):
        O.gen_kmesh([0, 3])
    with pytest.raises(O.OracleError):
        O.dos(RM.graphene_config1(make), [3, 3], 1.0, 1.0, 10, 0.1)
    with pytest.raises(O.OracleError):
        O.berry_curvature_at(RM.haldane(make), [0, 0], [[0, 0], [0, 1]], 1e-3)
    with pytest.raises(O.OracleError):
        RM.haldane(make).set_hop(1j, 0, 0, [0, 0])


def test_eigen_triangle_conventions():
    """For an H that is Hermitian only to rounding (a Wannier90 file printed to
    6 decimals), the reference's eigvalsh_v(UPLO::Upper) on the row-major
    buffer (solve_ham.rs:61) diagonalises the Hermitian matrix of H's lower
    triangle, and eigh(UPLO::Lower) (solve_ham.rs:122, response/traits.rs:168)
    that of its upper triangle: LAPACK reads the buffer column-major (H^T),
    which is also why eigh hands back conj(U) (lib.rs:1231-1277)."""
    rng = np.random.default_rng(12)
    A = rng.standard_normal((9, 9)) + 1j * rng.standard_normal((9, 9))
    H = A + A.conj().T + 1e-4 * (rng.standard_normal((9, 9)) + 1j * rng.standard_normal((9, 9)))
    low = np.tril(H, -1) + np.tril(H, -1).conj().T + np.diag(H.diagonal().real)
    up = np.triu(H, 1) + np.triu(H, 1).conj().T + np.diag(H.diagonal().real)
    assert np.abs(O.eigvalsh(H) - np.linalg.eigvalsh(low)).max() < 1e-12
    E, U = O.eigh(H)
    assert np.abs(E - np.linalg.eigvalsh(up)).max() < 1e-12
    assert np.abs(up @ U - U * E).max() < 1e-12
    assert np.abs(np.linalg.eigvalsh(low) - np.linalg.eigvalsh(up)).max() > 1e-6
    # column-major view of the row-major buffer: 'U' of H^T is H's lower triangle
    assert np.abs(np.linalg.eigvalsh(H.T, UPLO="U") - np.linalg.eigvalsh(low)).max() < 1e-12
    assert np.abs(np.linalg.eigvalsh(H.T, UPLO="L") - np.linalg.eigvalsh(up)).max() < 1e-12


def test_quantum_geometry_is_fubini_study():
    """quantum_geometry_at (quantum_geometry.rs:149-190): for isolated bands
    and eta -> 0 the metric Re G_n^{ab} is the Fubini-Study metric
    (1/2) Tr(d_a P_n d_b P_n) of the band projector, from central finite
    differences in Cartesian k (unit lattice, orbitals at the origin so the
    gauges coincide); the Berry part is berry_curvature_at's."""
    rng = np.random.default_rng(21)
    o = O.Model(np.eye(2), np.zeros((3, 2)), False)
    z = np.zeros(2, int)
    for i in range(3):
        o.add_element(rng.uniform(-2, 2), i, i, z)
    for R in ([1, 0], [0, 1], [1, 1], [1, -1]):
        for i in range(3):
            for j in range(3):
                o.add_element(0.4 * complex(*rng.standard_normal(2)), i, j, np.array(R))
    k = np.array([0.17, 0.31])
    h = 1e-5

    def projectors(kc):
        _, U = O.eigh(O.gen_ham(o, kc / (2 * np.pi), "atom"))
        return [np.outer(U[:, n], U[:, n].conj()) for n in range(3)]

    kc = 2 * np.pi * k
    dP = []
    for a in range(2):
        e = np.zeros(2)
        e[a] = h
        p1, p0 = projectors(kc + e), projectors(kc - e)
        dP.append([(x - y) / (2 * h) for x, y in zip(p1, p0)])
    for a, b in ((0, 1), (0, 0), (1, 1)):
        dirs = np.eye(2)[[a, b]]
        g, om, E = O.quantum_geometry_at(o, k, dirs, 1e-8)
        assert np.min(np.abs(np.diff(E))) > 0.1
        fs = np.array([0.5 * np.trace(dP[a][n] @ dP[b][n]).real for n in range(3)])
        assert np.abs(g - fs).max() <= 1e-6 * max(1.0, np.abs(fs).max())
        ob, _ = O.berry_curvature_batch(o, k[None, :], dirs, 1e-8)
        assert np.array_equal(om, ob[0]) or np.abs(om - ob[0]).max() <= 1e-12 * max(1.0, np.abs(ob).max())


def test_quantum_geometry_two_band_symmetry_and_direct_sum():
    """Two bands: g_0 = g_1 and Omega_0 = -Omega_1; the Direct integral's Berry
    part is hall_conductivity_direct (same formula, same normalisation)."""
    o = O.Model(np.eye(2), np.zeros((2, 2)), False)
    z = np.zeros(2, int)
    o.add_element(-1.0, 0, 0, z)
    o.add_element(1.0, 1, 1, z)
    o.add_element(-0.5j, 0, 1, np.array([1, 0]))
    o.add_element(0.5j, 0, 1, np.array([-1, 0]))
    o.add_element(-0.5, 0, 1, np.array([0, 1]))
    o.add_element(0.5, 0, 1, np.array([0, -1]))
    for d in ([1, 0], [0, 1]):
        o.add_element(0.5, 0, 0, np.array(d))
        o.add_element(-0.5, 1, 1, np.array(d))
    g, om, _ = O.quantum_geometry_at(o, np.array([0.1, 0.23]), np.eye(2), 1e-6)
    assert abs(g[0] - g[1]) <= 1e-12 * max(1.0, abs(g[0]))
    assert abs(om[0] + om[1]) <= 1e-12 * max(1.0, abs(om[0]))
    gm, bm = O.quantum_geometry_direct(o, [9, 9], np.eye(2), [0.0, 0.5], 0.0, 0.05)
    sig = O.hall_conductivity_direct(o, [9, 9], np.eye(2), [0.0, 0.5], 0.0, 0.05)
    assert np.abs(bm - sig).max() <= 1e-12 * max(1.0, np.abs(sig).max())
    gxx, _ = O.quantum_geometry_direct(o, [9, 9], np.eye(2)[[0, 0]], [0.0], 0.0, 0.05)
    assert gxx[0] > 0  # the metric of the occupied band is positive semi-definite


def _haldane_c3_broken(t2):
    """The Haldane model of build_haldane_2d (lib.rs:2356-2375) with one
    nearest-neighbour bond strengthened: C3 broken, so the dipole is nonzero."""
    m = RM.haldane(O.Model, t2_imag=t2, t=-1.0, delta=0.7)
    m.add_hop(-0.4, 0, 1, np.array([0, 0]))
    return m


def test_dipole_kernel_is_band_velocity_times_curvature():
    """berry_curvature_dipole_n_onek (mod.rs:89-163): kernel_n = d_c E_n *
    Omega_n^{ab}; d_c E_n equals the central finite difference of the band
    energy along Cartesian c (Hellmann-Feynman), Omega is berry_curvature_at."""
    m = _haldane_c3_broken(-0.3)
    k = np.array([0.13, 0.41])
    a, b, c = np.array([1.0, 0.0]), np.array([0.0, 1.0]), np.array([0.6, 0.8])
    kern, e = O.berry_curvature_dipole_onek(m, k, a, b, c, 1e-6)
    om, _ = O.berry_curvature_batch(m, k[None, :], np.array([a, b]), 1e-6)
    h = 1e-6
    # a Cartesian step dk along c moves the fractional k by lat @ dk / (2 pi)
    dkf = m.lat @ (h * c) / (2 * np.pi)
    ep = O.eigvalsh(O.gen_ham(m, k + dkf, "atom"))
    em = O.eigvalsh(O.gen_ham(m, k - dkf, "atom"))
    dE = (ep - em) / (2 * h)
    assert np.abs(kern - om[0] * dE).max() <= 1e-6 * max(1.0, np.abs(kern).max())
    assert np.abs(kern).max() > 1e-3


def test_dipole_time_reversal_even_and_symmetrization():
    """The dipole is even under time reversal (Haldane t2 -> -t2), as the
    reference's dipole_energy_cut_tr (lib.rs:2745-2790) asserts to 1e-10; the
    Direct sum keeps it on the k -> -k symmetric mesh.  Ordered with equal
    field rows equals Symmetrized."""
    mu = np.linspace(-1.5, 1.5, 7)
    d = np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 0.0]])
    D = O.extrinsic_nonlinear_hall_direct(_haldane_c3_broken(-0.3), [15, 15], d, mu, 600.0, 0.05)
    Dtr = O.extrinsic_nonlinear_hall_direct(_haldane_c3_broken(0.3), [15, 15], d, mu, 600.0, 0.05)
    assert np.abs(D).max() > 1e-4
    assert np.abs(D - Dtr).max() <= 1e-10 * max(1.0, np.abs(D).max())
    same = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 1.0]])
    s1 = O.extrinsic_nonlinear_hall_direct(_haldane_c3_broken(-0.3), [9, 9], same, mu, 600.0, 0.05, symmetrized=True)
    s2 = O.extrinsic_nonlinear_hall_direct(_haldane_c3_broken(-0.3), [9, 9], same, mu, 600.0, 0.05, symmetrized=False)
    assert np.array_equal(s1, s2)


def test_rmatrix_velocity_term():
    """The HasRMatrix term i[H, r_p] (velocity.rs:645-668): with position
    matrices diag(tau_cart) at R = 0 only, the Lattice-gauge velocity gains
    exactly i(tau_n - tau_m) H_mn, i.e. it becomes the Atom-gauge velocity of
    the plain model with the gauge factor removed (the Atom gauge zeroes the
    diagonal of r_p, so there the term vanishes); and the term keeps v
    Hermitian for Hermitian position matrices."""
    rng = np.random.default_rng(4)
    lat = np.array([[1.0, 0.2], [-0.1, 1.3]])
    orb = rng.random((3, 2))
    o = O.Model(lat, orb, False)
    z = np.zeros(2, int)
    for i in range(3):
        o.add_element(rng.uniform(-1, 1), i, i, z)
        for j in range(i + 1, 3):
            o.add_element(complex(*rng.standard_normal(2)), i, j, z)
    for R in ([1, 0], [0, 1], [1, -1]):
        for i in range(3):
            for j in range(3):
                o.add_element(0.3 * complex(*rng.standard_normal(2)), i, j, np.array(R))
    tau = orb @ lat
    rm = np.zeros((o.ham.shape[0], 2, 3, 3), complex)
    for d in range(2):
        rm[0, d] = np.diag(tau[:, d])
    orr = O.model_from_arrays(lat, orb, False, o.hamR, o.ham, rm)
    k = np.array([0.21, 0.37])
    dirs = np.array([[1.0, 0.0], [0.3, 0.7]])
    v_lat_r, _ = O.gen_v_projected(orr, k, "lattice", dirs)
    v_atom, _ = O.gen_v_projected(o, k, "atom", dirs)
    phi = np.exp(2j * np.pi * (orb @ k))
    gm = np.conj(phi)[:, None] * phi[None, :]
    assert np.abs(v_lat_r - v_atom / gm).max() <= 1e-13
    v_atom_r, _ = O.gen_v_projected(orr, k, "atom", dirs)
    assert np.abs(v_atom_r - v_atom).max() <= 1e-14
    rm2 = rng.standard_normal(rm.shape) + 1j * rng.standard_normal(rm.shape)
    idx = {tuple(R): i for i, R in enumerate(o.hamR)}
    for i, R in enumerate(o.hamR):
        rm2[idx[tuple(-R)]] = np.conj(np.swapaxes(rm2[i], 1, 2))
    rm2[0] = 0.5 * (rm2[0] + np.conj(np.swapaxes(rm2[0], 1, 2)))  # R = 0 is its own partner
    v2, _ = O.gen_v_projected(O.model_from_arrays(lat, orb, False, o.hamR, o.ham, rm2), k, "atom", dirs)
    assert np.abs(v2 - np.conj(np.swapaxes(v2, 1, 2))).max() <= 1e-12


def test_range_solves_half_open_interval():
    """solve_band_range_onek / solve_range_onek (zheevr RANGE = 'V'): the
    eigenvalues in (vl, vu] -- an eigenvalue at vl is excluded, one at vu
    included -- with the eigenvectors of the lower-triangle matrix as rows."""
    o = O.Model(np.eye(2), np.zeros((4, 2)), False)
    o.set_onsite([-1.0, 0.0, 1.0, 2.0])
    e = O.solve_band_range_onek(o, np.array([0.1, 0.2]), (0.0, 1.0), 1e-10)
    assert np.array_equal(e, [1.0])
    w, V = O.solve_range_onek(o, np.array([0.1, 0.2]), (-1.0, 2.0), 1e-10)
    assert np.array_equal(w, [0.0, 1.0, 2.0]) and V.shape == (3, 4)
    H = O.gen_ham(o, np.array([0.1, 0.2]), "atom")
    assert np.abs(H @ V.T - V.T * w).max() < 1e-14


# ------------------------------------------------------------ spin currents
def test_spin_matrix_algebra_and_layout():
    """build_spin_matrix (response/helpers.rs:11-39) is S = sigma/2 (x) I_norb
    with the spin index outer ([up block, down block], model_build.rs:71-98):
    Hermitian, S_a^2 = 1/4, [S_x, S_y] = i S_z and cyclic, S_z = +1/2 on the
    up block; and an onsite term J decorated with Pauli a by the builder's
    table (model_build.rs:118-145, pinned by the graphene band fixture for Z)
    is 2 J S_a."""
    for norb in (1, 3):
        S = {a: O.spin_matrix(norb, a) for a in "XYZ"}
        eye = np.eye(2 * norb)
        for a in "XYZ":
            assert np.array_equal(S[a], S[a].conj().T)
            assert np.array_equal(S[a] @ S[a], eye / 4)
            assert np.trace(S[a]) == 0
        for a, b, c in (("X", "Y", "Z"), ("Y", "Z", "X"), ("Z", "X", "Y")):
            assert np.array_equal(S[a] @ S[b] - S[b] @ S[a], 1j * S[c])
        assert np.array_equal(np.diag(S["Z"]).real, [0.5] * norb + [-0.5] * norb)
        assert np.array_equal(O.spin_matrix(norb, None), eye)
        orb = np.random.default_rng(norb).random((norb, 2))
        for a in "XYZ":
            m = O.Model(np.eye(2), orb, True)
            m.add_onsite([0.37] * norb, a)
            H = O.gen_ham(m, [0.1, 0.2], "lattice")  # only R = 0 is stored
            assert np.allclose(H, 2 * 0.37 * S[a], rtol=0, atol=1e-15)


def _kane_mele(make, axis, t2, delta=0.7, t=-1.0):
    """Honeycomb model H = 1 (x) H0 + sigma_axis (x) H_so: spin-independent
    nearest-neighbour hopping and staggered onsite (H0), and the Haldane
    next-neighbour term i t2 decorated with the Pauli matrix `axis` (H_so).
    It conserves S_axis; its S_axis = +-1/2 sectors are the spinless Haldane
    models with next-neighbour term +-t2 (H0 +- H_so)."""
    m = make(RM.HEX_LAT, RM.HEX_ORB, True)
    m.set_onsite([-delta, delta])
    for R in ([0, 0], [-1, 0], [0, -1]):
        m.add_hop(t, 0, 1, R)
    for R in ([1, 0], [-1, 1], [0, -1]):
        m.add_hop(complex(0.0, t2), 0, 0, R, axis)
    for R in ([-1, 0], [1, -1], [0, 1]):
        m.add_hop(complex(0.0, t2), 1, 1, R, axis)
    return m


@pytest.mark.parametrize("axis", ["X", "Y", "Z"])
def test_spin_berry_curvature_of_conserved_axis(axis):
    """The spin-current Berry curvature (J = {S, v_a}/2, traits.rs:170-175)
    of a model that conserves S_axis is +-1/2 times the spinless curvature of
    each sector: per k, the occupied sum equals (Omega_occ(H0 + H_so) -
    Omega_occ(H0 - H_so)) / 2 from the pinned spinless function; over the
    Brillouin zone 2 pi sigma_spin = (C_+ - C_-) / 2 = +-1.  A wrong sign of
    sigma_y, a missing 1/2 or the spin index taken inner fails this."""
    t2 = -0.3
    m = _kane_mele(make, axis, t2)
    up, dn = RM.haldane(make, t2), RM.haldane(make, -t2)
    d = np.eye(2)
    for k in np.random.default_rng(5).random((6, 2)):
        om, E = O.berry_curvature_at(m, k, d, 1e-3, spin=axis)
        ou, Eu = O.berry_curvature_at(up, k, d, 1e-3)
        od, Ed = O.berry_curvature_at(dn, k, d, 1e-3)
        assert np.allclose(E, np.sort(np.concatenate([Eu, Ed])), rtol=0, atol=1e-12)
        occ = om[E < 0].sum()
        ref = 0.5 * (ou[Eu < 0].sum() - od[Ed < 0].sum())
        assert abs(occ - ref) <= 1e-10 * max(1.0, abs(ref))
        # the charge curvature of the same model is the sum of the sectors
        oc, _ = O.berry_curvature_at(m, k, d, 1e-3)
        assert abs(oc[E < 0].sum() - (ou[Eu < 0].sum() + od[Ed < 0].sum())) <= 1e-10 * max(1.0, abs(ref))
    s = O.hall_conductivity_direct(m, [31, 31], d, [0.0], 0.0, 1e-3, spin=axis)[0]
    su = O.hall_conductivity_direct(up, [31, 31], d, [0.0], 0.0, 1e-3)[0]
    sd = O.hall_conductivity_direct(dn, [31, 31], d, [0.0], 0.0, 1e-3)[0]
    assert abs(s - 0.5 * (su - sd)) < 1e-10
    assert abs(abs(2 * math.pi * s) - 1.0) < 1e-3


def test_fermi_derivative_is_minus_df_dE():
    """fermi_derivative_from_width (thermodynamics.rs:122-126) against a
    central difference of the pinned fermi(), and its unit integral over E."""
    mu = 0.013
    for T in (10.0, 300.0, 2000.0):
        w = T * O.BOLTZMANN_CONSTANT_EV_PER_K
        E = mu + np.linspace(-30.0, 30.0, 601) * w
        h = 1e-5 * w
        fd = -(O.fermi(E + h, mu, T) - O.fermi(E - h, mu, T)) / (2.0 * h)
        d = O.fermi_derivative(E, mu, T)
        assert np.abs(d - fd).max() <= 1e-6 * d.max()
        assert abs(d.max() - 0.25 / w) <= 1e-12 / w  # at E = mu: f (1 - f) = 1/4
        Eg = mu + np.linspace(-60.0, 60.0, 200001) * w
        assert abs(np.trapezoid(O.fermi_derivative(Eg, mu, T), Eg) - 1.0) < 1e-8


# ------------------------------------------- the C++ restatement (checker)
def test_cpu_ref_matches_oracle():
    """oracle/cpu_ref.cpp (the benchmark's accuracy checker and CPU baseline)
    against the pinned oracle on a random spinful 3D model: eigenvalues of
    both modes and Omega (gauge-invariant, so LAPACK phase choices drop out)."""
    from oracle import cpu_ref

    try:
        cpu_ref.load()
    except (RuntimeError, OSError, subprocess.CalledProcessError) as e:
        pytest.skip(f"oracle/cpu_ref.cpp not buildable or no host LAPACK: {e}")
    rng = np.random.default_rng(17)
    lat = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    o = O.Model(lat, rng.random((3, 3)), True)
    n = o.nsta
    z = np.zeros(3, int)
    for i in range(n):
        o.add_element(rng.uniform(-1, 1), i, i, z)
        for j in range(i + 1, n):
            o.add_element(complex(*rng.standard_normal(2)), i, j, z)
    for _ in range(5):
        R = rng.integers(-2, 3, 3)
        if R.any():
            for i in range(n):
                for j in range(n):
                    o.add_element(complex(*rng.standard_normal(2)) * 0.4, i, j, R)
    ks = rng.random((12, 3))
    d = np.array([[1.0, 0.2, 0.0], [0.0, 1.0, 0.3]])
    om, E = cpu_ref.run(2, o.lat, o.orb, o.spin, o.hamR, o.ham, ks, d, 1e-3, threads=2)
    ob, Eb = O.berry_curvature_batch(o, ks, d, 1e-3)
    E0 = cpu_ref.run(0, o.lat, o.orb, o.spin, o.hamR, o.ham, ks, threads=2)
    scale = np.abs(Eb).max()
    assert np.abs(E - Eb).max() <= 1e-12 * scale
    assert np.abs(E0 - Eb).max() <= 1e-12 * scale
    assert np.abs(om - ob).max() <= 1e-9 * np.abs(ob).max()
