"""EnergyCut Hall conductivity (2D and 3D): the oracle pinned by closed forms
and the reference's own checks, the library's host part (band tracking +
Fermi cut, rtb_energy_cut_host / _3d, no device work) against the oracle on
the CPU, and the full device path (rtb_hall_energy_cut) against the oracle on
the GPU.

GPU tolerance: 1e-9 relative to max|sigma| (the vertex data differ from the
oracle's by FP64 rounding of the products and the eigensolver; band tracking
compares overlaps, which stay far from ties on the gapped test models)."""
import ctypes
import math

import numpy as np
import pytest

import rustb_amd as ra
from oracle import energy_cut_ref as EC
from oracle import oracle as O
from rustb_amd._lib import load
from tests import refmodels as RM

KB = O.BOLTZMANN_CONSTANT_EV_PER_K


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def haldane(t2=-0.3):
    return RM.haldane(O.Model, t2_imag=t2)


def haldane_broken(make):
    """The Haldane model with two extra hoppings that break its lattice
    symmetries (no shared vertex energies within a triangle)."""
    m = RM.haldane(make, t2_imag=-0.3)
    m.add_hop(0.09, 0, 0, [1, 1])
    m.add_hop(0.04 + 0.03j, 0, 1, [1, 0])
    return m


def soc_broken(make):
    """The spinful graphene example with three extra hoppings that break its
    lattice symmetries, so no two vertices of a triangle share an energy."""
    m = RM.graphene_soc_example(make)
    m.add_hop(0.13, 0, 0, [1, 1])
    m.add_hop(0.07, 1, 1, [1, -2])
    m.add_hop(0.05 + 0.02j, 0, 1, [1, 0])
    return m


def min_vertex_tie(pts, nx, ny):
    """Smallest |E_n(v) - E_n(v')| over the vertex pairs of every triangle.
    Where two vertices share an energy exactly, the reference's clip_triangle
    splits the occupied quadrilateral along a diagonal chosen by the sort order
    of the tied energies, so the result there hangs on their last bits (the two
    splits' 3-point rules differ at ~1e-6); parity cases keep clear of ties."""
    best = np.inf
    for ix in range(nx):
        for iy in range(ny):
            ixp, iyp = (ix + 1) % nx, (iy + 1) % ny
            i00, i10, i11, i01 = ix * ny + iy, ixp * ny + iy, ixp * ny + iyp, ix * ny + iyp
            for t in ((i00, i10, i01), (i11, i10, i01), (i00, i10, i11), (i00, i11, i01)):
                e = np.array([pts[v].band for v in t])
                for a in range(3):
                    for b in range(a + 1, 3):
                        best = min(best, float(np.abs(e[a] - e[b]).min()))
    return best


# ------------------------------------------------------------ oracle pins
@pytest.mark.parametrize("seed", range(6))
def test_clip_triangle_area_brute_force(seed):
    """The clipped sub-triangles cover exactly the region where the linear
    interpolation of the vertex energies is <= mu (counted on a fine grid)."""
    rng = np.random.default_rng(seed)
    e = list(rng.uniform(-1, 1, 3))
    c = [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)]
    N = 800
    i, j = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    keep = i + j < N - 1
    l1, l2 = (i[keep] + 1 / 3) / N, (j[keep] + 1 / 3) / N  # centroids of the lower sub-cells
    l0 = 1.0 - l1 - l2
    for mu in np.linspace(min(e) - 0.1, max(e) + 0.1, 9):
        area = sum(EC.sub_tri_area(c, t) for t in EC.clip_triangle(e, mu))
        frac = np.mean(l0 * e[0] + l1 * e[1] + l2 * e[2] <= mu)
        assert abs(area - 0.5 * frac) < 4e-3, (mu, area, 0.5 * frac)


def test_greedy_assign_recovers_permutations():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 9):
        for _ in range(5):
            p = rng.permutation(n)
            ov = np.zeros((n, n))
            ov[np.arange(n), p] = 1.0
            ov += 0.01 * rng.random((n, n))
            assert EC.greedy_assign(ov) == list(p)


def test_band_tracking_undoes_shuffled_band_order():
    """Vertex data whose band order is shuffled per vertex track back to the
    unshuffled order (up to the root vertex's permutation)."""
    m = haldane()
    mesh = (6, 5)
    pts = EC.vertices(m, mesh, np.eye(2))
    rng = np.random.default_rng(1)
    perms = [rng.permutation(2) for _ in pts]
    assert any(list(q) != [0, 1] for q in perms)
    shuf = [EC.permute_vertex(p, q) for p, q in zip(pts, perms)]
    q0 = perms[0]  # the root keeps its order; every vertex follows it
    EC.global_band_track(shuf, mesh)
    for a, b in zip(shuf, pts):
        np.testing.assert_array_equal(a.band, b.band[q0])
        np.testing.assert_array_equal(a.k_ab, b.k_ab[np.ix_(q0, q0)])


def test_sum_rule_all_bands_occupied():
    """sum_n Omega_n = 0 (K_mn = conj(K_nm)), so sigma vanishes above all bands."""
    m = haldane()
    s = EC.hall_conductivity_energy_cut(m, (8, 8), np.eye(2), [-9.0, 9.0], 0.0, 0.05)
    assert s[0] == 0.0 and abs(s[1]) < 1e-13


def test_thermal_convolution_closed_forms():
    """Constant and linear sigma_0 pass through (the window is symmetric and
    integrates to 1); a smooth sigma_0 gains (pi^2/6) width^2 sigma_0''.  As
    in the reference, the fine grid is a linspace whose step is a little below
    the width / 10 the interpolation index divides by, so sigma_0 is read at
    mu_lo + r (mu' - mu_lo), r = step / (width / 10).  The grid ends 12 widths
    beyond the mu range, so at the extreme mu the window's tail past x = 12
    (~e^-12 = 6e-6) is dropped; inside, only the tails past |x| = 18 (3e-8)."""
    width = 300.0 * KB
    mu = [-0.3, -0.05, 0.0, 0.2]
    lo, hi, d = min(mu) - 12 * width, max(mu) + 12 * width, 0.1 * width
    n_ext = math.ceil((hi - lo) / d) + 1
    r = (hi - lo) / (n_ext - 1) / d
    assert 0.99 < r < 1.0
    const = EC.thermal_convolution(mu, width, lambda ms: [2.5] * len(ms))
    np.testing.assert_allclose(const, 2.5, rtol=1e-5)
    np.testing.assert_allclose(const[1:3], 2.5, rtol=1e-7)
    lin = EC.thermal_convolution(mu, width, lambda ms: [0.7 + 1.3 * x for x in ms])
    want = [0.7 + 1.3 * (lo + r * (x - lo)) for x in mu]
    np.testing.assert_allclose(lin, want, rtol=1e-5, atol=1e-9)
    s = 0.5
    tanh = EC.thermal_convolution(mu, width, lambda ms: [math.tanh(x / s) for x in ms])
    for m_, v in zip(mu, tanh):
        t = math.tanh((lo + r * (m_ - lo)) / s)
        second = -2.0 * t * (1 - t * t) / (s * s)
        assert abs(v - (t + math.pi ** 2 / 6 * (r * width) ** 2 * second)) < 1e-4  # 4th order left


def test_matches_direct_and_chern_plateau():
    """lib.rs:2620-2660: EnergyCut within 3e-2 of Direct; 2 pi sigma = C in the gap."""
    m = haldane()
    mu = np.linspace(-3, 3, 13)
    ec = EC.hall_conductivity_energy_cut(m, (12, 12), np.eye(2), mu, 0.0, 0.05)
    dr = O.hall_conductivity_direct(m, (12, 12), np.eye(2), mu, 0.0, 0.05)
    assert np.abs(ec - dr).max() < 3e-2
    assert abs(2 * math.pi * ec[6] - 1.0) < 2e-3


def test_time_reversal_antisymmetry():
    """lib.rs:2662-2760: sigma(Haldane, t2) = -sigma(Haldane, -t2)."""
    mu = np.linspace(-3, 3, 7)
    a = EC.hall_conductivity_energy_cut(haldane(-0.3), (8, 8), np.eye(2), mu, 0.0, 0.05)
    b = EC.hall_conductivity_energy_cut(haldane(0.3), (8, 8), np.eye(2), mu, 0.0, 0.05)
    np.testing.assert_allclose(a, -b, atol=1e-12)


def test_oracle_errors():
    with pytest.raises(O.OracleError, match="sorted"):
        EC.hall_conductivity_energy_cut(haldane(), (4, 4), np.eye(2), [0.1, 0.0])
    chain = O.Model([[1.0]], [[0.0]], False)
    chain.add_hop(1.0, 0, 0, [1])
    with pytest.raises(O.OracleError, match="InvalidDimension"):
        EC.hall_conductivity_energy_cut(chain, (4,), [[1.0], [1.0]], [0.0])


# ---------------------------------------------- host part of the library
def vertex_arrays(pts, nx, ny):
    """E, K and the raw overlaps U_i^H U_j the device hands the host part
    (U = conj(evec): eigenvectors as columns)."""
    U = np.array([p.evec.conj() for p in pts])
    E = np.ascontiguousarray([p.band for p in pts])
    K = np.ascontiguousarray([p.k_ab for p in pts], dtype=np.complex128)
    ovx = np.ascontiguousarray(np.einsum("ian,iam->inm", U[:-ny].conj(), U[ny:]))
    ovy = np.ascontiguousarray(np.einsum("ian,iam->inm", U[:-1].conj(), U[1:]))
    return E, K, ovx, ovy


def host_sigma(E, K, ovx, ovy, nx, ny, n, mu, T, eta, det):
    mu = np.ascontiguousarray(mu, np.float64)
    out = np.zeros(mu.size)
    rc = load().rtb_energy_cut_host(nx, ny, n, P(E), P(K), P(ovx), P(ovy), P(mu), mu.size, T, eta, det, P(out))
    return rc, out


@pytest.mark.parametrize("T", [0.0, 300.0])
@pytest.mark.parametrize("shuffle", [False, True])
def test_host_part_matches_oracle(T, shuffle):
    """rtb_energy_cut_host on oracle vertex data equals the oracle bit for bit
    (same operation order on the host); with each vertex's bands shuffled the
    tracking must restore the same result."""
    m = haldane()
    nx, ny = 7, 6
    mu = np.linspace(-3, 3, 25)
    pts = EC.vertices(m, (nx, ny), np.eye(2))
    ref = EC.hall_conductivity_energy_cut(m, (nx, ny), np.eye(2), mu, T, 0.05)
    if shuffle:
        rng = np.random.default_rng(3)
        pts = [pts[0]] + [EC.permute_vertex(p, rng.permutation(2)) for p in pts[1:]]
    rc, out = host_sigma(*vertex_arrays(pts, nx, ny), nx, ny, 2, mu, T, 0.05, float(np.linalg.det(m.lat)))
    assert rc == 0
    np.testing.assert_array_equal(out, ref)


def test_host_part_undoes_three_cycles():
    """Every vertex but the first with three of its six bands rotated (a
    3-cycle, which is not its own inverse): tracking must restore the
    unshuffled result bit for bit, so a permutation applied as its inverse
    (or a transposed overlap) would fail here where 2-band swaps cannot."""
    m = wide_model(O.Model, 21, 6, 2)
    nx, ny = 6, 5
    mu = np.linspace(-3, 3, 13)
    pts = EC.vertices(m, (nx, ny), np.eye(2))
    assert min_vertex_tie(pts, nx, ny) > 1e-6
    ref = EC.hall_conductivity_energy_cut(m, (nx, ny), np.eye(2), mu, 0.0, 0.05)
    rng = np.random.default_rng(4)
    shuffled = [pts[0]]
    for q in pts[1:]:
        p = np.arange(6)
        idx = rng.choice(6, 3, replace=False)
        p[idx] = idx[[1, 2, 0]]
        assert np.any(p[p] != np.arange(6))
        shuffled.append(EC.permute_vertex(q, p))
    rc, out = host_sigma(*vertex_arrays(shuffled, nx, ny), nx, ny, 6, mu, 0.0, 0.05, 1.0)
    assert rc == 0 and np.abs(ref).max() > 1e-3
    np.testing.assert_array_equal(out, ref)


def test_host_part_spinful_matches_oracle():
    m = soc_broken(O.Model)
    nx, ny = 5, 4
    mu = np.linspace(-8, 8, 17)
    ref = EC.hall_conductivity_energy_cut(m, (nx, ny), np.eye(2), mu, 0.0, 0.02, spin="Z")
    pts = EC.vertices(m, (nx, ny), np.eye(2), "Z")
    rc, out = host_sigma(*vertex_arrays(pts, nx, ny), nx, ny, 4, mu, 0.0, 0.02, float(np.linalg.det(m.lat)))
    assert rc == 0
    assert np.abs(ref).max() > 1e-3  # a spin current that is not zero
    np.testing.assert_array_equal(out, ref)


def test_host_part_rejects_unsorted_mu():
    E = np.zeros((4, 2))
    K = np.zeros((4, 2, 2), np.complex128)
    ov = np.zeros((3, 2, 2), np.complex128)
    rc, _ = host_sigma(E, K, ov, ov, 2, 2, 2, [0.1, 0.0], 0.0, 0.01, 1.0)
    assert rc == -6


# ------------------------------------------------------------ GPU parity
def lib_pair(recipe, *a):
    return recipe(ra.Model.tb_model, *a), recipe(O.Model, *a)


@pytest.mark.gpu
@pytest.mark.parametrize("T", [0.0, 300.0])
def test_gpu_haldane_vs_oracle(T):
    m, o = lib_pair(haldane_broken)
    mesh = (10, 9)
    mu = np.linspace(-3, 3, 25)
    assert min_vertex_tie(EC.vertices(o, mesh, np.eye(2)), *mesh) > 1e-6
    p = ra.Parameters.rank2(mesh, [1.0, 0.0], [0.0, 1.0], mu, eta=0.05, integration="EnergyCut")
    p.with_temperature(T)
    got = m.hall_conductivity(p).conductivity
    ref = EC.hall_conductivity_energy_cut(o, mesh, np.eye(2), mu, T, 0.05)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * scale)


@pytest.mark.gpu
def test_gpu_spin_current_vs_oracle():
    m, o = lib_pair(soc_broken)
    mesh = (6, 6)
    mu = np.linspace(-8, 8, 17) + 0.0371
    assert min_vertex_tie(EC.vertices(o, mesh, np.eye(2), "Z"), *mesh) > 1e-6
    p = ra.Parameters.rank2(mesh, [1.0, 0.0], [0.0, 1.0], mu, eta=0.02, integration="EnergyCut", spin="Z")
    got = m.hall_conductivity(p).conductivity
    ref = EC.hall_conductivity_energy_cut(o, mesh, np.eye(2), mu, 0.0, 0.02, spin="Z")
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


@pytest.mark.gpu
def test_gpu_energy_cut_vs_direct_and_time_reversal():
    """The reference's checks on the library's own two forms (lib.rs:2620-2760)."""
    mu = np.linspace(-3, 3, 31)
    res = {}
    for t2 in (-0.3, 0.3):
        m = RM.haldane(ra.Model.tb_model, t2_imag=t2)
        for integ in ("Direct", "EnergyCut"):
            p = ra.Parameters.rank2((24, 24), [1.0, 0.0], [0.0, 1.0], mu, eta=0.05, integration=integ)
            res[t2, integ] = m.hall_conductivity(p).conductivity
    assert np.abs(res[-0.3, "Direct"] - res[-0.3, "EnergyCut"]).max() < 3e-2
    np.testing.assert_allclose(res[-0.3, "EnergyCut"], -res[0.3, "EnergyCut"], atol=1e-10)


@pytest.mark.gpu
def test_gpu_energy_cut_errors():
    m = RM.haldane(ra.Model.tb_model)
    p = ra.Parameters.rank2((4, 4), [1.0, 0.0], [0.0, 1.0], [0.2, 0.1], integration="EnergyCut")
    with pytest.raises(ra.TbError) as e:
        m.hall_conductivity(p)
    assert e.value.code == -6
    p = ra.Parameters.rank2((4, 4), [1.0, 0.0], [0.0, 1.0], [0.0], integration="Simplex")
    with pytest.raises(ra.TbError) as e:
        m.hall_conductivity(p)
    assert e.value.code == -6
    chain = ra.Model.tb_model([[1.0]], [[0.0]], False)
    chain.add_hop(1.0, 0, 0, [1])
    p = ra.Parameters.rank2((4,), [1.0], [1.0], [0.0], integration="EnergyCut")
    with pytest.raises(ra.TbError) as e:
        chain.hall_conductivity(p)
    assert e.value.code == -24
    cube = RM.qwz_stacked_3d(ra.Model.tb_model)
    p = ra.Parameters.rank2((3, 3, 3), [1.0, 0, 0], [0, 1.0, 0], [0.2, 0.1], integration="EnergyCut")
    with pytest.raises(ra.TbError) as e:
        cube.hall_conductivity(p)
    assert e.value.code == -6


RAGGED = [(1, 5), (5, 1), (1, 1), (2, 2), (3, 7)]


@pytest.mark.parametrize("mesh", RAGGED)
def test_host_part_ragged_meshes(mesh):
    """Meshes with one row, one column or one point: cells whose corners
    coincide give zero-area triangles, which the cut skips as the reference
    does, so sigma is 0 there (the (1, 5), (5, 1) and (1, 1) cases check the
    skipping and the walk, not the cut); tracking walks whatever neighbours
    exist."""
    m = haldane_broken(O.Model)
    nx, ny = mesh
    mu = np.linspace(-3, 3, 9)
    pts = EC.vertices(m, mesh, np.eye(2))
    ref = EC.hall_conductivity_energy_cut(m, mesh, np.eye(2), mu, 0.0, 0.05)
    E, K, ovx, ovy = vertex_arrays(pts, nx, ny)
    # empty overlap arrays still need a valid pointer
    ovx = ovx if ovx.size else np.zeros((1, 2, 2), np.complex128)
    ovy = ovy if ovy.size else np.zeros((1, 2, 2), np.complex128)
    rc, out = host_sigma(E, K, ovx, ovy, nx, ny, 2, mu, 0.0, 0.05, float(np.linalg.det(m.lat)))
    assert rc == 0
    np.testing.assert_array_equal(out, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("mesh", RAGGED)
def test_gpu_ragged_meshes_vs_oracle(mesh):
    m, o = lib_pair(haldane_broken)
    mu = np.linspace(-3, 3, 9)
    p = ra.Parameters.rank2(mesh, [1.0, 0.0], [0.0, 1.0], mu, eta=0.05, integration="EnergyCut")
    got = m.hall_conductivity(p).conductivity
    ref = EC.hall_conductivity_energy_cut(o, mesh, np.eye(2), mu, 0.0, 0.05)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * max(np.abs(ref).max(), 1e-3))



def wide_model(make, seed, n, dim):
    """A seeded random complex model of n orbitals (on-site block and four
    hopping vectors, no lattice or time-reversal symmetry), wide enough that
    the device cut splits each band sum over 64 lanes (n > 32) and, for
    n > 64, over several terms per lane."""
    rng = np.random.default_rng(seed)
    R = [[0, 0], [1, 0], [0, 1], [1, 1], [1, -1]] if dim == 2 else [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                                                                     [1, 1, 0]]
    w = np.array([1.0, 0.5, 0.5, 0.25, 0.25])[:, None, None]
    ham = (rng.standard_normal((5, n, n)) + 1j * rng.standard_normal((5, n, n))) * w
    ham[0] = 0.5 * (ham[0] + ham[0].conj().T)
    orb = rng.random((n, dim))
    if make is O.Model:
        return O.model_from_arrays(np.eye(dim), orb, False, np.array(R), ham)
    return ra.Model.from_arrays(np.eye(dim), orb, False, np.array(R), ham)


def overlap_margin(ovs):
    """Smallest gap between the two largest |<u_i|u_j>|^2 of a row over all
    neighbour overlaps: band tracking is decided far from ties above ~1e-8."""
    best = np.inf
    for ov in ovs:
        if ov.shape[0] == 0:
            continue
        a = np.sort(np.abs(ov) ** 2, axis=-1)
        best = min(best, float((a[..., -1] - a[..., -2]).min()))
    return best


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n", [(1, 40), (5, 72)])
def test_gpu_wide_model_vs_host_part(seed, n):
    """n = 40 and 72 against the host part (rtb_energy_cut_host, bit-exact
    with the oracle on the same vertex data) fed the oracle's vertex data."""
    m, o = wide_model(ra.Model, seed, n, 2), wide_model(O.Model, seed, n, 2)
    mesh = (7, 6)
    mu = np.linspace(-6, 6, 13)
    pts = EC.vertices(o, mesh, np.eye(2))
    assert min_vertex_tie(pts, *mesh) > 1e-5
    E, K, ovx, ovy = vertex_arrays(pts, *mesh)
    assert overlap_margin([ovx, ovy]) > 1e-8
    rc, ref = host_sigma(E, K, ovx, ovy, *mesh, n, mu, 0.0, 0.05, 1.0)
    assert rc == 0
    p = ra.Parameters.rank2(mesh, [1.0, 0.0], [0.0, 1.0], mu, eta=0.05, integration="EnergyCut")
    got = m.hall_conductivity(p).conductivity
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * np.abs(ref).max())

# ------------------------------------------------------------------ 3D
def qwz_broken_3d(make):
    """The stacked QWZ model with interlayer and diagonal hoppings, so that
    the bands depend on kz and no two vertices of a tetrahedron share an
    energy (the smallest gap between them is 1.4e-3 on the 5 x 4 x 3 mesh)."""
    m = RM.qwz_stacked_3d(make)
    m.add_hop(0.23, 0, 0, [0, 0, 1])
    m.add_hop(-0.11, 1, 1, [0, 0, 1])
    m.add_hop(0.07 + 0.05j, 0, 1, [1, 0, 1])
    m.add_hop(0.05, 0, 0, [1, 1, 0])
    m.add_hop(0.031 + 0.017j, 0, 1, [0, 1, 1])
    m.add_hop(0.043, 1, 1, [1, -1, 0])
    m.add_hop(0.029j, 0, 1, [1, 2, 0])
    return m


def min_vertex_tie_3d(pts, nx, ny, nz):
    best = np.inf
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                for vidx, _, _ in EC.build_tetrahedra_3d(ix, iy, iz, nx, ny, nz):
                    e = np.array([pts[v].band for v in vidx])
                    for a in range(4):
                        for b in range(a + 1, 4):
                            best = min(best, float(np.abs(e[a] - e[b]).min()))
    return best


def bloechl_occupied(V, e, mu):
    """Volume of {E <= mu} for E linear on a tetrahedron of volume V with
    vertex energies e (Bloechl, Jepsen and Andersen, PRB 49, 16223 (1994),
    eqs. B1-B4, the integrated number of states)."""
    e1, e2, e3, e4 = sorted(e)
    if mu <= e1:
        return 0.0
    if mu <= e2:
        return V * (mu - e1) ** 3 / ((e2 - e1) * (e3 - e1) * (e4 - e1))
    if mu <= e3:
        e21, e31, e41, e32, e42 = e2 - e1, e3 - e1, e4 - e1, e3 - e2, e4 - e2
        x = mu - e2
        return V / (e31 * e41) * (e21 ** 2 + 3 * e21 * x + 3 * x ** 2 - (e31 + e42) / (e32 * e42) * x ** 3)
    if mu < e4:
        return V * (1.0 - (e4 - mu) ** 3 / ((e4 - e1) * (e4 - e2) * (e4 - e3)))
    return V


@pytest.mark.parametrize("seed", range(8))
def test_tetrahedron_cut_volume_closed_form(seed):
    """The sub-tetrahedra of the cut (1, 2 or 3 vertices below mu) cover
    exactly the occupied volume of the linear interpolation."""
    rng = np.random.default_rng(seed)
    c = rng.random((4, 3))
    V = EC.tet_vol_from_pts(*c)
    e = list(rng.normal(size=4))
    for mu in np.linspace(min(e) + 1e-6, max(e) - 1e-6, 23):
        subs = EC.tet_cut(e, mu)
        got = 0.0
        for sign, sub in subs:
            got += EC.sub_tet_vol_3d(c, sub) if sign > 0 else V - EC.sub_tet_vol_3d(c, sub)
        assert abs(got - bloechl_occupied(V, e, mu)) <= 1e-12 * V


def test_3d_sum_rule_all_bands_occupied():
    m = qwz_broken_3d(O.Model)
    s = EC.hall_conductivity_energy_cut(m, (4, 3, 3), np.eye(3)[:2], [-0.2, 10.0], 0.0, 0.05)
    assert abs(s[1]) < 1e-12 and abs(s[0]) > 1e-3


def test_3d_qwz_plateau_and_direct():
    """lib.rs:3219-3270 (nk = 10): the EnergyCut plateau is 1/(2 pi) within
    3e-3 and EnergyCut is within 5e-2 of Direct, eta = 0.1."""
    m = RM.qwz_stacked_3d(O.Model)
    mu = np.array([-2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0])
    ec = EC.hall_conductivity_energy_cut(m, (10, 10, 4), np.eye(3)[:2], mu, 0.0, 0.1)
    dr = O.hall_conductivity_direct(m, (10, 10, 4), np.eye(3)[:2], mu, 0.0, 0.1)
    assert abs(ec[3] - 1 / (2 * math.pi)) < 3e-3
    assert np.abs(ec - dr).max() < 5e-2


def vertex_arrays_3d(pts, nx, ny, nz):
    U = np.array([p.evec.conj() for p in pts])
    E = np.ascontiguousarray([p.band for p in pts])
    K = np.ascontiguousarray([p.k_ab for p in pts], dtype=np.complex128)

    def ov(step):
        if len(pts) <= step:
            return np.zeros((1, E.shape[1], E.shape[1]), np.complex128)
        return np.ascontiguousarray(np.einsum("ian,iam->inm", U[:-step].conj(), U[step:]))

    return E, K, ov(ny * nz), ov(nz), ov(1)


def host_sigma_3d(E, K, ovx, ovy, ovz, mesh, n, mu, T, eta, det):
    mu = np.ascontiguousarray(mu, np.float64)
    out = np.zeros(mu.size)
    rc = load().rtb_energy_cut_host_3d(*mesh, n, P(E), P(K), P(ovx), P(ovy), P(ovz), P(mu), mu.size, T, eta, det,
                                       P(out))
    return rc, out


@pytest.mark.parametrize("T,shuffle", [(0.0, False), (0.0, True), (300.0, True)])
def test_host_part_3d_matches_oracle(T, shuffle):
    """rtb_energy_cut_host_3d on oracle vertex data equals the oracle bit for
    bit; shuffled band orders are restored by the tracking.  (At T > 0 a
    narrow mu window on a smaller mesh: the oracle's zero-width cut runs on a
    grid of width / 10 steps.)"""
    m = qwz_broken_3d(O.Model)
    mesh = (4, 3, 3) if T == 0.0 else (3, 3, 2)
    mu = np.linspace(-2.5, 2.5, 21) if T == 0.0 else np.linspace(-0.35, 0.35, 5)
    pts = EC.vertices(m, mesh, np.eye(3)[:2])
    ref = EC.hall_conductivity_energy_cut(m, mesh, np.eye(3)[:2], mu, T, 0.05)
    assert np.abs(ref).max() > 1e-3
    if shuffle:
        rng = np.random.default_rng(5)
        pts = [pts[0]] + [EC.permute_vertex(p, rng.permutation(2)) for p in pts[1:]]
    rc, out = host_sigma_3d(*vertex_arrays_3d(pts, *mesh), mesh, 2, mu, T, 0.05, float(np.linalg.det(m.lat)))
    assert rc == 0
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("mesh", [(1, 3, 2), (3, 1, 2), (2, 3, 1), (1, 1, 1)])
def test_host_part_3d_ragged_meshes(mesh):
    m = qwz_broken_3d(O.Model)
    mu = np.linspace(-2.5, 2.5, 9)
    pts = EC.vertices(m, mesh, np.eye(3)[:2])
    ref = EC.hall_conductivity_energy_cut(m, mesh, np.eye(3)[:2], mu, 0.0, 0.05)
    rc, out = host_sigma_3d(*vertex_arrays_3d(pts, *mesh), mesh, 2, mu, 0.0, 0.05, float(np.linalg.det(m.lat)))
    assert rc == 0
    np.testing.assert_array_equal(out, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("T", [0.0, 300.0])
def test_gpu_3d_vs_oracle(T):
    m, o = lib_pair(qwz_broken_3d)
    mesh = (5, 4, 3)
    mu = np.linspace(-2.5, 2.5, 21)
    assert min_vertex_tie_3d(EC.vertices(o, mesh, np.eye(3)[:2]), *mesh) > 1e-6
    p = ra.Parameters.rank2(mesh, [1.0, 0, 0], [0, 1.0, 0], mu, eta=0.05, integration="EnergyCut")
    p.with_temperature(T)
    got = m.hall_conductivity(p).conductivity
    ref = EC.hall_conductivity_energy_cut(o, mesh, np.eye(3)[:2], mu, T, 0.05)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


@pytest.mark.gpu
def test_gpu_3d_qwz_plateau_and_direct():
    """lib.rs:3219-3270 on the device, nk = 10 and 14."""
    m = RM.qwz_stacked_3d(ra.Model.tb_model)
    mu = np.linspace(-3.0, 3.0, 61)
    for nk in (10, 14):
        res = {}
        for integ in ("Direct", "EnergyCut"):
            p = ra.Parameters.rank2((nk, nk, 4), [1.0, 0, 0], [0, 1.0, 0], mu, eta=0.1, integration=integ)
            res[integ] = m.hall_conductivity(p).conductivity
        assert np.abs(res["Direct"] - res["EnergyCut"]).max() < 5e-2
        assert abs(res["EnergyCut"][30] - 1 / (2 * math.pi)) < 3e-3


@pytest.mark.gpu
@pytest.mark.parametrize("n,mesh,tol", [(40, (4, 3, 3), 1e-9), (100, (3, 2, 2), 1e-9), (136, (3, 2, 2), 1e-8)])
def test_gpu_3d_wide_model_vs_host_part(n, mesh, tol):
    """3D (tetrahedra, 64-lane band sums: 1, 2 and 3 terms per lane) against
    rtb_energy_cut_host_3d fed the oracle's vertex data.  Tolerance relative
    to max|sigma|: 1e-9, and 1e-8 at n = 136 where the vertex stage takes the
    two-stage eigensolver (more rounding than the n <= 128 route)."""
    m, o = wide_model(ra.Model, 3, n, 3), wide_model(O.Model, 3, n, 3)
    mu = np.linspace(-6, 6, 13)
    pts = EC.vertices(o, mesh, np.eye(3)[:2])
    assert min_vertex_tie_3d(pts, *mesh) > 1e-5
    E, K, ovx, ovy, ovz = vertex_arrays_3d(pts, *mesh)
    assert overlap_margin([ovx, ovy, ovz]) > 1e-8
    rc, ref = host_sigma_3d(E, K, ovx, ovy, ovz, mesh, n, mu, 0.0, 0.05, 1.0)
    assert rc == 0
    p = ra.Parameters.rank2(mesh, [1.0, 0, 0], [0, 1.0, 0], mu, eta=0.05, integration="EnergyCut")
    got = m.hall_conductivity(p).conductivity
    np.testing.assert_allclose(got, ref, rtol=0, atol=tol * np.abs(ref).max())


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n,mesh,slices", [(7, 32, (16, 12), None), (8, 24, (5, 4, 4), None),
                                                (9, 100, (6, 5), None), (10, 128, (4, 3), None),
                                                (16, 136, (4, 3), None), (17, 256, (3, 3), None),
                                                (18, 160, (3, 2, 2), None), (19, 144, (5, 4), "2")])
def test_gpu_device_walk_equals_host_walk(seed, n, mesh, slices, monkeypatch):
    """Band tracking on the device against the host walk (RTB_TRACK_HOST=1)
    on the same device vertex data: identical sigma, bit for bit.  Up to 128
    bands the overlap matrix is in LDS, above in a workspace slice per vertex
    (RTB_TRACK_SLICES=2: levels in launches of two vertices).  On these
    random models the walk reorders the bands of nearly every vertex (counted
    with the oracle's global_band_track)."""
    if slices is None:
        monkeypatch.delenv("RTB_TRACK_SLICES", raising=False)
    else:
        monkeypatch.setenv("RTB_TRACK_SLICES", slices)
    dim = len(mesh)
    pts = list(EC.vertices(wide_model(O.Model, seed, n, dim), mesh, np.eye(dim)[:2]))
    sorted_bands = [q.band.copy() for q in pts]
    EC.global_band_track(pts, mesh)
    assert sum(bool(np.any(np.diff(q.band) < 0)) for q in pts) > len(pts) // 2
    # the walk's permutation of each vertex (tracked band i = sorted band p[i]);
    # some are not involutions, so a permutation applied as its inverse fails
    perms = [np.searchsorted(sb, q.band) for sb, q in zip(sorted_bands, pts)]
    assert any(np.any(p[p] != np.arange(n)) for p in perms)
    m = wide_model(ra.Model, seed, n, dim)
    d = np.eye(dim)
    p = ra.Parameters.rank2(mesh, d[0], d[1], np.linspace(-6, 6, 25), eta=0.05, integration="EnergyCut")
    monkeypatch.delenv("RTB_TRACK_HOST", raising=False)
    dev = m.hall_conductivity(p).conductivity
    monkeypatch.setenv("RTB_TRACK_HOST", "1")
    host = m.hall_conductivity(p).conductivity
    assert np.abs(host).max() > 1e-3
    np.testing.assert_array_equal(dev, host)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,mesh", [(11, (1, 6, 4)), (12, (6, 1, 4)), (13, (6, 4, 1)), (14, (1, 9)), (15, (9, 1))])
def test_gpu_device_walk_ragged_meshes(seed, mesh, monkeypatch):
    """Meshes with a single point along an axis, where index steps of two
    axes coincide and the walk must take each pair's overlaps from the axis
    the host walk takes them from: the device walk equals the host walk
    (RTB_TRACK_HOST=1) bit for bit on a 24-band model whose bands the walk
    permutes at nearly every vertex (non-involutions included)."""
    dim, n = len(mesh), 24
    pts = list(EC.vertices(wide_model(O.Model, seed, n, dim), mesh, np.eye(dim)[:2]))
    sorted_bands = [q.band.copy() for q in pts]
    EC.global_band_track(pts, mesh)
    perms = [np.searchsorted(sb, q.band) for sb, q in zip(sorted_bands, pts)]
    assert any(np.any(p[p] != np.arange(n)) for p in perms)
    m = wide_model(ra.Model, seed, n, dim)
    d = np.eye(dim)
    p = ra.Parameters.rank2(mesh, d[0], d[1], np.linspace(-6, 6, 25), eta=0.05, integration="EnergyCut")
    monkeypatch.delenv("RTB_TRACK_HOST", raising=False)
    dev = m.hall_conductivity(p).conductivity
    monkeypatch.setenv("RTB_TRACK_HOST", "1")
    host = m.hall_conductivity(p).conductivity
    np.testing.assert_array_equal(dev, host)


@pytest.mark.gpu
def test_gpu_n136_two_stage_vs_host_part():
    """n = 136: the vertex stage takes the two-stage FP64 eigensolver and the
    walk keeps each vertex's overlap matrix in a workspace slice; against the
    host part fed the oracle's vertex data.  Tolerance 1e-8 of max|sigma|:
    the two-stage reduction rounds more than the n <= 128 route, and the
    smallest energy gap inside a triangle is 1.1e-4 here."""
    n = 136
    m, o = wide_model(ra.Model, 32, n, 2), wide_model(O.Model, 32, n, 2)
    mesh = (4, 3)
    mu = np.linspace(-6, 6, 13)
    pts = EC.vertices(o, mesh, np.eye(2))
    assert min_vertex_tie(pts, *mesh) > 5e-5
    E, K, ovx, ovy = vertex_arrays(pts, *mesh)
    assert overlap_margin([ovx, ovy]) > 1e-7
    rc, ref = host_sigma(E, K, ovx, ovy, *mesh, n, mu, 0.0, 0.05, 1.0)
    assert rc == 0
    p = ra.Parameters.rank2(mesh, [1.0, 0.0], [0.0, 1.0], mu, eta=0.05, integration="EnergyCut")
    got = m.hall_conductivity(p).conductivity
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-8 * np.abs(ref).max())
