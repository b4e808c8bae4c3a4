"""GPU parity of the mixed form (precision="mixed") on mesh entry points.

The mixed form evaluates the Bloch sum on the mesh by axis-separable DFT
passes in FP64 and solves the eigenproblem in FP32 with FP64 refinement.
Bars (from the north-star targets): eigenvalue error <= 1e-9 eV against
the FP64 zheev oracle, AHC relative error <= 1e-6, Chern numbers equal."""
import math

import numpy as np
import pytest

import rustb_amd as ra
from oracle import oracle as O
from tests import refmodels as RM
from tests.test_gpu_parity import random_pair

pytestmark = pytest.mark.gpu


def test_haldane_chern_mixed():
    m = RM.haldane(ra.Model.tb_model)
    mu = np.array([0.0, -2.0, 2.0])
    s = m.hall_conductivity(ra.Parameters.rank2([31, 31], [1, 0], [0, 1], mu, precision="mixed")).conductivity
    o = O.hall_conductivity_direct(RM.haldane(O.Model), [31, 31], np.eye(2), mu, 0.0, 1e-3)
    assert round(2 * math.pi * s[0]) == round(2 * math.pi * o[0]) == 1
    assert np.abs(s - o).max() <= 1e-6 * np.abs(o).max()


def test_qwz_3d_mixed():
    m = RM.qwz_stacked_3d(ra.Model.tb_model)
    p = ra.Parameters.rank2([14, 14, 4], [1, 0, 0], [0, 1, 0], [0.0, 0.5], eta=0.1, precision="mixed")
    s = m.hall_conductivity(p).conductivity
    o = O.hall_conductivity_direct(RM.qwz_stacked_3d(O.Model), [14, 14, 4], [[1, 0, 0], [0, 1, 0]], [0.0, 0.5],
                                   0.0, 0.1)
    assert np.abs(s - o).max() <= 1e-6 * np.abs(o).max()


@pytest.mark.parametrize("dim,kmesh,rmax", [(1, [17], 3), (2, [7, 9], 2), (3, [5, 4, 6], 3), (3, [3, 2, 3], 2)])
def test_random_models_mixed_vs_oracle(dim, kmesh, rmax):
    """Includes meshes smaller than the R extent (phases wrap mod N)."""
    m, o = random_pair(3, True, dim, 8, seed=dim * 7 + rmax, rmax=rmax)
    dirs = np.eye(dim)[:2] if dim > 1 else np.array([[1.0], [0.7]])
    dirs = dirs + 0.3 * np.roll(dirs, 1, axis=1) if dim > 1 else dirs
    mu = np.array([-0.5, 0.0, 0.8])
    p = ra.Parameters(kmesh=tuple(kmesh), direction=dirs, mu=mu, eta=0.05, precision="mixed")
    s = m.hall_conductivity(p).conductivity
    r = O.hall_conductivity_direct(o, kmesh, dirs, mu, 0.0, 0.05)
    assert np.abs(s - r).max() <= 1e-6 * max(np.abs(r).max(), 1e-3)
    p64 = ra.Parameters(kmesh=tuple(kmesh), direction=dirs, mu=mu, eta=0.05, precision="fp64")
    s64 = m.hall_conductivity(p64).conductivity
    assert np.abs(s64 - r).max() <= 1e-10 * max(np.abs(r).max(), 1e-3)


def test_dos_graphene_mixed():
    m, o = RM.graphene_config1(ra.Model.tb_model), RM.graphene_config1(O.Model)
    E, rho = m.dos([101, 101], -4.0, 4.0, 801, 0.02, precision="mixed")
    Er, rr = O.dos(o, [101, 101], -4.0, 4.0, 801, 0.02)
    assert np.abs(rho - rr).max() <= 1e-9 * rr.max()


def test_dos_random_3d_mixed():
    m, o = random_pair(4, False, 3, 6, seed=9)
    E, rho = m.dos([6, 5, 4], -6.0, 6.0, 97, 0.3, precision="mixed")
    Er, rr = O.dos(o, [6, 5, 4], -6.0, 6.0, 97, 0.3)
    assert np.abs(rho - rr).max() <= 1e-9 * rr.max()


@pytest.mark.parametrize("norb", [16, 32])
def test_mixed_eigenpairs_vs_zheev(norb):
    """FP32 eigensolve + FP64 refinement: |dE| <= 1e-9 eV, FP64-level vectors."""
    m, o = random_pair(norb, True, 3, 3, seed=norb + 5)
    ks = np.random.default_rng(8).random((40, 3))
    E, V = m.solve_all(ks, precision="mixed")
    H = O.gen_ham_batch(o, ks)
    for i in range(len(ks)):
        w = O.eigvalsh(H[i])
        assert np.abs(E[i] - w).max() <= 1e-9
        U = V[i].T
        assert np.abs(H[i] @ U - U * E[i]).max() <= 1e-9
        assert np.abs(U.conj().T @ U - np.eye(m.nsta)).max() <= 1e-10


def test_mixed_doubly_degenerate_spectrum():
    """Spin-doubled model (every level exactly twice): the cluster path runs
    at every k; eigenvalues and the gauge-invariant occupied Omega match."""
    rng = np.random.default_rng(21)
    lat, orb = np.eye(2), rng.random((6, 2))
    m = ra.Model.tb_model(lat, orb, True)
    o = O.Model(lat, orb, True)
    m.set_onsite(rng.uniform(-1, 1, 6)), o.set_onsite(m.ham[0].diagonal()[:6].real)
    for _ in range(12):
        R = rng.integers(-1, 2, 2)
        i, j = rng.integers(0, 6, 2)
        if i == j and not R.any():
            continue
        t = complex(*rng.standard_normal(2))
        m.add_hop(t, i, j, R), o.add_hop(t, i, j, R)
    assert np.array_equal(m.ham, o.ham)
    ks = rng.random((30, 2))
    E, V = m.solve_all(ks, precision="mixed")
    H = O.gen_ham_batch(o, ks)
    for i in range(len(ks)):
        assert np.abs(E[i] - O.eigvalsh(H[i])).max() <= 1e-9
        U = V[i].T
        assert np.abs(U.conj().T @ U - np.eye(12)).max() <= 1e-10
    p = ra.Parameters(kmesh=(1, 1), direction=np.eye(2), mu=np.array([0.0]), eta=1e-3, precision="mixed")
    occ = m.occupied_berry_curvature_on(ks, p)
    ob, Eb = O.berry_curvature_batch(o, ks, np.eye(2), 1e-3)
    ref = (ob * O.fermi(Eb, 0.0, 0.0)).sum(axis=1)
    assert np.abs(occ - ref).max() <= 1e-6 * max(1.0, np.abs(ref).max())


def test_mixed_berry_vs_oracle():
    m, o = random_pair(4, True, 3, 4, seed=44)
    ks = np.random.default_rng(9).random((24, 3))
    dirs = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    p = ra.Parameters(kmesh=(1, 1, 1), direction=dirs, mu=np.array([0.0]), eta=1e-3, precision="mixed")
    om, E = m.berry_curvature_on(ks, p)
    ob, Eb = O.berry_curvature_batch(o, ks, dirs, 1e-3)
    assert np.abs(E - Eb).max() <= 1e-9
    assert np.abs(om - ob).max() <= 1e-6 * np.abs(ob).max()
