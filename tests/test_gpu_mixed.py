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
