"""Host-side builder and synthetic-model checks (CPU only: no GPU call).

The C++ builder behind the C ABI must reproduce the reference's storage
(model_build.rs semantics, restated by the oracle) bit for bit."""
import numpy as np
import pytest

import rustb_amd as ra
from oracle import oracle as O
from tests import refmodels as RM

MODELS = [RM.gen_v_model, RM.haldane, RM.graphene_soc_example, RM.graphene_config1, RM.qwz_stacked_3d]


@pytest.mark.parametrize("recipe", MODELS, ids=lambda f: f.__name__)
def test_reference_models_storage(recipe):
    m = recipe(ra.Model.tb_model)
    o = recipe(O.Model)
    assert m.hamR.tolist() == o.hamR.tolist()
    assert np.array_equal(m.ham, o.ham)


def test_random_hop_sequences_match_oracle():
    rng = np.random.default_rng(7)
    for spin in (False, True):
        lat = np.eye(3) + 0.1 * rng.standard_normal((3, 3))
        orb = rng.random((3, 3))
        m = ra.Model.tb_model(lat, orb, spin)
        o = O.Model(lat, orb, spin)
        for _ in range(300):
            op = rng.integers(0, 4)
            R = rng.integers(-2, 3, 3)
            i, j = rng.integers(0, 3, 2)
            t = complex(*rng.standard_normal(2))
            pauli = [None, "X", "Y", "Z"][rng.integers(0, 4)]
            if i == j and not R.any():
                t = complex(t.real, 0.0)
            if op == 0:
                m.set_hop(t, i, j, R, pauli), o.set_hop(t, i, j, R, pauli)
            elif op == 1:
                m.add_hop(t, i, j, R, pauli), o.add_hop(t, i, j, R, pauli)
            elif op == 2:
                n = m.nsta
                a, b = rng.integers(0, n, 2)
                if a == b and not R.any():
                    t = complex(t.real, 0.0)
                m.add_element(t, a, b, R), o.add_element(t, a, b, R)
            else:
                v = rng.standard_normal(3)
                m.add_onsite(v, pauli), o.add_onsite(v, pauli)
        assert m.hamR.tolist() == o.hamR.tolist()
        assert np.array_equal(m.ham, o.ham)
        for iR, R in enumerate(m.hamR):  # ham[-R] = ham[R]^H
            inv = o.find_R(-R)
            assert np.array_equal(m.ham[inv], m.ham[iR].conj().T)


def test_builder_errors():
    m = ra.Model.tb_model(np.eye(2), [[0, 0], [0.5, 0.5]], False)
    with pytest.raises(ra.TbError) as e:
        m.set_hop(1j, 0, 0, [0, 0])
    assert e.value.kind == "OnsiteHoppingMustBeReal"
    with pytest.raises(ra.TbError) as e:
        m.set_hop(1.0, 0, 2, [1, 0])
    assert e.value.kind == "DimensionMismatch"
    with pytest.raises(ra.TbError) as e:
        m.add_element(1.0, 0, 5, [1, 0])
    assert e.value.kind == "DimensionMismatch"
    with pytest.raises(ra.TbError) as e:
        m.add_hop(1.0, 0, 1, [1, 0, 0])
    assert e.value.kind == "RVectorLengthError"
    # a rejected onsite write leaves the model unchanged (model_build.rs:500-506)
    before = m.ham.copy()
    with pytest.raises(ra.TbError):
        m.add_element(1 + 2j, 1, 1, [0, 0])
    assert np.array_equal(m.ham, before)


@pytest.mark.parametrize("make", [ra.workloads.random3d_n32, ra.workloads.slab2d_n256])
def test_synthetic_models_are_hermitian(make):
    m = make()
    hamR, ham = m.hamR, m.ham
    assert not hamR[0].any()
    index = {tuple(R): i for i, R in enumerate(hamR)}
    assert len(index) == hamR.shape[0]
    for i, R in enumerate(hamR):
        assert np.array_equal(ham[index[tuple(-R)]], ham[i].conj().T)
    assert np.all(ham[0].diagonal().imag == 0)


def test_synthetic_sizes_and_determinism():
    m = ra.workloads.random3d_n32()
    assert m.ham.shape == (125, 32, 32)
    m2 = ra.workloads.random3d_n32()
    assert np.array_equal(m.ham, m2.ham) and np.array_equal(m.orb, m2.orb)
    s = ra.workloads.slab2d_n256()
    assert s.ham.shape == (49, 256, 256)
    b, nk = ra.workloads.batch2d_model(0)
    assert b.nsta % 4 == 0 and 4 <= b.nsta <= 96 and 16 <= nk <= 128
    assert b.hamR.shape[0] % 2 == 1 and 51 <= b.hamR.shape[0] <= 501
    assert np.abs(b.hamR).max() <= 11


def test_kmesh_and_kpath_match_oracle():
    assert np.array_equal(ra.gen_kmesh([3, 4, 5]), O.gen_kmesh([3, 4, 5]))
    lat = [[3**0.5, -1.0], [3**0.5, 1.0]]
    path = [[0, 0], [1 / 3, 2 / 3], [2 / 3, 1 / 3], [0, 0]]
    a = ra.k_path(lat, path, 2001)
    b = O.k_path(lat, path, 2001)
    for x, y in zip(a, b):
        assert np.allclose(x, y, atol=1e-15, rtol=0)
