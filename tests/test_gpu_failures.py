"""Failure reporting and concurrent calls through the C ABI on the GPU.

* the mixed form's eigenpair check: forced to flag every k-point
  (RTB_MIXED_RESIDUAL=0), the FP64 re-solve must reproduce the plain form;
  with the re-solve switched off (RTB_MIXED_FALLBACK=0) the failure must
  reach the caller as EigenvalueComputationFailed (the reference's zheev
  info != 0, ndarray_lapack.rs:443-447; TbError from eigh, traits.rs:168);
* a model whose H(k) overflows (finite storage, hoppings of 1e300) fails
  with EigenvalueComputationFailed in both forms instead of returning
  numbers that are not finite;
* four threads calling one model at once agree bit for bit with serial calls.
"""
import os
import threading

import numpy as np
import pytest

import rustb_amd as ra
from tests.test_gpu_parity import random_pair

pytestmark = pytest.mark.gpu

EIGEN_FAILED = -8


class env:
    """Environment variables set for the duration of a block (the library
    reads its options per call)."""

    def __init__(self, **kv):
        self.kv = kv
        self.old = {}

    def __enter__(self):
        for k, v in self.kv.items():
            self.old[k] = os.environ.get(k)
            os.environ[k] = v

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def model():
    m, _ = random_pair(6, True, 3, 4, seed=12)
    return m


def hall_params(precision, mesh=(6, 5, 4)):
    d = np.array([[1.0, 0.0, 0.2], [0.0, 1.0, 0.0]])
    return ra.Parameters(kmesh=mesh, direction=d, mu=np.array([-0.4, 0.0, 0.6]), eta=0.02, precision=precision)


def test_unforced_mixed_check_passes_and_reports_residual(model):
    model.hall_conductivity(hall_params("mixed"))
    st = ra.last_stats()
    assert st["resolved"] == 0
    assert 0.0 < st["max_residual"] <= 1e-9 * 20.0


def test_forced_resolve_reproduces_fp64(model):
    with env(RTB_MIXED_RESIDUAL="0"):
        s = model.hall_conductivity(hall_params("mixed")).conductivity
        st = ra.last_stats()
    assert st["resolved"] == 6 * 5 * 4
    s64 = model.hall_conductivity(hall_params("fp64")).conductivity
    assert np.abs(s - s64).max() <= 1e-10 * max(np.abs(s64).max(), 1e-3)
    ks = np.random.default_rng(4).random((37, 3))
    with env(RTB_MIXED_RESIDUAL="0"):
        E, V = model.solve_all(ks, precision="mixed")
        assert ra.last_stats()["resolved"] == 37
    E64, V64 = model.solve_all(ks, precision="fp64")
    assert np.abs(E - E64).max() <= 1e-12 * np.abs(E64).max()
    # same eigenvector rows up to a phase: |<u, u64>| = 1 (isolated levels)
    ov = np.abs(np.einsum("kij,kij->ki", V.conj(), V64))
    gaps = np.diff(E64, axis=1).min(axis=1) > 1e-6
    assert np.abs(ov[gaps] - 1.0).max() <= 1e-10


def test_refinement_failure_reaches_the_abi(model):
    with env(RTB_MIXED_RESIDUAL="0", RTB_MIXED_FALLBACK="0"):
        with pytest.raises(ra.TbError) as e:
            model.hall_conductivity(hall_params("mixed"))
        assert e.value.code == EIGEN_FAILED
        with pytest.raises(ra.TbError) as e:
            model.solve_all(np.zeros((3, 3)), precision="mixed")
        assert e.value.code == EIGEN_FAILED
    # the same calls succeed again once the options are back
    model.hall_conductivity(hall_params("mixed"))


def test_overflowing_hamiltonian_fails_cleanly():
    m, _ = random_pair(4, False, 2, 2, seed=3)
    h = m.ham * 1e300  # finite storage; squares and sums overflow
    big = ra.Model.from_arrays(m.lat, m.orb, False, m.hamR, h)
    ks = np.random.default_rng(1).random((5, 2))
    for prec in ("fp64", "mixed"):
        with pytest.raises(ra.TbError) as e:
            big.solve_all(ks, precision=prec)
        assert e.value.code == EIGEN_FAILED
    with pytest.raises(ra.TbError) as e:
        big.solve_band_all(ks)
    assert e.value.code == EIGEN_FAILED
    p = ra.Parameters(kmesh=(4, 4), direction=np.eye(2), mu=np.array([0.0]), eta=0.01, precision="mixed")
    with pytest.raises(ra.TbError) as e:
        big.hall_conductivity(p)
    assert e.value.code == EIGEN_FAILED


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_concurrent_calls_match_serial(model, precision):
    """Four threads on one model with a 1 MB chunk budget (many chunks
    per call): every result equal to the serial one bit for bit."""
    p = hall_params(precision, mesh=(12, 10, 8))
    with env(RTB_WORKSPACE_MB="1"):
        ref = model.hall_conductivity(p).conductivity
        out = [None] * 4
        errs = []

        def run(i):
            try:
                for _ in range(3):
                    r = model.hall_conductivity(p).conductivity
                    assert np.array_equal(r, ref)
                out[i] = r
            except Exception as e:  # reported below
                errs.append(e)

        th = [threading.Thread(target=run, args=(i,)) for i in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    assert not errs, errs
    for r in out:
        assert np.array_equal(r, ref)
