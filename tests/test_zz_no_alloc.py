"""The allocation-free entry point rtb_hall_direct_noalloc through the C ABI,
against rtb_hall_direct on the same inputs: two calls in a row (the second
with a smaller mesh, fewer chemical potentials and a smaller model), both
precision forms, and a request too large for its fixed scratch.

Bars: FP64 1e-10 relative (the two calls sum the same terms in different
chunkings); mixed 1e-8 relative (the regular call takes the separable mesh
Fourier sum, the allocation-free one the dense sum)."""
import ctypes

import numpy as np
import pytest

from rustb_amd import _lib
from tests.test_gpu_parity import random_pair

pytestmark = pytest.mark.gpu

P = _lib.ptr


def hall(fn, model, mesh, mu, dirs, prec, spin=0, eta=0.02, T=0.0):
    km = np.asarray(mesh, dtype=np.uint64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    d = np.ascontiguousarray(dirs, dtype=np.float64)
    sig = np.zeros(mu.size)
    rc = fn(model._device(), P(km), P(mu), mu.size, T, eta, P(d), spin, prec, P(sig))
    return rc, sig


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_hall_direct_twice(prec):
    lib = _lib.load()
    big, _ = random_pair(8, True, 3, 5, seed=41)
    small, _ = random_pair(3, False, 3, 3, seed=42)
    dirs = np.array([[1.0, 0.0, 0.3], [0.0, 1.0, 0.0]])
    tol = 1e-10 if prec == 0 else 1e-8
    cases = [(big, (14, 12, 10), np.linspace(-1.0, 1.0, 7), 3),  # spin-z current on the spinful model
             (small, (5, 4, 3), np.array([0.1, -0.2]), 0)]
    for model, mesh, mu, spin in cases:
        rc, ref = hall(lib.rtb_hall_direct, model, mesh, mu, dirs, prec, spin)
        assert rc == 0, _lib.load().rtb_last_error()
        rc, out = hall(lib.rtb_hall_direct_noalloc, model, mesh, mu, dirs, prec, spin)
        assert rc == 0, _lib.load().rtb_last_error()
        assert np.all(np.isfinite(out))
        assert np.abs(out - ref).max() <= tol * max(np.abs(ref).max(), 1e-3), (mesh, out, ref)


def test_noalloc_rejects_what_does_not_fit():
    lib = _lib.load()
    model, _ = random_pair(2, False, 2, 2, seed=43)
    mu = np.zeros(12_000_000)  # three n_mu-long device arrays: 288 MB > 256 MiB
    rc, _ = hall(lib.rtb_hall_direct_noalloc, model, (4, 4), mu, np.eye(2), 0)
    assert rc == -14  # RTB_ERR_UNSUPPORTED
    msg = lib.rtb_last_error().decode()
    assert "allocation-free" in msg and "bytes" in msg
    # and the entry point works again right after
    rc, out = hall(lib.rtb_hall_direct_noalloc, model, (4, 4), np.array([0.0]), np.eye(2), 0)
    rc2, ref = hall(lib.rtb_hall_direct, model, (4, 4), np.array([0.0]), np.eye(2), 0)
    assert rc == 0 and rc2 == 0
    assert abs(out[0] - ref[0]) <= 1e-10 * max(abs(ref[0]), 1e-3)
