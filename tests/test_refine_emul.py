"""Host-side check of the FP64 cluster Rayleigh-Ritz used by the mixed form
(rustb_amd/csrc/refine_core.hpp, the exact per-k routine of the GPU kernel):
FP32-accurate eigenvectors of a matrix with near-degenerate clusters are
refined to FP64 eigenvalues and orthonormal vectors (CPU only)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg.lapack as L

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def rr(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("rr") / "librefine_emul.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", os.path.join(HERE, "native", "refine_emul.cpp"),
                    "-o", out], check=True)
    lib = ctypes.CDLL(out)

    def run(Z, HZ, E, gap_rel):
        Z = np.ascontiguousarray(Z, np.complex128)
        E = np.ascontiguousarray(E, np.float64)
        HZ = np.ascontiguousarray(HZ, np.complex128)
        lib.emul_cluster_rr(Z.shape[0], Z.ctypes.data_as(ctypes.c_void_p), HZ.ctypes.data_as(ctypes.c_void_p),
                            E.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(gap_rel))
        return Z, E

    return run


def fp32_like_pairs(rng, n, eig):
    """Hermitian H with spectrum `eig`; eigenvectors perturbed at FP32 level
    inside near-degenerate groups (mixing), then column-normalised; E the
    Rayleigh quotients (as the Rayleigh pass leaves them)."""
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    H = (Q * eig) @ Q.conj().T
    H = 0.5 * (H + H.conj().T)
    Z = Q + 3e-7 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    # mix close pairs strongly, as an FP32 solver does inside a cluster
    for j in range(n - 1):
        if eig[j + 1] - eig[j] < 1e-4:
            a = 0.6
            zj, zk = Z[:, j].copy(), Z[:, j + 1].copy()
            Z[:, j] = np.cos(a) * zj + np.sin(a) * zk
            Z[:, j + 1] = -np.sin(a) * zj + np.cos(a) * zk
    Z /= np.linalg.norm(Z, axis=0)
    HZ = H @ Z
    E = np.real(np.sum(Z.conj() * HZ, axis=0))
    order = np.argsort(E)
    return H, Z[:, order], HZ[:, order], E[order]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_cluster_rayleigh_ritz(rr, seed):
    rng = np.random.default_rng(seed)
    n = 40
    eig = np.sort(rng.uniform(-10, 10, n))
    eig[10] = eig[9] + 1e-7       # near-degenerate pair
    eig[20:23] = eig[20] + np.array([0.0, 3e-6, 8e-6])  # triple cluster
    eig[30] = eig[29]             # exact degeneracy
    eig = np.sort(eig)
    H, Z, HZ, E = fp32_like_pairs(rng, n, eig)
    w, _, _ = L.zheev(H)
    before = np.abs(E - w).max()
    Zr, Er = rr(Z, HZ, E, 1e-3)
    assert np.abs(Er - w).max() < 1e-11 * np.abs(w).max()
    assert before > 1e-8  # the mixing was visible before the refinement
    # columns inside each cluster orthonormal to FP64 precision; the rest keep
    # their FP32-level errors
    G = Zr.conj().T @ Zr
    thr = 1e-3 * max(1.0, np.abs(E).max())
    close = np.where(np.diff(E) <= thr)[0]
    assert len(close) >= 4
    for i in close:  # members of a cluster are orthonormal to FP64 precision
        assert np.abs(G[i:i + 2, i:i + 2] - np.eye(2)).max() < 1e-12
    assert np.abs(G - np.eye(n)).max() < 1e-5
    res = H @ Zr - Zr * Er
    assert np.abs(res).max() < 1e-4  # the injected FP32-level noise times ||H||


def test_no_cluster_is_identity(rr):
    rng = np.random.default_rng(3)
    n = 12
    eig = np.linspace(-5, 5, n)
    H, Z, HZ, E = fp32_like_pairs(rng, n, eig)
    Zr, Er = rr(Z.copy(), HZ, E.copy(), 1e-3)
    assert np.array_equal(Er, E) and np.array_equal(Zr, Z)
