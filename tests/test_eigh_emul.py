"""Host-side check of the eigensolver's team-generic device source.

tests/native/team_emul.cpp runs rustb_amd/csrc/eigh_core.hpp (the exact code
the GPU workgroups execute) with `nt` host threads and a barrier; this checks
it against LAPACK zheev for several sizes and team sizes, including
degenerate and block-diagonal spectra (CPU only)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg.lapack as L

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emul") / "libteam_emul.so")
    subprocess.run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", os.path.join(HERE, "native", "team_emul.cpp"),
                    "-o", out], check=True)
    lib = ctypes.CDLL(out)

    def run(H, nt, fp32=0, vectors=1):
        n = H.shape[0]
        Hc = np.ascontiguousarray(H, np.complex128)
        E = np.zeros(n)
        Z = np.zeros((n, n), np.complex128)
        lib.emul_eigh(n, nt, fp32, Hc.ctypes.data_as(ctypes.c_void_p), E.ctypes.data_as(ctypes.c_void_p),
                      Z.ctypes.data_as(ctypes.c_void_p), vectors)
        return E, Z

    def run_reg(H, fp32=0, rows_out=0):
        n = H.shape[0]
        Hc = np.ascontiguousarray(H, np.complex128)
        E = np.zeros(n)
        Z = np.zeros((n, n), np.complex128)
        rc = lib.emul_eigh_reg(n, fp32, Hc.ctypes.data_as(ctypes.c_void_p), E.ctypes.data_as(ctypes.c_void_p),
                               Z.ctypes.data_as(ctypes.c_void_p), 1, rows_out)
        assert rc == 0
        return E, (Z.T if rows_out else Z)

    run.reg = run_reg
    return run


def rand_herm(rng, n):
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return A + A.conj().T


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 33, 64])
@pytest.mark.parametrize("nt", [1, 4, 13, 64])
def test_fp64_matches_zheev(emul, n, nt):
    rng = np.random.default_rng(n * 100 + nt)
    H = rand_herm(rng, n)
    w, _, _ = L.zheev(H)
    E, Z = emul(H, nt)
    scale = max(1.0, np.abs(w).max())
    assert np.abs(E - w).max() <= 4e-15 * scale * n
    assert np.abs(H @ Z - Z * E).max() <= 4e-15 * scale * n
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= 1e-14 * n


def test_exact_two_by_two(emul):
    E, _ = emul(np.array([[0, 3], [3, 0]], complex), 4)
    assert np.array_equal(E, [-3.0, 3.0])


@pytest.mark.parametrize("case", ["degenerate", "block", "zero", "near_degenerate"])
def test_degenerate_spectra(emul, case):
    rng = np.random.default_rng(5)
    if case == "degenerate":
        H = np.diag([1.0] * 4 + [2.0] * 4 + [3.0] * 4).astype(complex)
    elif case == "block":
        H = np.kron(np.eye(3), rand_herm(rng, 4))
    elif case == "zero":
        H = np.zeros((9, 9), complex)
    else:
        H = np.kron(rand_herm(rng, 3), np.eye(4)) + 1e-9 * np.diag(rng.standard_normal(12))
    w, _, _ = L.zheev(H)
    E, Z = emul(H, 5)
    n = H.shape[0]
    assert np.abs(E - w).max() < 1e-13
    assert np.abs(H @ Z - Z * E).max() < 1e-13
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() < 1e-13


def test_fp32_variant(emul):
    rng = np.random.default_rng(11)
    H = rand_herm(rng, 32)
    w, _, _ = L.zheev(H)
    E, Z = emul(H, 8, fp32=1)
    assert np.abs(E - w).max() < 5e-5 * np.abs(w).max()
    assert np.abs(Z.conj().T @ Z - np.eye(32)).max() < 1e-4


# Register-resident route (tridiag_reg / backtrans_reg in eigh_reg.hpp): the
# slice shapes used on the GPU (32x2, 64x2, 128x2 FP32; 128x4 FP64) with
# padding (n < NB), both output layouts, and degenerate spectra.
@pytest.mark.parametrize("n", [1, 2, 3, 17, 32, 33, 64, 65, 100, 128])
def test_reg_fp64_matches_zheev(emul, n):
    rng = np.random.default_rng(n + 7)
    H = rand_herm(rng, n)
    w, _, _ = L.zheev(H)
    E, Z = emul.reg(H, rows_out=n % 2)
    scale = max(1.0, np.abs(w).max())
    assert np.abs(E - w).max() <= 4e-15 * scale * n
    assert np.abs(H @ Z - Z * E).max() <= 4e-15 * scale * n
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= 1e-14 * n


@pytest.mark.parametrize("n", [5, 40, 128])
def test_reg_fp32_matches_zheev(emul, n):
    rng = np.random.default_rng(n + 70)
    H = rand_herm(rng, n)
    w, _, _ = L.zheev(H)
    E, Z = emul.reg(H, fp32=1)
    scale = np.abs(w).max()
    assert np.abs(E - w).max() < 5e-6 * scale
    assert np.abs(H @ Z - Z * E).max() < 5e-6 * scale
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() < 1e-4


@pytest.mark.parametrize("case", ["degenerate", "block", "zero", "diagonal"])
def test_reg_special_spectra(emul, case):
    rng = np.random.default_rng(6)
    if case == "degenerate":
        H = np.diag([1.0] * 30 + [2.0] * 30 + [3.0] * 10).astype(complex)
    elif case == "block":
        H = np.kron(np.eye(9), rand_herm(rng, 8))
    elif case == "zero":
        H = np.zeros((70, 70), complex)
    else:
        H = np.diag(rng.standard_normal(66)).astype(complex)
    w, _, _ = L.zheev(H)
    E, Z = emul.reg(H)
    n = H.shape[0]
    assert np.abs(E - w).max() < 1e-13
    assert np.abs(H @ Z - Z * E).max() < 1e-13
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() < 1e-13
