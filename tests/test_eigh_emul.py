"""Host-side check of the eigensolver's team-generic device source.

tests/native/team_emul.cpp runs rustb_amd/csrc/eigh_core.hpp and
eigh_reg.hpp (the exact code the GPU workgroups execute: register-resident
Householder reduction, bisection, twisted factorisations / inverse
iteration, back-transformation) with host threads and a barrier, in the
slice shapes of the GPU dispatch, and checks it against LAPACK zheev for
several sizes, padding (n < NB), both output layouts, and degenerate,
block-diagonal and clustered spectra (CPU only)."""
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

    def run(H, fp32=0, rows_out=0):
        n = H.shape[0]
        Hc = np.ascontiguousarray(H, np.complex128)
        E = np.zeros(n)
        Z = np.zeros((n, n), np.complex128)
        rc = lib.emul_eigh_reg(n, fp32, Hc.ctypes.data_as(ctypes.c_void_p), E.ctypes.data_as(ctypes.c_void_p),
                               Z.ctypes.data_as(ctypes.c_void_p), 1, rows_out)
        assert rc == 0
        return E, (Z.T if rows_out else Z)

    return run


def rand_herm(rng, n):
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return A + A.conj().T

# Register-resident route (tridiag_reg / backtrans_reg in eigh_reg.hpp): the
# slice shapes used on the GPU (32x2, 64x2, 128x2 FP32; 128x4 FP64) with
# padding (n < NB), both output layouts, and degenerate spectra.
@pytest.mark.parametrize("n", [1, 2, 3, 17, 32, 33, 64, 65, 100, 128])
def test_reg_fp64_matches_zheev(emul, n):
    rng = np.random.default_rng(n + 7)
    H = rand_herm(rng, n)
    w, _, _ = L.zheev(H)
    E, Z = emul(H, rows_out=n % 2)
    scale = max(1.0, np.abs(w).max())
    assert np.abs(E - w).max() <= 4e-15 * scale * n
    assert np.abs(H @ Z - Z * E).max() <= 4e-15 * scale * n
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= 1e-14 * n


@pytest.mark.parametrize("n", [5, 40, 128])
def test_reg_fp32_matches_zheev(emul, n):
    rng = np.random.default_rng(n + 70)
    H = rand_herm(rng, n)
    w, _, _ = L.zheev(H)
    E, Z = emul(H, fp32=1)
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
    E, Z = emul(H)
    n = H.shape[0]
    assert np.abs(E - w).max() < 1e-13
    assert np.abs(H @ Z - Z * E).max() < 1e-13
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() < 1e-13


def test_reg_exact_two_by_two(emul):
    E, _ = emul(np.array([[0, 3], [3, 0]], complex))
    assert np.array_equal(E, [-3.0, 3.0])


@pytest.mark.parametrize("case", ["near_degenerate", "clusters_and_isolated", "few_members"])
def test_reg_clustered_spectra(emul, case):
    """Clusters (inverse iteration + MGS) next to isolated eigenvalues (twisted):
    three clusters of 4, clusters of 5, 3 and 4 among 40 isolated levels, and
    a pair and a triple among 30."""
    rng = np.random.default_rng(15)
    if case == "near_degenerate":
        H = np.kron(rand_herm(rng, 3), np.eye(4)) + 1e-9 * np.diag(rng.standard_normal(12))
    else:
        if case == "few_members":
            w = np.concatenate([[1.0, 1.0 + 1e-8], np.linspace(-3, 3, 30), np.full(3, -2.5)])
        else:
            w = np.concatenate([np.full(5, 1.0), np.linspace(-3, 3, 40), np.full(3, -2.5), 2.0 + 1e-7 * np.arange(4)])
        Q, _ = np.linalg.qr(rng.standard_normal((w.size, w.size)) + 1j * rng.standard_normal((w.size, w.size)))
        H = (Q * w) @ Q.conj().T
        H = 0.5 * (H + H.conj().T)
    n = H.shape[0]
    w, _, _ = L.zheev(H)
    E, Z = emul(H)
    assert np.abs(E - w).max() < 1e-13
    assert np.abs(H @ Z - Z * E).max() < 1e-13
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() < 1e-13


@pytest.mark.parametrize("n,fp32", [(9, 0), (64, 0), (128, 0), (100, 1)])
def test_reg_reads_lower_triangle_and_real_diagonal(emul, n, fp32):
    """An input Hermitian only to 1e-4 (imaginary diagonal included): the
    result is that of the Hermitian matrix of H's lower triangle with Re of
    its diagonal, which is what zheev('L') diagonalises."""
    rng = np.random.default_rng(n + 3)
    A = rand_herm(rng, n)
    H = A + 1e-4 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    low = np.tril(H, -1) + np.tril(H, -1).conj().T + np.diag(H.diagonal().real)
    w, _, _ = L.zheev(low)
    E, Z = emul(H, fp32=fp32)
    scale = np.abs(w).max()
    tol = 4e-15 * scale * n if not fp32 else 5e-6 * scale
    assert np.abs(E - w).max() <= tol
    assert np.abs(low @ Z - Z * E).max() <= tol


@pytest.mark.parametrize("san", [["-fsanitize=address,undefined", "-fno-sanitize-recover=all"],
                                 ["-fsanitize=thread"]], ids=["asan_ubsan", "tsan"])
def test_emulation_under_sanitizers(tmp_path, san):
    """The eigensolver workgroups and the cluster Rayleigh-Ritz (the device
    source run by host threads) under AddressSanitizer + UBSan and under
    ThreadSanitizer: tests/native/sanitize_main.cpp checks residuals and
    orthonormality itself; any sanitizer report fails the run."""
    exe = str(tmp_path / "sanitize_main")
    src = os.path.join(HERE, "native", "sanitize_main.cpp")
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", *san, src, "-o", exe, "-lpthread"],
                   check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", TSAN_OPTIONS="halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "FAIL" not in r.stdout and "Sanitizer" not in r.stderr, r.stderr[-3000:]


@pytest.fixture(scope="module")
def band_emul(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("band") / "libteam_emul.so")
    subprocess.run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", os.path.join(HERE, "native", "team_emul.cpp"),
                    "-o", out], check=True)
    lib = ctypes.CDLL(out)

    def run(H, fp32=0, rows_out=0):
        n = H.shape[0]
        Hc = np.ascontiguousarray(H, np.complex128)
        E = np.zeros(n)
        Z = np.zeros((n, n), np.complex128)
        rc = lib.emul_eigh_band(n, fp32, Hc.ctypes.data_as(ctypes.c_void_p), E.ctypes.data_as(ctypes.c_void_p),
                                Z.ctypes.data_as(ctypes.c_void_p), 1, rows_out)
        assert rc == 0
        return E, (Z.T if rows_out else Z)

    return run


# Two-stage route (eigh_band.hpp: dense -> band panels, bulge chasing, explicit
# tridiagonal vectors, back-transformation), which the GPU takes for
# 128 < n <= 256.  The routines are generic in n, so the host runs them at
# sizes where thread-barrier emulation is quick: one panel with a single row
# below it (17), two panels (33), several panels and partial windows (50),
# both output layouts and both precisions.
@pytest.mark.parametrize("n,fp32", [(17, 0), (33, 0), (50, 0), (50, 1), (41, 1)])
def test_band_route_matches_zheev(band_emul, n, fp32):
    rng = np.random.default_rng(n + 100 * fp32)
    H = rand_herm(rng, n)
    w, _, _ = L.zheev(H)
    E, Z = band_emul(H, fp32=fp32, rows_out=n % 2)
    scale = np.abs(w).max()
    tol = 4e-15 * scale * n if not fp32 else 5e-6 * scale
    assert np.abs(E - w).max() <= tol
    assert np.abs(H @ Z - Z * E).max() <= tol
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= (1e-14 * n if not fp32 else 1e-4)


@pytest.mark.parametrize("fp32", [0, 1])
def test_band_route_special_spectra(band_emul, fp32):
    """Degenerate, block-diagonal (zero band couplings: reflectors with tau = 0)
    and lower-triangle-only input (the imaginary diagonal and the upper
    triangle ignored, as zheev('L') on a row-major buffer); FP32 runs the
    chase with the band region in team memory."""
    rng = np.random.default_rng(9)
    cases = [np.diag([1.0] * 20 + [2.0] * 20).astype(complex), np.kron(np.eye(5), rand_herm(rng, 8))]
    A = rand_herm(rng, 45)
    cases.append(A + 1e-3 * (rng.standard_normal((45, 45)) + 1j * rng.standard_normal((45, 45))))
    for H in cases:
        n = H.shape[0]
        low = np.tril(H, -1) + np.tril(H, -1).conj().T + np.diag(H.diagonal().real)
        w, _, _ = L.zheev(low)
        E, Z = band_emul(H, fp32=fp32)
        scale = max(1.0, np.abs(w).max())
        tol, otol = (1e-13 * scale, 1e-13) if not fp32 else (5e-6 * scale, 1e-4)
        assert np.abs(E - w).max() <= tol
        assert np.abs(low @ Z - Z * E).max() <= tol
        assert np.abs(Z.conj().T @ Z - np.eye(n)).max() <= otol
