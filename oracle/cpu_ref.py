"""Loader for oracle/build/libcpu_ref.so, the reference-equivalent CPU path
(OpenMP over k + host LAPACK).  TEST AND CPU-BASELINE INFRASTRUCTURE ONLY: used by
tests/ and by bench.py's cpu_baseline leg, never by the product."""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libcpu_ref.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lapack_candidates():
    # MKL (sequential layer) first, the reference's default backend family
    # (Cargo.toml intel-mkl-* features); then scipy's bundled OpenBLAS.
    yield "/opt/conda/lib/libmkl_rt.so", ""
    for p in sorted(glob.glob("/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas-*.so")):
        yield p, "scipy_"
    try:
        import scipy

        for p in sorted(glob.glob(os.path.join(os.path.dirname(scipy.__file__) + ".libs", "libscipy_openblas-*.so"))):
            yield p, "scipy_"
    except Exception:
        pass


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.cpu_ref_init.restype = ctypes.c_int
        lib.cpu_ref_init.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        lib.cpu_ref_lib.restype = ctypes.c_char_p
        lib.cpu_ref_run.restype = ctypes.c_int
        lib.cpu_ref_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6 + [
            ctypes.c_long, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        ok = False
        for path, prefix in _lapack_candidates():
            if os.path.exists(path) and lib.cpu_ref_init(path.encode(), prefix.encode()) == 0:
                ok = True
                break
        if not ok:
            raise RuntimeError("no host LAPACK found for the CPU baseline")
        _lib = lib
    return _lib


def lapack_name() -> str:
    return os.path.basename(load().cpu_ref_lib().decode())


def _tables(lat, orb, spin, hamR, directions):
    lat = np.asarray(lat, float)
    dim = lat.shape[0]
    orb_sta = np.concatenate([orb, orb]) if spin else np.asarray(orb, float)
    R3 = np.zeros((hamR.shape[0], 3))
    R3[:, :dim] = hamR
    tau3 = np.zeros((orb_sta.shape[0], 3))
    tau3[:, :dim] = orb_sta
    d = np.atleast_2d(np.asarray(directions, float)) if directions is not None else np.zeros((2, dim))
    wR = np.ascontiguousarray(d @ (hamR.astype(float) @ lat).T)
    wtau = np.ascontiguousarray(d @ (orb_sta @ lat).T)
    return np.ascontiguousarray(R3), np.ascontiguousarray(tau3), wR, wtau


def run(mode, lat, orb, spin, hamR, ham, kvec, directions=None, eta=1e-3, threads=None):
    """mode 0: E [nk, n]; mode 2: (Omega [nk, n], E [nk, n])."""
    lib = load()
    ham = np.ascontiguousarray(ham, np.complex128)
    n = ham.shape[1]
    R3, tau3, wR, wtau = _tables(lat, orb, spin, np.asarray(hamR), directions)
    kvec = np.atleast_2d(np.asarray(kvec, float))
    k3 = np.zeros((kvec.shape[0], 3))
    k3[:, :kvec.shape[1]] = kvec
    nk = k3.shape[0]
    om = np.zeros((nk, n))
    E = np.zeros((nk, n))
    threads = threads or os.cpu_count()
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = lib.cpu_ref_run(mode, n, ham.shape[0], P(ham), P(R3), P(wR), P(tau3), P(wtau), P(k3), nk, float(eta),
                         int(threads), P(om), P(E))
    if rc != 0:
        raise RuntimeError(f"cpu_ref_run failed: {rc}")
    return (om, E) if mode == 2 else E
