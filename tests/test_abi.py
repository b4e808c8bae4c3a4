"""The C-ABI library loads without a GPU and exports every symbol declared in
include/rtb_hip.h (CPU only: no compute call)."""
import ctypes
import os
import re

from rustb_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(?:int|void|const char\*)\s+(rtb_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_hot_path():
    names = declared_functions()
    for must in ["rtb_model_create", "rtb_gen_ham_batch", "rtb_solve_band_all", "rtb_solve_all",
                 "rtb_berry_curvature", "rtb_hall_direct", "rtb_dos", "rtb_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers the whole header
    assert declared_functions() <= set(_lib.SIGNATURES)


def test_non_finite_hoppings_rejected_before_any_device_work():
    """rtb_model_create checks the hoppings first (no device is touched when
    they are rejected, so this runs without a GPU)."""
    import numpy as np

    import rustb_amd as ra

    for bad in (float("nan"), float("inf")):
        m = ra.Model.from_arrays(np.eye(2), [[0.0, 0.0], [0.5, 0.5]], False, np.array([[0, 0], [1, 0], [-1, 0]]),
                                 np.zeros((3, 2, 2), complex))
        h = m.ham.copy()
        h[1, 0, 1] = complex(bad, 0.0)
        m2 = ra.Model.from_arrays(m.lat, m.orb, False, m.hamR, h)
        try:
            m2._device()
        except ra.TbError as e:
            assert e.code == -11 and "finite" in str(e)  # RTB_ERR_INVALID_MODEL
        else:
            raise AssertionError("non-finite hopping accepted")
