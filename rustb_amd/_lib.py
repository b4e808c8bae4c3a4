"""ctypes binding of librtb_hip.so (the C ABI declared in include/rtb_hip.h).

The product path has no fallback: if the in-tree library is missing or fails
to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int64, c_size_t, c_uint64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librtb_hip.so")

RTB_OK = 0
ERRORS = {
    -1: "KVectorLengthMismatch",
    -2: "DimensionMismatch",
    -3: "InvalidKmeshDimensions",
    -4: "InvalidEnergyRange",
    -5: "InvalidDosParameter",
    -6: "InvalidResponseParameter",
    -7: "SpinNotAllowed",
    -8: "EigenvalueComputationFailed",
    -9: "OnsiteHoppingMustBeReal",
    -10: "RVectorLengthError",
    -11: "InvalidModelInvariant",
    -12: "InvalidArgument",
    -13: "Device",
    -14: "Unsupported",
    -15: "FileCreation",
    -16: "FileParse",
    -17: "InvalidOrbitalProjection",
    -18: "SpinMismatch",
    -19: "Io",
    -20: "NoOrbitals",
    -21: "Linalg",
    -22: "InvalidThermodynamicParameter",
    -23: "MissingHermitianConjugate",
    -24: "InvalidDimension",
}


class TbError(RuntimeError):
    """Error raised by the library; ``kind`` names the reference TbError variant."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = ERRORS.get(code, f"Unknown({code})")
        super().__init__(f"{self.kind}: {message}")


class SynthSpec(ctypes.Structure):
    _fields_ = [
        ("kind", c_int),
        ("dim", c_int),
        ("spin", c_int),
        ("norb", c_int),
        ("rmax", c_int),
        ("n_R_target", c_int),
        ("layers", c_int),
        ("scale", c_double),
        ("decay", c_double),
        ("seed", c_uint64),
    ]


class Stats(ctypes.Structure):
    _fields_ = [("ms", c_double * 10), ("launches", c_int * 10), ("flop", c_double * 10), ("nk", c_size_t),
                ("nsta", c_int), ("mflop", c_double * 10), ("max_residual", c_double), ("resolved", c_size_t)]


class BatchVariant(ctypes.Structure):
    _fields_ = [("nsta", c_int), ("fourier_tk", c_int), ("eigh_nb", c_int)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("n_models", c_int), ("n_passes", c_int), ("n_portions", c_int), ("n_tiles", c_int),
                ("n_queues", c_int), ("n_mesh_models", c_int), ("graphs_cached", c_int), ("captures", c_int),
                ("graph_nodes", c_int), ("nk_total", c_size_t), ("arena_bytes", c_size_t), ("graph_ms", c_double)]


_P = c_void_p
SIGNATURES = {
    "rtb_builder_new": (c_int, [c_int, c_int, c_int, _P, _P, POINTER(c_void_p)]),
    "rtb_builder_destroy": (None, [c_void_p]),
    "rtb_builder_set_hop": (c_int, [c_void_p, c_double, c_double, c_int, c_int, _P, c_int]),
    "rtb_builder_add_hop": (c_int, [c_void_p, c_double, c_double, c_int, c_int, _P, c_int]),
    "rtb_builder_add_element": (c_int, [c_void_p, c_double, c_double, c_int, c_int, _P]),
    "rtb_builder_set_onsite": (c_int, [c_void_p, _P, c_int]),
    "rtb_builder_add_onsite": (c_int, [c_void_p, _P, c_int]),
    "rtb_builder_n_R": (c_int, [c_void_p]),
    "rtb_builder_nsta": (c_int, [c_void_p]),
    "rtb_builder_export": (c_int, [c_void_p, _P, _P]),
    "rtb_synth_model": (c_int, [POINTER(SynthSpec), POINTER(c_int), POINTER(c_int), _P, _P, _P, _P]),
    "rtb_model_create": (c_int, [c_int, c_int, c_int, c_int, _P, _P, _P, _P, POINTER(c_void_p)]),
    "rtb_model_destroy": (None, [c_void_p]),
    "rtb_model_nsta": (c_int, [c_void_p]),
    "rtb_gen_ham_batch": (c_int, [c_void_p, _P, c_size_t, c_int, _P, c_int, _P, _P]),
    "rtb_solve_band_all": (c_int, [c_void_p, _P, c_size_t, c_int, _P]),
    "rtb_solve_all": (c_int, [c_void_p, _P, c_size_t, c_int, _P, _P]),
    "rtb_solve_band_range": (c_int, [c_void_p, _P, c_size_t, c_double, c_double, c_double, c_int, _P, _P]),
    "rtb_solve_range": (c_int, [c_void_p, _P, c_size_t, c_double, c_double, c_double, c_int, _P, _P, _P]),
    "rtb_berry_curvature": (c_int, [c_void_p, _P, c_size_t, _P, c_int, c_double, c_int, _P, _P]),
    "rtb_occupied_berry_curvature": (c_int, [c_void_p, _P, c_size_t, _P, c_int, c_double, c_double,
                                             c_double, c_int, _P]),
    "rtb_hall_direct": (c_int, [c_void_p, _P, _P, c_size_t, c_double, c_double, _P, c_int, c_int, _P]),
    "rtb_hall_direct_layers": (c_int, [c_void_p, _P, _P, c_size_t, c_double, c_double, _P, c_int, c_int, c_int, _P,
                                       _P, _P]),
    "rtb_hall_direct_noalloc": (c_int, [c_void_p, _P, _P, c_size_t, c_double, c_double, _P, c_int, c_int, _P]),
    "rtb_batch_create": (c_int, [_P, c_int, _P, _P, _P, c_int, POINTER(c_void_p)]),
    "rtb_batch_hall_direct": (c_int, [c_void_p, _P, c_size_t, c_double, c_double, c_int, _P]),
    "rtb_batch_set_graph": (c_int, [c_void_p, c_int]),
    "rtb_batch_info_get": (c_int, [c_void_p, POINTER(BatchInfo)]),
    "rtb_batch_destroy": (None, [c_void_p]),
    "rtb_hall_energy_cut": (c_int, [c_void_p, _P, _P, c_size_t, c_double, c_double, _P, c_int, _P]),
    "rtb_energy_cut_host": (c_int, [c_size_t, c_size_t, c_int, _P, _P, _P, _P, _P, c_size_t, c_double, c_double,
                                    c_double, _P]),
    "rtb_energy_cut_host_3d": (c_int, [c_size_t, c_size_t, c_size_t, c_int, _P, _P, _P, _P, _P, _P, c_size_t,
                                       c_double, c_double, c_double, _P]),
    "rtb_dos": (c_int, [c_void_p, _P, c_double, c_double, c_size_t, c_double, c_int, _P, _P]),
    "rtb_quantum_geometry": (c_int, [c_void_p, _P, c_size_t, _P, c_double, c_int, _P, _P, _P]),
    "rtb_quantum_geometry_direct": (c_int, [c_void_p, _P, _P, c_size_t, c_double, c_double, _P, c_int, _P, _P]),
    "rtb_nonlinear_hall_direct": (c_int, [c_void_p, _P, _P, c_size_t, c_double, c_double, _P, c_int, c_int, c_int,
                                          _P]),
    "rtb_w90_read": (c_int, [c_char_p, c_char_p, c_double, c_int, c_int, POINTER(c_void_p)]),
    "rtb_w90_read_with_rmatrix": (c_int, [c_char_p, c_char_p, c_double, c_int, c_int, POINTER(c_void_p)]),
    "rtb_w90_destroy": (None, [c_void_p]),
    "rtb_w90_export_rmatrix": (c_int, [c_void_p, _P]),
    "rtb_model_set_rmatrix": (c_int, [c_void_p, _P]),
    "rtb_w90_info": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_int),
                             POINTER(c_int)]),
    "rtb_w90_export": (c_int, [c_void_p, _P, _P, _P, _P, _P]),
    "rtb_w90_atom": (c_int, [c_void_p, c_int, POINTER(c_int), _P, POINTER(c_int), POINTER(c_int)]),
    "rtb_element_symbol": (c_char_p, [c_int]),
    "rtb_orbital_projection_name": (c_char_p, [c_int]),
    "rtb_w90_model": (c_int, [c_void_p, POINTER(c_void_p)]),
    "rtb_last_error": (c_char_p, []),
    "rtb_last_stats": (c_int, [POINTER(Stats)]),
    "rtb_workspace_release": (c_int, []),
    "rtb_workspace_bytes": (c_int, [POINTER(c_size_t), POINTER(c_int)]),
}

_lib = None


def load():
    """Load the in-tree library (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m rustb_amd.build` (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != RTB_OK:
        msg = load().rtb_last_error().decode(errors="replace")
        raise TbError(rc, msg)


def ptr(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


def last_stats() -> dict:
    s = Stats()
    check(load().rtb_last_stats(ctypes.byref(s)))
    return {"ms": list(s.ms), "launches": list(s.launches), "flop": list(s.flop), "nk": s.nk, "nsta": s.nsta,
            "mflop": list(s.mflop), "max_residual": s.max_residual, "resolved": s.resolved}


def workspace_bytes() -> tuple:
    """Device bytes held by the workspace slots of the current device, and
    their number."""
    b, n = c_size_t(), c_int()
    check(load().rtb_workspace_bytes(ctypes.byref(b), ctypes.byref(n)))
    return b.value, n.value


STAGES = ["fourier", "tridiag", "trieig", "backtransform", "gemm64", "gemm32", "refine", "kubo", "reduce", "total"]
