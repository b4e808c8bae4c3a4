"""rustb_amd — MI355X (gfx950) k-mesh hot path of the Rustb tight-binding crate.

The compute path is librtb_hip.so (hand-written HIP kernels behind a C ABI,
include/rtb_hip.h); this package is the host-side mirror of the reference's
`Model` surface for that path.
"""
from ._lib import TbError, last_stats, load
from .kpoints import gen_kmesh, k_path
from .model import (Atom, BandBerryCurvature, BandQuantumGeometry, HallConductivityResult, Model,
                    NonlinearHallResult, Parameters, QuantumGeometryMap, QuantumGeometryResult, synth_model)
from .batch import ModelBatch
from . import workloads

__all__ = [
    "Model",
    "ModelBatch",
    "Atom",
    "NonlinearHallResult",
    "BandQuantumGeometry",
    "QuantumGeometryMap",
    "QuantumGeometryResult",
    "Parameters",
    "BandBerryCurvature",
    "HallConductivityResult",
    "TbError",
    "gen_kmesh",
    "k_path",
    "synth_model",
    "last_stats",
    "load",
    "workloads",
]
