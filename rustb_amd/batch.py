"""Many models in one call: the reference's per-model Direct Hall loop
(`Model::hall_conductivity` once per model, response/linear/mod.rs:160-225)
as one planned device job (rtb_batch_* in include/rtb_hip.h): passes over the
models' concatenated k-points, work-queue Fourier launches across models,
eigensolve and Kubo band sum per matrix size, captured in a HIP graph per
parameter set and replayed.

The per-size kernel variants come from ``tune_table.json`` beside this file,
written by ``python -m rustb_amd.autotune`` (a GPU run).
"""
from __future__ import annotations

import ctypes
import json
import math
import os

import numpy as np

from ._lib import BatchInfo, BatchVariant, TbError, check, load, ptr
from .model import PRECISION

HERE = os.path.dirname(os.path.abspath(__file__))
TUNE_TABLE = os.path.join(HERE, "tune_table.json")


def load_table(path: str = TUNE_TABLE) -> list:
    """The variant table as [{"nsta", "fourier_tk", "eigh_nb"}, ...] ([] when
    the file is absent)."""
    if not os.path.exists(path):
        return []
    with open(path) as f:
        data = json.load(f)
    return [{"nsta": int(e["nsta"]), "fourier_tk": int(e["fourier_tk"]), "eigh_nb": int(e["eigh_nb"])}
            for e in data["table"]]


class ModelBatch:
    """models: Model objects (kept alive by the batch); kmeshes: one mesh per
    model (its dim entries); directions: per model the (2, dim) rows of the
    conductivity sigma_{ab} (default: the first two Cartesian axes); table:
    "default" (tune_table.json), None (built-in variants) or a list of
    {"nsta", "fourier_tk", "eigh_nb"} entries."""

    def __init__(self, models, kmeshes, directions=None, table="default"):
        self.models = list(models)
        nm = len(self.models)
        if nm == 0:
            raise TbError(-12, "a batch needs at least one model")
        if len(kmeshes) != nm:
            raise TbError(-2, "one k-mesh per model")
        km = np.ones((nm, 3), dtype=np.uint64)
        dirs = np.zeros((nm, 2, 3))
        for i, m in enumerate(self.models):
            mesh = np.atleast_1d(np.asarray(kmeshes[i], dtype=np.int64))
            if mesh.shape != (m.dim,):
                raise TbError(-3, f"model {i}: the k-mesh needs {m.dim} entries")
            if np.any(mesh <= 0):
                raise TbError(-3, f"model {i}: k-mesh entries must be positive")
            km[i, :m.dim] = mesh
            d = np.eye(max(m.dim, 2))[:2, :m.dim] if directions is None else np.asarray(directions[i], float)
            if d.shape != (2, m.dim):
                raise TbError(-2, f"model {i}: directions must have shape (2, {m.dim})")
            dirs[i, :, :m.dim] = d
        self.kmeshes = km
        self.nk = km.prod(axis=1).astype(np.int64)
        entries = load_table() if table == "default" else ([] if table is None else list(table))
        tab = (BatchVariant * max(len(entries), 1))()
        for j, e in enumerate(entries):
            tab[j] = BatchVariant(int(e["nsta"]), int(e["fourier_tk"]), int(e["eigh_nb"]))
        self.table = entries
        handles = (ctypes.c_void_p * nm)(*[m._device().value for m in self.models])
        h = ctypes.c_void_p()
        self._km = np.ascontiguousarray(km)
        self._dirs = np.ascontiguousarray(dirs)
        check(load().rtb_batch_create(ctypes.cast(handles, ctypes.c_void_p), nm, ptr(self._km), ptr(self._dirs),
                                      ctypes.cast(tab, ctypes.c_void_p) if entries else None, len(entries),
                                      ctypes.byref(h)))
        self._h = h

    def hall_conductivity(self, mu, T: float = 0.0, eta: float = 1e-3, precision: str = "fp64") -> np.ndarray:
        """sigma [n_models, n_mu]: each model's Direct AHC (sum / nk / det)."""
        if self._h is None:
            raise TbError(-12, "the batch is closed")
        mu = np.ascontiguousarray(np.atleast_1d(np.asarray(mu, np.float64)))
        out = np.zeros((len(self.models), mu.size))
        check(load().rtb_batch_hall_direct(self._h, ptr(mu), mu.size, float(T), float(eta), PRECISION[precision],
                                           ptr(out)))
        return out

    def chern_numbers(self, mu=0.0, T: float = 0.0, eta: float = 1e-3, precision: str = "fp64") -> np.ndarray:
        """round(2 pi sigma) per model at each mu (2D models: the integrated
        Berry curvature of the occupied bands over 2 pi)."""
        return np.rint(2.0 * math.pi * self.hall_conductivity(mu, T, eta, precision)).astype(np.int64)

    def set_graph(self, use_graph: bool) -> None:
        """True (default): capture once per parameter set and replay; False:
        launch directly each call, with per-stage times in last_stats()."""
        check(load().rtb_batch_set_graph(self._h, 1 if use_graph else 0))

    def info(self) -> dict:
        i = BatchInfo()
        check(load().rtb_batch_info_get(self._h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in BatchInfo._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            load().rtb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
