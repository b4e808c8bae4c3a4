"""k-space sampling helpers (host side).

gen_kmesh follows kpoints.rs:49-87 (first axis outermost, k_d = i_d / N_d);
k_path follows kpath.rs:17-70 including its truncating node-index cast.
The mesh entry points of the library generate the same points on the device.
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import TbError


def gen_kmesh(k_mesh):
    k_mesh = [int(x) for x in k_mesh]
    if len(k_mesh) == 0 or any(n <= 0 for n in k_mesh):  # kpoints.rs:39-47
        raise TbError(-3, f"invalid k-mesh {k_mesh}")
    axes = [np.arange(n, dtype=np.float64) / float(n) for n in k_mesh]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.reshape(-1) for g in grids], axis=1)


def k_path(lat, path, nk: int):
    lat = np.asarray(lat, dtype=np.float64)
    path = np.asarray(path, dtype=np.float64)
    if path.shape[1] != lat.shape[0]:
        raise TbError(-2, "path length mismatch")
    n_node = path.shape[0]
    k_metric = np.linalg.inv(lat @ lat.T)
    k_node = np.zeros(n_node)
    for n in range(1, n_node):
        dk = path[n] - path[n - 1]
        k_node[n] = k_node[n - 1] + math.sqrt(dk @ (k_metric @ dk))
    node_index = [0]
    for n in range(1, n_node - 1):
        node_index.append(int(k_node[n] / k_node[-1] * float(nk - 1)))
    node_index.append(nk - 1)
    k_dist = np.zeros(nk)
    k_vec = np.zeros((nk, path.shape[1]))
    k_vec[0] = path[0]
    for n in range(1, n_node):
        n_i, n_f = node_index[n - 1], node_index[n]
        j = np.arange(n_i, n_f + 1)
        frac = (j - n_i) / (n_f - n_i)
        k_dist[n_i:n_f + 1] = k_node[n - 1] + frac * (k_node[n] - k_node[n - 1])
        k_vec[n_i:n_f + 1] = (1.0 - frac)[:, None] * path[n - 1] + frac[:, None] * path[n]
    return k_vec, k_dist, k_node
