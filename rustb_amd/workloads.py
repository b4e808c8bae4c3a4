"""The benchmark workloads: seeded synthetic models and their k-meshes.

1 graphene DOS: spinless graphene, lattice rows (sqrt3,-1),(sqrt3,1), orbitals
  (0,0),(1/3,1/3), t=-2.85 at R=(0,0),(-1,0),(0,-1); 101x101 mesh; DOS on
  [-4,4] eV, 801 points, sigma=0.02.
2 random3d-n32: spinful 3D, 16 orbitals (n=32), |R_i|<=2 (125 R), amplitudes
  0.5 exp(-|R|) CN(0,1), onsite U[-1,1]; 64^3 mesh; eigenvalues+eigenvectors.
3 wannier3d-n128: spinful 3D, 64 orbitals (n=128), |R_i|<=5 (1331 R),
  amplitudes exp(-|R|/1.5) CN(0,1); 48^3 mesh; Omega_z and AHC at mu=0, T=0,
  eta=1e-4.
4 batch2d: 64 2D models, norb in {4..96}, N_R in [50,500] odd from |R_i|<=11,
  amplitudes exp(-|R|) CN(0,1); mesh n x n, n in [16,128]; Omega_z, Chern.
5 slab2d-n256: 32 layers x 8 orbitals, |R_i|<=3 (49 R); 200x200 mesh.
Generators are SplitMix64-seeded xoshiro256** (rtb_synth_model).
"""
from __future__ import annotations

import math

import numpy as np

from .model import Model, synth_model

SQ3 = math.sqrt(3.0)


def graphene() -> Model:
    m = Model.tb_model([[SQ3, -1.0], [SQ3, 1.0]], [[0.0, 0.0], [1.0 / 3.0, 1.0 / 3.0]], False)
    for R in ([0, 0], [-1, 0], [0, -1]):
        m.add_hop(-2.85, 0, 1, R)
    return m


def random3d_n32(seed: int = 2) -> Model:
    return synth_model(1, 3, True, 16, 2, seed, scale=0.5, decay=1.0)


def wannier3d_n128(seed: int = 3) -> Model:
    return synth_model(2, 3, True, 64, 5, seed, scale=1.0, decay=1.5)


def batch2d_model(i: int):
    """Model i of the 64-model batch and its mesh size (seed 4000 + i)."""
    rng = np.random.default_rng(4000 + i)
    norb = int(rng.choice(np.arange(4, 97, 4)))
    n_R = int(rng.integers(25, 251)) * 2 + 1  # odd in [51, 501]
    n_R = min(n_R, 529)
    nk = int(rng.integers(16, 129))
    m = synth_model(3, 2, False, norb, 11, 4000 + i, scale=1.0, decay=1.0, n_R_target=n_R)
    return m, nk


def slab2d_n256(seed: int = 5) -> Model:
    return synth_model(4, 2, False, 8, 3, seed, scale=1.0, decay=1.0, layers=32)


CONFIGS = {
    1: dict(name="graphene-dos", kmesh=(101, 101)),
    2: dict(name="random3d-n32", kmesh=(64, 64, 64)),
    3: dict(name="wannier3d-n128", kmesh=(48, 48, 48), eta=1e-4),
    4: dict(name="batch2d-64models"),
    5: dict(name="slab2d-n256", kmesh=(200, 200), eta=1e-4),
}
