"""Per-size kernel variants of the multi-model batch, measured on the GPU.

For each matrix size n of the 2D batch workload (spinless, norb 4..96 in
steps of 4) a seeded model of that size (|R_i| <= 11, 275 stored R, the
workload's amplitudes) on an nk x nk mesh runs as a one-model batch under
every candidate variant: Fourier tile depth fourier_tk in {1, 2, 4, 8}
(k-points per thread of the work-queue Fourier sum) and register-eigensolver
block rows eigh_nb in {32, 64, 128} (at least n).  Each candidate's graph
replay is timed (median of `reps` replays, device time from the graph's
event nodes) in FP64 and in the mixed form; the winner minimises the sum.
Writes the table and every candidate's times as JSON (default: tune_table.json
beside this file, which ModelBatch reads).

    python -m rustb_amd.autotune [--nk 40] [--reps 3] [--out PATH]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys

import numpy as np

from . import _lib
from .batch import TUNE_TABLE, ModelBatch
from .model import synth_model

TKS = (1, 2, 4, 8)


def candidates(n: int):
    nbs = [nb for nb in (32, 64, 128) if nb >= n] if n <= 128 else [0]
    return [(tk, nb) for tk in TKS for nb in nbs]


def time_variant(model, nk: int, tk: int, nb: int, reps: int) -> dict:
    b = ModelBatch([model], [(nk, nk)], table=[{"nsta": model.nsta, "fourier_tk": tk, "eigh_nb": nb}])
    out = {}
    mu = np.array([0.0])
    for precision in ("fp64", "mixed"):
        b.hall_conductivity(mu, precision=precision)  # plan + capture
        ms = []
        for _ in range(reps):
            b.hall_conductivity(mu, precision=precision)
            ms.append(_lib.last_stats()["ms"][-1])
        out[precision] = statistics.median(ms)
    b.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nk", type=int, default=40)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=TUNE_TABLE)
    ap.add_argument("--sizes", default=",".join(str(n) for n in range(4, 97, 4)))
    a = ap.parse_args(argv)
    sizes = [int(s) for s in a.sizes.split(",")]
    table, detail = [], []
    for n in sizes:
        m = synth_model(3, 2, False, n, 11, 9000 + n, scale=1.0, decay=1.0, n_R_target=275)
        best, res = None, {}
        for tk, nb in candidates(n):
            t = time_variant(m, a.nk, tk, nb, a.reps)
            res[f"tk{tk}_nb{nb}"] = t
            score = t["fp64"] + t["mixed"]
            if best is None or score < best[0]:
                best = (score, tk, nb)
        table.append({"nsta": n, "fourier_tk": best[1], "eigh_nb": best[2]})
        detail.append({"nsta": n, "ms": res})
        print(f"n={n}: fourier_tk={best[1]} eigh_nb={best[2]} ({best[0]:.3f} ms fp64+mixed)", flush=True)
    doc = {
        "workload": f"2D spinless, 275 R (|R_i| <= 11), {a.nk}x{a.nk} mesh, mu = 0, eta = 1e-3, one-model batch",
        "score": "median graph replay ms, FP64 + mixed",
        "command": "python -m rustb_amd.autotune " + " ".join(argv if argv is not None else sys.argv[1:]),
        "table": table,
        "candidates": detail,
    }
    tmp = a.out + ".tmp"
    with open(tmp, "w") as f:
        json.dump(doc, f, indent=1)
    os.replace(tmp, a.out)
    print(f"wrote {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
