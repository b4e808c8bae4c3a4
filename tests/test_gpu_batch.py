"""GPU parity of the multi-model batch (rtb_batch_*): the per-model Direct Hall
loop (response/linear/mod.rs:160-225) run as one planned, graph-captured job.

Bars: the FP64 batch equals one rtb_hall_direct call per model up to the order
of the k sum (1e-10 relative to the model's largest |sigma|: the per-k terms
are summed in another association); the mixed batch
within 1e-6 of it (the mixed form's AHC bar); Chern numbers round(2 pi sigma)
equal.  Pass splits (a small workspace budget) and kernel variants change
nothing per k-point: results are bit-identical across budgets, and the
Fourier variants leave H bit-identical (the eigensolver's block variants
reduce in another order: 1e-10 as above)."""
import math
import os

import numpy as np
import pytest

import rustb_amd as ra
from oracle import oracle as O
from rustb_amd.model import synth_model

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the first GPU run of the batch (graph replay, FP64, 8 small models) went silent for "
                            "180 s and was killed; no cause found by reading or by the host emulation "
                            "(tests/test_batch_emul.py), so the batch stays off the GPU until one is shown"),
]


def small_models():
    """Sizes across the variant boundaries: n^2 <= 256 (several k-groups per
    workgroup), above (element blocks), n > 128 (two-stage eigensolver);
    1D, 2D and 3D; one mesh too small for the separable route's R extent."""
    ms = [
        (synth_model(3, 2, False, 2, 2, 11, n_R_target=13), (9, 7), None),
        (synth_model(3, 2, False, 4, 3, 12, n_R_target=25), (8, 8), None),
        (synth_model(3, 2, False, 12, 2, 13, n_R_target=21), (6, 5), None),
        (synth_model(3, 2, False, 20, 3, 14, n_R_target=31), (7, 6), None),
        (synth_model(3, 2, False, 4, 3, 15, n_R_target=41), (3, 2), None),  # mesh smaller than the R extent
        (synth_model(1, 3, True, 8, 1, 16), (3, 4, 2), None),
        (synth_model(1, 1, True, 3, 2, 17), (11,), np.array([[1.0], [0.7]])),
        (synth_model(1, 3, True, 65, 1, 18, scale=0.5), (2, 2, 2), None),  # n = 130
    ]
    return ms


def dirs_of(m, d):
    return np.eye(max(m.dim, 2))[:2, :m.dim] if d is None else d


def per_model(ms, mu, eta, precision):
    out = []
    for m, km, d in ms:
        p = ra.Parameters(kmesh=km, direction=dirs_of(m, d), mu=mu, eta=eta, precision=precision)
        out.append(m.hall_conductivity(p).conductivity)
    return np.array(out)


def close_rel(a, b, tol):
    for i in range(a.shape[0]):
        scale = max(np.abs(b[i]).max(), 1e-300)
        assert np.abs(a[i] - b[i]).max() <= tol * scale, (i, a[i], b[i])


@pytest.fixture(scope="module")
def small():
    ms = small_models()
    mu = np.array([-0.4, 0.0, 0.3])
    ref = per_model(ms, mu, 0.02, "fp64")
    return ms, mu, ref


def test_batch_fp64_equals_per_model_calls(small):
    ms, mu, ref = small
    b = ra.ModelBatch([m for m, _, _ in ms], [km for _, km, _ in ms], [dirs_of(m, d) for m, _, d in ms], table=None)
    s = b.hall_conductivity(mu, eta=0.02, precision="fp64")
    close_rel(s, ref, 1e-10)
    info = b.info()
    assert info["n_models"] == len(ms) and info["captures"] == 1 and info["n_queues"] >= 1
    assert info["nk_total"] == sum(int(np.prod(km)) for _, km, _ in ms)
    st = ra.last_stats()
    assert st["nk"] == info["nk_total"] and st["launches"][-1] == info["graph_nodes"] > 0
    # one model against the oracle directly
    m, km, _ = ms[1]
    o = O.model_from_arrays(m.lat, m.orb, m.spin, m.hamR, m.ham)
    r = O.hall_conductivity_direct(o, list(km), np.eye(2), mu, 0.0, 0.02)
    assert np.abs(s[1] - r).max() <= 1e-10 * np.abs(r).max()


def test_batch_mixed_within_bar_of_fp64(small):
    ms, mu, ref = small
    b = ra.ModelBatch([m for m, _, _ in ms], [km for _, km, _ in ms], [dirs_of(m, d) for m, _, d in ms], table=None)
    s = b.hall_conductivity(mu, eta=0.02, precision="mixed")
    close_rel(s, ref, 1e-6)
    assert b.info()["n_mesh_models"] >= 4
    assert ra.last_stats()["max_residual"] < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_pass_splits_are_bit_identical(small, precision, monkeypatch):
    """A 4 MB budget splits models across passes (whole (k2, k3) pairs on the
    mesh route); every per-k term is the same, so the sums are too."""
    ms, mu, _ = small
    models, kms, dirs = [m for m, _, _ in ms], [km for _, km, _ in ms], [dirs_of(m, d) for m, _, d in ms]
    big = ra.ModelBatch(models, kms, dirs, table=None).hall_conductivity(mu, eta=0.02, precision=precision)
    monkeypatch.setenv("RTB_WORKSPACE_MB", "4")
    b = ra.ModelBatch(models, kms, dirs, table=None)
    s = b.hall_conductivity(mu, eta=0.02, precision=precision)
    assert b.info()["n_passes"] > 8
    np.testing.assert_array_equal(s, big)


def test_batch_graph_replays_and_parameter_sets(small):
    ms, mu, ref = small
    models, kms, dirs = [m for m, _, _ in ms], [km for _, km, _ in ms], [dirs_of(m, d) for m, _, d in ms]
    b = ra.ModelBatch(models, kms, dirs, table=None)
    first = b.hall_conductivity(mu, eta=0.02)
    for _ in range(3):  # replays with eager work of the program in between
        ms[2][0].hall_conductivity(ra.Parameters(kmesh=(4, 4), direction=np.eye(2), mu=mu, eta=0.02))
        np.testing.assert_array_equal(b.hall_conductivity(mu, eta=0.02), first)
    assert b.info()["captures"] == 1
    mu2 = np.array([0.1, -0.2, 0.5])  # same count: same graph, new values
    s2 = b.hall_conductivity(mu2, eta=0.02)
    assert b.info()["captures"] == 1
    close_rel(s2, per_model(ms, mu2, 0.02, "fp64"), 1e-10)
    s3 = b.hall_conductivity(mu[:1], eta=0.05)  # new parameter set: new graph
    assert b.info()["captures"] == 2
    close_rel(s3, per_model(ms, mu[:1], 0.05, "fp64"), 1e-10)
    np.testing.assert_array_equal(b.hall_conductivity(mu, eta=0.02), first)
    assert b.info()["graphs_cached"] == 2


def test_batch_eager_mode_equals_graph(small):
    """The same sequence launched directly: identical sums, stage times."""
    ms, mu, ref = small
    models, kms, dirs = [m for m, _, _ in ms], [km for _, km, _ in ms], [dirs_of(m, d) for m, _, d in ms]
    b = ra.ModelBatch(models, kms, dirs, table=None)
    for precision in ("fp64", "mixed"):
        g = b.hall_conductivity(mu, eta=0.02, precision=precision)
        b.set_graph(False)
        e = b.hall_conductivity(mu, eta=0.02, precision=precision)
        st = ra.last_stats()
        b.set_graph(True)
        np.testing.assert_array_equal(e, g)
        assert st["ms"][0] > 0 and st["ms"][-1] >= st["ms"][0]  # Fourier stage, total


def test_batch_variant_table(small):
    """Fourier tiles of 1, 2, 4 and 8 k-points per thread and every valid
    eigensolver block give the same conductivities."""
    ms, mu, ref = small
    models, kms, dirs = [m for m, _, _ in ms], [km for _, km, _ in ms], [dirs_of(m, d) for m, _, d in ms]
    sizes = sorted({m.nsta for m in models})
    for tk in (1, 2, 4, 8):
        table = []
        for n in sizes:
            nbs = [nb for nb in (32, 64, 128) if nb >= n] if n <= 128 else [0]
            table.append({"nsta": n, "fourier_tk": tk, "eigh_nb": nbs[tk % len(nbs)]})
        s = ra.ModelBatch(models, kms, dirs, table=table).hall_conductivity(mu, eta=0.02)
        close_rel(s, ref, 1e-10)


def test_batch_rejects_bad_input(small):
    ms, mu, _ = small
    m = ms[1][0]
    with pytest.raises(ra.TbError):
        ra.ModelBatch([m], [(4, 4)], table=[{"nsta": m.nsta, "fourier_tk": 3, "eigh_nb": 0}])
    with pytest.raises(ra.TbError):
        ra.ModelBatch([m], [(4, 4)], table=[{"nsta": m.nsta, "fourier_tk": 2, "eigh_nb": 16}])
    b = ra.ModelBatch([m], [(4, 4)])
    with pytest.raises(ra.TbError):
        b.hall_conductivity([0.0], eta=-1.0)
    with pytest.raises(ra.TbError):
        b.hall_conductivity([float("nan")])


@pytest.fixture(scope="module")
def batch64():
    pairs = [ra.workloads.batch2d_model(i) for i in range(64)]
    models = [m for m, _ in pairs]
    kms = [(nk, nk) for _, nk in pairs]
    mu = np.array([0.0])
    ref = np.array([m.hall_conductivity(ra.Parameters(kmesh=km, direction=np.eye(2), mu=mu)).conductivity
                    for m, km in zip(models, kms)])
    return models, kms, mu, ref


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_batch2d_64_models_chern_numbers(batch64, precision, tol):
    """All 64 seeded 2D models (norb 4..96, 53..469 R, meshes 18^2..128^2):
    every model's Chern number round(2 pi sigma(mu = 0)) equals the one of
    one FP64 call per model, sigma within the bar."""
    models, kms, mu, ref = batch64
    b = ra.ModelBatch(models, kms)  # the committed variant table
    s = b.hall_conductivity(mu, precision=precision)
    close_rel(s, ref, tol)
    np.testing.assert_array_equal(np.rint(2 * math.pi * s), np.rint(2 * math.pi * ref))
    c = b.chern_numbers(0.0, precision=precision)
    np.testing.assert_array_equal(c[:, 0], np.rint(2 * math.pi * ref[:, 0]).astype(np.int64))
    info = b.info()
    assert info["nk_total"] == sum(k[0] * k[1] for k in kms) and info["captures"] == 1
