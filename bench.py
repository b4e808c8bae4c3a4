"""Benchmark of the k-mesh hot path on one MI355X.

Default workload: wannier3d-n128 (spinful 3D model, 64 orbitals -> n = 128,
|R_i| <= 5 -> N_R = 1331, 48^3 = 110592 k-points): H(k) + dH/dk_{x,y} ->
eigh -> Kubo Omega_z -> AHC at mu = 0, T = 0, eta = 1e-4, all on the device
(one rtb_hall_direct call = one step; inputs resident in HBM, only the mu
values and sigma cross PCIe).

`--ingest` measures the input side instead: the same model written as
Wannier90 files (seedname_hr.dat in Wannier90's '(5I5,2F12.6)' layout, 21.8 M
lines, 1.09 GB) read by the library's loader and uploaded as a device model;
its JSON line compares the parse rate with the host's read ceiling for that
file (rtb_probe_read) and the oracle's line-by-line reading.

Prints ONE JSON line.  `value` is k-points/s of the whole job.  `roofline`
is the dominant kernel's algorithmic work per launch over its average launch
time, measured with HIP events inside the library on the kernel's stream
(rtb_last_stats: one event pair per launch group of that kernel).
`cpu_baseline` is the reference-equivalent CPU path (oracle/cpu_ref.cpp:
OpenMP over k, MKL/OpenBLAS zheev + zgemm, zaxpy-order Fourier sum) on a
bounded sample of the same mesh.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "k-points/s (H(k)+eigh+Berry Ω), mixed vs FP64 GPU form at ≤1e-9 eV eig error"

# published MI355X peaks
PEAK_FP64 = 78.6  # TFLOP/s (vector and matrix)
PEAK_FP32 = 157.3
PEAK_HBM = 8000.0  # GB/s


def device_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()


# Stages with a closed-form algorithmic flop count (the library reports the
# count per launch group in rtb_stats.flop): the kernels each one times, the
# arithmetic they run on and that unit's published peak.
ROOFLINE_STAGES = {
    "gemm64": ("zgemm3m_kernel<double, ...>", "FP64 matrix cores", PEAK_FP64),
    "gemm32": ("zgemm3m_kernel<float, ...>", "FP32 matrix cores", PEAK_FP32),
    "tridiag": ("eigh_tridiag_reg_kernel", "vector ALU", None),
    "backtransform": ("eigh_backtrans_reg_kernel", "vector ALU", None),
}
# above 128 orbitals the eigensolver's stages are the two-stage route's kernels
BAND_KERNELS = {"tridiag": "band_reduce_kernel + band_chase_kernel", "backtransform": "band_backtrans_kernel"}


def traffic_from_profiles(stage, workload="wannier3d-n128"):
    """HBM bytes per launch of the kernels of `stage` on `workload` measured
    with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel
    calibration), as committed under profiles/; None when no such
    measurement exists."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    key = "stage:" + stage if workload == "wannier3d-n128" else workload + "/stage:" + stage
    try:
        with open(path) as f:
            rec = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    return None if rec is None else rec.get("bytes_per_launch")


def mfma_ceiling():
    """Attainable FP64 and FP32 matrix-core rates of this device (TFLOP/s),
    measured by bench_probe/probe.hip: independent MFMA chains, no memory
    traffic.  None when the probe library is absent or fails."""
    try:
        lib = ctypes.CDLL(os.path.join(ROOT, "bench_probe", "librtb_probe.so"))
    except OSError:
        return None
    f64, f32 = ctypes.c_double(), ctypes.c_double()
    if lib.rtb_probe_mfma(ctypes.byref(f64), ctypes.byref(f32)) != 0:
        return None
    return {"fp64_mfma": round(f64.value, 2), "fp32_mfma": round(f32.value, 2)}


def copy_probe(nbytes=1 << 30, reps=3):
    """Device-to-device copies of a known size (HIP's blit kernel), the
    calibration run for the rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}))


def host_info():
    """CPU model and core counts of the host the CPU baseline runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "usable_cpus": usable}


def cpu_baseline(model, kmesh, dirs, eta, target_s):
    from oracle import cpu_ref

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    rng = np.random.default_rng(123)
    nk_total = int(np.prod(kmesh))
    done, t_used, batch = 0, 0.0, threads
    idx_all = []
    while t_used < target_s and done < nk_total:
        idx = rng.integers(0, nk_total, batch)
        ks = np.stack(np.unravel_index(idx, kmesh), axis=1) / np.asarray(kmesh, float)
        t0 = time.perf_counter()
        cpu_ref.run(2, model.lat, model.orb, model.spin, model.hamR, model.ham, ks, dirs, eta, threads)
        t_used += time.perf_counter() - t0
        done += batch
        idx_all.append(idx)
        batch = min(batch * 2, threads * 16)
    return {
        "value": done / t_used,
        "unit": "k-points/s",
        "cores": threads,
        "kind": "port",
        **host_info(),
        "sample": f"{done} seeded random k-points of the same {'x'.join(map(str, kmesh))} mesh "
                  f"(H + dH/dk zaxpy sum, zheev 'V', 4 zgemm, Kubo loop) in {t_used:.1f} s on {threads} threads, "
                  f"LAPACK {cpu_ref.lapack_name()}",
    }, np.concatenate(idx_all)


def accuracy_check(model, kmesh, idx, dirs, eta, precision):
    """GPU vs the CPU FP64 zheev path on the first sampled k-points."""
    import rustb_amd as ra
    from oracle import cpu_ref

    idx = idx[:64]
    ks = np.stack(np.unravel_index(idx, kmesh), axis=1) / np.asarray(kmesh, float)
    p = ra.Parameters(kmesh=tuple(kmesh), direction=np.asarray(dirs, float), mu=np.array([0.0]), eta=eta,
                      precision=precision)
    om, E = model.berry_curvature_on(ks, p)
    omr, Er = cpu_ref.run(2, model.lat, model.orb, model.spin, model.hamR, model.ham, ks, dirs, eta)
    occ = (om * (E < 0)).sum(axis=1)
    occr = (omr * (Er < 0)).sum(axis=1)
    return {
        "max_abs_dE_eV": float(np.abs(E - Er).max()),
        "omega_occ_max_rel_err": float(np.abs(occ - occr).max() / max(np.abs(occr).max(), 1e-300)),
        "k_points": int(len(idx)),
    }


def ingest_bench(args):
    """Wannier90 _hr.dat -> device model (Wannier90::from_hr, wannier90.rs:49-1228)."""
    import rustb_amd as ra
    from rustb_amd import _lib, workloads as W

    d = args.ingest_dir.rstrip("/") + "/"
    os.makedirs(d, exist_ok=True)
    hr = d + "wan_hr.dat"
    model = W.wannier3d_n128()
    n, nR = model.nsta, model.hamR.shape[0]
    expect = 3 + (nR + 14) // 15 + nR * n * n  # lines
    if not (os.path.exists(hr) and os.path.exists(d + "wan.win")
            and os.path.getsize(hr) >= 50 * nR * n * n):
        t = time.perf_counter()
        W.write_wannier90(model, d + "wan", centres=True)
        print(f"[ingest] wrote {hr} in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    size = os.path.getsize(hr)
    threads = min(16, os.cpu_count() or 1)
    lib = _lib.load()
    probe = ctypes.CDLL(os.path.join(ROOT, "bench_probe", "librtb_probe.so"))
    probe.rtb_probe_read.restype = ctypes.c_int
    probe.rtb_probe_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    ceil = ctypes.c_double()
    for _ in range(2):  # the second scan is from the page cache
        probe.rtb_probe_read(hr.encode(), threads, ctypes.byref(ceil))
    h = ctypes.c_void_p()
    dm = ctypes.c_void_p()
    for _ in range(args.warmup):
        _lib.check(lib.rtb_w90_read(d.encode(), b"wan", 0.0, -1, threads, ctypes.byref(h)))
        lib.rtb_w90_destroy(h)
    t_parse = t_up = 0.0
    for _ in range(args.steps):
        t0 = time.perf_counter()
        _lib.check(lib.rtb_w90_read(d.encode(), b"wan", 0.0, -1, threads, ctypes.byref(h)))
        t1 = time.perf_counter()
        _lib.check(lib.rtb_w90_model(h, ctypes.byref(dm)))
        device_sync()
        t2 = time.perf_counter()
        t_parse += t1 - t0
        t_up += t2 - t1
        lib.rtb_model_destroy(dm)
        lib.rtb_w90_destroy(h)
    t_parse /= args.steps
    t_up /= args.steps
    parse_gbps = size / t_parse / 1e9
    out = {
        "metric": "Wannier90 _hr.dat ingest (file -> device model), MB/s",
        "value": round(size / (t_parse + t_up) / 1e6, 1),
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round((t_parse + t_up) * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "text -> f64",
        "data": "synthetic",
        "config": {"workload": "wannier3d-n128 _hr.dat", "bytes": size, "lines": expect, "nsta": n, "n_R": nR,
                   "threads": threads, "files": "_hr.dat, .win, _centres.xyz"},
        "parse_ms": round(t_parse * 1e3, 1),
        "upload_ms": round(t_up * 1e3, 1),
        "roofline": {"bound": "host-read", "achieved": round(parse_gbps, 3), "peak": round(ceil.value, 3),
                     "unit": "GB/s", "frac": round(parse_gbps / ceil.value, 4) if ceil.value else None,
                     "traffic": None, "kernel": "rtb_w90_read (host threads)",
                     "peak_is": f"newline scan of the page-cached file on {threads} threads (rtb_probe_read)"},
    }
    if not args.no_cpu:
        # the oracle (line-by-line, one core) on a bounded sample: the same
        # model cut to |R_i| <= 1 (27 R, 442k lines)
        from oracle import wannier90_ref as OW
        keep = np.abs(model.hamR).max(axis=1) <= 1
        small = ra.Model.from_arrays(model.lat, model.orb, model.spin, model.hamR[keep], model.ham[keep])
        W.write_wannier90(small, d + "sample", centres=True)
        ss = os.path.getsize(d + "sample_hr.dat")
        t0 = time.perf_counter()
        OW.from_hr(d, "sample", 0.0)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(ss / tc / 1e6, 2), "unit": "MB/s", "cores": 1, "kind": "port", **host_info(),
                               "sample": f"oracle/wannier90_ref.py on the |R_i|<=1 cut ({ss} bytes, {tc:.1f} s)"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def energy_cut_workload(name="random2d-n32"):
    """random2d-n32: a random 2D model, 32 orbitals spinless (n = 32), |R_i|
    <= 3 (49 R), on a 128 x 128 mesh; random2d-n256: 256 orbitals spinless,
    |R_i| <= 2 (25 R), on a 32 x 32 mesh (the two-stage eigensolver and the
    band tracking with its matrices in global memory).  Both: 41 chemical
    potentials in [-2, 2] eV, eta = 1e-3, T = 0."""
    from rustb_amd.model import synth_model

    if name == "random2d-n256":
        model = synth_model(3, 2, False, 256, 2, 7, scale=1.0, decay=1.0, n_R_target=25)
        return model, (32, 32), np.linspace(-2.0, 2.0, 41), 1e-3
    model = synth_model(3, 2, False, 32, 3, 7, scale=1.0, decay=1.0, n_R_target=49)
    return model, (128, 128), np.linspace(-2.0, 2.0, 41), 1e-3


def energy_cut_bench(args):
    """hall_conductivity with Integration::EnergyCut (response/linear/mod.rs:226-268):
    device vertex stage, band tracking on the device (one launch per level of
    the walk's schedule), the triangle Fermi cut on the device."""
    import rustb_amd as ra
    from rustb_amd import _lib

    model, mesh, mu, eta = energy_cut_workload(args.ec_workload)
    nk = mesh[0] * mesh[1]
    p = ra.Parameters.rank2(mesh, [1.0, 0.0], [0.0, 1.0], mu, eta=eta, integration="EnergyCut")
    for _ in range(args.warmup):
        model.hall_conductivity(p)
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sig = model.hall_conductivity(p).conductivity
    wall = (time.perf_counter() - t0) / args.steps
    st = _lib.last_stats()
    names = _lib.STAGES
    # device time = the sum of the stage timers
    dev_ms = sum(st["ms"][i] for i in range(len(names)) if names[i] != "total")
    g = names.index("gemm64")
    tf = st["flop"][g] / (st["ms"][g] * 1e-3) / 1e12 if st["ms"][g] > 0 else 0.0
    out = {
        "metric": "k-points/s (EnergyCut AHC: vertex kernels, band tracking, triangle Fermi cut)",
        "value": round(nk / wall, 2), "unit": "k-points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.ec_workload + "-energycut", "nsta": model.nsta, "n_R": int(model.hamR.shape[0]),
                   "kmesh": list(mesh), "n_mu": int(mu.size), "eta": eta, "T": 0.0},
        "device_ms": round(dev_ms, 2), "host_ms": round(wall * 1e3 - dev_ms, 2),
        "stage_ms_per_step": {names[i]: round(st["ms"][i], 2) for i in range(len(names))},
        "sigma_mu0": float(sig[mu.size // 2]),
        "roofline": {"bound": "mfma", "achieved": round(tf, 3), "peak": 78.6, "unit": "TFLOP/s",
                     "frac": round(tf / 78.6, 4), "traffic": None, "kernel": "zgemm3m_kernel<double, ...>",
                     "stage": "gemm64", "note": "the overlap and Kubo products; host_ms = the walk schedule, launches and synchronisations"},
    }
    if not args.no_cpu and args.ec_workload == "random2d-n32":
        # the oracle (pure-Python loops, one core) on an 8 x 8 mesh of the same model
        from oracle import energy_cut_ref as EC
        from oracle import oracle as O

        o = O.model_from_arrays(model.lat, model.orb, model.spin, model.hamR, model.ham)
        t0 = time.perf_counter()
        EC.hall_conductivity_energy_cut(o, (8, 8), np.eye(2), mu, 0.0, eta)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(64 / tc, 2), "unit": "k-points/s", "cores": 1, "kind": "port", **host_info(),
                               "sample": f"oracle/energy_cut_ref.py on an 8 x 8 mesh of the same model ({tc:.1f} s)"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--workload", default="wannier3d-n128", choices=["wannier3d-n128", "slab2d-n256"],
                    help="wannier3d-n128: the headline configuration; slab2d-n256: 32 layers x 8 orbitals "
                         "(n = 256, 49 R) on 200 x 200")
    ap.add_argument("--kmesh", default=None)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ingest", action="store_true",
                    help="measure Wannier90 _hr.dat ingestion of the same model instead of the k-mesh path")
    ap.add_argument("--ingest-dir", default="/tmp/rtb_ingest")
    ap.add_argument("--energy-cut", action="store_true",
                    help="measure the EnergyCut Hall conductivity of a 2D model instead of the k-mesh path")
    ap.add_argument("--ec-workload", default="random2d-n32", choices=["random2d-n32", "random2d-n256"],
                    help="the EnergyCut workload (with --energy-cut); random2d-n256 has no CPU leg (the "
                         "oracle's pure-Python loops take minutes on any mesh of that model)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 3 device copies of 1 GiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(k-points shard trivially; see the sharding notes). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe()
        return
    if args.ingest:
        ingest_bench(args)
        return
    if args.energy_cut:
        energy_cut_bench(args)
        return

    import rustb_amd as ra
    from rustb_amd import _lib

    eta = 1e-4
    if args.workload == "slab2d-n256":
        kmesh = (200, 200)
        dirs = [[1.0, 0.0], [0.0, 1.0]]
        model = ra.workloads.slab2d_n256()
    else:
        kmesh = (48, 48, 48)
        dirs = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
        model = ra.workloads.wannier3d_n128()
    if args.kmesh:
        kmesh = tuple(int(x) for x in args.kmesh.split(","))
    n, nR = model.nsta, model.hamR.shape[0]
    nk = int(np.prod(kmesh))
    params = ra.Parameters(kmesh=kmesh, direction=np.array(dirs), mu=np.array([0.0]), eta=eta,
                           precision=args.precision)
    model._device()  # upload outside the timed region

    for _ in range(args.warmup):
        model.hall_conductivity(params)
    device_sync()
    names = _lib.STAGES
    ns = len(names)
    stage_ms = np.zeros(ns)
    stage_launch = np.zeros(ns)
    stage_flop = np.zeros(ns)
    stage_mflop = np.zeros(ns)
    t0 = time.perf_counter()
    sigma = None
    step_s = []
    ws_gb = []
    for _ in range(args.steps):
        ts = time.perf_counter()
        sigma = model.hall_conductivity(params).conductivity[0]  # synchronous call
        step_s.append(time.perf_counter() - ts)
        st = _lib.last_stats()
        stage_ms += st["ms"]
        stage_launch += st["launches"]
        stage_flop += st["flop"]
        stage_mflop += st["mflop"]
        ws_gb.append(_lib.workspace_bytes()[0] / 1e9)
    device_sync()
    elapsed = time.perf_counter() - t0
    value = nk * args.steps / elapsed

    # dominant kernel among the stages with an algorithmic flop count
    cand = [names.index(k) for k in ROOFLINE_STAGES if stage_launch[names.index(k)] > 0]
    dom = max(cand, key=lambda i: stage_ms[i])
    dname = names[dom]
    kernel, unit_name, peak = ROOFLINE_STAGES[dname]
    if n > 128 and dname in BAND_KERNELS:
        kernel = BAND_KERNELS[dname]
    f32 = args.precision == "mixed" and dname in ("tridiag", "backtransform")
    if peak is None:  # vector ALU: FP32 in the mixed form, FP64 otherwise
        peak = PEAK_FP32 if f32 else PEAK_FP64
    per_launch_ms = stage_ms[dom] / stage_launch[dom]
    flop_per_launch = stage_flop[dom] / stage_launch[dom]
    achieved = flop_per_launch / (per_launch_ms * 1e-3) / 1e12
    roofline = {
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": round(achieved / peak, 4),
        "traffic": traffic_from_profiles(dname, args.workload),
        "kernel": kernel,
        "stage": dname,
        "arithmetic": unit_name,
        "per_launch_ms": round(per_launch_ms, 4),
        "launches_per_step": int(round(stage_launch[dom] / args.steps)),
        "algorithmic_flop_per_launch": flop_per_launch,
        "dominant_stage_by_time": names[int(np.argmax(stage_ms[:ns - 1]))],
    }
    if stage_mflop[dom] > 0:
        # what the matrix cores executed (3M products, padded tiles) against
        # the rate this device attains on independent MFMA chains
        ceil = mfma_ceiling()
        executed = stage_mflop[dom] / stage_launch[dom] / (per_launch_ms * 1e-3) / 1e12
        roofline["mfma_executed"] = round(executed, 3)
        if ceil:
            att = ceil["fp64_mfma" if dname == "gemm64" else "fp32_mfma"]
            roofline["attainable_peak"] = att
            roofline["frac_of_attainable"] = round(executed / att, 4)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "k-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "nsta": n, "n_R": nR, "kmesh": list(kmesh), "n_k": nk,
                   "precision": args.precision, "eta": eta, "mu": [0.0], "T": 0.0, "outputs": "E, Omega_z, AHC"},
        "stage_ms_per_step": {names[i]: round(stage_ms[i] / args.steps, 2) for i in range(ns)},
        "step_ms": {"min": round(min(step_s) * 1e3, 2), "median": round(float(np.median(step_s)) * 1e3, 2),
                    "max": round(max(step_s) * 1e3, 2)},
        "workspace_gb_after_each_step": [round(x, 2) for x in ws_gb],
        "eigenpair_check": {"max_residual_eV": st["max_residual"], "resolved_in_fp64": int(st["resolved"]),
                            "of_k": nk},
        "sigma_xy_mu0": float(sigma),
        "roofline": roofline,
    }
    if args.precision == "mixed" and not args.no_fp64_compare:
        # the plain FP64 GPU form on the same workload, same process
        p64 = ra.Parameters(kmesh=kmesh, direction=np.array(dirs), mu=np.array([0.0]), eta=eta, precision="fp64")
        model.hall_conductivity(p64)
        device_sync()
        t0 = time.perf_counter()
        s64 = model.hall_conductivity(p64).conductivity[0]
        device_sync()
        t64 = time.perf_counter() - t0
        out["fp64_plain"] = {"value": round(nk / t64, 2), "ms_per_step": round(t64 * 1e3, 2),
                             "stage_ms": {names[i]: round(_lib.last_stats()["ms"][i], 2) for i in range(ns)}}
        out["mixed_over_fp64"] = round(value / (nk / t64), 3)
        out["ahc_rel_err_mixed_vs_fp64"] = float(abs(sigma - s64) / abs(s64))
    if not args.no_cpu:
        cb, idx = cpu_baseline(model, kmesh, dirs, eta, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(value / cb["value"], 1)
        out["accuracy_vs_fp64_zheev"] = accuracy_check(model, kmesh, idx, dirs, eta, args.precision)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
