#!/usr/bin/env python3
"""Throughput of the LBGK hot path on one MI355X: D3Q19 MLUPS and HBM roofline.

Workload (the configuration the headline metric is quoted on): D3Q19 LBGK
channel flow past a sphere, 256x128x128 cells, FP32, tau = 0.56 (nu = 0.02),
sphere of radius 16 at (64, 64, 64) from a normalised level-set map
(no-slip), velocity inlet u = 0.04 on X0, pressure outlet rho = 1 on XE,
no-slip side walls; initial u = (0.04, 0, 0) + 1e-3 * U[-1, 1] per cell and
component from a seeded generator, u = 0 inside the sphere.  A step is one
fused collide+stream pass over the whole lattice with every BC applied.

The step kernel runs with the library's default launch settings (reported
under roofline.launch), so that the rocprofv3 summaries of the same command
time the same launches; --autotune lets the library's autotuner choose them
first (every workgroup size and XCD run length timed over 20 real steps).
The warm-up steps are followed by untimed runs of --steps steps that capture
the hipGraphs the timed region replays (100-step chunks and the remainder,
for both starting parities when --steps is odd).  The timed region then holds
exactly --steps steps, every one a graph replay, with the populations already
in HBM; it is bracketed by stream synchronisations.  ``value`` is MLUPS over
computation nodes (solid and ghost cells are not counted) from the wall clock
around it.  The roofline entry is the step kernel's algorithmic HBM bytes per
launch over its average launch duration, measured after the timed region
with HIP events recorded on the lattice's stream around each of 400
back-to-back eager launches (minimum and median reported too); its traffic
figure is read from the committed rocprofv3 counter summary.
``cpu_baseline`` times the CPU oracle (the C++ restatement of ASL's LBGK) on
a bounded sample of the same lattice with every CPU the process may use (no
more than an OMP_NUM_THREADS cap of the environment); the host's CPU facts
are recorded with it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--autotune] [--remap G]
       python bench.py --copy-calibration BYTES   (one plain device copy, for
       scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters)
       python bench.py --pattern-probe   (device copy and the bare D3Q19 AA
           pattern in several layouts: the ceilings the step kernel is judged against)
       python bench.py --precision-study N [--snapshots 500,2000,5000]
           (D3Q27 periodic Taylor-Green vortex on N^3, tau = 0.5005: FP32, FP16
           and BF16 storage against FP64; one JSON line per storage type)
       python bench.py --closure-study N [--tau-fine T] [--chunk C] [--train-snapshots K]
           [--eval-chunks E] [--closure-target total|filtered] [--ce-start none|coarse|both]
           [--closure-shifts S1,S2,...]
           (learned closure: N^3
           coarse vs (4N)^3 fine decaying turbulence; KE(t) of fine, coarse and
           coarse+closure, wall times)
       python bench.py --cavity-study   (the 64^3 lid-driven cavity: GPU and CPU-oracle MLUPS)
       python bench.py --storage-study  (the sphere channel with FP32, FP16, BF16, FP64 storage)
       python bench.py --batch-study [--steps K]
           (64 lid-driven cavities of log-uniform sizes 32..160 and tau in
           [0.51, 1]: one batched launch per step vs one launch per lattice)
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "D3Q19 MLUPS & % HBM roofline; coarse+closure vs plain fine at equal KE error"
HBM_PEAK_GBS = 8000.0
SIZE = (256, 128, 128)
NU = (0.56 - 0.5) / 3.0
U_IN = 0.04
SEED = 20261202
KERNEL_LAUNCHES = 400  # eager step launches timed one by one for the roofline entry


def splitmix64_uniform(seed, n):
    """Outputs 0..n-1 of the splitmix64 generator seeded with seed, as floats in [-1, 1).

    splitmix64 (S. Vigna, public domain): output k is mix(seed + (k+1) *
    0x9E3779B97F4A7C15).  The top 24 bits give an exact float32 in [0, 1),
    mapped to 2u - 1.  Counter-based, so any host can rebuild the field.
    """
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(n, dtype=np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    u01 = (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
    return np.float32(2.0) * u01 - np.float32(1.0)


def setup_fields():
    """Sphere map and the perturbed initial velocity (float32, ASL layout).

    u = (U_IN, 0, 0) + 1e-3 * xi, xi in [-1, 1) from splitmix64 output
    3 * cell + component (cell: ghost-inclusive linear index), u = 0 where
    the map is <= 0 (inside the sphere).
    """
    from asl_amd import geometry
    mapv = geometry.normalize(geometry.sphere(SIZE, 16.0, (64.0, 64.0, 64.0)))
    S = tuple(s + 2 for s in SIZE)
    ncell = int(np.prod(S))
    xi = splitmix64_uniform(SEED, 3 * ncell).reshape(S + (3,))
    u = np.zeros((3,) + S, dtype=np.float32)
    u[0] = U_IN
    u += np.float32(1e-3) * np.moveaxis(xi, -1, 0)
    u[:, mapv <= 0] = 0.0
    return mapv, u


def tgv_fields(n, u0=0.04):
    """3D Taylor-Green vortex on a periodic n^3 lattice (ASL layout, float64).

    u = u0 sin kx cos ky cos kz, v = -u0 cos kx sin ky cos kz, w = 0 with
    k = 2 pi / n at x = index - 1, and the matching pressure
    p = u0^2 / 16 (cos 2kx + cos 2ky)(cos 2kz + 2), rho = 1 + 3 p.
    """
    k = 2.0 * np.pi / n
    x = (np.arange(n + 2, dtype=np.float64) - 1.0) * k
    s, c = np.sin(x), np.cos(x)
    c2 = np.cos(2 * x)
    # separable products, evaluated left to right as u0 * f(x) * g(y) * h(z)
    u = np.empty((3, n + 2, n + 2, n + 2))
    u[0] = (u0 * s)[:, None, None] * c[None, :, None] * c[None, None, :]
    u[1] = (-u0 * c)[:, None, None] * s[None, :, None] * c[None, None, :]
    u[2] = 0.0
    rho = 1.0 + (3.0 * (u0 * u0 / 16.0)) * (c2[:, None, None] + c2[None, :, None]) * (c2 + 2.0)[None, None, :]
    return rho, u


def build_gpu(mapv, u, storage="f32"):
    import asl_amd
    from asl_amd import _lib
    L = asl_amd.Lattice(SIZE, 19, dtype=storage, nu=NU)
    L.init_equilibrium(1.0, u)
    bcs = [L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.X0, u=(U_IN, 0.0, 0.0)),
           L.add_slice_bc(_lib.BC_CONSTANT_PRESSURE, _lib.XE, p=1.0),
           L.add_slice_bc(_lib.BC_NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
           L.add_map_bc(_lib.BC_NOSLIP_MAP, mapv)]
    for b in bcs:
        L.execute_bc(b)
    L.finalize()
    return L


def host_cpu():
    """What the host offers the CPU baseline: logical CPUs, the ones this
    process may run on, the OpenMP thread cap set in the environment, and the
    CPU model, sockets and cores per socket from lscpu."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = os.cpu_count()
    if os.environ.get("OMP_NUM_THREADS"):
        info["omp_num_threads_env"] = os.environ["OMP_NUM_THREADS"]
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.strip()] = v.strip()
    except (OSError, ValueError):
        pass
    return info


def baseline_threads(host):
    """All CPUs this process may use, but no more than an OpenMP cap the
    environment sets (the share of the host granted to one GPU's job)."""
    n = host.get("affinity") or 1
    cap = host.get("omp_num_threads_env", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return n


def cpu_baseline(mapv, u, budget_s=15.0):
    """The CPU oracle (FP32, ASL's algorithm) on the same lattice for ~budget_s seconds."""
    import oracle
    from oracle.lbgk import set_threads
    host = host_cpu()
    threads = set_threads(baseline_threads(host))
    O = oracle.OracleLattice(SIZE, 19, fp64=False, nu=NU)
    O.init_f(1.0, u.astype(np.float64))
    bcs = [O.add_slice_bc(O.CONSTANT_VELOCITY, 1, u=(U_IN, 0.0, 0.0)),
           O.add_slice_bc(O.CONSTANT_PRESSURE, 2, p=1.0),
           O.add_slice_bc(O.NOSLIP, 4 | 8 | 16 | 32),
           O.add_map_bc(O.NOSLIP_MAP, mapv)]
    for b in bcs:
        O.execute_bc(b)
    O.run(1, bcs)  # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        O.run(1, bcs)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 2000:
            break
    counted = int(((mapv > 0)[1:-1, 1:-1, 1:-1]).sum())
    return {"value": round(counted * steps / el / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "CPU oracle (C++/OpenMP, FP32, ASL's pool-shift LBGK + BCs) on the same 256x128x128 "
                      "sphere lattice, %d steps in %.1f s on %d threads" % (steps, el, threads),
            "host": host}


def copy_calibration(nbytes):
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print("copied %d bytes device to device" % (a.numel() * 4), flush=True)


def pattern_probe(reps=60):
    """Memory-system ceilings behind the step kernel (asl_amd/libaslprobe.so).

    1. A plain device copy of 1 GiB (one and four floats per lane): the
       bandwidth a read+write stream reaches on this device.
    2. The bare D3Q19 AA pattern on the sphere channel's 256x128x128 interior
       (19 FP32 reads + 19 FP32 writes per cell, no boundary logic): ASL's
       layout and rows padded to 128-byte lines, 1 / 2 / 4 cells per lane,
       256- and 512-lane workgroups.  Algorithmic TB/s = 152 B x cells over
       the mean of the even and odd launch medians.
    """
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "asl_amd", "libaslprobe.so"))
    dp = ctypes.POINTER(ctypes.c_double)
    lib.aslprobe_copy.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_int, dp]
    lib.aslprobe_aa.argtypes = [ctypes.c_int] * 8 + [dp]
    nbytes = 1 << 30
    for vec in (1, 4):
        ms = np.zeros(reps)
        rc = lib.aslprobe_copy(nbytes, vec, reps, ms.ctypes.data_as(dp))
        if rc:
            sys.exit("pattern probe: copy failed (%d)" % rc)
        med = float(np.median(ms[2:]))
        print(json.dumps({"probe": "device copy", "bytes_copied": nbytes, "floats_per_lane": vec,
                          "ms_median": round(med, 5), "tbps_read_plus_write": round(2 * nbytes / med / 1e9, 3),
                          "frac_of_8tbps": round(2 * nbytes / med / 1e9 / 8.0, 4)}), flush=True)
    nx, ny, nz = SIZE
    cells = nx * ny * nz
    for layout, lname in ((0, "ASL rows (nz + 2)"), (1, "rows padded to 128-byte lines")):
        for cpl in (1, 2, 4):
            for block in (256, 512):
                for remap in (4, 16):
                    ms = np.zeros(reps)
                    rc = lib.aslprobe_aa(nx, ny, nz, layout, cpl, block, remap, reps, ms.ctypes.data_as(dp))
                    if rc:
                        sys.exit("pattern probe: AA failed (%d)" % rc)
                    ev, od = float(np.median(ms[2::2])), float(np.median(ms[3::2]))
                    tb = 152.0 * cells / ((ev + od) / 2) / 1e9
                    print(json.dumps({"probe": "D3Q19 AA pattern, 256x128x128 interior", "layout": lname,
                                      "cells_per_lane": cpl, "workgroup": block, "xcd_run": remap,
                                      "ms_even": round(ev, 5), "ms_odd": round(od, 5),
                                      "tbps_algorithmic": round(tb, 3), "frac_of_8tbps": round(tb / 8.0, 4)}),
                          flush=True)


def precision_study(n, snapshots, u0=0.04, tau=0.5005):
    """Storage precision vs FP64 on the D3Q27 Taylor-Green vortex (one GPU).

    The FP64 run goes first and keeps u at every snapshot on the host; each
    other storage type is then run and compared: max|u - u64| / u0, the
    relative KE drift against FP64, and KE(t) / KE(0).
    """
    import asl_amd
    nu = (tau - 0.5) / 3.0
    rho, u = tgv_fields(n, u0)
    inner = (slice(None),) + (slice(1, -1),) * 3
    ref = {}
    for dt in ("f64", "f32", "f16", "bf16"):
        L = asl_amd.Lattice((n, n, n), 27, dtype=dt, periodic=True, nu=nu)
        L.init_equilibrium(rho, u)
        ke0 = L.conservation()[4]
        done, t0, rows = 0, time.perf_counter(), []
        for snap in snapshots:
            L.step(snap - done)
            done = snap
            ke = L.conservation()[4]
            v = L.velocity()[inner].astype(np.float64)
            if dt == "f64":
                ref[snap] = (v, ke / ke0)
                rows.append({"step": snap, "ke_ratio": ke / ke0})
            else:
                rows.append({"step": snap, "ke_ratio": ke / ke0,
                             "max_du_over_u0": float(np.abs(v - ref[snap][0]).max() / u0),
                             "rms_du_over_u0": float(np.sqrt(np.mean((v - ref[snap][0]) ** 2)) / u0),
                             "ke_ratio_minus_fp64": ke / ke0 - ref[snap][1]})
            del v
            print("# %s step %d done (%.1f s)" % (dt, snap, time.perf_counter() - t0), flush=True)
        del L
        print(json.dumps({"study": "storage precision, D3Q27 periodic Taylor-Green vortex", "n": n, "tau": tau,
                          "u0": u0, "storage": dt, "snapshots": rows}), flush=True)


def cavity_study(no_cpu=False, steps=1000):
    """64^3 lid-driven cavity (tau 0.6, lid 0.05 on ZE executed first, no-slip
    elsewhere), FP32: GPU MLUPS and step-kernel bandwidth, CPU oracle MLUPS."""
    n = 64
    L = batch_cavity(n, 0.6)
    block, remap = L.autotune(20)
    L.step(100, write_macro=False, graph=True)
    L.synchronize()
    t0 = time.perf_counter()
    L.time_steps(steps, write_macro=False, graph=True)
    L.synchronize()
    wall = time.perf_counter() - t0
    _, ms_kernel = L.time_steps(steps, write_macro=False)
    cells = n ** 3
    kern_s = ms_kernel / steps / 1e3
    line = {"study": "64^3 lid-driven cavity, D3Q19 FP32", "steps": steps,
            "gpu_mlups": round(cells * steps / wall / 1e6, 1), "ms_per_step": round(wall * 1e3 / steps, 5),
            "kernel_ms_avg": round(kern_s * 1e3, 5),
            "kernel_gbps": round((152.0 * cells + cells) / kern_s / 1e9, 1),
            "launch": {"workgroup": block, "xcd_run": remap}}
    if not no_cpu:
        import oracle
        from oracle.lbgk import set_threads
        threads = set_threads(min(16, os.cpu_count() or 1))
        O = oracle.OracleLattice((n, n, n), 19, fp64=False, nu=(0.6 - 0.5) / 3.0)
        O.init_f(1.0, (0.0, 0.0, 0.0))
        bcs = [O.add_slice_bc(O.CONSTANT_VELOCITY, 32, u=(0.05, 0.0, 0.0)), O.add_slice_bc(O.NOSLIP, 31)]
        for b in bcs:
            O.execute_bc(b)
        O.run(5, bcs)
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and k < steps:
            O.run(10, bcs)
            k += 10
        el = time.perf_counter() - t0
        line["cpu_oracle_mlups"] = round(cells * k / el / 1e6, 2)
        line["cpu_cores"] = threads
        line["cpu_sample_steps"] = k
    print(json.dumps(line), flush=True)


def storage_study(steps=2000):
    """The sphere channel with each population storage type and step-kernel
    variant (LDS-staged tiles, one cell per lane, runs of 4 cells per lane):
    step-kernel time, hipGraph MLUPS and algorithmic bandwidth (FP32 collision
    throughout; FP64 for the f64 lattice; BF16/FP16 store f - 1)."""
    from asl_amd import _lib
    mapv, u = setup_fields()
    for storage, bpc in (("f32", 152.0), ("f16", 76.0), ("bf16", 76.0), ("f64", 304.0)):
        L = build_gpu(mapv, u, storage)
        counted = L.stats()["counted"]
        for variant, lds, cpl in (("tiles", 1, 0), ("tiles_even_steps", 2, 0), ("one_cell_per_lane", 0, 1),
                                  ("runs_of_4", 0, 4)):
            L.set_tuning(_lib.TUNE_LDS, lds)
            L.set_tuning(_lib.TUNE_CPL, cpl)
            L.step(200, write_macro=False, graph=True)
            L.synchronize()
            ms_graph, _ = L.time_steps(steps, write_macro=False, graph=True)
            _, ms_kernel = L.time_steps(500, write_macro=False)
            kern = ms_kernel / 500
            byts = bpc * counted + 1.0 * int(np.prod(SIZE))
            print(json.dumps({"study": "population storage, 256x128x128 sphere channel", "storage": storage,
                              "variant": variant, "workgroup": L.tuning_value(_lib.TUNE_BLOCK),
                              "bytes_per_cell_update": bpc, "kernel_ms": round(kern, 5),
                              "graph_mlups": round(counted * steps / ms_graph / 1e3, 1),
                              "kernel_tbps_algorithmic": round(byts / (kern / 1e3) / 1e12, 3),
                              "frac_of_8tbps": round(byts / (kern / 1e3) / 8e12, 4)}), flush=True)
        del L


def launch_study():
    """Step-kernel duration (HIP events, median of 200 launches) on the sphere
    channel for cells per lane x XCD run length x workgroup size."""
    from asl_amd import _lib
    mapv, u = setup_fields()
    L = build_gpu(mapv, u)
    L.step(20, write_macro=False)
    counted = L.stats()["counted"]
    byts = 152.0 * counted + 1.0 * int(np.prod(SIZE))
    # cells per lane 1, 2, 4 (direct), then LDS-staged tiles of 256 cells on
    # every step ("lds") and on even steps only ("lds_even", odd steps direct)
    for cpl in (1, 2, 4, "lds", "lds_even"):
        for remap in (0, 4, 16, 64):
            for block in (256, 512, 1024):
                if cpl in (2, 4) and block > 512 or cpl == "lds" and block != 256:
                    continue
                L.set_tuning(_lib.TUNE_LDS, {"lds": 1, "lds_even": 2}.get(cpl, 0))
                L.set_tuning(_lib.TUNE_CPL, cpl if cpl in (1, 2, 4) else 0)
                L.set_tuning(_lib.TUNE_XCD_REMAP, remap)
                L.set_tuning(_lib.TUNE_BLOCK, block)
                L.time_steps(20, write_macro=False)
                _, _, each = L.time_steps(200, write_macro=False, each=True)
                med = float(np.median(each))
                print(json.dumps({"study": "step-kernel launch settings, 256x128x128 sphere channel",
                                  "cells_per_lane": L.tuning_value(_lib.TUNE_CPL),
                                  "lds_tiles": L.tuning_value(_lib.TUNE_LDS), "xcd_run": remap,
                                  "workgroup": L.tuning_value(_lib.TUNE_BLOCK), "kernel_ms_median": round(med, 5),
                                  "kernel_ms_even": round(float(np.median(each[0::2])), 5),
                                  "kernel_ms_odd": round(float(np.median(each[1::2])), 5),
                                  "tbps_algorithmic": round(byts / med / 1e9, 3),
                                  "frac_of_8tbps": round(byts / med / 1e9 / 8.0, 4)}), flush=True)


def batch_specs(n=64, seed=SEED):
    """(L, tau) of the batched-cavities workload: L = round(exp(U[ln 32, ln 160])), tau ~ U[0.51, 1]."""
    rng = np.random.default_rng(seed)
    L = np.rint(np.exp(rng.uniform(np.log(32.0), np.log(160.0), n))).astype(int)
    tau = rng.uniform(0.51, 1.0, n)
    return [(int(a), float(b)) for a, b in zip(L, tau)]


def batch_cavity(n, tau, lid=0.05):
    """Lid-driven cavity n^3 (lid u = (lid, 0, 0) on ZE executed first, no-slip elsewhere)."""
    import asl_amd
    from asl_amd import _lib
    L = asl_amd.Lattice((n, n, n), 19, dtype="f32", nu=(tau - 0.5) / 3.0)
    L.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    bcs = [L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(lid, 0.0, 0.0)),
           L.add_slice_bc(_lib.BC_NOSLIP, _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0)]
    for b in bcs:
        L.execute_bc(b)
    return L


def batch_study(steps, warmup, group_mb=None):
    """MLUPS of 64 uneven cavities: one batched launch per step (hipGraph chunks)
    against the same lattices stepped one launch per lattice and step."""
    import asl_amd
    specs = batch_specs()
    cells = sum(n ** 3 for n, _ in specs)
    lat = [batch_cavity(n, t) for n, t in specs]
    B = asl_amd.Batch(lat)
    B.step(max(warmup, 1), write_macro=False)
    block = B.autotune(10)
    if group_mb is None:
        group_mb = B.autotune_groups(2)
    else:
        from asl_amd import _lib
        B.set_tuning(_lib.TUNE_GROUP_MB, group_mb)
    ms_batch = B.time_steps(steps, graph=True)
    # the two step kernels for the whole batch (LDS-staged tiles; direct, one
    # cell per lane in the workgroup the autotuner last timed for it)
    from asl_amd import _lib
    kernels = {}
    for name, lds in (("tiles", 1), ("direct", 0)):
        B.set_tuning(_lib.TUNE_LDS, lds)
        B.step(100, write_macro=False, graph=True)  # captures this variant's chunk
        kernels[name] = round(cells * steps / B.time_steps(steps, graph=True) / 1e3, 1)
    solo = [batch_cavity(n, t) for n, t in specs]
    for L in solo:
        L.step(max(warmup, 1), write_macro=False)
        L.autotune(4)
    for L in solo:
        L.synchronize()
    t0 = time.perf_counter()
    for L in solo:
        L.step(steps, write_macro=False)
    for L in solo:
        L.synchronize()
    ms_solo = (time.perf_counter() - t0) * 1e3
    # sequential single-lattice launches, one lattice after the other (ASL's pattern, without its waits)
    t0 = time.perf_counter()
    for _ in range(min(steps, 100)):
        for L in solo:
            L.step(1, write_macro=False)
    for L in solo:
        L.synchronize()
    ms_interleaved = (time.perf_counter() - t0) * 1e3 * steps / min(steps, 100)
    print(json.dumps({"study": "batched lattices", "lattices": len(specs), "cells": cells,
                      "sizes": [n for n, _ in specs], "steps": steps,
                      "batch_graph_mlups": round(cells * steps / ms_batch / 1e3, 1), "batch_block": block,
                      "batch_graph_mlups_by_kernel": kernels,
                      "batch_group_mb": group_mb,
                      "batch_gbps": round(152.0 * cells * steps / ms_batch / 1e6, 1),
                      "per_lattice_streams_mlups": round(cells * steps / ms_solo / 1e3, 1),
                      "per_lattice_interleaved_mlups": round(cells * steps / ms_interleaved / 1e3, 1)}), flush=True)


def closure_study(nc, tau_fine, chunk, snaps, eval_chunks, target="total", ce_init=False, shifts=(0.0,), seed=SEED):
    """Coarse N^3 with and without the learned closure against the 4x finer run
    (``asl_amd.closure.study``): one JSON line."""
    from asl_amd import closure
    r = closure.study(nc, tau_fine, chunk=chunk, snapshots=snaps, eval_chunks=eval_chunks, target=target,
                      seed=seed, ce_init=ce_init, shifts=shifts, log=lambda m: print(m, flush=True))
    print(json.dumps(r), flush=True)


def variant_study(n, steps=500):
    """Step-kernel cost of the per-cell relaxation variants on a periodic n^3
    D3Q19 FP32 Taylor-Green lattice (tau 0.56): plain LBGK, LBGKTurbulence, the
    learned closure on the matrix cores (BF16 MFMA) and in FP32 on the vector
    ALUs (random network, output scale 0.1).  Kernel time from HIP events on
    the lattice stream; MLUPS from hipGraph-replayed steps."""
    import asl_amd
    from asl_amd import _lib, closure
    rho, u = tgv_fields(n)
    nu = (0.56 - 0.5) / 3.0
    for name in ("plain", "turbulence", "closure_mfma", "closure_fp32"):
        L = asl_amd.Lattice((n, n, n), 19, periodic=True, nu=nu)
        L.init_equilibrium(rho, u)
        if name == "turbulence":
            L.set_turbulence(True)
        elif name.startswith("closure"):
            L.set_closure(closure.init_params(1), None, None, 0.1)
            L.set_closure_mode(_lib.CLOSURE_BF16_MFMA if name == "closure_mfma" else _lib.CLOSURE_FP32)
        L.step(100, write_macro=False, graph=True)
        L.synchronize()
        ms_graph, _ = L.time_steps(steps, write_macro=False, graph=True)
        _, ms_kernel = L.time_steps(steps, write_macro=False)
        print(json.dumps({"study": "per-cell relaxation variants, %d^3 periodic D3Q19 FP32" % n, "variant": name,
                          "kernel_ms": round(ms_kernel / steps, 5),
                          "graph_mlups": round(n ** 3 * steps / ms_graph / 1e3, 1)}), flush=True)
        del L


TRAFFIC_PROFILE = os.path.join("profiles", "lbgk_step_d3q19_f32_sphere_256x128x128_traffic.md")


def profile_traffic():
    """(HBM bytes per launch of the step kernel, where the figure comes from):
    read from the committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated); (None, None) if it is absent."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None, None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    if not m:
        return None, None
    return float(m.group(1)), TRAFFIC_PROFILE + " (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline sample")
    ap.add_argument("--remap", type=int, default=-1, metavar="G",
                    help="XCD run length of the step kernel's chunks (default: the library's)")
    ap.add_argument("--cpl", type=int, default=-1, metavar="C",
                    help="cells per lane of the step kernel (1, 2, 4; default: the library's choice)")
    ap.add_argument("--lds", type=int, default=-1, metavar="B",
                    help="1: the step kernel stages tiles of 256 cells through LDS on every step; 2: on even "
                         "steps only; 0: direct (default: the library's choice)")
    ap.add_argument("--launch-study", action="store_true",
                    help="step-kernel duration on the sphere channel for every cells-per-lane x XCD run x workgroup")
    ap.add_argument("--autotune", action="store_true",
                    help="let the library's autotuner choose the launch settings before timing")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES bytes device to device once and exit (reference kernel of known "
                         "traffic for the HBM counters)")
    ap.add_argument("--precision-study", type=int, default=0, metavar="N",
                    help="only run the storage-precision study on an N^3 D3Q27 Taylor-Green vortex")
    ap.add_argument("--snapshots", default="500,2000,5000", help="steps at which the precision study compares")
    ap.add_argument("--closure-study", type=int, default=0, metavar="N",
                    help="only run the learned-closure study with an N^3 coarse lattice")
    ap.add_argument("--tau-fine", type=float, default=0.52)
    ap.add_argument("--chunk", type=int, default=50, help="coarse steps between snapshots")
    ap.add_argument("--train-snapshots", type=int, default=16)
    ap.add_argument("--eval-chunks", type=int, default=16)
    ap.add_argument("--closure-target", choices=("total", "filtered"), default="total",
                    help="closure study: choose the output scale against the fine run's kinetic-energy decay "
                         "or against the box-filtered fine kinetic energy")
    ap.add_argument("--ce-start", choices=("none", "coarse", "both"), default="none",
                    help="closure study: start the coarse lattice (or both) from equilibrium plus the first "
                         "Chapman-Enskog term of the field instead of equilibrium")
    ap.add_argument("--closure-seed", type=int, default=SEED,
                    help="closure study: seed of the evaluation field (the training field uses seed + 1)")
    ap.add_argument("--closure-shifts", default="-1.2,-0.9,-0.6,-0.3,0,0.3",
                    help="closure study: output-bias shifts tried with each output scale (comma-separated)")
    ap.add_argument("--variant-study", type=int, default=0, metavar="N",
                    help="time plain / LBGKTurbulence / closure step kernels on a periodic N^3 lattice")
    ap.add_argument("--cavity-study", action="store_true",
                    help="only measure the 64^3 lid-driven cavity (GPU and CPU oracle)")
    ap.add_argument("--storage-study", action="store_true",
                    help="the sphere channel with FP32, FP16, BF16 and FP64 population storage")
    ap.add_argument("--group-mb", type=int, default=-1, metavar="MB",
                    help="batch study: lattice-group size instead of the group autotuner's choice")
    ap.add_argument("--batch-study", action="store_true",
                    help="only run the batched-lattices study (64 cavities of uneven sizes)")
    ap.add_argument("--pattern-probe", action="store_true",
                    help="only run the memory-system probes: device copy and the bare AA access pattern")
    args = ap.parse_args()
    if args.pattern_probe:
        pattern_probe()
        return
    if args.cavity_study:
        cavity_study(args.no_cpu)
        return
    if args.storage_study:
        storage_study()
        return
    if args.variant_study:
        variant_study(args.variant_study)
        return
    if args.closure_study:
        closure_study(args.closure_study, args.tau_fine, args.chunk, args.train_snapshots, args.eval_chunks,
                      args.closure_target, {"none": False, "coarse": "coarse", "both": True}[args.ce_start],
                      tuple(float(v) for v in args.closure_shifts.split(",")), args.closure_seed)
        return
    if args.batch_study:
        batch_study(args.steps if args.steps != 5000 else 1000, args.warmup if args.warmup != 200 else 20,
                    args.group_mb if args.group_mb >= 0 else None)
        return
    if args.precision_study:
        precision_study(args.precision_study, [int(x) for x in args.snapshots.split(",")])
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.launch_study:
        launch_study()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this LBGK build runs on one GPU only; --gpus %d is not supported "
              "(the multi-GPU slab decomposition is a design, not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        sys.exit("bench.py: --steps must be >= 1 and --warmup >= 0")

    mapv, u = setup_fields()
    L = build_gpu(mapv, u)
    from asl_amd import _lib
    # the library's autotuner picks the workgroup size and XCD run length from
    # timed real steps (they advance the flow like the warm-up steps)
    if args.cpl >= 0:
        L.set_tuning(_lib.TUNE_CPL, args.cpl)
    if args.lds >= 0:
        L.set_tuning(_lib.TUNE_LDS, args.lds)
    block, remap = L.autotune(20) if args.autotune else L.tuning()
    if args.remap >= 0:
        L.set_tuning(_lib.TUNE_XCD_REMAP, args.remap)
        remap = args.remap
    cpl = L.tuning_value(_lib.TUNE_CPL)
    st = L.stats()
    counted = st["counted"]
    interior = int(np.prod(SIZE))
    if args.warmup:
        L.step(args.warmup, write_macro=False, graph=True)
    # Capture the graphs the timed steps replay (a chunk of 100 and the
    # remainder), for both starting parities when the step count is odd, so
    # that the timed region only replays them.
    for _ in range(2 if args.steps % 2 else 1):
        L.step(args.steps, write_macro=False, graph=True)
    L.synchronize()
    # the timed steps: every one a hipGraph replay
    t0 = time.perf_counter()
    ms_graph, _ = L.time_steps(args.steps, write_macro=False, graph=True)
    L.synchronize()
    wall = time.perf_counter() - t0
    mlups = counted * args.steps / wall / 1e6
    # the step kernel's own duration: HIP events on the lattice's stream
    # around each of KERNEL_LAUNCHES back-to-back eager launches
    _, _, each = L.time_steps(KERNEL_LAUNCHES, write_macro=False, each=True)

    # algorithmic bytes per launch: 19 FP32 reads + 19 FP32 writes per
    # computation node, plus the one-byte cell type of every interior cell
    bytes_per_launch = 152.0 * counted + 1.0 * interior
    kern_s = float(np.mean(each)) / 1e3
    achieved = bytes_per_launch / kern_s / 1e9
    traffic, traffic_src = profile_traffic()
    line = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "D3Q19 LBGK channel flow past a sphere, 256x128x128, FP32, tau=0.56, "
                               "sphere r=16 at (64,64,64), inlet u=0.04 (X0), outlet rho=1 (XE), "
                               "no-slip walls", "lattice": "D3Q19", "cells": interior,
                   "computation_nodes": counted, "storage": "FP32 SoA, in-place AA pattern",
                   "timed_steps_graph_replayed": args.steps, "timed_steps_eager": 0},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel_ms_avg": round(kern_s * 1e3, 5), "kernel_ms_min": round(float(np.min(each)), 5),
                     "kernel_ms_median": round(float(np.median(each)), 5), "kernel_launches_timed": KERNEL_LAUNCHES,
                     "bytes_per_launch": bytes_per_launch, "graph_ms_per_step": round(ms_graph / args.steps, 5),
                     "launch": {"workgroup": L.tuning_value(_lib.TUNE_BLOCK), "xcd_run": remap,
                                "cells_per_lane": cpl, "lds_tiles": L.tuning_value(_lib.TUNE_LDS)}},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(mapv, u)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
