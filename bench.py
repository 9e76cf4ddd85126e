#!/usr/bin/env python3
"""Benchmark of the MD-Bench short-range force step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line.  A step is one velocity-Verlet MD timestep of the whole system (integrate,
ghost update, force, integrate; a neighbour rebuild every reneigh_every steps), with every array
already resident in HBM.  `value` = atoms x steps / wall time of the K timed steps.

Workloads (seeded fcc lattices, synthetic velocities):
  cp-lj-64   (default) cluster-pair LJ, i-clusters of 4 against j-clusters of 4, FP32 pair
             math with FP64 positions/velocities, full list, 64^3 cells = 1,048,576 atoms,
             rho* 0.8442, r_c 2.5, skin 0.3, T 1.44, dt 0.005, reneighbour every 20 steps.
  cp-lj-64-4x8, cp-lj-64-half: the same with j-clusters of 8 / a half list.
  vl-lj-64   Verlet-list LJ, FP64, full list, same lattice.
             vl-lj-64-sort100: the same with the atoms re-sorted by bin every 100 steps
             (MD-Bench's SORT_ATOMS option).
  eam-32     Verlet-list EAM, FP64: 32^3 Cu cells (131,072 atoms), a 3.615 A, synthetic funcfl
             tables (nr = nrho = 2000, rc 4.95 A), seeded Maxwell-Boltzmann velocities at
             1600 K, dt 1 fs, reneighbour every 10 steps.
             eam-64: the same at 64^3 cells (1,048,576 atoms); eam-64-sort100 with the atoms
             re-sorted by bin every 100 steps.
  ka-128     Kob-Andersen 80/20 binary LJ mixture, FP32-mixed cluster pair 4x4: 128^3 cells
             (8,388,608 atoms), rho 1.2, T 0.8, per-pair cutoff 2.5 sigma_ab.
  eam-surrogate-64  the learned EAM surrogate: trained on 2000 plain-EAM snapshots of 4000-atom
             boxes, then run on 64^3 Cu cells (1,048,576 atoms) beside plain EAM.
  stub-seq, stub-fix, stub-rand  MD-Bench's synthetic-pattern force benchmark: 1M atoms on a
             line, 76 neighbours in a sequential, fixed or random pattern, force kernel only.
  hbm-copy   a plain 256 MiB streaming kernel: the scale factor for the HBM byte counters.

`roofline` covers the dominant kernel (the force kernel): algorithmic bytes per launch over its
average launch time, both measured inside the timed region with HIP events on the library stream.
`cpu_baseline` times MD-Bench itself (built from its sources into oracle/_ref) on the host cores
on a bounded sample of the same workload.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
METRIC = "atom-steps/s (LJ, EAM) + surrogate/plain EAM speedup at fixed force RMSE"
HBM_PEAK_GBS = 8000.0
# HIP events bracket every EVENT_EVERY-th launch of the reported kernel inside the timed region
# (each event pair costs ~5 us of idle GPU between kernels, 2 % of a 1M-atom cluster step);
# runs of fewer than 100 steps sample every 5th launch, so that a 20-step run has 4 samples
EVENT_EVERY = 10


def event_every(steps):
    return EVENT_EVERY if steps >= 100 else 5
# EAM workloads start from seeded Maxwell-Boltzmann velocities at 1600 K
MAXWELL_SEED = 1600


def binding():
    import importlib.util
    path = os.path.join(REPO, "md-bench_amd", "mdbench_gfx950.py")
    spec = importlib.util.spec_from_file_location("mdbench_gfx950", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mdbench_gfx950"] = mod
    spec.loader.exec_module(mod)
    return mod


def host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def host_cores():
    """Threads for the CPU baseline: every CPU this process may run on, up to 16.  A one-GPU job
    on the GPU hosts gets 16 CPUs (its OMP_NUM_THREADS is 16 there) whatever the machine shows,
    so the baseline uses that share rather than every physical core of the socket."""
    return max(1, min(host_threads(), 16))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_flags():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except Exception:
        pass
    return set()


def run_reference(binary, args, cores):
    """Time MD-Bench's own binary; returns atom-steps/s from its 'atom updates/us' line."""
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_SCHEDULE="static",
               OMP_PROC_BIND="close")
    out = subprocess.run([binary] + args, env=env, capture_output=True, text=True, timeout=300,
                         cwd="/tmp", errors="replace")
    m = re.search(r"([0-9.]+) atom updates/us", out.stdout)
    if out.returncode != 0 or not m:
        raise RuntimeError("reference run failed: %s" % out.stdout[-500:])
    return float(m.group(1)) * 1e6


def cpu_baseline_vl(nx, steps):
    cores = host_cores()
    ref = os.path.join(REPO, "oracle", "_ref", "mdbench-vl-fast")
    sample = "%d^3 cells (%d atoms), %d timesteps, full list, MD-Bench -Ofast scalar FP64" % (
        nx, 4 * nx ** 3, steps)
    if os.path.exists(ref):
        v = run_reference(ref, ["-nx", str(nx), "-ny", str(nx), "-nz", str(nx), "-n",
                                str(steps)], cores)
        return {"value": v, "unit": "atom-steps/s", "cores": cores, "kind": "reference",
                "sample": sample}
    return None


def bench_vl(args, mod, resort=0):
    nx = 64
    sim = mod.VLSim(nx=nx, ny=nx, nz=nx, ntimes=args.steps + args.warmup, resort_every=resort)
    try:
        natoms = sim.atom.Natoms
        sim.run(args.warmup)
        L = mod.lib()
        L.mdgfx950Synchronize()
        L.mdgfx950TimeOnly(b"force_lj")  # events around the reported kernel only
        L.mdgfx950TimeEvery(event_every(args.steps))
        L.mdgfx950EnableTiming(1)
        t0 = time.perf_counter()
        sim.run(args.steps)
        L.mdgfx950Synchronize()
        t1 = time.perf_counter()
        import ctypes as C
        ms = C.c_double()
        nl = C.c_longlong()
        L.mdgfx950KernelTime(b"force_lj", C.byref(ms), C.byref(nl))
        L.mdgfx950EnableTiming(0)
        L.mdgfx950TimeOnly(None)
        L.mdgfx950TimeEvery(1)
        st = sim.state()
        sum_nn = int(st["numneigh"].sum())
        nlocal = st["nlocal"]
        nghost = int(st["nghost"])
        elapsed = t1 - t0
        # HBM roofline on the launch's compulsory bytes: per atom x_i (24 B), f_i (24 B),
        # numneigh (4 B) and the integration the run loop's launches also do (v read and
        # written, the new position written: 72 B); per neighbour its index (4 B); each ghost
        # position once (24 B).  A neighbour's position is a re-read of that set, served by L2
        # and the Infinity Cache.  MD-Bench's data-volume model (src/verletlist/stats.c:56-59)
        # counts it as 24 B per neighbour; its rate is reported beside, as a gather rate, not
        # an HBM figure (it exceeds the HBM peak).
        bytes_launch = nlocal * (52.0 + 72.0) + sum_nn * 4.0 + nghost * 24.0
        model_bytes = nlocal * (52.0 + 72.0) + sum_nn * 28.0
        avg_ms = ms.value / max(nl.value, 1)
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        res = {
            "metric": METRIC, "value": natoms * args.steps / elapsed, "unit": "atom-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "vl-lj-64%s: Verlet-list LJ FP64 full list, 64^3 fcc cells "
                                   "(%d atoms), rho 0.8442, rc 2.5, skin 0.3, reneigh 20%s"
                                   % ("-sort%d" % resort if resort else "", natoms,
                                      ", atoms re-sorted by bin every %d steps" % resort
                                      if resort else ""),
                       "atoms": natoms, "avg_neighbors": sum_nn / nlocal},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "k_force_lj (fast, 4 lanes/atom)",
                         "avg_launch_ms": avg_ms, "launches": nl.value,
                         "bytes_per_launch": bytes_launch, "bytes_basis": "compulsory",
                         "model_bytes_per_launch": model_bytes,
                         "model_gather_gbs": model_bytes / (avg_ms * 1e-3) / 1e9},
        }
        return res, nx
    finally:
        sim.close()


def cpu_baseline_cp(nx, steps):
    cores = host_cores()
    flags = cpu_flags()
    name = "mdbench-cp-avx512" if "avx512f" in flags else "mdbench-cp-avx2"
    ref = os.path.join(REPO, "oracle", "_ref", name)
    if not os.path.exists(ref):
        return None
    v = run_reference(ref, ["-nx", str(nx), "-ny", str(nx), "-nz", str(nx), "-n", str(steps)],
                      cores)
    return {"value": v, "unit": "atom-steps/s", "cores": cores, "kind": "reference",
            "sample": "%d^3 cells (%d atoms), %d timesteps, MD-Bench cluster-pair 4xN SIMD "
                      "(%s, -Ofast, 'DP' with approximate reciprocal), full list"
                      % (nx, 4 * nx ** 3, steps, name.split("-")[-1])}


def bench_cp(args, mod, N=4, half=0):
    import ctypes as C
    nx = 64
    sim = mod.CPSim(nx=nx, ny=nx, nz=nx, ntimes=args.steps + args.warmup, cluster_m=4,
                    cluster_n=N, half_neigh=half, precision=0)
    try:
        natoms = 4 * nx ** 3
        sim.run(args.warmup)
        # total energy per atom (potential from one energy-on force call at the same positions,
        # kinetic 1/2 sum m v^2) before and after the timed steps, outside the timed region
        def etot():
            return (sim.energy_virial()[0] + 0.5 * sim.kinetic_sum()) / natoms
        e_start = etot()
        L = mod.lib()
        L.mdgfx950Synchronize()
        L.mdgfx950TimeOnly(b"force_cp")  # events around the reported kernel only
        L.mdgfx950TimeEvery(event_every(args.steps))
        L.mdgfx950EnableTiming(1)
        t0 = time.perf_counter()
        sim.run(args.steps)
        L.mdgfx950Synchronize()
        t1 = time.perf_counter()
        ms = C.c_double()
        nl = C.c_longlong()
        L.mdgfx950KernelTime(b"force_cp", C.byref(ms), C.byref(nl))
        L.mdgfx950EnableTiming(0)
        L.mdgfx950TimeOnly(None)
        L.mdgfx950TimeEvery(1)
        e_end = etot()
        info = sim.info()
        pairs = int(sim.get("numneigh").astype(np.int64).sum())
        nci = info["nci"]
        elapsed = t1 - t0
        # Algorithmic bytes of one FP32-mixed cluster-pair force launch, the per-unit figures of
        # the gather-bound model: per i-cluster the list count (4 B), 4 FP32 positions (48 B) and
        # 4 FP32 forces (48 B); per cluster pair the j-cluster index (4 B) and N FP32 positions
        # (12 N B); a half list adds the read-modify-write of N j-forces (24 N B) per pair.  The
        # full-list launches of the run loop also integrate (final kick, next half kick and
        # drift): per atom the FP64 velocity read and written and the new position written
        # (72 B).  (The kernel reads 16-byte FP32 records: cell-relative offsets and the cell
        # index and type.  The tiled half-list pass sums the j-forces of a tile in LDS and
        # writes them once per tile; a gather pass adds them per atom and integrates.)
        bytes_launch = nci * (4.0 + 48.0 + 48.0) + pairs * (4.0 + 12.0 * N + (24.0 * N if half else 0.0))
        if not half:
            bytes_launch += 72.0 * natoms
        avg_ms = ms.value / max(nl.value, 1)
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        wl = "cp-lj-64%s%s" % ("-4x8" if N == 8 else "", "-half" if half else "")
        res = {
            "metric": METRIC, "value": natoms * args.steps / elapsed, "unit": "atom-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 pair math, f64 state",
            "data": "synthetic",
            "config": {"workload": "%s: LJ cluster-pair 4x%d %s list, FP32-mixed, 64^3 fcc "
                                   "cells (%d atoms), rho 0.8442, rc 2.5, skin 0.3, reneigh 20, plain "
                                   "(uncompensated) FP32 force sums"
                                   % (wl, N, "half" if half else "full", natoms),
                       "atoms": natoms, "i_clusters": nci,
                       "j_clusters_per_i_cluster": pairs / nci,
                       "compensated": 0,
                       "energy_per_atom": {"start": e_start, "end": e_end,
                                           "drift": e_end - e_start,
                                           "over_steps": args.steps}},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "kernel": ("k_cp_force_half_tiled + k_cp_gather_half (j-forces "
                                    "of a tile summed in LDS, no atomics)" if half else
                                    "k_cp_force_f32_full (four i-clusters per wavefront, one "
                                    "j-atom per lane)"),
                         "avg_launch_ms": avg_ms, "launches": nl.value,
                         "bytes_per_launch": bytes_launch},
        }
        return res, nx
    finally:
        sim.close()


def cpu_baseline_eam(nx, steps):
    cores = host_cores()
    ref = os.path.join(REPO, "oracle", "_ref", "mdbench-vl-fast")
    if not os.path.exists(ref):
        return None
    import tempfile
    mod = binding()
    path = os.path.join(tempfile.gettempdir(), "mdb_bench_cu_%d.eam" % os.getpid())
    mod.lib().mdbWriteFuncflSynthetic(path.encode())
    try:
        v = run_reference(ref, ["-f", "eam", "-e", path, "-nx", str(nx), "-ny", str(nx), "-nz",
                                str(nx), "-n", str(steps)], cores)
    finally:
        os.unlink(path)
    return {"value": v, "unit": "atom-steps/s", "cores": cores, "kind": "reference",
            "sample": "%d^3 Cu cells (%d atoms), %d timesteps, MD-Bench EAM -Ofast scalar FP64 "
                      "(its own EAM setup: a 3.6148 A, 600 K), same synthetic funcfl table"
                      % (nx, 4 * nx ** 3, steps)}


def bench_eam(args, mod, nx=32, resort=0):
    import ctypes as C
    sim = mod.VLSim(nx=nx, ny=nx, nz=nx, force_field=1, lattice_a=3.615, temp=1600.0,
                    maxwell_seed=MAXWELL_SEED, reneigh_every=10,
                    ntimes=args.steps + args.warmup, resort_every=resort)
    try:
        natoms = sim.atom.Natoms
        sim.run(args.warmup)
        L = mod.lib()
        L.mdgfx950Synchronize()
        L.mdgfx950TimeOnly(b"force_eam")  # events around the reported kernel only
        L.mdgfx950TimeEvery(event_every(args.steps))
        L.mdgfx950EnableTiming(1)
        t0 = time.perf_counter()
        sim.run(args.steps)
        L.mdgfx950Synchronize()
        t1 = time.perf_counter()
        ms = C.c_double()
        nl = C.c_longlong()
        L.mdgfx950KernelTime(b"force_eam", C.byref(ms), C.byref(nl))
        L.mdgfx950EnableTiming(0)
        L.mdgfx950TimeOnly(None)
        L.mdgfx950TimeEvery(1)
        st = sim.state()
        sum_nn = int(st["numneigh"].sum())
        nlocal = st["nlocal"]
        elapsed = t1 - t0
        # Algorithmic bytes of one EAM force call (two passes over the list; spline tables stay in
        # cache and are not counted).  Density: per atom x_i, count, fp_i write (36 B), per listed
        # neighbour index + x_j (28 B).  Force: per atom x_i, fp_i, count, f_i write (60 B), per
        # listed neighbour index + x_j + fp_j (36 B).  The run loop's force pass also integrates
        # (per atom v read and written, the new position written: 72 B).
        bytes_launch = nlocal * (36.0 + 60.0 + 72.0) + sum_nn * (28.0 + 36.0)
        avg_ms = ms.value / max(nl.value, 1)
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        res = {
            "metric": METRIC, "value": natoms * args.steps / elapsed, "unit": "atom-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "eam-%d%s: Verlet-list EAM FP64, %d^3 Cu fcc cells (%d atoms), "
                                   "a 3.615 A, synthetic funcfl nr=nrho=2000 rc 4.95 A, Maxwell 1600 K, "
                                   "dt 1 fs, reneigh 10%s"
                                   % (nx, "-sort%d" % resort if resort else "", nx, natoms,
                                      ", atoms re-sorted by bin every %d steps" % resort
                                      if resort else ""),
                       "atoms": natoms, "avg_neighbors": sum_nn / nlocal},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "k_eam_density_tab + k_eam_force_tab (4 lanes/atom, spline rows in LDS)",
                         "avg_launch_ms": avg_ms, "launches": nl.value,
                         "bytes_per_launch": bytes_launch},
        }
        return res, nx
    finally:
        sim.close()


KA = dict(ntypes=2, pair_epsilon=[1.0, 1.5, 1.5, 0.5], pair_sigma=[1.0, 0.8, 0.8, 0.88],
          pair_cutoff=[2.5, 2.0, 2.0, 2.2], rho=1.2, temp=0.8, type1_fraction=0.2, type_seed=4711)


def cpu_baseline_ka(nx, steps):
    """The oracle's FP64 Verlet-list loop (one core) on a sub-box of the mixture: MD-Bench itself
    has no per-pair sigma table, so it cannot run this mixture."""
    import ctypes as C
    so = os.path.join(REPO, "oracle", "liboracle_md.so")
    if not os.path.exists(so):
        return None
    mod = binding()
    L = C.CDLL(so)
    x, v, t = mod.create_lattice(nx=nx, ny=nx, nz=nx, rho=1.2, temp=0.8)
    rng = np.random.default_rng(4711)
    t = (rng.random(t.size) < 0.2).astype(np.int32)
    eps = np.array(KA["pair_epsilon"])
    sig6 = np.array(KA["pair_sigma"]) ** 6
    cutf = np.array(KA["pair_cutoff"]) ** 2
    cutn = np.full(4, 2.8 ** 2)
    lat = (4.0 / 1.2) ** (1.0 / 3.0)
    f = np.zeros_like(x)
    ks = np.zeros(steps + 1)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    d = lambda a: a.ctypes.data_as(dp)  # noqa: E731
    L.orc_vl_run.restype = C.c_int
    t0 = time.perf_counter()
    rc = L.orc_vl_run(C.c_int(nx), C.c_int(nx), C.c_int(nx), C.c_double(lat), C.c_int(t.size),
                      d(x), d(v), t.ctypes.data_as(ip), C.c_int(2), d(eps), d(sig6), d(cutf),
                      d(cutn), C.c_double(2.8), C.c_int(0), C.c_int(0), None, C.c_double(0.005),
                      C.c_double(0.0025), C.c_int(steps), C.c_int(20), C.c_double(1.0), d(f),
                      d(ks))
    t1 = time.perf_counter()
    if rc != 0:
        return None
    return {"value": t.size * steps / (t1 - t0), "unit": "atom-steps/s", "cores": 1,
            "kind": "port",
            "sample": "%d^3 cells (%d atoms) of the same mixture, %d timesteps, the FP64 "
                      "Verlet-list oracle (strict IEEE C, one core)" % (nx, t.size, steps)}


def bench_ka(args, mod, nx=128, N=4):
    import ctypes as C
    sim = mod.CPSim(nx=nx, ny=nx, nz=nx, ntimes=args.steps + args.warmup, cluster_m=4,
                    cluster_n=N, precision=0, **KA)
    try:
        natoms = 4 * nx ** 3
        sim.run(args.warmup)
        L = mod.lib()
        L.mdgfx950Synchronize()
        L.mdgfx950TimeOnly(b"force_cp")  # events around the reported kernel only
        L.mdgfx950TimeEvery(event_every(args.steps))
        L.mdgfx950EnableTiming(1)
        t0 = time.perf_counter()
        sim.run(args.steps)
        L.mdgfx950Synchronize()
        t1 = time.perf_counter()
        ms = C.c_double()
        nl = C.c_longlong()
        L.mdgfx950KernelTime(b"force_cp", C.byref(ms), C.byref(nl))
        L.mdgfx950EnableTiming(0)
        L.mdgfx950TimeOnly(None)
        L.mdgfx950TimeEvery(1)
        info = sim.info()
        pairs = int(sim.get("numneigh").astype(np.int64).sum())
        nci = info["nci"]
        elapsed = t1 - t0
        # as for cp-lj-64 (integration included), plus a 1-byte type per atom: i (4 B per
        # cluster) and j (N B per pair)
        bytes_launch = nci * (4.0 + 48.0 + 48.0 + 4.0) + pairs * (4.0 + 13.0 * N) + 72.0 * natoms
        avg_ms = ms.value / max(nl.value, 1)
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        res = {
            "metric": METRIC, "value": natoms * args.steps / elapsed, "unit": "atom-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 pair math, f64 state",
            "data": "synthetic",
            "config": {"workload": "ka-%d: Kob-Andersen 80/20 LJ mixture, cluster-pair 4x%d full "
                                   "list, FP32-mixed, %d^3 fcc cells (%d atoms), rho 1.2, T 0.8, "
                                   "rc 2.5 sigma_ab, skin 0.3, reneigh 20" % (nx, N, nx, natoms),
                       "atoms": natoms, "i_clusters": nci,
                       "j_clusters_per_i_cluster": pairs / nci},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "k_cp_force_f32_full (type-pair table in LDS)",
                         "avg_launch_ms": avg_ms, "launches": nl.value,
                         "bytes_per_launch": bytes_launch},
        }
        return res, 16
    finally:
        sim.close()


def bench_surrogate(args, mod, nx=64):
    """Learned EAM surrogate at 64^3 Cu cells: trained here on plain-EAM snapshots (2000 frames of
    10^3-cell boxes at 300-2000 K, every 50 steps, 90/10 split), then `steps` MD steps with the
    surrogate as the force call and a plain-EAM judge every 50 steps (fallback above 30 meV/A),
    beside the same number of plain-EAM steps on the same system."""
    import ctypes as C
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "eam_surrogate", os.path.join(REPO, "md-bench_amd", "eam_surrogate.py"))
    es = importlib.util.module_from_spec(spec)
    sys.modules["eam_surrogate"] = es
    spec.loader.exec_module(es)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    model, info, data = es.train_surrogate(temps=np.linspace(300.0, 2000.0, 20), per_temp=100,
                                           every=50, nx=10, epochs=args.sur_epochs,
                                           batch_frames=16, log=log)
    frame = data[2][0]
    del data
    L = mod.lib()
    sim = mod.VLSim(nx=nx, ny=nx, nz=nx, force_field=1, lattice_a=3.615, temp=1600.0,
                    maxwell_seed=MAXWELL_SEED, reneigh_every=10, ntimes=0)
    try:
        natoms = sim.atom.Natoms
        # plain EAM on this system
        sim.run(args.warmup)
        L.mdgfx950Synchronize()
        t0 = time.perf_counter()
        sim.run(args.steps)
        L.mdgfx950Synchronize()
        plain = natoms * args.steps / (time.perf_counter() - t0)
        # the surrogate, judged every 50 steps
        es.use_in(sim, model, judge_every=50, tol=0.030)
        sim.run(args.warmup)
        L.mdgfx950Synchronize()
        L.mdgfx950TimeOnly(b"surrogate")  # events around the reported kernel only
        L.mdgfx950TimeEvery(event_every(args.steps))
        L.mdgfx950EnableTiming(1)
        t0 = time.perf_counter()
        sim.run(args.steps)
        L.mdgfx950Synchronize()
        t1 = time.perf_counter()
        ms = C.c_double()
        nl = C.c_longlong()
        L.mdgfx950KernelTime(b"surrogate", C.byref(ms), C.byref(nl))
        L.mdgfx950EnableTiming(0)
        L.mdgfx950TimeOnly(None)
        L.mdgfx950TimeEvery(1)
        judge = es.judge_stats(sim)
        elapsed = t1 - t0
        value = natoms * args.steps / elapsed
        st = sim.state()
        nlocal = st["nlocal"]
        xr = st["x"].reshape(-1, 3)
        nn = st["numneigh"].astype(np.int64)
        ii = np.repeat(np.arange(nlocal), nn)
        d = xr[st["neighbors"][:nn.sum()]] - xr[ii]
        pairs = int(((d * d).sum(1) < es.RC ** 2).sum())
        # Algorithmic bytes of one surrogate force call (descriptor, network and force passes
        # together), the per-unit figures of the gather-bound model for this workload: per
        # listed neighbour one gather of its index and FP32 position (16 B, one pass over the
        # list); per atom its position (24 B), 32 descriptors and 36 moments written and read
        # back (2 x 272 B), its coefficient row written and read (2 x 128 B) and its force
        # (24 B).  The implementation reads more than this (the list and positions in both
        # passes, FP64 positions, a neighbour's coefficient row per pair inside rc); that
        # excess is what keeps it below the bound.
        per_atom = 24 + 2 * 4 * (32 + 36) + 2 * 128 + 24
        bytes_launch = nn.sum() * 16.0 + nlocal * per_atom
        impl_bytes = (nn.sum() * 56.0 + pairs * (8.0 + 128.0)
                      + nlocal * (48 + 2 * 4 * (32 + 36) + 2 * 128 + 24))
        regions = max(nl.value, 1)
        avg_ms = 3.0 * ms.value / regions  # three timed launches per force call
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        res = {
            "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 pairs, bf16 MFMA network",
            "data": "synthetic",
            "config": {"workload": "eam-surrogate-64: learned EAM surrogate (24 radial + 8 angular "
                                   "descriptors, MLP 32-64-64-1) on %d^3 Cu cells (%d atoms), "
                                   "Maxwell 1600 K start, plain-EAM judge every 50 steps" % (nx, natoms),
                       "atoms": natoms,
                       "plain_eam_atom_steps_per_s": plain,
                       "speedup_vs_plain_eam": value / plain,
                       "judge": judge,
                       "train": {k: v for k, v in info.items() if k != "history"}},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "k_sur_desc + k_sur_mlp + k_sur_force",
                         "avg_launch_ms": avg_ms, "launches": regions // 3,
                         "bytes_per_launch": bytes_launch,
                         "implementation_bytes_per_launch": impl_bytes,
                         "pairs_inside_rc": pairs, "listed_neighbours": int(nn.sum())},
        }
        if not args.no_cpu_baseline:
            res["cpu_baseline"] = cpu_baseline_surrogate(model, frame)
        return res, None
    finally:
        sim.close()


def cpu_baseline_surrogate(model, frame, reps=3):
    """The oracle's FP64 numpy statement of the same model (descriptors, network, forces) on one
    training frame, timed as force evaluations per second x atoms."""
    sys.path.insert(0, REPO)
    from oracle import surrogate as so
    p = so.Params()
    p.set_flat(model.theta.astype(np.float64))
    p.d_mean, p.d_std = model.d_mean.astype(np.float64), model.d_std.astype(np.float64)
    p.e0, p.s_e = model.e0, model.s_e
    p.r_core, p.k_core = model.r_core, model.k_core
    t0 = time.perf_counter()
    for _ in range(reps):
        so.energy_and_forces(p, frame["x"], frame["nlocal"], frame["owner"], frame["numneigh"],
                             frame["neighbors"])
    dt = (time.perf_counter() - t0) / reps
    return {"value": frame["nlocal"] / dt, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": "%d-atom Cu frame, %d force evaluations of the FP64 numpy oracle of the "
                      "same model (force call only)" % (frame["nlocal"], reps)}


def bench_stub(args, mod, pattern="seq"):
    """MD-Bench's synthetic-pattern benchmark (main-stub): 1,048,576 atoms on a line, 76
    neighbours each in the given pattern, the FP64 LJ force kernel alone."""
    natoms, nn = 1 << 20, 76
    r = mod.stub_run(pattern, natoms, nn, 1, ntimes=max(args.steps // 5, 20), strict=0)
    # HBM roofline on the compulsory bytes: per atom 6 doubles (x read, f written) and
    # (nneighs + 2) ints; the neighbour positions are re-reads of x (L2/Infinity Cache).
    # MD-Bench's data-volume model (src/verletlist/main-stub.c:244-249, stats.c:56-59) adds
    # 24 B per neighbour position; its rate is reported beside, as a gather rate.
    bytes_call = natoms * (6 * 8.0 + (nn + 2) * 4.0)
    model_bytes = bytes_call + natoms * nn * 24.0
    achieved = bytes_call / (r["ms_per_call"] * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": r["atom_updates_per_s"], "unit": "atom-steps/s",
        "n_gpus": 1, "steps": max(args.steps // 5, 20), "warmup": 1,
        "ms_per_step": r["ms_per_call"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "stub-%s: synthetic neighbour pattern '%s', %d atoms x %d "
                               "neighbours, force kernel only" % (pattern, pattern, natoms, nn),
                   "atoms": natoms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_force_lj (fast, 4 lanes/atom)",
                     "avg_launch_ms": r["ms_per_call"], "bytes_per_launch": bytes_call,
                     "bytes_basis": "compulsory", "model_bytes_per_launch": model_bytes,
                     "model_gather_gbs": model_bytes / (r["ms_per_call"] * 1e-3) / 1e9},
    }
    return res, None


def bench_copy(args, mod):
    """Plain streaming kernel of known size, the scale for HBM counters: each launch of
    torch.mul(src, 1.0, out=dst) reads and writes 256 MiB (2^26 FP32 values each way)."""
    import torch
    n = 1 << 26
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(max(args.warmup, 1)):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        torch.mul(src, 1.0, out=dst)
    ev1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = ev0.elapsed_time(ev1) / args.steps
    nbytes = 2.0 * 4 * n
    gbs = nbytes / (ms * 1e-3) / 1e9
    res = {"metric": "HBM copy bandwidth (counter scale)", "value": gbs, "unit": "GB/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * (t1 - t0) / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "hbm-copy: 2^26 FP32 read + written per launch",
                      "bytes_per_launch": nbytes},
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "traffic": nbytes}}
    return res, None


import numpy as np  # noqa: E402


def attach_traffic(res, workload):
    """HBM bytes per launch of the reported kernel, from the two counter passes summarised in
    profiles/ (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads for a plain copy
    of known size); stays null for a workload without such a measurement."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    roof = res.get("roofline")
    if not roof or not os.path.exists(path):
        return
    with open(path) as fh:
        t = json.load(fh).get(workload)
    if t:
        roof["traffic"] = t["bytes"]
        roof["traffic_source"] = "profiles/" + t["source"]


WORKLOADS = {
    "cp-lj-64": bench_cp,
    "cp-lj-64-4x8": lambda a, m: bench_cp(a, m, N=8),
    "cp-lj-64-half": lambda a, m: bench_cp(a, m, half=1),
    "vl-lj-64": bench_vl,
    "vl-lj-64-sort100": lambda a, m: bench_vl(a, m, resort=100),
    "eam-32": bench_eam,
    "eam-64": lambda a, m: bench_eam(a, m, nx=64),
    "eam-64-sort100": lambda a, m: bench_eam(a, m, nx=64, resort=100),
    "ka-128": bench_ka,
    "eam-surrogate-64": bench_surrogate,
    "stub-seq": lambda a, m: bench_stub(a, m, "seq"),
    "stub-fix": lambda a, m: bench_stub(a, m, "fix"),
    "stub-rand": lambda a, m: bench_stub(a, m, "rand"),
    "hbm-copy": bench_copy,
}
DEFAULT = "cp-lj-64"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sur-epochs", type=int, default=300,
                    help="training epochs of eam-surrogate-64")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; a multi-GPU (spatial domain "
              "decomposition) run is not part of it. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    mod = binding()
    res, nx = WORKLOADS[args.workload](args, mod)
    attach_traffic(res, args.workload)
    cb = None
    if not args.no_cpu_baseline and nx is not None:
        try:
            if args.workload.startswith("vl"):
                cb = cpu_baseline_vl(nx, 20)
            elif args.workload.startswith("eam"):
                cb = cpu_baseline_eam(min(nx, 32), 20)
            elif args.workload.startswith("ka"):
                cb = cpu_baseline_ka(nx, 100)
            else:
                cb = cpu_baseline_cp(nx, 20)
        except Exception as exc:  # the GPU figure stands on its own
            print("cpu baseline failed: %s" % exc, file=sys.stderr)
    res.setdefault("cpu_baseline", cb)
    if res["cpu_baseline"] is None:
        res["cpu_baseline"] = cb
    if res["cpu_baseline"]:
        res["cpu_baseline"]["cpu_model"] = cpu_model()
        res["cpu_baseline"]["host_threads_visible"] = host_threads()
        res["cpu_baseline"]["cores_basis"] = (
            "the CPU share of a one-GPU job on the GPU host: at most 16 threads (the host "
            "sets OMP_NUM_THREADS to 16 there), not every core of the socket")
    print(json.dumps(res, default=float))


if __name__ == "__main__":
    main()
