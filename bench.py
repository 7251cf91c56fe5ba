#!/usr/bin/env python3
"""Benchmark of the short-range MD step on one MI355X.

Workload: Lennard-Jones fluid, N = 4,194,304 (a 128^3 simple-cubic lattice,
jittered U(-0.05, 0.05)), density 0.8442, eps = sigma = 1, rc = 2.5 (shift auto),
skin 0.3, dt = 0.005.  Equilibration: 1000 Langevin steps at T = 0.72, gamma = 1
(untimed, always), then W untimed NVE steps (--warmup), then K timed NVE steps.
A step is one velocity-Verlet step: integrate1, skin check, rebuild when
triggered, pair forces, integrate2.  NVE energy drift is the fitted slope of the
total energy over a separate run of --drift-steps NVE steps (default 10^4), for
both the mixed and the FP64 path.

Prints one JSON line: particle-steps/s of the mixed-precision path (value),
the all-FP64 GPU path on the same state, force accuracy of the mixed path
against the FP64 path, NVE energy drift, the HBM roofline fraction of the
dominant kernel (the pair-force kernel that also finishes the step; HIP events
on the library's stream) and a CPU baseline (the FP64
oracle, OpenMP over the host cores, on a bounded sample of the same system).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def lattice(n, rho, seed, jitter=0.05):
    L = (n / rho) ** (1.0 / 3.0)
    a = int(round(n ** (1.0 / 3.0)))
    if a ** 3 < n:
        a += 1
    sp = L / a
    g = np.arange(a, dtype=np.float64)
    ii, jj, kk = np.meshgrid(g, g, g, indexing="ij")
    sites = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)[:n]
    rng = np.random.default_rng(seed)
    return (sites + 0.5) * sp + rng.uniform(-jitter, jitter, size=(n, 3)), (L, L, L)


def velocities(n, T, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, math.sqrt(T), size=(n, 3))
    v -= v.mean(axis=0)
    v *= math.sqrt(T / (np.sum(v * v) / (3 * n)))
    return v


def make_system(espressopp, box, cells, rc, skin, precision, pos, vel):
    system = espressopp.System(precision=precision)
    system.rng = espressopp.esutil.RNG(2024)
    system.bc = espressopp.bc.OrthorhombicBC(system.rng, box)
    system.skin = skin
    system.storage = espressopp.storage.DomainDecomposition(system, (1, 1, 1), cells)
    n = len(pos)
    system.storage.addParticlesArray(np.arange(n, dtype=np.int32), pos, None, vel)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=rc)
    lj = espressopp.interaction.VerletListLennardJones(vl)
    lj.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, rc, "auto"))
    system.addInteraction(lj)
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    return system, vl, lj, integ


def total_energy(system, lj, n):
    T = system.storage._h.out_dbl("espp_kinetic")
    return lj.computeEnergy() + 0.5 * T


def timed_run(espressopp, system, integ, steps):
    lib = espressopp._lib.load()
    espressopp._lib.check(lib.espp_device_synchronize())
    t0 = time.perf_counter()
    integ.run(steps)
    espressopp._lib.check(lib.espp_device_synchronize())
    return time.perf_counter() - t0


def kernel_stats(h, name):
    import ctypes
    from espressopp_amd import _lib
    cnt, ms = ctypes.c_int64(0), ctypes.c_double(0.0)
    _lib.check(h.lib.espp_prof_stats(h.h, _lib.KERNEL[name], ctypes.byref(cnt), ctypes.byref(ms)))
    return cnt.value, ms.value


def vl_entries(h, vl):
    import ctypes
    from espressopp_amd import _lib
    e, cap, mx = ctypes.c_int64(0), ctypes.c_int32(0), ctypes.c_int32(0)
    _lib.check(h.lib.espp_vl_stats(h.h, vl._id, ctypes.byref(e), ctypes.byref(cap), ctypes.byref(mx)))
    return e.value, cap.value, mx.value


def drift_run(system, integ, lj, n, steps, chunk=500):
    """NVE energy drift: total energy every `chunk` steps over `steps` steps, least-squares
    slope, per particle per 1e5 steps (and the RMS of the residuals, per particle)."""
    t, e = [0], [total_energy(system, lj, n)]
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        integ.run(k)
        done += k
        t.append(done)
        e.append(total_energy(system, lj, n))
    t, e = np.asarray(t, dtype=np.float64), np.asarray(e)
    slope, icpt = np.polyfit(t, e, 1)
    resid = e - (slope * t + icpt)
    return dict(steps=steps, samples=len(t), per_particle_per_1e5_steps=slope / n * 1e5,
                fluctuation_rms_per_particle=float(np.sqrt(np.mean(resid ** 2)) / n),
                e_first_per_particle=float(e[0] / n), e_last_per_particle=float(e[-1] / n))


def cpu_baseline(pos, vel, box, cells, rc, skin, steps, threads):
    from oracle.oracle import OracleSim
    n = len(pos)
    s = OracleSim(box, cells, skin, threads=threads)
    s.add_particles(np.arange(n, dtype=np.int32), pos, vel=vel)
    s.decompose()
    ia = s.lj_create(s.vl_create(rc))
    s.lj_set(ia, 0, 0, 1.0, 1.0, rc, "auto")
    s.run(0, dt=0.005)                      # list build and first forces, untimed
    t0 = time.perf_counter()
    s.run(steps)
    dt = time.perf_counter() - t0
    return n * steps / dt, dt


def calibrate_copy(mib):
    """A plain elementwise copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel read
    against its mib MiB in and out give the scale of those counters on this GPU."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=0,
                    help="untimed NVE steps after the 1000-step Langevin equilibration")
    ap.add_argument("--equil", type=int, default=1000,
                    help="Langevin equilibration steps at T=0.72 (the workload's definition: 1000)")
    ap.add_argument("--drift-steps", type=int, default=10000,
                    help="NVE steps of the separate energy-drift run of each path (0: skip)")
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--prof-steps", type=int, default=100,
                    help="NVE steps after the timed run with per-kernel HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (counter calibration) and exit")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the short-range MD path here runs on a single MI355X "
              "(--gpus %d requested). Multi-GPU domain decomposition is not built." % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    import espressopp_amd as espressopp

    if espressopp._lib.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    n, rho, rc, skin = args.n, 0.8442, 2.5, 0.3
    pos, box = lattice(n, rho, seed=2024)
    vel = velocities(n, 0.72, seed=2025)
    cells = [int(b / (rc + skin)) for b in box]
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731

    # mixed path: Langevin equilibration (untimed), then timed NVE
    t_setup = time.perf_counter()
    system, vl, lj, integ = make_system(espressopp, box, cells, rc, skin, "mixed", pos, vel)
    log("setup mixed: %.2fs" % (time.perf_counter() - t_setup))
    lang = espressopp.integrator.LangevinThermostat(system)
    lang.gamma = 1.0
    lang.temperature = 0.72
    integ.addExtension(lang)
    t = time.perf_counter()
    integ.run(args.equil)
    log("equilibration %d Langevin steps: %.2fs" % (args.equil, time.perf_counter() - t))
    lang.disconnect()
    if args.warmup:
        integ.run(args.warmup)
    h = system.storage._h
    state = system.storage.getParticleArrays()

    # force accuracy of the mixed path against the FP64 path on the same state
    err = None
    fp64 = None
    if not args.no_fp64:
        sys64, vl64, lj64, integ64 = make_system(espressopp, box, cells, rc, skin, "fp64",
                                                 state["pos"] + state["image"] * np.asarray(box), state["v"])
        h.call("espp_compute_forces")
        sys64.storage._h.call("espp_compute_forces")
        f32 = system.storage.getParticleArrays()["f"]
        f64 = sys64.storage.getParticleArrays()["f"]
        nref = np.linalg.norm(f64, axis=1)
        floor = 0.1 * math.sqrt(np.mean(nref ** 2))
        rel = np.linalg.norm(f32 - f64, axis=1) / np.maximum(nref, floor)
        e32, e64 = lj.computeEnergy(), lj64.computeEnergy()
        w32, w64 = lj.computeVirial(), lj64.computeVirial()
        err = dict(force_rel_max=float(rel.max()), force_rel_p999=float(np.quantile(rel, 0.999)),
                   force_rel_median=float(np.median(rel)), force_floor=float(floor),
                   energy_rel=abs(e32 - e64) / abs(e64), virial_rel=abs(w32 - w64) / abs(w64))
        log("accuracy:", err)
        del f32, f64

    dt_run = timed_run(espressopp, system, integ, args.steps)
    value = n * args.steps / dt_run
    # per-kernel durations from HIP events on the library's stream, in a separate NVE run
    # (events are kept out of the timed run)
    h.call("espp_prof_reset")
    h.call("espp_prof_enable", 1)
    integ.run(args.prof_steps)
    h.call("espp_prof_enable", 0)
    # The dominant kernel: the LJ pair kernel that also finishes the step (every step of
    # a run but the last).  Algorithmic bytes per particle: own mirror 16 + 4 per list
    # entry (int32 index) + x, v, mass/type read 56 + x, v, mirror written 64; the
    # force itself stays in registers.
    launches, pair_ms = kernel_stats(h, "pair_step")
    entries, cap, maxcount = vl_entries(h, vl)
    nbar = entries / n
    b_force = 16 + 4 * nbar + 56 + 64
    avg_s = pair_ms / max(launches, 1) / 1e3
    achieved = b_force * n / avg_s / 1e9
    stats = {k: kernel_stats(h, k) for k in ("pair", "nbr_build", "integrate1", "integrate2", "binning")}
    resorts = integ.getNumResorts()
    import ctypes
    tile, region = ctypes.c_int32(0), ctypes.c_int32(0)
    espressopp._lib.check(h.lib.espp_tile_stats(h.h, ctypes.byref(tile), ctypes.byref(region)))
    log("tile plan: %d particles per tile, largest staged region %d particles" % (tile.value, region.value))
    lb, wb = ctypes.c_int64(0), ctypes.c_int64(0)
    espressopp._lib.check(h.lib.espp_vl_batch_stats(h.h, vl._id, 16, ctypes.byref(lb), ctypes.byref(wb)))
    log("16-entry batches of the pair loop: %.3f per lane, %.3f per wavefront (its longest lane)" %
        (lb.value / n, wb.value * 64.0 / n))
    log("mixed: %.3f ms/step, %.4g particle-steps/s, pair+step kernel %.3f ms avg (%d launches), nbar %.2f, "
        "cap %d, resorts %d" % (dt_run / args.steps * 1e3, value, avg_s * 1e3, launches, nbar, cap, resorts))
    for k, (c, ms) in stats.items():
        log("  %-11s %5d launches, %.3f ms avg" % (k, c, ms / max(c, 1)))
    drift = None
    if args.drift_steps:
        t = time.perf_counter()
        drift = drift_run(system, integ, lj, n, args.drift_steps)
        log("mixed drift over %d NVE steps (%.1fs): %s" % (args.drift_steps, time.perf_counter() - t, drift))

    if not args.no_fp64:
        k64 = args.fp64_steps or args.steps
        dt64 = timed_run(espressopp, sys64, integ64, k64)
        sys64.storage._h.call("espp_prof_reset")
        sys64.storage._h.call("espp_prof_enable", 1)
        integ64.run(min(args.prof_steps, k64))
        sys64.storage._h.call("espp_prof_enable", 0)
        l64, ms64 = kernel_stats(sys64.storage._h, "pair")
        fp64 = dict(value=n * k64 / dt64, ms_per_step=dt64 / k64 * 1e3, pair_kernel_ms=ms64 / max(l64, 1),
                    steps=k64)
        log("fp64: %.3f ms/step, %.4g particle-steps/s, pair kernel %.3f ms" %
            (fp64["ms_per_step"], fp64["value"], fp64["pair_kernel_ms"]))
        if args.drift_steps:
            t = time.perf_counter()
            fp64["drift"] = drift_run(sys64, integ64, lj64, n, args.drift_steps)
            log("fp64 drift over %d NVE steps (%.1fs): %s" % (args.drift_steps, time.perf_counter() - t,
                                                              fp64["drift"]))
        del sys64, vl64, lj64, integ64

    cpu = None
    if not args.no_cpu:
        # every host core this process may use: the affinity set, capped by the OpenMP thread
        # count the host grants one GPU's job (OMP_NUM_THREADS, when set)
        affinity = len(os.sched_getaffinity(0))
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = min(affinity, omp) if omp > 0 else affinity
        cval, ctime = cpu_baseline(state["pos"] + state["image"] * np.asarray(box), state["v"], box, cells, rc,
                                   skin, args.cpu_steps, threads)
        cpu = {"value": cval, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "host_nproc": os.cpu_count(), "affinity_cpus": affinity, "omp_num_threads": omp or None,
               "sample": "same 4,194,304-particle LJ state after equilibration, %d NVE steps of the FP64 "
                         "oracle (half Verlet list, Newton-3, OpenMP on %d threads) = %.1f s"
                         % (args.cpu_steps, threads, ctime)}
        log("cpu:", cpu)

    out = {
        "metric": "particle-steps/s, LJ 4M: mixed vs FP64 GPU path at force rel err <=1e-5",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "equilibration_steps": args.equil,
        "ms_per_step": dt_run / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32-fixed/fp64",
        "data": "synthetic: jittered simple-cubic lattice, Maxwell-Boltzmann velocities, Langevin equilibration",
        "config": {"workload": "LJ fluid N=%d rho=0.8442 rc=2.5 skin=0.3 dt=0.005 NVE after %d-step "
                               "Langevin equilibration at T=0.72 (+%d untimed NVE steps)"
                               % (n, args.equil, args.warmup),
                   "n": n, "box": box[0], "cells": cells, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_lj_tiled_mixed (step variant)", "bytes_per_particle": b_force, "nbar": nbar,
                     "avg_launch_ms": avg_s * 1e3, "launches": launches},
        "fp64_path": fp64,
        "speedup_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "accuracy": err,
        "nve_drift_per_particle_per_1e5_steps": drift["per_particle_per_1e5_steps"] if drift else None,
        "nve_drift": drift,
        "resorts": resorts,
        "cpu_baseline": cpu,
    }
    traffic_file = os.path.join(ROOT, "profiles", "lj_mixed_4M_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            out["roofline"]["traffic"] = json.load(f).get("bytes_per_launch")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
