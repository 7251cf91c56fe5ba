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


def prune_stats(h, ia):
    import ctypes
    from espressopp_amd import _lib
    p, w, e = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    _lib.check(h.lib.espp_prune_stats(h.h, ia._id, ctypes.byref(p), ctypes.byref(w), ctypes.byref(e)))
    return p.value, w.value, e.value


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


def c4_batched_pt(steps=500, every=100, seed=2026):
    """64 parallel-tempering replicas, N uniform in [2048, 65536] (seeded), density 0.8442,
    temperatures geometric 0.7-1.5, rc 2.5, skin 0.3, dt 0.005, Langevin gamma 1: every
    replica advances in the same graph-replayed launches; an exchange every `every` steps."""
    import random
    from espressopp_amd.batch import BatchedParallelTempering, ReplicaBatch
    rng = np.random.default_rng(seed)
    ns = [int(x) for x in rng.integers(2048, 65536, size=64, endpoint=True)]
    temps = 0.7 * (1.5 / 0.7) ** (np.arange(64) / 63.0)
    states = []
    for k, n in enumerate(ns):
        pos, box = lattice(n, 0.8442, seed=100 + k)
        states.append((pos, velocities(n, temps[k], seed=200 + k), box[0]))
    b = ReplicaBatch(ns, [s_[2] for s_ in states], rc=2.5, skin=0.3, gamma=1.0, temperature=temps,
                     seed=1000 + np.arange(64))
    for r, (pos, vel, _) in enumerate(states):
        b.load(r, np.arange(len(pos), dtype=np.int32), pos, vel)
    tune = b.autotune(reps=10)
    b.run(1000, 0.005)                       # equilibration (untimed)
    pt = BatchedParallelTempering(b, random.Random(seed), dt=0.005)
    total = sum(ns)
    lib = __import__("espressopp_amd")._lib.load()
    lib.espp_device_synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(every, steps - done)
        b.run(k, 0.005)
        done += k
        pt.exchange()
    lib.espp_device_synchronize()
    dt = time.perf_counter() - t0
    # the same steps with the kernels launched directly instead of replayed from graphs
    b.use_graph(False)
    lib.espp_device_synchronize()
    t1 = time.perf_counter()
    b.run(every, 0.005)
    lib.espp_device_synchronize()
    direct_ms = (time.perf_counter() - t1) / every * 1e3
    b.use_graph(True)
    t1 = time.perf_counter()
    b.run(every, 0.005)
    lib.espp_device_synchronize()
    graph_ms = (time.perf_counter() - t1) / every * 1e3
    force_ms = b.force_time(20)
    build_ms = b.build_time(10)
    accepted = sum(1 for e in pt.log if e[4])
    resorts = sum(b.stats(r)["resorts"] for r in range(64))
    return {"workload": "64 LJ replicas, N uniform in [2048, 65536] (seed %d), %d particles, rho 0.8442, "
                        "T geometric 0.7-1.5, Langevin gamma 1, dt 0.005, exchange every %d steps"
                        % (seed, total, every),
            "value": total * steps / dt, "unit": "particle-steps/s (all replicas)", "steps": steps,
            "ms_per_step": dt / steps * 1e3, "ms_per_step_graph": graph_ms, "ms_per_step_direct_launch": direct_ms,
            "force_launch_ms": force_ms, "build_launch_ms_all_replicas": build_ms, "autotune": {"lpp_by_replica": tune["lpp"].tolist(),
                                                      "table_ms_class_by_lpp_1_2_4_8": tune["table_ms"].tolist(),
                                                      "before_ms": tune["before_ms"], "after_ms": tune["after_ms"]},
            "exchanges_attempted": len(pt.log), "exchanges_accepted": accepted, "resorts": resorts,
            "n_min": min(ns), "n_max": max(ns)}


def c1_single_replica(steps=2000, seed=2024):
    """The 32,000-particle LJ fluid (jittered lattice, rho 0.8442, Langevin T = 1,
    gamma = 1, dt 0.005) as a batch of one replica: every step on the device, graph
    replayed, no host round trip per step."""
    from espressopp_amd.batch import ReplicaBatch
    n = 32000
    pos, box = lattice(n, 0.8442, seed=seed)
    b = ReplicaBatch([n], [box[0]], rc=2.5, skin=0.3, gamma=1.0, temperature=[1.0], seed=[seed])
    b.load(0, np.arange(n, dtype=np.int32), pos, velocities(n, 1.0, seed + 1))
    b.autotune(reps=10)
    b.run(500, 0.005)
    lib = __import__("espressopp_amd")._lib.load()
    lib.espp_device_synchronize()
    t0 = time.perf_counter()
    b.run(steps, 0.005)
    lib.espp_device_synchronize()
    dt = time.perf_counter() - t0
    return {"workload": "LJ fluid N=32000 rho 0.8442 Langevin T=1 gamma=1 dt 0.005, one replica of the batched "
                        "engine", "us_per_step": dt / steps * 1e6, "value": n * steps / dt,
            "unit": "particle-steps/s", "lpp": int(b.lpp[0]), "resorts": b.stats(0)["resorts"]}


def c1_per_system(steps=2000, seed=2024):
    """The 32,000-particle LJ fluid (jittered lattice, rho 0.8442, Langevin T = 1,
    gamma = 1, dt 0.005) on the per-system path, mixed precision: steps launched in
    windows with the skin check on the device (the default), and for comparison
    with a host wait after every step."""
    import espressopp_amd as espressopp
    n = 32000
    pos, box = lattice(n, 0.8442, seed=seed)
    vel = velocities(n, 1.0, seed + 1)
    cells = [int(b / 2.8) for b in box]
    out = {"workload": "LJ fluid N=32000 rho 0.8442 Langevin T=1 gamma=1 dt 0.005, mixed path, per-system handle"}
    for window, key in ((32, "windowed"), (0, "host_wait_per_step")):
        system, vl, lj, integ = make_system(espressopp, box, cells, 2.5, 0.3, "mixed", pos, vel)
        system.storage._h.call("espp_set_step_window", window)
        lang = espressopp.integrator.LangevinThermostat(system)
        lang.gamma = 1.0
        lang.temperature = 1.0
        integ.addExtension(lang)
        integ.run(500)
        dt = timed_run(espressopp, system, integ, steps)
        out[key] = {"us_per_step": dt / steps * 1e6, "value": n * steps / dt, "resorts": integ.getNumResorts()}
    out["value"] = out["windowed"]["value"]
    out["unit"] = "particle-steps/s"
    return out


def c5_learned_cg(n_sites=250000, reps=5, seed=7):
    """Throughput of the learned pair force's BF16 MFMA kernels on a coarse-grained
    system of n_sites sites at the 4:1-mapped melt density 0.85 / 4 (random positions,
    the same pair count per site), cutoff 3.0 (+ skin 0.3): inference (forces) and one
    force-matching step (forward + backward gradient) per frame."""
    import espressopp_amd as espressopp
    from espressopp_amd.tools.decomp import cellGrid
    rho = 0.85 / 4
    L = (n_sites / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, L, (n_sites, 3))
    box = (L, L, L)
    system = espressopp.System(precision="fp64")
    system.rng = espressopp.esutil.RNG(seed)
    system.bc = espressopp.bc.OrthorhombicBC(system.rng, box)
    system.skin = 0.3
    system.storage = espressopp.storage.DomainDecomposition(system, (1, 1, 1), cellGrid(box, (1, 1, 1), 3.0, 0.3))
    system.storage.addParticlesArray(np.arange(n_sites, dtype=np.int32), pos)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=3.0)
    mlp = espressopp.interaction.VerletListPairMLP(vl, cutoff=3.0, densityRadii=(1.5, 3.0),
                                                   densityNorm=((1.0, 1.0), (10.0, 3.0)))
    mlp.setParameters(espressopp.interaction.VerletListPairMLP.initialParameters(seed))
    system.addInteraction(mlp)
    h = system.storage._h
    lib = espressopp._lib.load()
    h.call("espp_compute_forces")
    fref = rng.normal(0, 1.0, (n_sites, 3))
    mlp.trainFrame(fref)
    lib.espp_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        h.call("espp_compute_forces")
    lib.espp_device_synchronize()
    t_inf = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        mlp.trainFrame(fref)
    t_train = (time.perf_counter() - t0) / reps
    entries, cap, _ = vl_entries(h, vl)
    # pairs inside the model cutoff (the kernels evaluate every listed entry, skin shell included)
    within = entries * (3.0 / 3.3) ** 3
    # algorithmic FLOP per evaluated pair: forward 2 (34*64 + 64*64 + 64*64 + 64); the
    # training pass recomputes the forward, back-propagates through W3 and W2 and forms
    # the three weight gradients: 2 (64*64 + 64*64) + 2 (34*64 + 64*64 + 64*64) more
    f_fwd = 2.0 * (34 * 64 + 64 * 64 + 64 * 64 + 64)
    f_train = f_fwd + 2.0 * (64 * 64 * 2) + 2.0 * (34 * 64 + 64 * 64 * 2)
    return {"workload": "learned CG pair force, %d sites at density %.4f, cutoff 3.0, list %.1f entries per site"
                        % (n_sites, rho, entries / n_sites),
            "inference_ms": t_inf * 1e3, "train_frame_ms": t_train * 1e3,
            "pair_evaluations_per_s_inference": entries / t_inf, "pairs_within_cutoff": within,
            "mfma_tflops_inference": entries * f_fwd / t_inf / 1e12,
            "mfma_tflops_train": entries * f_train / t_train / 1e12,
            "mfma_peak_tflops_bf16": 2500.0,
            "epoch_s_500_frames": 500 * t_train}


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
    ap.add_argument("--prune-shell", type=float, default=None,
                    help="shell width of the pruned LJ lists (0: none; default: the library's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="skip the batched parallel-tempering measurement")
    ap.add_argument("--no-c5", action="store_true", help="skip the learned CG force measurement")
    ap.add_argument("--only", choices=["c1", "c1s", "c4", "c5"], default=None,
                    help="run only one companion workload and print its JSON (profiling runs)")
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
    if args.only:
        fn = {"c1": c1_single_replica, "c1s": c1_per_system, "c4": c4_batched_pt, "c5": c5_learned_cg}[args.only]
        print(json.dumps(fn()), flush=True)
        return

    n, rho, rc, skin = args.n, 0.8442, 2.5, 0.3
    pos, box = lattice(n, rho, seed=2024)
    vel = velocities(n, 0.72, seed=2025)
    cells = [int(b / (rc + skin)) for b in box]
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731

    # mixed path: Langevin equilibration (untimed), then timed NVE
    t_setup = time.perf_counter()
    system, vl, lj, integ = make_system(espressopp, box, cells, rc, skin, "mixed", pos, vel)
    if args.prune_shell is not None:
        system.storage._h.call("espp_set_prune_shell", float(args.prune_shell))
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
    # a run but the last), over all its launches: those that walk the full list and
    # write its pruned copy, and those that walk the pruned list.  Algorithmic bytes
    # per particle, the reference's full-list walk: own mirror 16 + 4 per list entry
    # (int32 index) + x, v, mass/type read 56 + x, v, mirror written 64; the force
    # itself stays in registers.
    l_full, ms_full = kernel_stats(h, "pair_step")
    l_pruned, ms_pruned = kernel_stats(h, "pair_step_pruned")
    launches, pair_ms = l_full + l_pruned, ms_full + ms_pruned
    entries, cap, maxcount = vl_entries(h, vl)
    nbar = entries / n
    b_force = 16 + 4 * nbar + 56 + 64
    avg_s = pair_ms / max(launches, 1) / 1e3
    achieved = b_force * n / avg_s / 1e9
    prunes, pruned_walks, pruned_entries = prune_stats(h, lj)
    # the same with the entries the launches walked: the full list on pruning launches,
    # the current pruned list (a sample of its length) on the others
    nbar_walked = (l_full * nbar + l_pruned * pruned_entries / n) / max(launches, 1)
    b_walked = 16 + 4 * nbar_walked + 56 + 64
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
    log("  full-list walks (pruning) %d, %.3f ms avg; pruned-list walks %d, %.3f ms avg; pruned list %.2f "
        "entries per particle" % (l_full, ms_full / max(l_full, 1), l_pruned, ms_pruned / max(l_pruned, 1),
                                   pruned_entries / n))
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

    c4 = None
    if not args.no_c4:
        t = time.perf_counter()
        c4 = c4_batched_pt()
        log("c4 batched PT (%.1fs): %s" % (time.perf_counter() - t, json.dumps(c4)))
    c1 = None
    c1s = None
    if not args.no_c4:
        t = time.perf_counter()
        c1 = c1_single_replica()
        log("c1 via the batched engine (%.1fs): %s" % (time.perf_counter() - t, json.dumps(c1)))
        t = time.perf_counter()
        c1s = c1_per_system()
        log("c1 per system (%.1fs): %s" % (time.perf_counter() - t, json.dumps(c1s)))
    c5 = None
    if not args.no_c5:
        t = time.perf_counter()
        c5 = c5_learned_cg()
        log("c5 learned CG (%.1fs): %s" % (time.perf_counter() - t, json.dumps(c5)))

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
                     "avg_launch_ms": avg_s * 1e3, "launches": launches,
                     "full_list_launches": l_full, "full_list_avg_ms": ms_full / max(l_full, 1),
                     "pruned_list_launches": l_pruned, "pruned_list_avg_ms": ms_pruned / max(l_pruned, 1),
                     "pruned_nbar": pruned_entries / n, "walked_nbar": nbar_walked,
                     "walked_bytes_per_particle": b_walked,
                     "walked_frac": b_walked * n / avg_s / 1e9 / HBM_PEAK_GBS},
        "fp64_path": fp64,
        "speedup_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "accuracy": err,
        "nve_drift_per_particle_per_1e5_steps": drift["per_particle_per_1e5_steps"] if drift else None,
        "nve_drift": drift,
        "resorts": resorts,
        "cpu_baseline": cpu,
        "c4_batched_parallel_tempering": c4,
        "c1_single_replica_batched": c1,
        "c1_per_system": c1s,
        "c5_learned_cg_force": c5,
    }
    traffic_file = os.path.join(ROOT, "profiles", "lj_mixed_4M_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            out["roofline"]["traffic"] = json.load(f).get("bytes_per_launch")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
