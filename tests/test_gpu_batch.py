"""Batched replicas (espp_batch) on the GPU.

Parity: each replica's pair set equals the oracle's bit for bit, forces and
energies match the FP64 oracle, short NVE and Langevin trajectories of small
replicas follow the oracle, and every replica of a 64-replica batch at the
parallel-tempering workload's sizes (N uniform in [2048, 65536], seeded,
temperatures geometric 0.7-1.5) is bit-identical to the same replica advanced
alone.  Graph-replayed steps equal directly launched ones.
"""
import math
import random

import numpy as np
import pytest

from tests._systems import cells_for, force_rel_err, jittered_lattice, mb_velocities

pytestmark = pytest.mark.gpu

RHO, RC, SKIN, DT = 0.8442, 2.5, 0.3, 0.005


def replica_state(n, seed, T=1.0):
    pos, box = jittered_lattice(n, RHO, seed=seed)
    vel = mb_velocities(n, T, seed=seed + 1)
    # ids deliberately not in position order
    ids = np.random.default_rng(seed + 2).permutation(n).astype(np.int32)
    return ids, pos, vel, box[0]


def make_batch(states, temps, gamma=1.0, seeds=None, lpp=None):
    from espressopp_amd.batch import ReplicaBatch
    n = [len(s[0]) for s in states]
    box = [s[3] for s in states]
    seeds = seeds if seeds is not None else [1000 + k for k in range(len(states))]
    b = ReplicaBatch(n, box, rc=RC, skin=SKIN, gamma=gamma, temperature=temps, seed=seeds)
    for r, (ids, pos, vel, _) in enumerate(states):
        b.load(r, ids, pos, vel)
    if lpp is not None:
        b.lpp = lpp
    return b


def oracle_for(state, seed, gamma=0.0, T=1.0):
    from oracle.oracle import OracleSim
    ids, pos, vel, L = state
    box = (L, L, L)
    ref = OracleSim(box, cells_for(box, RC, SKIN), SKIN, seed=seed)
    ref.add_particles(ids, pos, vel=vel)
    ref.decompose()
    vl = ref.vl_create(RC)
    ia = ref.lj_create(vl)
    ref.lj_set(ia, 0, 0, 1.0, 1.0, RC, "auto")
    if gamma > 0:
        ref.langevin(gamma, T)
    return ref, vl, ia


def by_id(ref, field):
    ids = ref.ids()
    order = np.argsort(ids)
    return ref.get(field)[order]


@pytest.mark.parametrize("lpp", [1, 4])
def test_batch_pairs_forces_energy_match_oracle(gpu, lpp):
    # the list build's prefilter: FP32 at N = 2048 (box too small for the 16-bit copies),
    # 16-bit copies with one period per box (3375, 5000) and with two (24000)
    sizes = (2048, 3375, 5000, 24000)
    states = [replica_state(n, 10 + k) for k, n in enumerate(sizes)]
    b = make_batch(states, [1.0] * len(sizes), gamma=0.0, lpp=[lpp] * len(sizes))
    b.run(0, DT)
    e, w = b.energy()
    for r, st in enumerate(states):
        ref, vl, ia = oracle_for(st, seed=1000 + r)
        ref.calc_forces()
        rp = ref.vl_pairs(vl)
        rp = np.sort(np.asarray(rp, dtype=np.int32).reshape(-1, 2), axis=1)
        rp = rp[np.lexsort((rp[:, 1], rp[:, 0]))]
        gp = b.pairs(r)
        print("replica %d: %d pairs (oracle %d)" % (r, len(gp), len(rp)), flush=True)
        assert np.array_equal(gp, rp)
        # positions enter the pair kernel as 32-bit fixed point (L / 2^32 ~ 3e-9 at these
        # boxes): the mixed path's bound is 1e-5 per particle (floor 0.1 RMS |F|), 1e-6 energy
        f_ref = by_id(ref, "f")
        f = b.get(r, "f")
        err = force_rel_err(f, f_ref, 0.1)
        e_ref = ref.energy(ia)
        w_ref, _ = ref.virial(ia)
        print("  force rel err max %.2e, energy rel %.2e, virial rel %.2e" %
              (err.max(), abs(e[r] - e_ref) / abs(e_ref), abs(w[r] - w_ref) / abs(w_ref)), flush=True)
        assert err.max() < 1e-5
        assert abs(e[r] - e_ref) <= 1e-6 * abs(e_ref)
        assert abs(w[r] - w_ref) <= 1e-6 * abs(w_ref)


def test_batch_nve_and_langevin_follow_oracle(gpu):
    states = [replica_state(2048, 21), replica_state(4096, 23)]
    for gamma, steps, tol in ((0.0, 60, 1e-7), (1.0, 30, 1e-7)):
        b = make_batch(states, [1.0, 1.3], gamma=gamma, seeds=[77, 78])
        b.run(steps, DT)
        for r, st in enumerate(states):
            ref, _, _ = oracle_for(st, seed=77 + r, gamma=gamma, T=[1.0, 1.3][r])
            ref.run(steps, dt=DT)
            L = st[3]
            rp = by_id(ref, "pos") + by_id(ref, "image") * L
            gp = b.get(r, "pos") + b.get(r, "image") * L
            rv = by_id(ref, "v")
            gv = b.get(r, "v")
            print("gamma %.1f replica %d: max |dx| %.2e, max |dv| %.2e, resorts %d (oracle %d)" %
                  (gamma, r, np.abs(gp - rp).max(), np.abs(gv - rv).max(), b.stats(r)["resorts"], ref.resorts),
                  flush=True)
            assert np.abs(gp - rp).max() < tol
            assert np.abs(gv - rv).max() < 10 * tol
            assert b.stats(r)["resorts"] == ref.resorts


def test_batch_graph_equals_direct_launches(gpu):
    states = [replica_state(n, 30 + k) for k, n in enumerate((2048, 6000, 9000))]
    runs = []
    for graph in (True, False):
        b = make_batch(states, [0.8, 1.0, 1.2])
        b.use_graph(graph)
        out = []
        for chunk in (7, 10, 23):          # several replays with host reads in between
            b.run(chunk, DT)
            out.append([b.get(r, "pos").copy() for r in range(3)])
        runs.append(out)
    for a, c in zip(*runs):
        for x, y in zip(a, c):
            assert np.array_equal(x, y)


def test_batch_build_time_keeps_pair_sets(gpu):
    # the build timing rebins every replica and builds its lists several times from
    # that binning: the lists keep their pair sets, and a run continues afterwards
    states = [replica_state(n, 40 + k) for k, n in enumerate((2048, 5000, 24000))]
    b = make_batch(states, [1.0, 1.1, 1.2], gamma=0.0)
    b.run(0, DT)
    before = [b.pairs(r) for r in range(3)]
    ms = b.build_time(3)
    print("build of 3 replicas: %.3f ms" % ms, flush=True)
    assert ms > 0.0
    for r in range(3):
        assert np.array_equal(b.pairs(r), before[r])
    b.run(10, DT)
    assert all(np.isfinite(b.get(r, "pos")).all() for r in range(3))


def c4_replicas(count=64, seed=2026):
    rng = np.random.default_rng(seed)
    ns = rng.integers(2048, 65536, size=count, endpoint=True)
    temps = 0.7 * (1.5 / 0.7) ** (np.arange(count) / (count - 1))
    return [int(x) for x in ns], temps


def test_batch_c4_replicas_bit_identical_to_alone(gpu):
    ns, temps = c4_replicas()
    print("64 replicas, N from %d to %d, %d particles" % (min(ns), max(ns), sum(ns)), flush=True)
    states = [replica_state(n, 100 + k, T=temps[k]) for k, n in enumerate(ns)]
    b = make_batch(states, temps)
    tune = b.autotune(reps=5)
    print("autotune: %s, %.3f -> %.3f ms" % (tune["lpp"].tolist(), tune["before_ms"], tune["after_ms"]), flush=True)
    lpp = tune["lpp"]
    b.run(100, DT)
    # one exchange with the reference's rule, then 100 more steps
    from espressopp_amd.batch import BatchedParallelTempering
    pt = BatchedParallelTempering(b, random.Random(5), dt=DT)
    swaps = pt.exchange()
    print("swaps:", swaps, flush=True)
    T_after = [b.temperature(r) for r in range(len(ns))]
    b.run(100, DT)
    order = np.argsort(ns)
    picks = sorted({int(order[0]), int(order[-1]), int(order[len(ns) // 2]), 17})
    for r in picks:
        alone = make_batch([states[r]], [temps[r]], seeds=[1000 + r], lpp=[int(lpp[r])])
        alone.run(100, DT)
        alone.set_temperature(0, T_after[r])
        alone.run(100, DT)
        for field in ("pos", "v", "image"):
            assert np.array_equal(alone.get(0, field), b.get(r, field)), (r, field)
        assert alone.stats(0) == b.stats(r)
        print("replica %d (N=%d, lpp %d): bit-identical alone" % (r, ns[r], lpp[r]), flush=True)


def test_batch_exchange_rule(gpu):
    states = [replica_state(2048, 50 + k) for k in range(4)]
    temps = [0.7, 0.9, 1.1, 1.3]
    b = make_batch(states, temps)
    b.run(50, DT)
    from espressopp_amd.batch import BatchedParallelTempering
    from espressopp_amd.parallel_tempering import ParallelTempering
    pt = BatchedParallelTempering(b, random.Random(3), dt=DT)
    e, _ = b.energy()
    tm = b.measured_temperature()
    before = [b.temperature(r) for r in range(4)]
    swaps = pt.exchange()
    # first call pairs (1, 2) (odd), as src/ParallelTempering.py:176-180 does
    assert [(m, n) for (m, n, *_rest) in pt.log] == [(1, 2)]
    m, n, d, draw, yes = pt.log[0]
    assert d == pytest.approx(ParallelTempering.delta(tm[1], tm[2], e[1], e[2], 2048, 2048), rel=1e-12)
    assert yes == (d <= 0 or math.exp(-d) > draw)
    after = [b.temperature(r) for r in range(4)]
    if yes:
        assert swaps == [(1, 2)] and after[1] == pytest.approx(tm[2]) and after[2] == pytest.approx(tm[1])
    else:
        assert swaps == [] and after == before
