"""The benchmark workload at its full size against the oracle.

N = 4,194,304 (a jittered 128^3 lattice at density 0.8442, rc 2.5, skin 0.3)
after 200 Langevin steps at T = 0.72 on the GPU: the Verlet list of both GPU
paths holds exactly the oracle's pairs (count and an order-free hash of the pair
set), and the forces, energy and virial of both paths match the oracle's (FP64:
1e-11 per particle, 1e-12 relative energy and virial; mixed: 1e-5 per particle
with the 0.1 RMS |F| floor, 1e-6 relative energy and virial).  The FP64 bound
is the summation order: the GPU sums each particle's full list, the oracle its
Newton-3 half list, and over 4M particles a few reach 1.5e-12.
"""
import os

import numpy as np
import pytest

from tests._systems import cells_for, force_rel_err, gpu_system, jittered_lattice, mb_velocities

pytestmark = pytest.mark.gpu

RHO, RC, SKIN = 0.8442, 2.5, 0.3


def pair_hash(pairs, n):
    """Sum over pairs of a 64-bit mix of (min id, max id), wrapping: equal for
    equal pair sets in any order and orientation."""
    p = np.asarray(pairs).reshape(-1, 2)
    a = np.minimum(p[:, 0], p[:, 1]).astype(np.uint64)
    b = np.maximum(p[:, 0], p[:, 1]).astype(np.uint64)
    k = a * np.uint64(n) + b
    with np.errstate(over="ignore"):
        k ^= k >> np.uint64(30)
        k *= np.uint64(0xBF58476D1CE4E5B9)
        k ^= k >> np.uint64(27)
        k *= np.uint64(0x94D049BB133111EB)
        k ^= k >> np.uint64(31)
        return int(k.sum(dtype=np.uint64)), len(p)


def lj_system(pos, vel, box, precision):
    import espressopp_amd as espressopp
    n = len(pos)
    system = gpu_system(box, cells_for(box, RC, SKIN), SKIN, precision)
    system.storage.addParticlesArray(np.arange(n, dtype=np.int32), pos, None, vel)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=RC)
    ia = espressopp.interaction.VerletListLennardJones(vl)
    ia.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, RC, "auto"))
    system.addInteraction(ia)
    return system, vl, ia


def test_full_size_pairs_forces_energy_virial_match_oracle(gpu):
    import espressopp_amd as espressopp
    from oracle.oracle import OracleSim
    n = 128 ** 3
    pos, box = jittered_lattice(n, RHO, seed=2024)
    vel = mb_velocities(n, 0.72, seed=2025)
    # a liquid state: 200 Langevin steps on the mixed path
    system, vl, ia = lj_system(pos, vel, box, "mixed")
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    lang = espressopp.integrator.LangevinThermostat(system)
    lang.gamma = 1.0
    lang.temperature = 0.72
    integ.addExtension(lang)
    integ.run(200)
    a = system.storage.getParticleArrays()
    order = np.argsort(a["id"])
    state = (a["pos"] + a["image"] * np.asarray(box))[order]
    v = a["v"][order]
    del system, vl, ia, integ, lang, a

    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    ref = OracleSim(box, cells_for(box, RC, SKIN), SKIN, threads=threads)
    ref.add_particles(np.arange(n, dtype=np.int32), state, vel=v)
    ref.decompose()
    rvl = ref.vl_create(RC)
    ria = ref.lj_create(rvl)
    ref.lj_set(ria, 0, 0, 1.0, 1.0, RC, "auto")
    ref.calc_forces()
    h_ref = pair_hash(ref.vl_pairs(rvl), n)
    ids = ref.ids()
    f_ref = ref.get("f")[np.argsort(ids)]
    e_ref = ref.energy(ria)
    w_ref, _ = ref.virial(ria)
    print("oracle: %d pairs, E/N %.9f, W/N %.9f" % (h_ref[1], e_ref / n, w_ref / n), flush=True)
    del ref

    for precision, ftol, etol in (("fp64", 1e-11, 1e-12), ("mixed", 1e-5, 1e-6)):
        system, vl, ia = lj_system(state, v, box, precision)
        espressopp.integrator.VelocityVerlet(system).run(0)
        h = pair_hash(vl.pairs_array(), n)
        g = system.storage.getParticleArrays()
        f = g["f"][np.argsort(g["id"])]
        err = force_rel_err(f, f_ref, 0.1)
        e, w = ia.computeEnergy(), ia.computeVirial()
        print("%s: %d pairs, hash %s; force rel err max %.2e; energy rel %.2e; virial rel %.2e"
              % (precision, h[1], "equal" if h == h_ref else "DIFFERENT", err.max(),
                 abs(e - e_ref) / abs(e_ref), abs(w - w_ref) / abs(w_ref)), flush=True)
        assert h == h_ref
        assert err.max() <= ftol
        assert abs(e - e_ref) <= etol * abs(e_ref)
        assert abs(w - w_ref) <= etol * abs(w_ref)
        del system, vl, ia, g
