"""Pruned Lennard-Jones lists of the mixed path.

A force walk of the full Verlet list (rc + skin) also writes the pairs within
rc + w at its positions as a pruned list; the following walks take the pruned
list while twice the displacement bound accumulated since (the skin check's
maxDist) stays below w.  The entries it drops are beyond the cutoff window and
add exactly zero, so:

- forces from a pruned list equal forces from the full list at the same
  positions, bit for bit;
- trajectories with and without pruning agree to the rounding order of the
  rare pairs inside the cutoff window (their contributions are added after
  their batch of entries, and batches differ between the two lists);
- the pruned list holds exactly the pairs within rc + w (k-d tree count).
"""
import ctypes

import numpy as np
import pytest

from tests._systems import cells_for, gpu_system, jittered_lattice, mb_velocities

pytestmark = pytest.mark.gpu

RC, SKIN = 2.5, 0.3


def prune_stats(system, ia):
    h = system.storage._h
    p, w, e = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    h.call("espp_prune_stats", ia._id, ctypes.byref(p), ctypes.byref(w), ctypes.byref(e))
    return p.value, w.value, e.value


def set_shell(system, width):
    system.storage._h.call("espp_set_prune_shell", float(width))


def lj_fluid(n, T, seed, shell, gamma=1.0):
    import espressopp_amd as espressopp
    pos, box = jittered_lattice(n, 0.8442, seed=seed)
    vel = mb_velocities(n, T, seed=seed + 1)
    system = gpu_system(box, cells_for(box, RC, SKIN), SKIN, "mixed", seed=seed + 2)
    set_shell(system, shell)
    system.storage.addParticlesArray(np.arange(n, dtype=np.int32), pos, None, vel)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=RC)
    ia = espressopp.interaction.VerletListLennardJones(vl)
    ia.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, RC, "auto"))
    system.addInteraction(ia)
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    if gamma > 0:
        lang = espressopp.integrator.LangevinThermostat(system)
        lang.gamma = gamma
        lang.temperature = T
        integ.addExtension(lang)
    return system, vl, ia, integ, box


def unfolded(system, box):
    a = system.storage.getParticleArrays()
    return a["pos"] + a["image"] * np.asarray(box), a["v"], a["f"]


def test_pruned_trajectory_follows_the_full_list(gpu):
    n, steps = 32000, 300
    runs = {}
    for shell in (-1.0, 0.0):          # skin / 3, and no pruning (the default)
        system, vl, ia, integ, box = lj_fluid(n, 1.5, seed=7, shell=shell)
        integ.run(steps)
        runs[shell] = (unfolded(system, box), integ.getNumResorts(), prune_stats(system, ia))
    (xa, va, _), resorts_a, (prunes, walks, _) = runs[-1.0]
    (xb, vb, _), resorts_b, (prunes_b, walks_b, _) = runs[0.0]
    dx, dv = np.abs(xa - xb).max(), np.abs(va - vb).max()
    print("N=%d, %d steps at T=1.5: resorts %d / %d, prunes %d, pruned walks %d; max |dx| %.2e, max |dv| %.2e"
          % (n, steps, resorts_a, resorts_b, prunes, walks, dx, dv), flush=True)
    assert prunes_b == 0 and walks_b == 0
    assert resorts_a == resorts_b
    # every list build is pruned, the pruned lists are walked, and they expire and are
    # taken again between builds
    assert prunes > resorts_a + 1 and walks > 0
    assert dx < 1e-9 and dv < 1e-8


def test_pruned_forces_equal_full_list_forces(gpu):
    system, vl, ia, integ, box = lj_fluid(20000, 1.2, seed=11, shell=-1.0, gamma=0.0)
    integ.run(57)
    h = system.storage._h
    _, walks0, _ = prune_stats(system, ia)
    h.call("espp_compute_forces")
    _, walks1, _ = prune_stats(system, ia)
    f_pruned = system.storage.getParticleArrays()["f"]
    set_shell(system, 0.0)
    h.call("espp_compute_forces")
    f_full = system.storage.getParticleArrays()["f"]
    print("pruned walks %d -> %d; max |dF| %.2e" % (walks0, walks1, np.abs(f_pruned - f_full).max()), flush=True)
    assert walks1 == walks0 + 1          # the first evaluation walked the pruned list
    assert np.array_equal(f_pruned, f_full)


@pytest.mark.parametrize("width", [0.05, 0.1, 0.2])
def test_pruned_list_holds_the_pairs_within_rc_plus_shell(gpu, width):
    from scipy.spatial import cKDTree
    system, vl, ia, integ, box = lj_fluid(16000, 1.0, seed=13, shell=width, gamma=0.0)
    integ.run(20)
    # a fresh full list at these positions (it holds every pair within rc + skin);
    # the next force walk prunes it here
    system.storage.decompose()
    system.storage._h.call("espp_compute_forces")
    prunes, _, entries = prune_stats(system, ia)
    a = system.storage.getParticleArrays()
    L = np.asarray(box)
    tree = cKDTree(np.mod(a["pos"], L), boxsize=L)
    pairs = len(tree.query_pairs(RC + width))
    print("shell %.2f: pruned entries %d, 2 x pairs within rc + shell %d, full list %d"
          % (width, entries, 2 * pairs, vl.totalSize() * 2), flush=True)
    assert abs(entries - 2 * pairs) <= 2
    assert entries < 2 * vl.totalSize()
