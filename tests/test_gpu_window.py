"""Windows of steps launched without a host wait (espp_set_step_window).

In a window every kernel of a step first reads a device flag that the previous
step's displacement reduction raises once the summed displacement bound passes
skin/2 (the reference's skin check), and returns at once when it is set; the
host then reads how many steps ran and rebuilds.  The trajectory must be the
step-by-step loop's bit for bit: same positions, images, velocities, forces,
resorts and thermostat stream, for an LJ fluid (the force kernel finishes the
step) and a bonded melt (WCA over an untiled list, then FENE finishing the
step).
"""
import ctypes

import numpy as np
import pytest

from tests._systems import cells_for, gpu_system, jittered_lattice, mb_velocities, melt_chains

pytestmark = pytest.mark.gpu


def window_stats(system):
    w, l, r = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    system.storage._h.call("espp_step_window_stats", ctypes.byref(w), ctypes.byref(l), ctypes.byref(r))
    return w.value, l.value, r.value


def lj_run(window, n=32000, steps=300, T=1.0):
    import espressopp_amd as espressopp
    pos, box = jittered_lattice(n, 0.8442, seed=31)
    vel = mb_velocities(n, T, seed=32)
    system = gpu_system(box, cells_for(box, 2.5, 0.3), 0.3, "mixed", seed=33)
    system.storage._h.call("espp_set_step_window", int(window))
    system.storage.addParticlesArray(np.arange(n, dtype=np.int32), pos, None, vel)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=2.5)
    ia = espressopp.interaction.VerletListLennardJones(vl)
    ia.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, 2.5, "auto"))
    system.addInteraction(ia)
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    lang = espressopp.integrator.LangevinThermostat(system)
    lang.gamma = 1.0
    lang.temperature = T
    integ.addExtension(lang)
    for chunk in (1, 2, 97, steps - 100):        # runs of every length, the last one long
        integ.run(chunk)
    return system.storage.getParticleArrays(), integ.getNumResorts(), integ.step, window_stats(system)


def test_window_lj_fluid_equals_step_by_step(gpu):
    a, ra, sa, (wins, launched, ran) = lj_run(32)
    b, rb, sb, (wins0, _, _) = lj_run(0)
    print("windows %d, steps launched %d, ran %d; resorts %d / %d" % (wins, launched, ran, ra, rb), flush=True)
    assert wins > 0 and wins0 == 0
    assert ran <= launched
    assert ra == rb and sa == sb
    for field in ("pos", "image", "v", "f", "id"):
        assert np.array_equal(a[field], b[field]), field


def test_window_melt_with_fene_last_equals_step_by_step(gpu):
    import espressopp_amd as espressopp
    pos, box, bonds = melt_chains(30, 40, 0.5, seed=42, min_dist=0.95)
    n = len(pos)
    vel = mb_velocities(n, 1.0, seed=43)
    rc, skin = 2 ** (1 / 6), 0.3
    out = {}
    for window in (16, 0):
        system = gpu_system(box, cells_for(box, rc, skin), skin, "mixed", seed=44)
        system.storage._h.call("espp_set_step_window", int(window))
        system.storage.addParticlesArray(np.arange(n, dtype=np.int32), pos, None, vel)
        system.storage.decompose()
        vl = espressopp.VerletList(system, cutoff=rc)
        ia = espressopp.interaction.VerletListLennardJones(vl)
        ia.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, rc, "auto"))
        system.addInteraction(ia)
        fpl = espressopp.FixedPairList(system.storage)
        fpl.addBonds(bonds.tolist())
        system.addInteraction(espressopp.interaction.FixedPairListFENE(system, fpl,
                                                                       espressopp.interaction.FENE(30.0, 0.0, 1.5)))
        integ = espressopp.integrator.VelocityVerlet(system)
        integ.dt = 0.003
        lang = espressopp.integrator.LangevinThermostat(system)
        lang.gamma = 1.0
        lang.temperature = 1.0
        integ.addExtension(lang)
        integ.run(150)
        out[window] = (system.storage.getParticleArrays(), integ.getNumResorts(), window_stats(system))
    (a, ra, (wins, launched, ran)), (b, rb, _) = out[16], out[0]
    print("melt: windows %d, launched %d, ran %d; resorts %d / %d" % (wins, launched, ran, ra, rb), flush=True)
    assert wins > 0
    assert ra == rb
    for field in ("pos", "image", "v", "f"):
        assert np.array_equal(a[field], b[field]), field


def test_window_reports_a_lost_particle(gpu):
    import espressopp_amd as espressopp
    pos, box = jittered_lattice(4000, 0.8442, seed=51)
    system = gpu_system(box, cells_for(box, 2.5, 0.3), 0.3, "mixed", seed=52)
    system.storage.addParticlesArray(np.arange(4000, dtype=np.int32), pos, None, mb_velocities(4000, 1.0, seed=53))
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=2.5)
    ia = espressopp.interaction.VerletListLennardJones(vl)
    ia.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, 2.5, "auto"))
    system.addInteraction(ia)
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    integ.run(20)
    # a velocity of NaN makes the particle's next position NaN inside a window
    p = system.storage.getParticle(17)
    p.v = espressopp.Real3D(float("nan"), 0.0, 0.0)
    with pytest.raises(RuntimeError, match="outer space"):
        integ.run(50)
