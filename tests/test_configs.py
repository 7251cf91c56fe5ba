"""The benchmark's companion workloads at their real sizes, against the oracle.

LJ fluid: N = 32,000 on a jittered simple-cubic lattice, density 0.8442, rc 2.5
(shift auto), skin 0.3, dt 0.005, Langevin T = 1, gamma = 1.
Kremer-Grest melt: 1000 random-walk chains of 200 beads (N = 200,000), density
0.85, bond 0.97, no two beads closer than 0.7 (the overlap push-off), WCA
(rc = 2^(1/6)) + FENE (K = 30, R0 = 1.5): forces, energies and virial.
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim

from _systems import cells_for, force_rel_err, gpu_system, jittered_lattice, mb_velocities, melt_chains

pytestmark = pytest.mark.gpu

import espressopp_amd as espressopp  # noqa: E402


def _lj_fluid(precision, pos, vel, box, cells):
    system = gpu_system(box, cells, 0.3, precision, seed=2024)
    system.storage.addParticlesArray(np.arange(len(pos)), pos, None, vel)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=2.5)
    lj = espressopp.interaction.VerletListLennardJones(vl)
    lj.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, 2.5, "auto"))
    system.addInteraction(lj)
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    lang = espressopp.integrator.LangevinThermostat(system)
    lang.gamma = 1.0
    lang.temperature = 1.0
    integ.addExtension(lang)
    return system, lj, integ


def test_lj_fluid_32000_langevin(gpu):
    n = 32000
    pos, box = jittered_lattice(n, 0.8442, seed=2024)
    vel = mb_velocities(n, 1.0, seed=2025)
    cells = cells_for(box, 2.5, 0.3)
    ref = OracleSim(box, cells, 0.3, seed=2024, threads=8)
    ref.add_particles(np.arange(n), pos, vel=vel)
    ref.decompose()
    ria = ref.lj_create(ref.vl_create(2.5))
    ref.lj_set(ria, 0, 0, 1.0, 1.0, 2.5, "auto")
    ref.langevin(1.0, 1.0)
    ref.run(200, dt=0.005)
    rT, rP = ref.temperature(), ref.pressure()
    rp = ref.get("pos") + ref.get("image") * np.asarray(box)
    for precision in ("fp64", "mixed"):
        system, lj, integ = _lj_fluid(precision, pos, vel, box, cells)
        integ.run(200)
        a = system.storage.getParticleArrays()
        gp = a["pos"] + a["image"] * np.asarray(box)
        T = espressopp.analysis.Temperature(system).compute()
        P = espressopp.analysis.Pressure(system).compute()
        if precision == "fp64":
            # same noise stream, same FP64 arithmetic up to summation order: one trajectory
            assert np.abs(gp - rp).max() < 1e-7
            assert T == pytest.approx(rT, rel=1e-9) and P == pytest.approx(rP, rel=1e-8)
        else:
            # 200 thermostatted steps from the same start: the same state to within the
            # growth of ~1e-6 force differences; T and P agree closely
            assert np.abs(gp - rp).max() < 1e-3
            assert T == pytest.approx(rT, rel=1e-3) and P == pytest.approx(rP, rel=1e-3)


def test_kremer_grest_melt_200k(gpu):
    pos, box, bonds = melt_chains(1000, 200, 0.85, seed=2026)
    n = len(pos)
    rc, skin = 2 ** (1 / 6), 0.3
    cells = cells_for(box, rc, skin)
    ref = OracleSim(box, cells, skin, threads=8)
    ref.add_particles(np.arange(n), pos)
    ref.decompose()
    wca = ref.lj_create(ref.vl_create(rc))
    ref.lj_set(wca, 0, 0, 1.0, 1.0, rc, "auto")
    rfpl = ref.fpl_create()
    ref.fpl_add(rfpl, bonds)
    fene = ref.fene_create(rfpl, 30.0, 0.0, 1.5)
    ref.calc_forces()
    for precision in ("fp64", "mixed"):
        system = gpu_system(box, cells, skin, precision, seed=2027)
        system.storage.addParticlesArray(np.arange(n), pos)
        system.storage.decompose()
        vl = espressopp.VerletList(system, cutoff=rc)
        lj = espressopp.interaction.VerletListLennardJones(vl)
        lj.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, rc, "auto"))
        system.addInteraction(lj)
        fpl = espressopp.FixedPairList(system.storage)
        fpl.addBonds(bonds.tolist())
        fb = espressopp.interaction.FixedPairListFENE(system, fpl, espressopp.interaction.FENE(30.0, 0.0, 1.5))
        system.addInteraction(fb)
        integ = espressopp.integrator.VelocityVerlet(system)
        integ.dt = 0.01
        integ.run(0)
        f = system.storage.getParticleArrays()["f"]
        tol, etol = (1e-11, 1e-11) if precision == "fp64" else (1e-5, 1e-6)
        assert force_rel_err(f, ref.get("f"), 0.1).max() <= tol, precision
        assert lj.computeEnergy() == pytest.approx(ref.energy(wca), rel=etol)
        assert fb.computeEnergy() == pytest.approx(ref.energy(fene), rel=etol)
        assert lj.computeVirial() == pytest.approx(ref.virial(wca)[0], rel=10 * etol)


def test_thermostatted_statistics_mixed_vs_fp64(gpu):
    # Langevin LJ fluid, N = 8000, T = 1: the mixed and FP64 paths run with independent
    # noise, and their temperature, pressure and g(r) agree within the statistical error
    # of the samples (block means, samples 200 steps apart); both hold T at the target
    n = 20 ** 3
    pos, box = jittered_lattice(n, 0.8442, seed=61)
    vel = mb_velocities(n, 1.0, seed=62)
    cells = cells_for(box, 2.5, 0.3)
    nsamp, gap, nbins = 30, 200, 60
    stats = {}
    for precision, seed in (("fp64", 101), ("mixed", 202)):
        system = gpu_system(box, cells, 0.3, precision, seed=seed)
        system.storage.addParticlesArray(np.arange(n), pos, None, vel)
        system.storage.decompose()
        vl = espressopp.VerletList(system, cutoff=2.5)
        lj = espressopp.interaction.VerletListLennardJones(vl)
        lj.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, 2.5, "auto"))
        system.addInteraction(lj)
        integ = espressopp.integrator.VelocityVerlet(system)
        integ.dt = 0.005
        lang = espressopp.integrator.LangevinThermostat(system)
        lang.gamma = 1.0
        lang.temperature = 1.0
        integ.addExtension(lang)
        integ.run(1000)
        T, P, G = [], [], []
        for _ in range(nsamp):
            integ.run(gap)
            T.append(espressopp.analysis.Temperature(system).compute())
            P.append(espressopp.analysis.Pressure(system).compute())
            G.append(espressopp.analysis.RadialDistrF(system).compute(nbins))
        stats[precision] = (np.array(T), np.array(P), np.array(G))

    def mean_se(a):
        # standard error from means of 5 blocks of consecutive samples
        blocks = a.reshape(5, -1, *a.shape[1:]).mean(axis=1)
        return a.mean(axis=0), blocks.std(axis=0, ddof=1) / np.sqrt(5)

    for precision in stats:
        mt, st = mean_se(stats[precision][0])
        assert abs(mt - 1.0) < max(5 * st, 0.01), (precision, mt, st)
    for q in range(3):
        m1, s1 = mean_se(stats["fp64"][q])
        m2, s2 = mean_se(stats["mixed"][q])
        z = np.abs(m1 - m2) / np.maximum(np.sqrt(s1 ** 2 + s2 ** 2), 1e-3)
        if q < 2:
            assert z < 5.0, (q, m1, m2, s1, s2)
        else:
            mask = m1 > 0.2                        # bins inside the first shell and beyond
            assert z[mask].max() < 6.0, (z[mask].max(), int(np.argmax(z * mask)))
