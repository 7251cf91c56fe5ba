"""The reference's espressopp.vec names and standard systems on the GPU path.

test_vec_lennard_jones_matches_plain_names follows testsuite/vec/test_vec_lennard_jones.py
(the uncapped case: the capped potential is not on this path): the 2048-particle
fixture run for 10 steps through espressopp.vec.* and through the plain names
gives the same positions; here both are also checked against the FP64 oracle.
"""
import numpy as np
import pytest

import espressopp_amd as espressopp
from tests._systems import FIXTURE_2048

pytestmark = pytest.mark.gpu


def generate_md(use_vec):
    if use_vec:
        Default = espressopp.vec.standard_system.Default
        VerletList = espressopp.vec.VerletList
        VerletListLennardJones = espressopp.vec.interaction.VerletListLennardJones
        LennardJones = espressopp.vec.interaction.LennardJones
    else:
        Default = espressopp.standard_system.Default
        VerletList = espressopp.VerletList
        VerletListLennardJones = espressopp.interaction.VerletListLennardJones
        LennardJones = espressopp.interaction.LennardJones
    rc, skin, timestep = 2.5, 0.3, 0.005
    pid, typ, x, y, z, vx, vy, vz, Lx, Ly, Lz = espressopp.tools.readxyz(FIXTURE_2048)
    box = (Lx, Ly, Lz)
    n = len(pid)
    system, integrator = Default(box=box, rc=rc, skin=skin, dt=timestep, temperature=None)
    parts = [[i + 1, 0, 1.0, espressopp.Real3D(x[i], y[i], z[i]), espressopp.Real3D(vx[i], vy[i], vz[i])]
             for i in range(n)]
    system.storage.addParticles(parts, "id", "type", "mass", "pos", "v")
    system.storage.decompose()
    vl = VerletList(system, cutoff=rc)
    inter = VerletListLennardJones(vl)
    inter.setPotential(type1=0, type2=0, potential=LennardJones(epsilon=1.0, sigma=1.0, cutoff=rc, shift=0))
    system.addInteraction(inter)
    integrator.run(10)
    conf = espressopp.analysis.Configurations(system, pos=True, vel=True, force=True)
    conf.gather()
    return [conf[0][i] for i in range(n + 1)], box, (pid, x, y, z, vx, vy, vz)


def test_vec_lennard_jones_matches_plain_names(gpu):
    pos0, box, data = generate_md(True)
    pos1, _, _ = generate_md(False)
    assert len(pos0) == len(pos1)
    assert max((pos0[i] - pos1[i]).sqr() for i in range(len(pos1))) == 0.0
    assert pos0[0].sqr() == 0.0                        # id 0 is not stored (the reference's map default)
    from oracle.oracle import OracleSim
    from tests._systems import cells_for
    pid, x, y, z, vx, vy, vz = data
    ref = OracleSim(box, cells_for(box, 2.5, 0.3), 0.3)
    ref.add_particles(np.arange(1, len(pid) + 1), np.stack([x, y, z], axis=1), vel=np.stack([vx, vy, vz], axis=1))
    ref.decompose()
    ia = ref.lj_create(ref.vl_create(2.5))
    ref.lj_set(ia, 0, 0, 1.0, 1.0, 2.5, 0.0)
    ref.run(10, dt=0.005)
    rp = ref.get("pos")[np.argsort(ref.ids())]
    gp = np.array([list(pos0[i]) for i in range(1, len(pid) + 1)])
    assert np.abs(gp - rp).max() < 1e-9


def test_standard_systems_build_and_run(gpu):
    # random positions overlap (as in the reference's builder): forces only
    system, integ = espressopp.standard_system.LennardJones(500, box=(9.0, 9.0, 9.0), rc=1.12246,
                                                            temperature=1.0)
    assert espressopp.analysis.NPart(system).compute() == 500
    integ.run(0)
    # from the reference's fixture the builder gives a system that runs
    system, integ = espressopp.standard_system.LennardJones(0, rc=2.5, temperature=1.0, xyzfilename=FIXTURE_2048)
    assert espressopp.analysis.NPart(system).compute() == 2048
    integ.run(20)
    T = espressopp.analysis.Temperature(system).compute()
    assert np.isfinite(system.storage.getParticleArrays()["pos"]).all() and 0.1 < T < 10
    # random walks without a push-off may overlap (as in the reference): forces only
    system, integ = espressopp.standard_system.PolymerMelt(10, 20, box=(12.0, 12.0, 12.0), temperature=1.0)
    fene = system.getInteraction(1)
    assert fene.getFixedPairList().totalSize() == 10 * 19
    integ.run(0)
    assert np.isfinite(system.storage.getParticleArrays()["f"]).all()
    system, integ = espressopp.standard_system.Minimal(100, (6.0, 6.0, 6.0))
    assert espressopp.analysis.NPart(system).compute() == 100
    system, integ = espressopp.vec.standard_system.Default((8.0, 8.0, 8.0), rc=2.5, skin=0.3)
    assert system.vectorization.storageVec is system.storage
