"""Host-side names of the drop-in (no GPU): espressopp.vec aliases, standard
systems, topology.polymerRW, bc.getRandomPos, analysis.Configuration."""
import math

import numpy as np

import espressopp_amd as espressopp


def test_vec_names_are_the_gpu_path():
    v = espressopp.vec
    assert v.VerletList is espressopp.VerletList and v.FixedPairList is espressopp.FixedPairList
    assert v.integrator.VelocityVerlet is espressopp.integrator.VelocityVerlet
    assert v.integrator.LangevinThermostat is espressopp.integrator.LangevinThermostat
    assert v.interaction.VerletListLennardJones is espressopp.interaction.VerletListLennardJones
    assert v.interaction.LennardJones is espressopp.interaction.LennardJones
    assert v.interaction.FixedPairListFENE is espressopp.interaction.FixedPairListFENE
    assert issubclass(v.storage.DomainDecomposition, espressopp.storage.DomainDecomposition)
    for name in ("Default", "Minimal", "LennardJones", "PolymerMelt", "KGMelt"):
        assert hasattr(espressopp.standard_system, name)


def test_polymer_random_walk_bonds():
    pos, bonds = espressopp.tools.topology.polymerRW(5, (1.0, 2.0, 3.0), 50, 0.97)
    assert len(pos) == 50 and bonds[0] == (5, 6) and bonds[-1] == (53, 54)
    d = [math.sqrt(sum((pos[k + 1][c] - pos[k][c]) ** 2 for c in range(3))) for k in range(49)]
    assert max(abs(x - 0.97) for x in d) < 1e-12
    out = espressopp.tools.topology.polymerRW(0, (0, 0, 0), 30, 1.0, return_angles=True, mindist=1.2)
    pos, bonds, angles, avecos = out
    assert len(angles) == 28
    # no step folds back closer than mindist to the bead two back
    p = np.array([list(x) for x in pos])
    assert (np.linalg.norm(p[2:] - p[:-2], axis=1) > 1.2 - 1e-12).all()


def test_random_position_inside_the_box():
    rng = espressopp.esutil.RNG(3)
    bc = espressopp.bc.OrthorhombicBC(rng, (4.0, 5.0, 6.0))
    for _ in range(100):
        p = bc.getRandomPos()
        assert 0 <= p[0] < 4 and 0 <= p[1] < 5 and 0 <= p[2] < 6


def test_configuration_lookup_by_id():
    c = espressopp.analysis.Configuration([3, 7], pos=np.array([[1.0, 2, 3], [4, 5, 6]]))
    assert list(c[7]) == [4, 5, 6] and list(c.getCoordinates(3)) == [1, 2, 3]
    assert list(c[5]) == [0, 0, 0]
    assert c.size == 2 and c.getIds() == [3, 7]
