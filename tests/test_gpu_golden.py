"""The reference's own golden values, reproduced by the HIP path.

testsuite/tabulated/result.dat (lines 14 and 22) prints the step-0 pair and bond energies of
polymer_melt_tabulated.py on polymer_melt.start, with plain LJ + FENE and with Akima tables
of both, and the pressure and Pxy of the plain run.  Both GPU precisions compute them through
the package's Python API: the FP64 path must land within half a unit of the printed digit,
the mixed path within that plus 1e-6 relative (its energy and pressure bar).
"""
import numpy as np
import pytest

import _melt_golden as mg

pytestmark = pytest.mark.gpu

import espressopp_amd as espressopp  # noqa: E402


def _write_tab(path, rows):
    r, e, f = rows
    with open(path, "w") as fh:
        for a, b, c in zip(r, e, f):
            fh.write("%15.8g %15.8g %15.8g\n" % (a, b, c))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("run,hci", [("plain", 1), ("tab", 1), ("plain", 2), ("tab", 2)])
def test_melt_golden_step0(gpu, tmp_path, precision, run, hci):
    # the reference runs this script with full cells and with half cells
    # (testsuite/tabulated/CMakeLists.txt: polymer_melt_tabulated.py 1 and 2)
    ids, pos, box, bonds, angles = mg.load(tmp_path)
    golden = mg.GOLDEN_PLAIN if run == "plain" else mg.GOLDEN_TAB
    cells = espressopp.tools.decomp.cellGrid(box, (1, 1, 1), mg.RC, mg.SKIN, hci)
    system = espressopp.System(precision=precision)
    system.rng = espressopp.esutil.RNG(54321)
    system.bc = espressopp.bc.OrthorhombicBC(system.rng, box)
    system.skin = mg.SKIN
    system.storage = espressopp.storage.DomainDecomposition(system, (1, 1, 1), cells, hci)
    system.storage.addParticles([[int(i), espressopp.Real3D(*p)] for i, p in zip(ids, pos)], "id", "pos")
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=mg.RC)
    fpl = espressopp.FixedPairList(system.storage)
    fpl.addBonds([tuple(b) for b in bonds.tolist()])
    if run == "plain":
        pair = espressopp.interaction.VerletListLennardJones(vl)
        pair.setPotential(0, 0, espressopp.interaction.LennardJones(epsilon=1.0, sigma=1.0, shift=False,
                                                                    cutoff=mg.RC))
        bond = espressopp.interaction.FixedPairListFENE(system, fpl,
                                                       espressopp.interaction.FENE(K=30.0, r0=0.0, rMax=1.5))
    else:
        fl, ff = str(tmp_path / "pot-lj.txt"), str(tmp_path / "pot-fene.txt")
        _write_tab(fl, mg.lj_table())
        _write_tab(ff, mg.fene_table())
        pair = espressopp.interaction.VerletListTabulated(vl)
        pair.setPotential(0, 0, espressopp.interaction.Tabulated(itype=2, filename=fl, cutoff=mg.RC))
        bond = espressopp.interaction.FixedPairListTabulated(system, fpl,
                                                            espressopp.interaction.Tabulated(itype=2, filename=ff))
    system.addInteraction(pair)
    system.addInteraction(bond)
    rel = 0.0 if precision == "fp64" else 1e-6
    for name, val in (("epair", pair.computeEnergy()), ("ebond", bond.computeEnergy())):
        assert abs(val - golden[name]) <= mg.HALF_UNIT[name] + rel * abs(golden[name]), (name, val)
    if run == "plain":
        V = float(np.prod(box))
        P = espressopp.analysis.Pressure(system).compute()
        assert abs(P - golden["P"]) <= mg.HALF_UNIT["P"] + rel * golden["P"], P
        # result.dat's angle term follows K (1 - cos(theta - theta0)), the opposite sign of
        # today's Cosine.hpp (the printed eangle confirms it, tests/test_oracle.py)
        _, wxy_ang = mg.cosine_virial(pos, box, angles)
        pxy = espressopp.analysis.PressureTensor(system).compute()[3] - wxy_ang / V
        assert abs(pxy - golden["Pxy"]) <= mg.HALF_UNIT["Pxy"] + rel * golden["Pxy"], pxy
