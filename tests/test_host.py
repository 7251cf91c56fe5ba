"""Host-side logic of the drop-in package (no GPU needed)."""
import math

import numpy as np
import pytest

import espressopp_amd as espressopp
from espressopp_amd.tools import decomp, lattice, velocities

from _systems import FIXTURE_2048


def test_cell_grid_truncates():
    # src/tools/decomp.py:322-342 -> Int3D(float) truncation
    assert tuple(decomp.cellGrid((13.4367695311,) * 3, (1, 1, 1), 2.5, 0.4)) == (4, 4, 4)
    assert tuple(decomp.cellGrid((170.636,) * 3, (1, 1, 1), 2.5, 0.3)) == (60, 60, 60)
    with pytest.raises(decomp.Error):
        decomp.cellGrid((2.0, 10.0, 10.0), (1, 1, 1), 2.5, 0.3)


def test_node_grid_single_gpu():
    assert tuple(decomp.nodeGrid(1, (10, 10, 10), 2.5, 0.3)) == (1, 1, 1)


def test_int3d_real3d():
    assert tuple(espressopp.Int3D(4.9, 3.2, 1.0)) == (4, 3, 1)
    r = espressopp.Real3D(1.0, 2.0, 2.0)
    assert r.abs() == 3.0 and r.sqr() == 9.0
    assert (r - espressopp.Real3D(1.0)) * espressopp.Real3D(0, 1, 1) == 2.0


def test_readxyz_fixture():
    pid, t, x, y, z, vx, vy, vz, Lx, Ly, Lz = espressopp.tools.readxyz(FIXTURE_2048)
    assert len(pid) == 2048 and Lx == 13.4367695311
    assert abs(2048 / (Lx * Ly * Lz) - 0.8442) < 1e-4


def test_create_cubic_lattice():
    x, y, z, L, _, _ = lattice.createCubic(1000, 0.8442, perfect=True)
    assert len(x) == 1000 and abs(L - (1000 / 0.8442) ** (1 / 3)) < 1e-12
    assert min(x) > 0 and max(x) < L


def test_gaussian_velocities_exact_temperature():
    vx, vy, vz = velocities.gaussian(1.5, 4000, zero_momentum=False)
    v2 = np.sum(np.array(vx) ** 2 + np.array(vy) ** 2 + np.array(vz) ** 2)
    assert v2 / (3 * 4000) == pytest.approx(1.5, rel=1e-12)


def test_potential_objects():
    lj = espressopp.interaction.LennardJones(1.0, 1.0, cutoff=2.5)
    assert lj.shift == pytest.approx(4 * (2.5 ** -12 - 2.5 ** -6))
    assert lj.computeEnergy(2.6) == 0.0
    lj2 = espressopp.interaction.LennardJones(1.0, 1.0, cutoff=math.inf)
    assert lj2.computeEnergy(2 ** (1 / 6)) == pytest.approx(-1.0)
    fene = espressopp.interaction.FENE(K=30.0, r0=0.0, rMax=1.5)
    assert fene.computeEnergy(0.97) == pytest.approx(18.279, rel=1e-4)
    assert tuple(fene.computeForce((0.9, 0.1, 0.1))) == pytest.approx((-42.782, -4.7535, -4.7535), rel=1e-4)


def test_folding_helper():
    bc = espressopp.bc.OrthorhombicBC(None, (10.0, 10.0, 10.0))
    p, img = bc.getFoldedPosition((12.5, -0.5, 3.0))
    assert tuple(p) == pytest.approx((2.5, 9.5, 3.0)) and img == [1, -1, 0]
    assert tuple(bc.getMinimumImageVector((9.5, 0, 0), (0.5, 0, 0))) == pytest.approx((-1.0, 0, 0))


def test_run_requires_integer_steps():
    class Dummy(espressopp.integrator.VelocityVerlet):
        def __init__(self):
            self._dt = 0.005
    with pytest.raises(ValueError):
        Dummy().run(1.5)


def test_read_table_follows_sscanf_rule(tmp_path):
    # InterpolationTable::readFile: a line counts when sscanf("%lg %lg %lg") reads three numbers
    from espressopp_amd.interaction import read_table
    fn = tmp_path / "t.tab"
    fn.write_text("# header\n\n0.5 1.0 2.0\n1.0e0 -2 +3.5 trailing words\n1.5 2.0\n2.0 x 3.0\n"
                  "  2.5\t-1e-3   4E2\n3.0 1.0 2.0abc\n")
    r, e, f = read_table(str(fn))
    assert list(r) == [0.5, 1.0, 2.5, 3.0]
    assert list(e) == [1.0, -2.0, -1e-3, 1.0]
    assert list(f) == [2.0, 3.5, 400.0, 2.0]


def test_tabulated_potential_checks_its_table(tmp_path):
    import espressopp_amd as espressopp
    fn = tmp_path / "short.tab"
    fn.write_text("0.5 1.0 2.0\n1.0 0.5 1.0\n1.5 0.2 0.4\n")
    espressopp.interaction.Tabulated(1, str(fn), 2.0)
    espressopp.interaction.Tabulated(3, str(fn), 2.0)
    with pytest.raises(RuntimeError):
        espressopp.interaction.Tabulated(2, str(fn), 2.0)     # Akima needs four rows
    with pytest.raises(ValueError):
        espressopp.interaction.Tabulated(4, str(fn), 2.0)


_LAMMPS_BONDED = """LAMMPS data file, a short chain

3 atoms
2 bonds
0 angles
0 dihedrals
0 impropers

1 atom types
1 bond types
0 angle types

0.0 10.0 xlo xhi
-1.0 9.0 ylo yhi
2.0 14.0 zlo zhi

Atoms

1 1 1 1.0 2.0 3.0
2 1 1 1.5 2.5 3.5
3 1 1 2.0 3.0 4.5

Bonds

1 1 1 2
2 1 2 3
"""

_LAMMPS_TYPES = """LAMMPS data file, charges

2 atoms
0 bonds
0 angles
0 dihedrals
0 impropers

2 atom types

0.0 5.0 xlo xhi
0.0 6.0 ylo yhi
0.0 7.0 zlo zhi

Atoms

1 1 2 -0.5 1.0 1.0 1.0
2 1 1 0.5 2.0 2.0 2.0
"""


def test_lammps_read(tmp_path):
    # src/tools/lammps.py:33-156: return tuples by content
    from espressopp_amd.tools import lammps
    fn = tmp_path / "bonded.data"
    fn.write_text(_LAMMPS_BONDED)
    bonds, x, y, z, Lx, Ly, Lz = lammps.read(str(fn))
    assert bonds == [(1, 2), (2, 3)]
    assert x == [1.0, 1.5, 2.0] and z == [3.0, 3.5, 4.5]
    assert (Lx, Ly, Lz) == (10.0, 10.0, 12.0)
    fn2 = tmp_path / "types.data"
    fn2.write_text(_LAMMPS_TYPES)
    p_type, bonds, angles, q, x, y, z, Lx, Ly, Lz = lammps.read(str(fn2))
    assert p_type == [2, 1] and q == [-0.5, 0.5] and bonds == [] and angles == []
    assert (x, Lx, Ly, Lz) == ([1.0, 2.0], 5.0, 6.0, 7.0)
