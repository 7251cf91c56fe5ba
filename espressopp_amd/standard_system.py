"""espressopp.standard_system: ready-made systems (src/standard_system/*.py).

Default and Minimal (src/standard_system/Default.py:46-66, Minimal.py): a system
with RNG, OrthorhombicBC, skin, a DomainDecomposition on the cell grid of
tools.decomp.cellGrid and a VelocityVerlet integrator (Langevin gamma 1 at
`temperature` when given).  LennardJones (LennardJones.py): the same with a
VerletListLennardJones and num_particles particles (random positions, or an xyz
file).  PolymerMelt (PolymerMelt.py): random-walk chains (tools.topology.polymerRW)
with WCA and FENE(K=30, r0=0, rMax=1.5) bonds.  Everything runs on the GPU path.
"""
from . import bc as _bc
from . import esutil as _esutil
from . import integrator as _integrator
from . import interaction as _interaction
from . import storage as _storage
from . import tools as _tools
from ._vec3 import Real3D
from .system import MPI, FixedPairList, System, VerletList


def _base(box, rc, skin, dt, temperature, halfCellInt=1, precision="fp64"):
    system = System(precision=precision)
    system.rng = _esutil.RNG()
    system.bc = _bc.OrthorhombicBC(system.rng, box)
    system.skin = skin
    nodeGrid = _tools.decomp.nodeGrid(MPI.COMM_WORLD.size, box, rc, skin)
    cellGrid = _tools.decomp.cellGrid(box, nodeGrid, rc, skin, halfCellInt)
    system.storage = _storage.DomainDecomposition(system, nodeGrid, cellGrid, halfCellInt)
    integ = _integrator.VelocityVerlet(system)
    integ.dt = dt
    if temperature is not None:
        thermostat = _integrator.LangevinThermostat(system)
        thermostat.gamma = 1.0
        thermostat.temperature = temperature
        integ.addExtension(thermostat)
    return system, integ


def Default(box, rc=1.12246, skin=0.3, dt=0.005, temperature=None, halfCellInt=1, precision="fp64"):
    """System and integrator, no interactions, no particles."""
    return _base(box, rc, skin, dt, temperature, halfCellInt, precision)


def _random_particles(system, num_particles):
    props = ["id", "type", "mass", "pos", "v"]
    zero = Real3D(0.0, 0.0, 0.0)
    parts = [[pid, 0, 1.0, system.bc.getRandomPos(), zero] for pid in range(1, num_particles + 1)]
    system.storage.addParticles(parts, *props)
    system.storage.decompose()


def Minimal(num_particles, box, rc=1.12246, skin=0.3, dt=0.005, temperature=None, precision="fp64"):
    """Default plus num_particles particles at random positions, at rest."""
    system, integ = _base(box, rc, skin, dt, temperature, 1, precision)
    _random_particles(system, num_particles)
    return system, integ


def _from_xyz(filename):
    pid, typ, x, y, z, vx, vy, vz, Lx, Ly, Lz = _tools.readxyz(filename)
    return pid, typ, x, y, z, vx, vy, vz, (Lx, Ly, Lz)


def LennardJones(num_particles, box=(0, 0, 0), rc=1.12246, skin=0.3, dt=0.005, epsilon=1.0, sigma=1.0, shift="auto",
                 temperature=None, xyzfilename=None, precision="fp64"):
    """LJ system of num_particles particles (random positions) or of an xyz file."""
    data = None
    if xyzfilename:
        data = _from_xyz(xyzfilename)
        box = data[-1]
    elif box[0] <= 0 or box[1] <= 0 or box[2] <= 0:
        box = (100.0, 100.0, 100.0)
    system, integ = _base(box, rc, skin, dt, temperature, 1, precision)
    lj = _interaction.VerletListLennardJones(VerletList(system, cutoff=rc))
    lj.setPotential(type1=0, type2=0, potential=_interaction.LennardJones(epsilon, sigma, rc, shift))
    system.addInteraction(lj)
    if data is None:
        _random_particles(system, num_particles)
    else:
        pid, typ, x, y, z, vx, vy, vz, _ = data
        parts = [[pid[k], typ[k], 1.0, Real3D(x[k], y[k], z[k]), Real3D(vx[k], vy[k], vz[k])]
                 for k in range(len(pid))]
        system.storage.addParticles(parts, "id", "type", "mass", "pos", "v")
        system.storage.decompose()
    return system, integ


def PolymerMelt(num_chains, monomers_per_chain, box=(0, 0, 0), bondlen=0.97, rc=1.12246, skin=0.3, dt=0.005,
                epsilon=1.0, sigma=1.0, shift="auto", temperature=None, xyzfilename=None, precision="fp64"):
    """Chains of WCA beads with FENE bonds; random walks unless an xyz file is given
    (chain k holds particles k*monomers_per_chain .. of the file)."""
    data = None
    if xyzfilename:
        data = _from_xyz(xyzfilename)
        box = data[-1]
    elif box[0] <= 0 or box[1] <= 0 or box[2] <= 0:
        box = (100.0, 100.0, 100.0)
    system, integ = _base(box, rc, skin, dt, temperature, 1, precision)
    lj = _interaction.VerletListLennardJones(VerletList(system, cutoff=rc))
    lj.setPotential(type1=0, type2=0, potential=_interaction.LennardJones(epsilon, sigma, rc, shift))
    system.addInteraction(lj)
    bondlist = FixedPairList(system.storage)
    props = ["id", "type", "mass", "pos", "v"]
    zero = Real3D(0.0, 0.0, 0.0)
    if data is None:
        pid, typ = 1, 0
        for _ in range(num_chains):
            positions, bonds = _tools.topology.polymerRW(pid, system.bc.getRandomPos(), monomers_per_chain, bondlen)
            system.storage.addParticles([[pid + k, typ, 1.0, positions[k], zero] for k in range(monomers_per_chain)],
                                        *props)
            bondlist.addBonds(bonds)
            pid += monomers_per_chain
            typ += 1
    else:
        p, t, x, y, z, vx, vy, vz, _ = data
        for i in range(num_chains):
            chain, bonds = [], []
            for k in range(monomers_per_chain):
                idx = i * monomers_per_chain + k
                chain.append([p[idx], t[idx], 1.0, Real3D(x[idx], y[idx], z[idx]), Real3D(vx[idx], vy[idx], vz[idx])])
                if k > 0:
                    bonds.append((p[idx - 1], p[idx]))
            system.storage.addParticles(chain, *props)
            bondlist.addBonds(bonds)
    system.storage.decompose()
    fene = _interaction.FixedPairListFENE(system, bondlist, _interaction.FENE(K=30.0, r0=0.0, rMax=1.5))
    system.addInteraction(fene)
    return system, integ


class KGMelt:
    """Box of a Kremer-Grest melt (src/standard_system/KGMelt.py): density 0.85."""

    def __init__(self, num_chains, chain_len):
        self._num_chains = num_chains
        self._chain_len = chain_len
        self._num_particles = num_chains * chain_len
        self._density = 0.85
        self._L = pow(self._num_particles / self._density, 1.0 / 3.0)
        self._box = (self._L, self._L, self._L)
        self._system = System()
