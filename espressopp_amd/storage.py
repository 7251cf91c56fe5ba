"""espressopp.storage.DomainDecomposition on one GPU.

Reference: src/storage/DomainDecomposition.py:59-114 (argument checks),
src/storage/Storage.py:233-430 (particle insertion and access),
src/storage/DomainDecomposition.cpp:70-170 (grid set-up).  The node grid must
be (1, 1, 1): one GPU owns the whole box.
"""
import numpy as np

from . import _lib
from ._vec3 import Int3D, Real3D, toInt3DFromVector

_PROPS = ("id", "pos", "type", "mass", "v", "f", "q", "modepos", "modemom", "fm", "radius",
          "fradius", "vradius", "varmass", "adrat", "lambda_adr", "lambda_adrd", "state", "pib")
# the properties the GPU storage keeps; the others of the reference's Particle
# (charge, radius, AdResS and path-integral fields, ...) belong to no part of this
# path and are refused rather than dropped
_STORED = ("id", "pos", "type", "mass", "v")


class Particle:
    """Particle accessor (src/Particle.py): each property reads the device copy."""

    def __init__(self, pid, storage):
        self.pid = int(pid)
        self.storage = storage

    def _row(self, field):
        # one particle by id through the device's id -> slot map: O(1) in the particle count
        try:
            return self.storage._h.get_by_id(field, [self.pid])[0]
        except ValueError:
            raise RuntimeError("particle %d does not exist" % self.pid) from None

    @property
    def id(self):
        return self.pid

    @property
    def pos(self):
        return Real3D(self._row("pos"))

    @pos.setter
    def pos(self, v):
        self.storage._h.set("pos", [self.pid], np.asarray(list(v), dtype=np.float64))

    @property
    def v(self):
        return Real3D(self._row("v"))

    @v.setter
    def v(self, val):
        self.storage._h.set("v", [self.pid], np.asarray(list(val), dtype=np.float64))

    @property
    def f(self):
        return Real3D(self._row("f"))

    @property
    def imageBox(self):
        return Int3D(*[int(x) for x in self._row("image")])

    @property
    def type(self):
        return int(self._row("type"))

    @type.setter
    def type(self, t):
        self.storage._h.set("type", [self.pid], [float(t)])

    @property
    def mass(self):
        return float(self._row("mass"))

    @mass.setter
    def mass(self, m):
        self.storage._h.set("mass", [self.pid], [float(m)])


class DomainDecomposition:
    def __init__(self, system, nodeGrid="auto", cellGrid="auto", halfCellInt="auto", nocheck=False):
        if system.bc is None:
            raise RuntimeError("system has no boundary conditions")
        nodeGrid = Int3D(1, 1, 1) if nodeGrid == "auto" else toInt3DFromVector(nodeGrid)
        if tuple(nodeGrid) != (1, 1, 1):
            # NodeGridMismatch (src/storage/DomainDecomposition.cpp:94-97): one GPU owns the box
            raise RuntimeError("node grid %s does not match the number of GPU ranks (1)" % (tuple(nodeGrid),))
        cellGrid = Int3D(2, 2, 2) if cellGrid == "auto" else toInt3DFromVector(cellGrid)
        halfCellInt = 1 if halfCellInt == "auto" else int(halfCellInt)
        if not nocheck:
            for k in range(3):
                if nodeGrid[k] * cellGrid[k] == 1:
                    print("Warning! cellGrid[{}] has been adjusted to 2 (was={})".format(k, cellGrid[k]))
                    cellGrid[k] = 2
        self.system = system
        self.nodeGrid = nodeGrid
        self.cellGrid = cellGrid
        self.halfCellInt = halfCellInt
        seed = system.rng.get_seed() if system.rng is not None else 12345
        precision = {"fp64": _lib.ESPP_FP64, "mixed": _lib.ESPP_MIXED}[system.precision]
        self._h = _lib.Handle(system.bc.boxL, cellGrid, system.skin, seed, precision, halfCellInt)

    def getCellGrid(self):
        return Int3D(self.cellGrid)

    def getNodeGrid(self):
        return Int3D(self.nodeGrid)

    # -------------------------------------------------------------- particles
    def addParticle(self, pid, pos):
        self.addParticles([(pid, pos)], "id", "pos")
        return Particle(pid, self)

    def addParticles(self, particleList, *properties):
        """Storage.addParticles (src/storage/Storage.py:233-330)."""
        if not properties:
            properties = ("id", "pos")
        idx = {}
        for k, p in enumerate(properties):
            key = p.lower()
            if key not in _PROPS:
                raise SyntaxError("unknown particle property: %s" % p)
            idx[key] = k
        unstored = [p for p in idx if p not in _STORED]
        if unstored:
            raise NotImplementedError("particle properties %s are not stored by the GPU storage (it keeps %s)"
                                      % (", ".join(unstored), ", ".join(_STORED)))
        if "id" not in idx:
            raise Exception("particle property id is mandatory")
        if "pos" not in idx:
            raise Exception("particle property pos is mandatory")
        plist = list(particleList)
        existing = set(self._h.ids().tolist()) if self._h.n else set()
        for part in plist:
            if int(part[idx["id"]]) in existing:
                print("WARNING: Particle ", part[idx["id"]], " already exists")
                print("WARNING: Some particles already exist. The list of particles was not added.")
                return
        for part in plist:
            if len(part) != len(properties):
                raise SyntaxError("particle has %d entries, but %d expected" % (len(part), len(properties)))
        n = len(plist)
        if n == 0:
            return
        ids = np.fromiter((int(p[idx["id"]]) for p in plist), dtype=np.int32, count=n)
        pos = np.array([list(p[idx["pos"]]) for p in plist], dtype=np.float64).reshape(n, 3)
        types = (np.fromiter((int(p[idx["type"]]) for p in plist), dtype=np.int32, count=n)
                 if "type" in idx else None)
        vel = (np.array([list(p[idx["v"]]) for p in plist], dtype=np.float64).reshape(n, 3)
               if "v" in idx else None)
        mass = (np.fromiter((float(p[idx["mass"]]) for p in plist), dtype=np.float64, count=n)
                if "mass" in idx else None)
        self.addParticlesArray(ids, pos, types, vel, mass)

    def addParticlesArray(self, ids, pos, types=None, vel=None, mass=None):
        """Bulk insertion from arrays (Storage::addParticlesFromArray, src/storage/Storage.cpp:1052)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        t = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        self._h.call("espp_add_particles", len(ids), _lib.ptr(ids), _lib.ptr(t), _lib.ptr(pos),
                     _lib.ptr(v), _lib.ptr(m))

    def decompose(self):
        self._h.call("espp_decompose")

    def particleExists(self, pid):
        ids = self._h.ids()
        k = int(np.searchsorted(ids, pid))
        return k < len(ids) and ids[k] == pid

    def getParticle(self, pid):
        return Particle(pid, self)

    def getRealParticleIDs(self):
        return [self._h.ids().tolist()]

    def getNRealParticles(self):
        return self._h.n

    def getParticleArrays(self):
        """ids (ascending) with positions, velocities, forces, images, types and masses."""
        h = self._h
        return dict(id=h.ids(), pos=h.get("pos"), v=h.get("v"), f=h.get("f"),
                    image=h.get("image").astype(np.int64), type=h.get("type").astype(np.int64),
                    mass=h.get("mass"))

    def setParticleArrays(self, field, ids, values):
        self._h.set(field, ids, values)

    def printRealParticles(self):
        a = self.getParticleArrays()
        for pid, t, p in zip(a["id"], a["type"], a["pos"]):
            print("CPU %-3i ID %-5i TYPE %-3i POS(%8.3f, %8.3f, %8.3f)" % (0, pid, t, p[0], p[1], p[2]))


def Storage(*args, **kwargs):
    raise TypeError("use storage.DomainDecomposition")
