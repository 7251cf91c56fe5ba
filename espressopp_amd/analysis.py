"""espressopp.analysis: Temperature, Pressure, PressureTensor, NPart, RadialDistrF,
Configurations.

Reference: src/analysis/Temperature.hpp:60-125 (T = sum m v^2 / (3 N), no
centre-of-mass correction), src/analysis/Pressure.cpp:41-121
(P = (sum m v^2 + sum_interactions virial) / (3 V)),
src/analysis/PressureTensor.hpp:51-82.  The sums are FP64 reductions on the GPU.
"""
import ctypes

import numpy as np

from . import _lib
from ._vec3 import Real3D


class Observable:
    def __init__(self, system):
        self.system = system

    def _h(self):
        return self.system.storage._h


class Temperature(Observable):
    def __init__(self, system):
        Observable.__init__(self, system)
        self._types = None

    def add_type(self, type_id):
        if self._types is None:
            self._types = set()
        self._types.add(int(type_id))

    def remove_type(self, type_id):
        if self._types is None or int(type_id) not in self._types:
            print('Warning, type {} not found'.format(type_id))
            return
        self._types.discard(int(type_id))

    def compute(self):
        h = self._h()
        mask = 0xFFFFFFFFFFFFFFFF
        if self._types is not None:
            if any(t >= 64 for t in self._types):
                raise ValueError("type filters support types below 64")
            mask = sum(1 << t for t in self._types)
        mv2 = ctypes.c_double(0.0)
        cnt = ctypes.c_int64(0)
        h.call("espp_kinetic_tensor", mask, ctypes.byref(mv2), ctypes.byref(cnt), None)
        return mv2.value / (3.0 * cnt.value)

    def getEkin(self):
        return 0.5 * self._h().out_dbl("espp_kinetic")


class Pressure(Observable):
    def compute(self):
        h = self._h()
        L = self.system.bc.boxL
        tripleV = 3.0 * L[0] * L[1] * L[2]
        mv2 = h.out_dbl("espp_kinetic")
        w = sum(ia.computeVirial() for ia in self.system._interactions)
        return (mv2 + w) / tripleV


class PressureTensor(Observable):
    """Kinetic + virial tensor over V, components xx, yy, zz, xy, xz, yz."""

    def compute(self):
        h = self._h()
        L = self.system.bc.boxL
        V = L[0] * L[1] * L[2]
        kin = np.zeros(6)
        h.call("espp_kinetic_tensor", 0xFFFFFFFFFFFFFFFF, None, None, _lib.ptr(kin))
        w = np.zeros(6)
        for ia in self.system._interactions:
            w += np.asarray(ia.computeVirialTensor())
        return tuple((kin + w) / V)


class NPart(Observable):
    def compute(self):
        return self._h().n


class RadialDistrF(Observable):
    """g(r) over all pairs up to half the box (src/analysis/RadialDistrF.cpp:47-140,
    src/analysis/RadialDistrF.py): compute(rdfN) returns rdfN values for bins of
    width (L_y / 2) / rdfN.  The pair histogram is a GPU kernel."""

    def __init__(self, system):
        Observable.__init__(self, system)
        self.print_progress = False

    def compute(self, rdfN):
        out = np.zeros(int(rdfN))
        self._h().call("espp_rdf", int(rdfN), out.ctypes.data_as(ctypes.c_void_p))
        return out.tolist()


class Configuration:
    """One gathered snapshot (src/analysis/Configuration.cpp:103-140): coordinates,
    velocities and forces by particle id; an id that is not stored gives (0, 0, 0),
    as the reference's map lookup does."""

    def __init__(self, ids, pos=None, vel=None, force=None):
        self._index = {int(p): k for k, p in enumerate(ids)}
        self._ids = [int(p) for p in ids]
        self._pos, self._vel, self._force = pos, vel, force

    def _get(self, arr, what, pid):
        if arr is None:
            print("Error: This configuration has no information about %s" % what)
            return Real3D(0.0, 0.0, 0.0)
        k = self._index.get(int(pid))
        return Real3D(0.0, 0.0, 0.0) if k is None else Real3D(arr[k])

    def getCoordinates(self, pid):
        return self._get(self._pos, "coordinates", pid)

    __getitem__ = getCoordinates

    def getVelocities(self, pid):
        return self._get(self._vel, "velocities", pid)

    def getForces(self, pid):
        return self._get(self._force, "forces", pid)

    def getIds(self):
        return list(self._ids)

    @property
    def size(self):
        return len(self._ids)


class Configurations(Observable):
    """analysis.Configurations(system, pos, vel, force, radius, folded)
    (src/analysis/Configurations.py:76-98, Configurations.cpp:120-190): gather()
    appends a snapshot of the particles; [k] is the k-th snapshot.  folded=False
    stores unfolded positions (position + image * box)."""

    def __init__(self, system, pos=True, vel=False, force=False, radius=False, folded=True):
        Observable.__init__(self, system)
        self._flags = (pos, vel, force)
        self._folded = folded
        self._snaps = []
        self.capacity = 0

    def gather(self):
        a = self.system.storage.getParticleArrays()
        p = a["pos"] if self._folded else a["pos"] + a["image"] * np.asarray(list(self.system.bc.boxL))
        snap = Configuration(a["id"], p if self._flags[0] else None, a["v"] if self._flags[1] else None,
                             a["f"] if self._flags[2] else None)
        if self.capacity and len(self._snaps) >= self.capacity:
            self._snaps.pop(0)
        self._snaps.append(snap)

    def __getitem__(self, k):
        return self._snaps[k]

    def __iter__(self):
        return iter(self._snaps)

    def back(self):
        return self._snaps[-1]

    def clear(self):
        self._snaps = []

    @property
    def size(self):
        return len(self._snaps)
