"""espressopp.analysis: Temperature, Pressure, PressureTensor, NPart, RadialDistrF.

Reference: src/analysis/Temperature.hpp:60-125 (T = sum m v^2 / (3 N), no
centre-of-mass correction), src/analysis/Pressure.cpp:41-121
(P = (sum m v^2 + sum_interactions virial) / (3 V)),
src/analysis/PressureTensor.hpp:51-82.  The sums are FP64 reductions on the GPU.
"""
import ctypes

import numpy as np

from . import _lib


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
