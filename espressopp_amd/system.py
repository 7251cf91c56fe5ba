"""espressopp.System, VerletList and FixedPairList over the C ABI.

Reference: src/System.py:111-232 (System), src/VerletList.py:66-126,
src/FixedPairList.py:79-148.  One process drives one GPU, so the PMI
controller/worker split of the reference collapses to plain objects; calls the
reference marks ``pmiinvoke`` still return one result per rank (a list of one).
"""
import numpy as np

from . import _lib


class _Comm:
    size = 1
    rank = 0


class _MPI:
    COMM_WORLD = _Comm()


MPI = _MPI()


class System:
    """Global container: rng, bc, skin, storage and the short-range interactions
    (src/System.hpp:46-100).  ``precision`` selects the pair path of the GPU
    storage created afterwards: 'fp64' (default, reference arithmetic) or
    'mixed' (FP32 pair math with FP64 integration)."""

    def __init__(self, precision="fp64"):
        self.rng = None
        self.bc = None
        self._skin = 0.0
        self.storage = None
        self.integrator = None
        self.comm = MPI.COMM_WORLD
        self.maxCutoff = 0.0
        self.precision = precision
        self._interactions = []
        self._interaction2id = {}
        self._interaction_pid = 0

    @property
    def skin(self):
        return self._skin

    @skin.setter
    def skin(self, s):
        if s < 0:
            raise ValueError("skin must be non-negative")
        self._skin = float(s)
        if self.storage is not None:
            self.storage._h.call("espp_set_skin", float(s))

    def _handle(self):
        if self.storage is None:
            raise RuntimeError("system has no storage")
        return self.storage._h

    def addInteraction(self, interaction, name=None):
        # System::addInteraction (src/System.cpp:75-85): order defines force summation order
        if interaction._system_attached is not None and interaction._system_attached is not self:
            raise RuntimeError("interaction belongs to another system")
        interaction._attach(self)
        self._interactions.append(interaction)
        try:
            self.maxCutoff = max(self.maxCutoff, interaction.getMaxCutoff())
        except AttributeError:
            pass
        if name is not None:
            if name in self._interaction2id:
                raise RuntimeError("Interaction with name {} already defined.".format(name))
            self._interaction2id[name] = self._interaction_pid
        self._interaction_pid += 1
        return True

    def getNumberOfInteractions(self):
        return len(self._interactions)

    def getInteraction(self, number):
        if not self._interactions:
            raise RuntimeError("interaction list of system is empty")
        if 0 <= number < len(self._interactions):
            return self._interactions[number]
        raise RuntimeError("Interaction number %i does not exist" % number)

    def getInteractionByName(self, name):
        return self.getInteraction(self._interaction2id[name])

    def getAllInteractions(self):
        return {k: self.getInteraction(v) for k, v in self._interaction2id.items()}

    def removeInteraction(self, number):
        """System::removeInteraction (src/System.cpp:108-135): drop the interaction at
        `number` of the short-range list and recompute maxCutoff if it held the largest."""
        if not 0 <= number < len(self._interactions):
            raise RuntimeError("Probably you are trying to remove the interaction %d which does not exist. "
                               "Check your script!" % number)
        ia = self._interactions[number]
        self._handle().call("espp_remove_interaction", int(number))
        del self._interactions[number]
        ia._system_attached = None
        try:
            cut = ia.getMaxCutoff()
        except AttributeError:
            cut = 0.0
        if cut >= self.maxCutoff:
            self.maxCutoff = 0.0
            for other in self._interactions:
                try:
                    self.maxCutoff = max(self.maxCutoff, other.getMaxCutoff())
                except AttributeError:
                    pass

    def removeInteractionByName(self, name):
        """src/System.py:150-161: remove by name and renumber the later names."""
        if name not in self._interaction2id:
            raise RuntimeError("Interaction {} not found".format(name))
        iid = self._interaction2id[name]
        self.removeInteraction(iid)
        self._interaction2id = {k: v if v < iid else v - 1 for k, v in self._interaction2id.items() if v != iid}
        self._interaction_pid = max(self._interaction2id.values()) + 1 if self._interaction2id else 0

    def setTrace(self, switch):
        pass


class VerletList:
    """Verlet list with skin (src/VerletList.py:66-126, src/VerletList.cpp:44-117).

    Holds every pair with d^2 <= (cutoff + skin)^2 and is rebuilt on every
    decompose.  The device keeps it as a full (both-direction) list; sizes and
    pair dumps count each pair once, as the reference's half list does."""

    def __init__(self, system, cutoff, exclusionlist=[], useBuffers=True, useSOA=False):
        if system.storage is None:
            raise RuntimeError("system has no storage")
        self.system = system
        self.cutoff = float(cutoff)
        h = system.storage._h
        self._h = h
        self._id = h.out_int("espp_vl_create", float(cutoff))
        if exclusionlist:
            self.exclude(exclusionlist)

    @property
    def builds(self):
        return self._h.out_i64("espp_vl_builds", self._id)

    def totalSize(self):
        return self._h.out_i64("espp_vl_size", self._id)

    def localSize(self):
        return self.totalSize()

    def exclude(self, exclusionlist):
        pairs = np.asarray(list(exclusionlist), dtype=np.int32).reshape(-1, 2)
        self._h.call("espp_vl_exclude", self._id, len(pairs), _lib.ptr(pairs))

    def rebuild(self):
        self._h.call("espp_vl_rebuild", self._id)

    def pairs_array(self):
        """All pairs as an (n, 2) int32 array of ids, first id < second id."""
        total = self.totalSize()
        out = np.zeros((max(total, 1), 2), dtype=np.int32)
        n = self._h.out_i64("espp_vl_pairs", self._id, _lib.ptr(out), total)
        return out[:n]

    def getAllPairs(self):
        return [[tuple(int(x) for x in p) for p in self.pairs_array()]]

    def getPair(self, i):
        pairs = self.pairs_array()
        if i <= 0 or i > len(pairs):
            return ()
        return tuple(int(x) for x in pairs[i - 1])

    def getVerletCutoff(self):
        return self._h.out_dbl("espp_vl_cutoff", self._id)

    def connect(self):
        pass

    def disconnect(self):
        pass

    def resetTimers(self):
        pass

    def getTimers(self):
        return [(0.0,)]


class FixedPairList:
    """Bond list (src/FixedPairList.py:79-148, src/FixedPairList.cpp:88-141)."""

    def __init__(self, storage):
        self.storage = storage
        self._h = storage._h
        self._id = self._h.out_int("espp_fpl_create")
        self._bonds = []

    def add(self, pid1, pid2):
        p = np.asarray([[pid1, pid2]], dtype=np.int32)
        self._h.call("espp_fpl_add", self._id, 1, _lib.ptr(p))
        self._bonds.append((int(pid1), int(pid2)))
        return True

    def addBonds(self, bondlist):
        p = np.asarray(list(bondlist), dtype=np.int32).reshape(-1, 2)
        if len(p):
            self._h.call("espp_fpl_add", self._id, len(p), _lib.ptr(p))
            self._bonds.extend((int(a), int(b)) for a, b in p)

    def size(self):
        return self._h.out_i64("espp_fpl_size", self._id)

    def totalSize(self):
        return self.size()

    def getBonds(self):
        return [list(self._bonds)]

    def getAllBonds(self):
        return list(self._bonds)
