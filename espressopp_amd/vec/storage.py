"""espressopp.vec.storage.DomainDecomposition(system, nodeGrid, cellGrid, halfCellInt=1)
(src/vec/storage/DomainDecomposition.py:11-40): the same storage as
espressopp.storage.DomainDecomposition on the GPU path."""
from ..storage import DomainDecomposition as _DD


class DomainDecomposition(_DD):
    def __init__(self, system, nodeGrid, cellGrid, halfCellInt=1):
        _DD.__init__(self, system, nodeGrid, cellGrid, halfCellInt)


StorageVec = DomainDecomposition
