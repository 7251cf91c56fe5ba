"""espressopp.vec: the names of the reference's vectorised path (src/vec/__init__.py).

In the reference, espressopp.vec is a second, SIMD-vectorised implementation of
the same short-range step (vec.Vectorization holding a particle array, a CSR
Verlet list, LennardJones/FENE kernels, VelocityVerlet and LangevinThermostat).
Here both name sets drive the same GPU path, so a script written against
espressopp.vec runs unchanged: Vectorization, VerletList, FixedPairList,
storage.DomainDecomposition, integrator.VelocityVerlet / LangevinThermostat,
interaction.LennardJones / VerletListLennardJones / FENE / FixedPairListFENE and
standard_system.Default.
"""
from ..system import FixedPairList, VerletList
from . import integrator, interaction, standard_system, storage


class Vectorization:
    """vec.Vectorization(system, integrator=None) (src/vec/Vectorization.py:19-45).
    The reference keeps a SIMD particle array beside the storage; the GPU path's
    storage is already a structure of arrays, so storageVec is the storage."""

    def __init__(self, system, integrator=None):
        self.system = system
        self.integrator = integrator

    @property
    def storageVec(self):
        return self.system.storage


__all__ = ["Vectorization", "VerletList", "FixedPairList", "storage", "integrator", "interaction",
           "standard_system"]
