"""espressopp_amd -- MI355X-native short-range MD step of ESPResSo++.

A drop-in for the reference's Python API on this path: ``import
espressopp_amd as espressopp`` and the classes System, storage.DomainDecomposition,
VerletList, FixedPairList, interaction.LennardJones / VerletListLennardJones /
FENE / FixedPairListFENE, integrator.VelocityVerlet / LangevinThermostat and
analysis.Temperature / Pressure run the step on one GPU through the C ABI of
libespp_amd.so (include/espp_amd.h); interaction.Tabulated and
ParallelTempering cover the tabulated potentials and replica exchange;
espressopp_amd.vec and standard_system carry the reference's vectorised-path
names and ready-made systems, batch the graph-replayed replica batch, and
interaction.VerletListPairMLP the learned coarse-grained pair force.
"""
from . import _lib
from ._vec3 import Int3D, Real3D, toInt3DFromVector, toReal3DFromVector
from .system import MPI, FixedPairList, System, VerletList
from . import analysis, bc, esutil, integrator, interaction, storage, tools
from .parallel_tempering import MultiSystem, ParallelTempering
from . import batch, standard_system, vec

infinity = float("inf")
nan = float("nan")

__all__ = ["System", "VerletList", "FixedPairList", "ParallelTempering", "MultiSystem", "Real3D", "Int3D", "MPI",
           "analysis", "bc", "batch", "standard_system", "vec",
           "esutil", "integrator", "interaction", "storage", "tools", "infinity",
           "toReal3DFromVector", "toInt3DFromVector"]
