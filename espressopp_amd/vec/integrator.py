"""espressopp.vec.integrator (src/vec/integrator/*.py): VelocityVerlet and
LangevinThermostat of the vectorised path, the same GPU integrator here."""
from ..integrator import Extension, LangevinThermostat, VelocityVerlet

MDIntegratorVec = VelocityVerlet
VelocityVerletBase = VelocityVerlet

__all__ = ["VelocityVerlet", "VelocityVerletBase", "MDIntegratorVec", "LangevinThermostat", "Extension"]
