"""espressopp.vec.standard_system.Default (src/vec/standard_system/Default.py):
system, integrator and Vectorization context, no interactions, no particles;
Langevin gamma 1 at `temperature` when given."""
from .. import standard_system as _ss


def Default(box, rc=1.12246, skin=0.3, dt=0.005, temperature=None, halfCellInt=1, precision="fp64"):
    from . import Vectorization
    system, integ = _ss.Default(box, rc, skin, dt, temperature, halfCellInt, precision)
    system.vectorization = Vectorization(system, integ)
    return system, integ
