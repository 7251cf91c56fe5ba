"""espressopp.vec.interaction (src/vec/interaction/LennardJones.py, FENE.py):
LennardJones, VerletListLennardJones, FENE and FixedPairListFENE of the
vectorised path, the same GPU interactions here."""
from ..interaction import FENE, FixedPairListFENE, LennardJones, VerletListLennardJones

VerletListLennardJonesBase = VerletListLennardJones

__all__ = ["LennardJones", "VerletListLennardJones", "VerletListLennardJonesBase", "FENE", "FixedPairListFENE"]
