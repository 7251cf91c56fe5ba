"""espressopp.tools subset used by the short-range MD path: decomp, lattice,
velocities, readxyz, replicate, analyse, timers."""
from . import analyse, decomp, lattice, timers, velocities
from .io import readxyz, writexyz
from .replicate import replicate
