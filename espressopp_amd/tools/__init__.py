"""espressopp.tools subset used by the short-range MD path: decomp, lattice,
velocities, readxyz/writexyz, lammps.read, replicate, analyse, timers and
checkpoint (save/restore)."""
from . import analyse, checkpoint, decomp, lammps, lattice, timers, velocities
from .io import readxyz, writexyz
from .replicate import replicate
