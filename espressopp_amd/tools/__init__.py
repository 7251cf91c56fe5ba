"""espressopp.tools subset used by the short-range MD path: decomp, lattice,
velocities, readxyz/writexyz, lammps.read, replicate, analyse, timers,
topology.polymerRW, checkpoint (save/restore) and coarse_grain."""
from . import analyse, checkpoint, decomp, lammps, lattice, timers, topology, velocities
from .io import readxyz, writexyz
from .replicate import replicate
