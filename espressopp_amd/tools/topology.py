"""espressopp.tools.topology.polymerRW (src/tools/topology.py:27-140): a random-walk
chain of fixed bond length, each step uniform on the sphere (z uniform in
[-b, b], azimuth uniform), optionally rejecting steps that fold back closer than
mindist to the bead two back."""
import math
import random

from .._vec3 import Real3D


def polymerRW(pid, startpos, numberOfMonomers, bondlength, return_angles=False, return_dihedrals=False,
              mindist=None, rng=None):
    draw = rng if rng is not None else (lambda: random.uniform(0, 1))
    x, y, z = float(startpos[0]), float(startpos[1]), float(startpos[2])
    positions = [Real3D(x, y, z)]
    bonds, angles, dihedrals = [], [], []
    avecostheta = 0.0
    for i in range(numberOfMonomers - 1):
        while True:
            nz = (2.0 * draw() - 1.0) * bondlength
            phi = 2.0 * math.pi * draw()
            rr = math.sqrt(bondlength * bondlength - nz * nz)
            nx, ny = rr * math.cos(phi), rr * math.sin(phi)
            if not (mindist and i > 0):
                break
            a = [positions[i][k] - positions[i - 1][k] for k in range(3)]
            b = [-nx, -ny, -nz]
            la = math.sqrt(sum(v * v for v in a))
            lb = math.sqrt(sum(v * v for v in b))
            if math.sqrt(sum((a[k] - b[k]) ** 2 for k in range(3))) > mindist:
                avecostheta += -sum(a[k] * b[k] for k in range(3)) / (la * lb)
                break
        x, y, z = x + nx, y + ny, z + nz
        positions.append(Real3D(x, y, z))
        bonds.append((pid + i, pid + i + 1))
        if return_angles and i < numberOfMonomers - 2:
            angles.append((pid + i, pid + i + 1, pid + i + 2))
        if return_dihedrals and i < numberOfMonomers - 3:
            dihedrals.append((pid + i, pid + i + 1, pid + i + 2, pid + i + 3))
    out = [positions, bonds]
    if return_angles:
        out.append(angles)
    if return_dihedrals:
        out.append(dihedrals)
    if mindist:
        out.append(avecostheta / max(numberOfMonomers - 2, 1))
    return tuple(out)
