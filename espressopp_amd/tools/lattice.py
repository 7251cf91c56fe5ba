"""Simple-cubic lattice (src/tools/lattice.py:24-71)."""
import random as _random


def createCubic(N, rho, perfect=True, RNG=None):
    L = (N / rho) ** (1.0 / 3.0)
    a = int(N ** (1.0 / 3.0))
    if a ** 3 < N:
        a = a + 1
    lattice_spacing = L / a

    def rnd(magn_):
        rand = _random.random() if RNG is None else RNG()
        return magn_ * (2.0 * rand - 1.0)

    magn = 0.0 if perfect else lattice_spacing / 10.0
    ct = 0
    x, y, z = [], [], []
    for i in range(a):
        for j in range(a):
            for k in range(a):
                if ct < N:
                    x.append(0.5 * lattice_spacing + i * lattice_spacing + rnd(magn))
                    y.append(0.5 * lattice_spacing + j * lattice_spacing + rnd(magn))
                    z.append(0.5 * lattice_spacing + k * lattice_spacing + rnd(magn))
                    ct += 1
    return x, y, z, L, L, L
