"""The reference's 40,000-bead polymer-melt start and the step-0 values its test prints.

Fixture: testsuite/tabulated/polymer_melt.start (LAMMPS data file: 40,000 beads, 40,000
FENE bonds, 40,000 cosine angles; gzip-compressed here), and testsuite/tabulated/result.dat
(the reference's own output of testsuite/tabulated/polymer_melt_tabulated.py).

Set-up of that script (polymer_melt_tabulated.py:45-200): ids 1..N, rc = 1.12, skin 0.3,
LJ(eps = sigma = 1, cutoff rc, shift = False) over a Verlet list, FENE(K = 30, r0 = 0,
rMax = 1.5) over the bonds, Cosine(K = 1.5, theta0 = 3.1415926) over the angles, and for the
tabulated run the same potentials as Akima tables written by writeTabFile (N = 257 rows,
"%15.8g" columns; LJ on [0.01, 1.12], FENE on [0.0001, 1.49]).  At step 0 all velocities
are zero, so P = W / (3V) and Pxy = W_xy / V, W being the summed virial of all three
interactions.

The script hands the Verlet list cutoff = rc + skin (so the list radius is rc + 2 skin,
larger than its cell size); only pairs within rc = 1.12 carry energy, so a list of radius
rc + skin over cells of at least that size gives the same energies and virials.
"""
import gzip
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
START = os.path.join(HERE, "golden", "polymer_melt.start.gz")

RC, SKIN = 1.12, 0.3
# result.dat lines 14 (plain potentials) and 22 (tabulated): printed with
# '%5d %8.4f %10.5f %8.5f %12.3f %12.3f %12.3f %12.3f %12.3f'
GOLDEN_PLAIN = dict(P=4.08575, Pxy=0.01663, epair=38010.180, ebond=722365.068, eangle=32301.755)
GOLDEN_TAB = dict(P=3.80758, Pxy=0.01577, epair=37865.384, ebond=722336.920, eangle=32301.056)
HALF_UNIT = dict(P=0.5e-5, Pxy=0.5e-5, epair=0.5e-3, ebond=0.5e-3, eangle=0.5e-3)


def load(tmp_path):
    """(ids, pos, box, bonds, angles) of the melt; ids and bond/angle ids are 1-based."""
    from espressopp_amd.tools import lammps
    path = os.path.join(str(tmp_path), "polymer_melt.start")
    with gzip.open(START, "rb") as src, open(path, "wb") as dst:
        dst.write(src.read())
    bonds, angles, x, y, z, Lx, Ly, Lz = lammps.read(path)
    n = len(x)
    return (np.arange(1, n + 1, dtype=np.int32), np.stack([x, y, z], axis=1), (Lx, Ly, Lz),
            np.asarray(bonds, dtype=np.int32), np.asarray(angles, dtype=np.int32))


def _table_rows(energy, force, N, low, high):
    """writeTabFile (polymer_melt_tabulated.py:94-107): rows r = low + i delta printed as
    '%15.8g %15.8g %15.8g', read back as the reference's table reader does."""
    delta = (high - low) / (N - 1)
    rows = []
    for i in range(N):
        r = low + i * delta
        line = "%15.8g %15.8g %15.8g" % (r, energy(r), force(r))
        rows.append([float(v) for v in line.split()])
    a = np.asarray(rows)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()


def lj_table():
    # LennardJones(1, 1, cutoff=1.12, shift=False): E = 4((1/r^2)^6 - (1/r^2)^3), F_x at (r, 0, 0)
    # (src/interaction/LennardJones.hpp:109-136); the table ends at the cutoff
    def e(r):
        f2 = 1.0 / (r * r)
        f6 = f2 * f2 * f2
        return 4.0 * (f6 * f6 - f6)

    def f(r):
        f2 = 1.0 / (r * r)
        f6 = f2 * f2 * f2
        return r * (f6 * (48.0 * f6 - 24.0) * f2)
    return _table_rows(e, f, 257, 0.01, RC)


def fene_table():
    # FENE(K=30, r0=0, rMax=1.5) (src/interaction/FENE.hpp:91-112)
    K, R = 30.0, 1.5

    def e(r):
        return -0.5 * R * R * K * math.log(1.0 - (r / R) ** 2)

    def f(r):
        return r * (-K / (1.0 - r * r / (R * R)))
    return _table_rows(e, f, 257, 0.0001, 1.49)


def cosine_virial(pos, box, angles, ids_base=1, K=1.5, theta0=3.1415926):
    """Scalar virial and xy virial of the Cosine angle potential over the triples
    (FixedTripleListInteractionTemplate::computeVirial / computeVirialTensor with
    Cosine::_computeForceRaw, src/interaction/Cosine.hpp:79-118): w = d12.F12 + d32.F32,
    d12 = p1 - p2 and d32 = p3 - p2 in the minimum image.  The angle term is not part of
    the rebuilt path; it is restated here only to close the pressure sum."""
    L = np.asarray(box)
    p1, p2, p3 = (pos[angles[:, k] - ids_base] for k in range(3))
    d12 = p1 - p2
    d12 -= L * np.round(d12 / L)
    d32 = p3 - p2
    d32 -= L * np.round(d32 / L)
    s12 = np.sum(d12 * d12, axis=1)
    s32 = np.sum(d32 * d32, axis=1)
    m12, m32 = np.sqrt(s12), np.sqrt(s32)
    c = np.clip(np.sum(d12 * d32, axis=1) / (m12 * m32), -1.0, 1.0)
    s = np.sqrt(1.0 - c * c)
    a11, a12, a22 = c / s12, -1.0 / (m12 * m32), c / s32
    pref = np.where(s > 1e-9, K * math.cos(theta0) - K * math.sin(theta0) * c / np.where(s > 1e-9, s, 1.0), 0.0)
    F12 = pref[:, None] * (a11[:, None] * d12 + a12[:, None] * d32)
    F32 = pref[:, None] * (a22[:, None] * d32 + a12[:, None] * d12)
    w = float(np.sum(d12 * F12) + np.sum(d32 * F32))
    wxy = float(np.sum(d12[:, 0] * F12[:, 1]) + np.sum(d32[:, 0] * F32[:, 1]))
    return w, wxy


def cosine_energy(pos, box, angles, ids_base=1, K=1.5, theta0=3.1415926):
    """Cosine angle energy K (1 + cos(theta - theta0)) summed over the triples (eangle column)."""
    L = np.asarray(box)
    p1, p2, p3 = (pos[angles[:, k] - ids_base] for k in range(3))
    d12 = p1 - p2
    d12 -= L * np.round(d12 / L)
    d32 = p3 - p2
    d32 -= L * np.round(d32 / L)
    c = np.clip(np.sum(d12 * d32, axis=1) / np.sqrt(np.sum(d12 * d12, axis=1) * np.sum(d32 * d32, axis=1)), -1, 1)
    return float(np.sum(K * (1.0 + np.cos(np.arccos(c) - theta0))))
