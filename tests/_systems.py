"""Shared set-up of identical systems on the GPU package and the oracle."""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE_2048 = os.path.join(HERE, "golden", "lennard_jones_fluid_10000_2048.xyz")


def read_fixture():
    """The reference's 2048-particle LJ fluid (testsuite/potential_array/lennard_jones_fluid_10000_2048.xyz)."""
    from espressopp_amd.tools import readxyz
    pid, types, x, y, z, vx, vy, vz, Lx, Ly, Lz = readxyz(FIXTURE_2048)
    return (np.asarray(pid, dtype=np.int32), np.stack([x, y, z], axis=1), np.stack([vx, vy, vz], axis=1),
            (Lx, Ly, Lz))


def jittered_lattice(n, rho, seed, jitter=0.05):
    """First n sites of the smallest simple-cubic lattice holding n, jittered U(-j, j) (sigma units)."""
    L = (n / rho) ** (1.0 / 3.0)
    a = int(round(n ** (1.0 / 3.0)))
    if a ** 3 < n:
        a += 1
    sp = L / a
    g = np.arange(a)
    ii, jj, kk = np.meshgrid(g, g, g, indexing="ij")
    sites = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)[:n]
    rng = np.random.default_rng(seed)
    pos = (sites + 0.5) * sp + rng.uniform(-jitter, jitter, size=(n, 3))
    return pos, (L, L, L)


def mb_velocities(n, T, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, math.sqrt(T), size=(n, 3))
    v -= v.mean(axis=0)
    v *= math.sqrt(T / (np.sum(v * v) / (3 * n)))
    return v


def random_walk_chains(nchains, nbeads, rho, seed, bond=0.97):
    """Random-walk chains of fixed bond length, folded into a cubic box of density rho."""
    n = nchains * nbeads
    L = (n / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 3))
    bonds = []
    for c in range(nchains):
        start = c * nbeads
        pos[start] = rng.uniform(0, L, 3)
        for b in range(1, nbeads):
            while True:
                u = rng.normal(size=3)
                u /= np.linalg.norm(u)
                cand = pos[start + b - 1] + bond * u
                if b < 2 or np.linalg.norm(cand - pos[start + b - 2]) > 1.02:
                    break
            pos[start + b] = cand
            bonds.append((start + b - 1, start + b))
    return pos, (L, L, L), np.asarray(bonds, dtype=np.int32)


def cells_for(box, rc, skin):
    return [max(int(b / (rc + skin)), 1) for b in box]


def gpu_system(box, cells, skin, precision="fp64", seed=12345):
    import espressopp_amd as espressopp
    system = espressopp.System(precision=precision)
    system.rng = espressopp.esutil.RNG(seed)
    system.bc = espressopp.bc.OrthorhombicBC(system.rng, box)
    system.skin = skin
    system.storage = espressopp.storage.DomainDecomposition(system, (1, 1, 1), cells, nocheck=True)
    return system


def force_rel_err(f, f_ref, floor_frac=0.0):
    """Per-particle |dF| / max(|F|, floor), floor = floor_frac * RMS |F|."""
    nref = np.linalg.norm(f_ref, axis=1)
    floor = floor_frac * math.sqrt(np.mean(nref ** 2))
    return np.linalg.norm(f - f_ref, axis=1) / np.maximum(nref, max(floor, 1e-300))


def pair_set(pairs):
    p = np.sort(np.asarray(pairs, dtype=np.int64).reshape(-1, 2), axis=1)
    return set(map(tuple, p.tolist()))


def melt_chains(nchains, nbeads, rho, seed, bond=0.97, min_dist=0.7, tries=500):
    """Random-walk chains of fixed bond length in a cubic box of density rho whose
    beads keep min_dist from every other bead (the overlap push-off of a
    Kremer-Grest start), found by rejection on a cell grid."""
    n = nchains * nbeads
    L = (n / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    nc = max(int(L / min_dist), 3)
    cs = L / nc
    grid = {}
    pos = np.zeros((n, 3))
    md2 = min_dist * min_dist

    def free(p, skip):
        c = np.floor(p / cs).astype(int)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                        if j == skip:
                            continue
                        d = pos[j] - p
                        d -= L * np.round(d / L)
                        if d @ d < md2:
                            return False
        return True

    def put(k, p):
        pos[k] = p
        c = np.floor(p / cs).astype(int) % nc
        grid.setdefault((c[0], c[1], c[2]), []).append(k)

    def take(k):
        c = np.floor(pos[k] / cs).astype(int) % nc
        grid[(c[0], c[1], c[2])].remove(k)

    bonds = []
    for ch in range(nchains):
        start = ch * nbeads
        for _attempt in range(1000):            # a trapped chain is removed and regrown
            while True:
                p = rng.uniform(0, L, 3)
                if free(p, -1):
                    break
            put(start, p)
            grown = 1
            for b in range(1, nbeads):
                for _ in range(tries):
                    u = rng.normal(size=3)
                    cand = (pos[start + b - 1] + bond * u / np.linalg.norm(u)) % L
                    if free(cand, start + b - 1):
                        break
                else:
                    break
                put(start + b, cand)
                grown += 1
            if grown == nbeads:
                break
            for k in range(start, start + grown):
                take(k)
        else:
            raise RuntimeError("chain growth trapped")
        bonds.extend((start + b - 1, start + b) for b in range(1, nbeads))
    # beads are stored folded; bond vectors use the minimum image
    return pos, (L, L, L), np.asarray(bonds, dtype=np.int32)
