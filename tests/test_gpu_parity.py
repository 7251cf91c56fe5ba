"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Bars: neighbour pair sets bit-exact; FP64 path forces within 1e-12 relative;
mixed path per-particle relative force error <= 1e-5, relative to
max(|F_i|, 0.1 * RMS|F|) (the floor covers particles whose net force nearly
cancels: an FP32 pair force carries ~6e-8 of its own magnitude, up to ~1e2 in a
dense LJ liquid), energies and virials within 1e-6 relative.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim

from _systems import (cells_for, force_rel_err, gpu_system, jittered_lattice, mb_velocities, melt_chains,
                      pair_set, random_walk_chains, read_fixture)

pytestmark = pytest.mark.gpu

import espressopp_amd as espressopp  # noqa: E402


def build_lj(pos, box, cells, rc, skin, precision, ids=None, types=None, vel=None, lj=((0, 0, 1.0, 1.0),),
             shift="auto"):
    n = len(pos)
    ids = np.arange(n, dtype=np.int32) if ids is None else np.asarray(ids, dtype=np.int32)
    system = gpu_system(box, cells, skin, precision)
    system.storage.addParticlesArray(ids, pos, types, vel)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=rc)
    ia = espressopp.interaction.VerletListLennardJones(vl)
    for t1, t2, eps, sig in lj:
        ia.setPotential(t1, t2, espressopp.interaction.LennardJones(eps, sig, rc, shift))
    system.addInteraction(ia)
    ref = OracleSim(box, cells, skin)
    ref.add_particles(ids, pos, types=types, vel=vel)
    ref.decompose()
    rvl = ref.vl_create(rc)
    ria = ref.lj_create(rvl)
    for t1, t2, eps, sig in lj:
        ref.lj_set(ria, t1, t2, eps, sig, rc, shift)
    return system, vl, ia, ref, rvl, ria


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pairs_fixture(gpu, precision):
    pid, pos, _, box = read_fixture()
    rc, skin = 2.5, 0.3
    cells = cells_for(box, rc, skin)
    system, vl, _, ref, rvl, _ = build_lj(pos, box, cells, rc, skin, precision, ids=pid)
    got = pair_set(vl.pairs_array())
    want = pair_set(ref.vl_pairs(rvl))
    assert vl.totalSize() == ref.vl_size(rvl) == len(want)
    assert got == want


@pytest.mark.parametrize("cut,count", [(0.0, 0), (1.0, 3 * 216), (math.sqrt(2.0), 9 * 216),
                                       (math.sqrt(3.0), 13 * 216)])
def test_pairs_lattice_counts(gpu, cut, count):
    # testsuite/unittest/testVerletList.py:38-109
    N, SIZE = 6, 6.0
    system = gpu_system((SIZE,) * 3, (3, 3, 3), 0.001)
    g = (np.arange(N) + 0.5) / N * SIZE
    pos = np.array([[x, y, z] for x in g for y in g for z in g])
    system.storage.addParticlesArray(np.arange(N ** 3), pos)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cut)
    assert vl.totalSize() == count


@pytest.mark.parametrize("seed,box,rc,skin", [(11, (9.0, 9.0, 9.0), 2.5, 0.3),
                                              (12, (8.7, 11.3, 14.9), 2.5, 0.4),
                                              (13, (6.1, 6.1, 9.3), 1.12, 0.3),
                                              (14, (20.0, 7.5, 33.0), 1.5, 0.2)])
def test_pairs_random_boxes(gpu, seed, box, rc, skin):
    rng = np.random.default_rng(seed)
    n = 3000
    pos = rng.uniform(0, 1, (n, 3)) * np.asarray(box)
    # particles on the box faces and one ulp inside them exercise folding and the ghost shift
    pos[:6] = 0.0
    pos[6:12] = np.nextafter(np.asarray(box), 0.0)
    pos[12, 0] = -1e-300
    cells = cells_for(box, rc, skin)
    system, vl, _, ref, rvl, _ = build_lj(pos, box, cells, rc, skin, "fp64")
    assert pair_set(vl.pairs_array()) == pair_set(ref.vl_pairs(rvl))


def test_pairs_borderline_distances(gpu):
    # pairs placed at (rc+skin) +- a few ulp across the periodic boundary: membership must
    # follow the reference's FP64 test in the reference's orientation
    box = (10.0, 10.0, 10.0)
    rc, skin = 2.5, 0.3
    cut = rc + skin
    pos = []
    for k in range(40):
        y = 0.2 + 0.24 * k
        x0 = 0.05 + 0.001 * k
        for m, dz in enumerate((-2, -1, 0, 1, 2)):
            z = 0.3 + 2.0 * m
            d = cut + dz * 4e-16
            pos.append([x0, y, z])
            pos.append([x0 - d + 10.0, y, z])
    pos = np.asarray(pos)
    system, vl, _, ref, rvl, _ = build_lj(pos, box, (3, 3, 3), rc, skin, "fp64")
    assert pair_set(vl.pairs_array()) == pair_set(ref.vl_pairs(rvl))


@pytest.mark.parametrize("L", [20.0, 40.0])
def test_pairs_borderline_distances_16bit(gpu, L):
    # Cubic grids wide enough for the list build's 16-bit position copies (delta = L / (2^16 m),
    # m >= 1 below; a box of 10 is too small and takes the FP32 prefilter): pairs at (rc+skin)
    # times 1 +- a few ulp (the reference's FP64 test decides), 1 +- 1e-4 (inside the 16-bit
    # window) and 1 +- 1e-3 (outside it), along axes and diagonals, inside the box and across
    # its periodic boundary, among random particles
    rc, skin = 2.5, 0.3
    cut = rc + skin
    c = int(L // cut)
    cs = L / c
    bx, byz = cs * 1.0012 + cs / 8.0, 2.0 * cs * 1.0001
    dmin = max(max(bx, byz) / 32000.0, math.sqrt(bx * bx + 2.0 * byz * byz) / 46000.0)
    assert math.floor(L / (65536.0 * dmin)) >= 1
    rng = np.random.default_rng(int(L))
    dirs = [np.asarray(v, dtype=float) / np.linalg.norm(v)
            for v in ((1, 0, 0), (0, 1, 0), (0, 0, -1), (1, 1, 0), (1, -1, 1), (0.3, 0.7, -0.2))]
    pos = [rng.uniform(0.0, L, 3) for _ in range(1500)]
    for d in dirs:
        for rel in (-1e-3, -1e-4, -4e-16, -2e-16, 0.0, 2e-16, 4e-16, 1e-4, 1e-3):
            for edge in (False, True):
                a = rng.uniform(0.0, L, 3)
                if edge:                       # the partner lies across the periodic boundary
                    a = np.where(d > 0, L - 0.2 * rng.uniform(0.0, 1.0, 3), np.where(d < 0, 0.2 * rng.uniform(0.0, 1.0, 3), a))
                b = a + d * (cut * (1.0 + rel))
                pos.append(a)
                pos.append(b - np.floor(b / L) * L)
    pos = np.asarray(pos)
    system, vl, _, ref, rvl, _ = build_lj(pos, (L, L, L), (c, c, c), rc, skin, "mixed")
    assert pair_set(vl.pairs_array()) == pair_set(ref.vl_pairs(rvl))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("cells", [(2, 2, 2), (3, 3, 3)])
def test_forces_borderline_cutoff(gpu, precision, cells):
    # pairs at rc +- a few ulp, inside the box and across its periodic boundary: each falls on
    # the reference's side of the cut test (|d|^2 > rc^2, d rounded as the reference forms it),
    # untiled (2 cells) and tiled (3 cells) kernels alike
    box = (10.0, 10.0, 10.0)
    rc, skin = 2.5, 0.3
    pos = []
    for k in range(30):
        y = 0.2 + 0.32 * k
        x0 = 0.05 + 0.001 * k
        for m, dz in enumerate((-2, -1, 0, 1, 2)):
            z = 0.3 + 2.0 * m
            d = rc + dz * 4e-16
            pos.append([x0, y, z])
            pos.append([x0 - d + 10.0 if k % 2 == 0 else x0 + d, y, z])
    pos = np.asarray(pos)
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells, rc, skin, precision, shift=0.0)
    espressopp.integrator.VelocityVerlet(system).run(0)
    ref.calc_forces()
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert force_rel_err(system.storage.getParticleArrays()["f"], ref.get("f"), 0.1).max() <= tol
    assert ia.computeEnergy() == pytest.approx(ref.energy(ria), rel=1e-11 if precision == "fp64" else 1e-6)


def _half_cell_border_positions():
    # pairs at (rc+skin) +- a few ulp across the x boundary whose partners sit in the same
    # full cell but in neighbouring half cells in y: the reference's orientation of d (and
    # so the rounding of the list test) depends on halfCellInt for them
    cut = 2.8
    pos = []
    for k in range(40):
        z = 0.2 + 0.24 * k
        x0 = 0.05 + 0.001 * k
        for m, dz in enumerate(range(-6, 7)):
            yi = 1.40 + 0.002 * m
            yj = yi + 0.05 + 0.001 * m
            d = cut + dz * 2e-16
            dx = math.sqrt(d * d - (yj - yi) ** 2)
            pos.append([x0, yi, z])
            pos.append([x0 - dx + 10.0, yj, z])
    return np.asarray(pos)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("hci,cells", [(1, (3, 3, 3)), (2, (7, 7, 7)), (3, (10, 10, 10))])
def test_pairs_half_cells(gpu, precision, hci, cells):
    # DomainDecomposition(..., halfCellInt): the GPU bins on cells of at least rc+skin and
    # takes each pair's orientation from the caller's grid; pair sets equal the oracle's on
    # the same half-cell grid, including borderline pairs whose orientation changes with it
    box, rc, skin = (10.0, 10.0, 10.0), 2.5, 0.3
    for pos in (_half_cell_border_positions(), np.random.default_rng(70 + hci).uniform(0, 10.0, (3000, 3))):
        system = espressopp.System(precision=precision)
        system.rng = espressopp.esutil.RNG(5)
        system.bc = espressopp.bc.OrthorhombicBC(system.rng, box)
        system.skin = skin
        system.storage = espressopp.storage.DomainDecomposition(system, (1, 1, 1), cells, hci)
        system.storage.addParticlesArray(np.arange(len(pos), dtype=np.int32), pos)
        system.storage.decompose()
        vl = espressopp.VerletList(system, cutoff=rc)
        ref = OracleSim(box, cells, skin, half_cell_int=hci)
        ref.add_particles(np.arange(len(pos)), pos)
        ref.decompose()
        assert pair_set(vl.pairs_array()) == pair_set(ref.vl_pairs(ref.vl_create(rc)))


def test_forces_fp64_fixture(gpu):
    pid, pos, _, box = read_fixture()
    rc, skin = 2.5, 0.3
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells_for(box, rc, skin), rc, skin, "fp64", ids=pid)
    system.integrator = espressopp.integrator.VelocityVerlet(system)
    system.integrator.run(0)
    ref.calc_forces()
    f = system.storage.getParticleArrays()["f"]
    assert force_rel_err(f, ref.get("f"), 1e-3).max() < 1e-12
    assert ia.computeEnergy() == pytest.approx(ref.energy(ria), rel=1e-12)
    w, t = ref.virial(ria)
    assert ia.computeVirial() == pytest.approx(w, rel=1e-12)
    assert np.allclose(ia.computeVirialTensor(), t, rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("n,seed", [(2048, 0), (32000, 1)])
def test_forces_mixed(gpu, n, seed):
    pos, box = jittered_lattice(n, 0.8442, seed)
    rc, skin = 2.5, 0.3
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells_for(box, rc, skin), rc, skin, "mixed")
    espressopp.integrator.VelocityVerlet(system).run(0)
    ref.calc_forces()
    f = system.storage.getParticleArrays()["f"]
    err = force_rel_err(f, ref.get("f"), 0.1)
    print("mixed n=%d: max %.3e  p99.9 %.3e  median %.3e" % (n, err.max(), np.quantile(err, 0.999), np.median(err)))
    assert err.max() <= 1e-5
    assert ia.computeEnergy() == pytest.approx(ref.energy(ria), rel=1e-6)
    assert ia.computeVirial() == pytest.approx(ref.virial(ria)[0], rel=1e-6)


def sparse_positions(n, box, min_dist, seed):
    """n points in the box, each at least min_dist from the others (periodic), by rejection."""
    rng = np.random.default_rng(seed)
    L = np.asarray(box)
    pos = np.zeros((0, 3))
    while len(pos) < n:
        p = rng.uniform(0, 1, 3) * L
        d = pos - p
        d -= L * np.round(d / L)
        if len(pos) == 0 or np.min(np.einsum("ij,ij->i", d, d)) >= min_dist * min_dist:
            pos = np.vstack([pos, p])
    return pos


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_forces_sparse_box(gpu, precision):
    # A dilute, elongated box: a cell holds about one particle, so a tile's staged
    # region is made of many short runs, several of them starting within one
    # wavefront's 64 entries.
    box = (20.0, 7.5, 33.0)
    pos = sparse_positions(1200, box, 0.9, 21)
    rc, skin = 1.5, 0.2
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells_for(box, rc, skin), rc, skin, precision)
    assert pair_set(vl.pairs_array()) == pair_set(ref.vl_pairs(rvl))
    espressopp.integrator.VelocityVerlet(system).run(0)
    ref.calc_forces()
    f = system.storage.getParticleArrays()["f"]
    if precision == "fp64":
        assert force_rel_err(f, ref.get("f"), 1e-3).max() < 1e-12
        assert ia.computeEnergy() == pytest.approx(ref.energy(ria), rel=1e-12)
    else:
        assert force_rel_err(f, ref.get("f"), 0.1).max() <= 1e-5
        assert ia.computeEnergy() == pytest.approx(ref.energy(ria), rel=1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multitype_forces(gpu, precision):
    pid, pos, _, box = read_fixture()
    rc, skin = 2.5, 0.4
    types = (pid % 3).astype(np.int32)
    lj = ((0, 0, 0.1, 1.0), (0, 1, (0.1 * 0.3) ** 0.5, 0.8), (1, 1, 0.9, 0.6), (1, 2, 1.0, 1.1), (2, 2, 0.5, 1.0))
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells_for(box, rc, skin), rc, skin, precision,
                                             ids=pid, types=types, lj=lj, shift=0.0)
    espressopp.integrator.VelocityVerlet(system).run(0)
    ref.calc_forces()
    f = system.storage.getParticleArrays()["f"]
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert force_rel_err(f, ref.get("f"), 0.1).max() <= tol
    assert ia.computeEnergy() == pytest.approx(ref.energy(ria), rel=tol if precision == "fp64" else 1e-6)


def test_golden_nve_temperature_fp64(gpu):
    # testsuite/potential_array/potential_array_test.py:37-163 on the GPU FP64 path
    pid, pos, _, box = read_fixture()
    rc, skin = 2.5, 0.4
    cells = espressopp.tools.decomp.cellGrid(box, (1, 1, 1), rc, skin)
    system = gpu_system(box, cells, skin, "fp64")
    system.storage.addParticles([[int(p), int(p) % 2, espressopp.Real3D(*x)] for p, x in zip(pid, pos)],
                                "id", "type", "pos")
    system.storage.decompose()
    integrator = espressopp.integrator.VelocityVerlet(system)
    integrator.dt = 0.005
    vl = espressopp.VerletList(system, rc)
    ia = espressopp.interaction.VerletListLennardJones(vl)
    LJ = espressopp.interaction.LennardJones
    ia.setPotential(type1=0, type2=0, potential=LJ(epsilon=0.1, sigma=1.0, cutoff=rc, shift=0.0))
    ia.setPotential(type1=0, type2=1, potential=LJ(epsilon=(0.1 * 0.3) ** 0.5, sigma=0.8, cutoff=rc, shift=0.0))
    ia.setPotential(type1=1, type2=1, potential=LJ(epsilon=0.9, sigma=0.6, cutoff=rc, shift=0.0))
    system.addInteraction(ia)
    integrator.run(100)
    n = len(pid)
    system.storage.addParticles([[n + k, 6, espressopp.Real3D(*pos[k])] for k in range(10)], "id", "type", "pos")
    system.storage.decompose()
    integrator.run(100)
    T = espressopp.analysis.Temperature(system).compute()
    assert abs(T - 0.160074388303) < 1.0e-8


def test_nve_trajectory_matches_oracle(gpu):
    pos, box = jittered_lattice(4000, 0.8442, seed=4)
    vel = mb_velocities(4000, 1.0, seed=5)
    rc, skin = 2.5, 0.3
    cells = cells_for(box, rc, skin)
    out = {}
    for precision in ("fp64", "mixed"):
        system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells, rc, skin, precision, vel=vel)
        integ = espressopp.integrator.VelocityVerlet(system)
        integ.dt = 0.005
        integ.run(50)
        out[precision] = (system.storage.getParticleArrays(), integ.getNumResorts(), vl.builds)
    ref.run(50, dt=0.005)
    rp = ref.get("pos") + ref.get("image") * np.asarray(box)
    for precision, tol in (("fp64", 1e-9), ("mixed", 1e-4)):
        a, resorts, builds = out[precision]
        gp = a["pos"] + a["image"] * np.asarray(box)
        assert np.abs(gp - rp).max() < tol, precision
    assert out["fp64"][1] == ref.resorts


def test_langevin_matches_oracle_and_exclusion(gpu):
    pos, box = jittered_lattice(1000, 0.8442, seed=6)
    rc, skin = 2.5, 0.3
    cells = cells_for(box, rc, skin)
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells, rc, skin, "fp64")
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    lang = espressopp.integrator.LangevinThermostat(system)
    lang.gamma = 1.0
    lang.temperature = 1.0
    lang.addExclusions([3])
    integ.addExtension(lang)
    integ.run(20)
    ref.langevin(1.0, 1.0, [3])
    ref.run(20, dt=0.005)
    a = system.storage.getParticleArrays()
    assert np.abs(a["pos"] - ref.get("pos")).max() < 1e-9
    assert np.abs(a["v"] - ref.get("v")).max() < 1e-8


def test_langevin_excluded_particle_does_not_move(gpu):
    # testsuite/LangevinThermostat/testLangevinThermostat.py:84-92
    box = (10.0, 10.0, 10.0)
    system = gpu_system(box, (5, 5, 5), 0.3)
    system.storage.addParticles([(1, 1, espressopp.Real3D(5.5, 5.0, 5.0), 1.0),
                                 (2, 1, espressopp.Real3D(6.5, 5.0, 5.0), 1.0),
                                 (3, 1, espressopp.Real3D(7.5, 5.0, 5.0), 1.0)], "id", "type", "pos", "mass")
    system.storage.decompose()
    espressopp.VerletList(system, cutoff=1.5)
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.01
    lang = espressopp.integrator.LangevinThermostat(system)
    lang.gamma = 1.0
    lang.temperature = 1.0
    lang.addExclusions([1])
    integ.addExtension(lang)
    before = [system.storage.getParticle(i).pos[j] for i in range(1, 4) for j in range(3)]
    integ.run(10)
    after = [system.storage.getParticle(i).pos[j] for i in range(1, 4) for j in range(3)]
    assert before[:3] == after[:3]
    assert all(b != a for b, a in zip(before[3:], after[3:]))


@pytest.mark.parametrize("precision,r0", [("fp64", 0.0), ("mixed", 0.0), ("fp64", 0.2), ("mixed", 0.2)])
def test_fene_wca_melt(gpu, precision, r0):
    # r0 > 0 takes FENE's other force branch (src/interaction/FENE.hpp:100-104) and arrives
    # through FixedPairListFENE.setPotential (src/interaction/FENE.py:114-116)
    pos, box, bonds = random_walk_chains(40, 50, 0.85, seed=7)
    rc, skin = 2 ** (1 / 6), 0.3
    cells = cells_for(box, rc, skin)
    n = len(pos)
    system = gpu_system(box, cells, skin, precision)
    system.storage.addParticlesArray(np.arange(n), pos)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=rc)
    wca = espressopp.interaction.VerletListLennardJones(vl)
    wca.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, rc, "auto"))
    system.addInteraction(wca)
    fpl = espressopp.FixedPairList(system.storage)
    fpl.addBonds(bonds.tolist())
    fene = espressopp.interaction.FixedPairListFENE(system, fpl, espressopp.interaction.FENE(30.0, 0.0, 1.5))
    system.addInteraction(fene)
    if r0 > 0.0:
        fene.setPotential(espressopp.interaction.FENE(30.0, r0, 1.5))
    espressopp.integrator.VelocityVerlet(system).run(0)
    ref = OracleSim(box, cells, skin)
    ref.add_particles(np.arange(n), pos)
    ref.decompose()
    ria = ref.lj_create(ref.vl_create(rc))
    ref.lj_set(ria, 0, 0, 1.0, 1.0, rc, "auto")
    rfpl = ref.fpl_create()
    ref.fpl_add(rfpl, bonds)
    rfene = ref.fene_create(rfpl, 30.0, r0, 1.5)
    ref.calc_forces()
    f = system.storage.getParticleArrays()["f"]
    tol = 1e-11 if precision == "fp64" else 1e-5
    assert force_rel_err(f, ref.get("f"), 0.1).max() <= tol
    etol = 1e-11 if precision == "fp64" else 1e-6
    assert fene.computeEnergy() == pytest.approx(ref.energy(rfene), rel=etol)
    assert wca.computeEnergy() == pytest.approx(ref.energy(ria), rel=etol)
    assert fene.computeVirial() == pytest.approx(ref.virial(rfene)[0], rel=etol)


def test_remove_interaction_and_bond_list_swap(gpu):
    # System.removeInteraction (src/System.cpp:108-135) and FixedPairListFENE.setFixedPairList
    # (src/interaction/FENE.py:122-124): forces follow the remaining interactions and the new bonds
    pos, box, bonds = random_walk_chains(10, 30, 0.85, seed=17)
    rc, skin = 2 ** (1 / 6), 0.3
    cells = cells_for(box, rc, skin)
    n = len(pos)
    system = gpu_system(box, cells, skin, "fp64")
    system.storage.addParticlesArray(np.arange(n), pos)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=rc)
    wca = espressopp.interaction.VerletListLennardJones(vl)
    wca.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, rc, "auto"))
    fpl = espressopp.FixedPairList(system.storage)
    fpl.addBonds(bonds.tolist())
    half = espressopp.FixedPairList(system.storage)
    half.addBonds(bonds[::2].tolist())
    fene = espressopp.interaction.FixedPairListFENE(system, fpl, espressopp.interaction.FENE(30.0, 0.0, 1.5))
    system.addInteraction(wca)
    system.addInteraction(fene)
    system.removeInteraction(0)
    assert system.getNumberOfInteractions() == 1 and system.getInteraction(0) is fene
    fene.setFixedPairList(half)
    espressopp.integrator.VelocityVerlet(system).run(0)
    ref = OracleSim(box, cells, skin)
    ref.add_particles(np.arange(n), pos)
    ref.decompose()
    rfpl = ref.fpl_create()
    ref.fpl_add(rfpl, bonds[::2])
    rfene = ref.fene_create(rfpl, 30.0, 0.0, 1.5)
    ref.calc_forces()
    f = system.storage.getParticleArrays()["f"]
    assert force_rel_err(f, ref.get("f"), 0.1).max() <= 1e-12
    assert fene.computeEnergy() == pytest.approx(ref.energy(rfene), rel=1e-12)
    with pytest.raises(RuntimeError):
        system.removeInteraction(3)


def test_particle_access_by_id(gpu):
    # getParticle(pid) reads one particle through the device's id -> slot map; setters write one
    pos, box = jittered_lattice(3000, 0.8, seed=3)
    system = gpu_system(box, cells_for(box, 2.5, 0.3), 0.3, "fp64")
    ids = np.arange(10, 3010, dtype=np.int32)[::-1].copy()
    system.storage.addParticlesArray(ids, pos)
    system.storage.decompose()
    arrays = system.storage.getParticleArrays()
    for k in (0, 17, 2999):
        pid = int(arrays["id"][k])
        p = system.storage.getParticle(pid)
        assert list(p.pos) == list(arrays["pos"][k])
        assert p.mass == 1.0 and p.type == 0
    p = system.storage.getParticle(1234)
    p.v = espressopp.Real3D(1.0, -2.0, 0.5)
    p.pos = espressopp.Real3D(1.5, 2.5, 3.5)
    assert list(p.v) == [1.0, -2.0, 0.5] and list(p.pos) == [1.5, 2.5, 3.5]
    with pytest.raises(RuntimeError):
        system.storage.getParticle(5).pos
    with pytest.raises(NotImplementedError):
        system.storage.addParticles([(9000, espressopp.Real3D(1.0, 1.0, 1.0), 0.5)], "id", "pos", "q")


def _lattice_box(box, rho, seed, jitter=0.05):
    # jittered simple-cubic lattice filling an orthorhombic box at about density rho
    box = np.asarray(box, dtype=float)
    a = (1.0 / rho) ** (1.0 / 3.0)
    m = np.maximum((box / a).astype(int), 1)
    g = [(np.arange(k) + 0.5) * (L / k) for k, L in zip(m, box)]
    pos = np.stack(np.meshgrid(*g, indexing="ij"), axis=-1).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    return pos + rng.uniform(-jitter, jitter, pos.shape)


@pytest.mark.parametrize("case", ["two_types_orthorhombic", "fene_then_wca", "wca_then_fene", "short_rows"])
def test_step_variants_match_oracle(gpu, case):
    # Multi-step trajectories through the force kernel that also finishes the step: several
    # types in an orthorhombic box, an LJ interaction that accumulates onto FENE forces, and
    # FENE bonds accumulating onto WCA forces (the order of the reference's melt benchmark)
    if case == "two_types_orthorhombic":
        box = (14.2, 16.5, 19.3)
        pos = _lattice_box(box, 0.7, seed=41)
        n = len(pos)
        types = (np.arange(n) % 3 == 0).astype(np.int32)
        rc, skin = 2.5, 0.3
        lj = ((0, 0, 1.0, 1.0), (0, 1, 1.2, 1.05), (1, 1, 0.8, 0.95))
        bonds = None
        dt = 0.005
    elif case == "short_rows":
        # rows of three cells: the largest tiles span more than three rows, so every
        # rebuild's tentative tile plan fails first and the smaller tiles are used
        box = (8.5, 30.0, 30.0)
        pos = _lattice_box(box, 0.8442, seed=44)
        n = len(pos)
        types = None
        rc, skin = 2.5, 0.3
        lj = ((0, 0, 1.0, 1.0),)
        bonds = None
        dt = 0.005
    else:
        # a looser start than a dense melt, so 40 plain steps stay well resolved
        pos, box, bonds = melt_chains(30, 40, 0.5, seed=42, min_dist=0.95)
        n = len(pos)
        types = None
        rc, skin = 2 ** (1 / 6), 0.3
        lj = ((0, 0, 1.0, 1.0),)
        dt = 0.003
    vel = mb_velocities(n, 1.0, seed=43)
    cells = cells_for(box, rc, skin)
    steps = 40
    fene_last = case == "wca_then_fene"
    out = {}
    for precision in ("fp64", "mixed"):
        system = gpu_system(box, cells, skin, precision)
        system.storage.addParticlesArray(np.arange(n), pos, types, vel)
        system.storage.decompose()
        fene = None
        if bonds is not None:
            fpl = espressopp.FixedPairList(system.storage)
            fpl.addBonds(bonds.tolist())
            fene = espressopp.interaction.FixedPairListFENE(system, fpl, espressopp.interaction.FENE(30.0, 0.0, 1.5))
            if not fene_last:
                system.addInteraction(fene)
        vl = espressopp.VerletList(system, cutoff=rc)
        ia = espressopp.interaction.VerletListLennardJones(vl)
        for t1, t2, eps, sig in lj:
            ia.setPotential(t1, t2, espressopp.interaction.LennardJones(eps, sig, rc, "auto"))
        system.addInteraction(ia)
        if fene is not None and fene_last:
            system.addInteraction(fene)
        integ = espressopp.integrator.VelocityVerlet(system)
        integ.dt = dt
        integ.run(steps)
        out[precision] = (system.storage.getParticleArrays(), integ.getNumResorts())
    ref = OracleSim(box, cells, skin)
    ref.add_particles(np.arange(n), pos, types=types, vel=vel)
    ref.decompose()
    if bonds is not None and not fene_last:
        rfpl = ref.fpl_create()
        ref.fpl_add(rfpl, bonds)
        ref.fene_create(rfpl, 30.0, 0.0, 1.5)
    ria = ref.lj_create(ref.vl_create(rc))
    for t1, t2, eps, sig in lj:
        ref.lj_set(ria, t1, t2, eps, sig, rc, "auto")
    if bonds is not None and fene_last:
        rfpl = ref.fpl_create()
        ref.fpl_add(rfpl, bonds)
        ref.fene_create(rfpl, 30.0, 0.0, 1.5)
    ref.run(steps, dt=dt)
    rp = ref.get("pos") + ref.get("image") * np.asarray(box)
    rv = ref.get("v")
    for precision, tol in (("fp64", 1e-9), ("mixed", 1e-4)):
        a, resorts = out[precision]
        gp = a["pos"] + a["image"] * np.asarray(box)
        assert np.abs(gp - rp).max() < tol, precision
        assert np.abs(a["v"] - rv).max() < 100 * tol, precision
        # the forces a caller reads after the run are those of the last step
        assert force_rel_err(a["f"], ref.get("f"), 0.1).max() <= (1e-9 if precision == "fp64" else 1e-4)
    assert out["fp64"][1] == ref.resorts


def test_pressure_and_temperature(gpu):
    pid, pos, vel, box = read_fixture()
    rc, skin = 2.5, 0.3
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells_for(box, rc, skin), rc, skin, "fp64", ids=pid, vel=vel)
    assert espressopp.analysis.Temperature(system).compute() == pytest.approx(ref.temperature(), rel=1e-12)
    assert espressopp.analysis.Pressure(system).compute() == pytest.approx(ref.pressure(), rel=1e-12)


def test_empty_system_and_errors(gpu):
    system = gpu_system((10.0, 10.0, 10.0), (3, 3, 3), 0.3)
    system.storage.decompose()
    vl = espressopp.VerletList(system, 2.5)
    assert vl.totalSize() == 0
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.run(5)
    with pytest.raises(ValueError):
        integ.dt = 0.0
    system.storage.addParticlesArray([0, 1], [[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]])
    system.storage.decompose()
    system.storage.setParticleArrays("pos", [1], [[float("nan"), 2.0, 2.0]])
    with pytest.raises(RuntimeError):
        system.storage.decompose()
    with pytest.raises(ValueError):
        gpu_system((10.0, 10.0, 10.0), (5, 5, 5), 0.3)
        espressopp.VerletList(gpu_system((10.0, 10.0, 10.0), (5, 5, 5), 0.3), 2.5)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lost_particle_is_reported(gpu, precision):
    # an infinite coordinate, and a particle sent off by a runaway velocity, end in the
    # reference's "moved to outer space" error at the next binning instead of a long
    # binning of one overfull cell or a silent run on NaNs
    pos, box = jittered_lattice(4000, 0.8442, seed=71)
    cells = cells_for(box, 2.5, 0.3)
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells, 2.5, 0.3, precision)
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    integ.run(5)
    system.storage.setParticleArrays("v", [17], [[1e200, 0.0, 0.0]])
    with pytest.raises(RuntimeError, match="outer space"):
        integ.run(5)
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells, 2.5, 0.3, precision)
    system.storage.setParticleArrays("pos", [5], [[float("inf"), 1.0, 1.0]])
    with pytest.raises(RuntimeError, match="outer space"):
        system.storage.decompose()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_free_flight_gpu(gpu, precision):
    # testsuite/integrator/PTestVelocityVerlet.cpp:222-256: free particles cross the box and return
    pos, box = jittered_lattice(125, 0.5, seed=3, jitter=0.0)
    L = box[0]
    system = gpu_system(box, (2, 2, 2), 0.3, precision)
    system.storage.addParticlesArray(np.arange(125), pos, vel=np.tile([1.0, -0.5, 0.25], (125, 1)))
    system.storage.decompose()
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = L / 400.0
    integ.run(1600)                          # x moves 4 L, y -2 L, z 1 L
    a = system.storage.getParticleArrays()
    p = a["pos"] + a["image"] * np.asarray(box)
    assert np.allclose(p, pos + np.array([4 * L, -2 * L, L]), atol=1e-6)
    assert np.all((a["pos"] >= -1e-9) & (a["pos"] <= L + 1e-9))


def test_decompose_folds_and_conserves(gpu):
    # testsuite/storage/PTestDomainDecomposition.cpp:117-238 pattern: every particle is kept,
    # folded into the box with its image count, and fields survive reordering into cells
    rng = np.random.default_rng(21)
    box = (9.0, 11.0, 13.0)
    n = 5000
    ids = rng.permutation(np.arange(3, 3 * n + 3, 3)).astype(np.int32)   # sparse, unordered ids
    pos = rng.uniform(-2.5, 3.5, (n, 3)) * np.asarray(box)              # up to a few boxes away
    vel = rng.normal(size=(n, 3))
    mass = rng.uniform(0.5, 2.0, n)
    types = rng.integers(0, 4, n).astype(np.int32)
    system = gpu_system(box, (3, 3, 4), 0.3)
    system.storage.addParticlesArray(ids, pos, types, vel, mass)
    system.storage.decompose()
    ref = OracleSim(box, (3, 3, 4), 0.3)
    ref.add_particles(ids, pos, types=types, vel=vel, mass=mass)
    ref.decompose()
    a = system.storage.getParticleArrays()
    order = np.argsort(ids)
    assert np.array_equal(a["id"], ids[order])
    assert np.array_equal(a["id"], ref.ids())
    assert np.array_equal(a["pos"], ref.get("pos"))           # the fold is FP64 and bit-exact
    assert np.array_equal(a["image"], ref.get("image"))
    assert np.array_equal(a["v"], vel[order])
    assert np.array_equal(a["mass"], mass[order])
    assert np.array_equal(a["type"], types[order])
    assert np.all((a["pos"] >= -1e-9) & (a["pos"] <= np.asarray(box) + 1e-9))
    # set/get by id after decompose
    pick = ids[order][::7]
    newv = rng.normal(size=(len(pick), 3))
    system.storage.setParticleArrays("v", pick, newv)
    assert np.array_equal(system.storage.getParticleArrays()["v"][::7], newv)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pairs_and_forces_262k(gpu, precision):
    # 64^3 particles at the benchmark density: the full pair set and the forces against the oracle
    n = 64 ** 3
    pos, box = jittered_lattice(n, 0.8442, seed=22)
    rc, skin = 2.5, 0.3
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells_for(box, rc, skin), rc, skin, precision)
    got = np.sort(vl.pairs_array().reshape(-1, 2).astype(np.int64), axis=1)
    want = np.sort(ref.vl_pairs(rvl).reshape(-1, 2).astype(np.int64), axis=1)
    assert len(got) == len(want) == vl.totalSize()
    kg, kw = np.sort(got[:, 0] * n + got[:, 1]), np.sort(want[:, 0] * n + want[:, 1])
    assert np.array_equal(kg, kw)
    espressopp.integrator.VelocityVerlet(system).run(0)
    ref.calc_forces()
    err = force_rel_err(system.storage.getParticleArrays()["f"], ref.get("f"), 0.1)
    assert err.max() <= (1e-12 if precision == "fp64" else 1e-5)
    assert ia.computeEnergy() == pytest.approx(ref.energy(ria), rel=1e-12 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_verlet_list_built_after_a_run(gpu, precision):
    # A list built between runs sees cells whose x order dates from the last binning while
    # the particles have moved since (no resort): its pair set must still be every pair
    # within cutoff + skin at the current positions (periodic k-d tree over the folded
    # positions as the check)
    from scipy.spatial import cKDTree
    n = 6400
    pos, box = jittered_lattice(n, 0.8, seed=31)
    rc, skin = 1.5, 1.0
    cells = cells_for(box, rc, skin)
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells, rc, skin, precision,
                                             vel=mb_velocities(n, 1.5, seed=32))
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    integ.run(10)
    assert integ.getNumResorts() == 0
    for cut in (1.2, 1.5):
        vl2 = espressopp.VerletList(system, cutoff=cut)
        a = system.storage.getParticleArrays()
        x = np.mod(a["pos"], np.asarray(box))
        x[x >= np.asarray(box)] = 0.0
        tree = cKDTree(x, boxsize=np.asarray(box))
        want = tree.query_pairs(cut + skin, output_type="ndarray")
        ids = a["id"].astype(np.int64)
        want = np.sort(ids[want], axis=1)
        got = np.sort(vl2.pairs_array().reshape(-1, 2).astype(np.int64), axis=1)
        assert len(got) == len(want)
        assert np.array_equal(np.sort(got[:, 0] * n + got[:, 1]), np.sort(want[:, 0] * n + want[:, 1]))


def _write_table(path, r, e, f, junk=True):
    with open(path, "w") as fh:
        if junk:
            fh.write("# r energy force\n\n")
        for a, b, c in zip(r, e, f):
            fh.write("%.17g %.17g %.17g\n" % (a, b, c))


def _lj_rows(N, lo, hi, eps, sig, rc):
    from oracle import oracle
    r = np.linspace(lo, hi, N)
    e = np.array([oracle.lj_energy(eps, sig, rc, 0.0, x) for x in r])
    f = np.array([oracle.lj_force(eps, sig, float("inf"), [x, 0.0, 0.0])[0] for x in r])
    return r, e, f


@pytest.mark.parametrize("precision,itype,cells", [("fp64", 1, None), ("fp64", 2, None), ("fp64", 3, None),
                                                   ("mixed", 3, None), ("mixed", 1, None), ("fp64", 3, (2, 2, 2)),
                                                   ("mixed", 2, (2, 2, 2))])
def test_verlet_list_tabulated(gpu, tmp_path, precision, itype, cells):
    # VerletListTabulated against the oracle's Tabulated on the same table files, two types
    pid, pos, _, box = read_fixture()
    rc, skin = 2.5, 0.3
    cells = cells or cells_for(box, rc, skin)
    types = (pid % 2).astype(np.int32)
    tabs = {}
    for (t1, t2, eps, sig) in ((0, 0, 1.0, 1.0), (0, 1, 0.8, 0.9), (1, 1, 1.2, 1.1)):
        r, e, f = _lj_rows(513, 0.6, rc, eps, sig, rc)
        fn = str(tmp_path / ("t%d%d.tab" % (t1, t2)))
        _write_table(fn, r, e, f)
        tabs[(t1, t2)] = (fn, r, e, f)
    system = gpu_system(box, cells, skin, precision)
    system.storage.addParticlesArray(pid, pos, types)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=rc)
    ia = espressopp.interaction.VerletListTabulated(vl)
    for (t1, t2), (fn, *_r) in tabs.items():
        ia.setPotential(t1, t2, espressopp.interaction.Tabulated(itype, fn, rc))
    system.addInteraction(ia)
    espressopp.integrator.VelocityVerlet(system).run(0)
    ref = OracleSim(box, cells, skin)
    ref.add_particles(pid, pos, types=types)
    ref.decompose()
    ria = ref.tab_create(ref.vl_create(rc))
    for (t1, t2), (fn, r, e, f) in tabs.items():
        ref.tab_set(ria, t1, t2, itype, r, e, f, cutoff=rc)
    ref.calc_forces()
    f = system.storage.getParticleArrays()["f"]
    # mixed: distances from the 32-bit fixed-point mirror (quantisation ~L/2^32), the LJ bars
    tol, etol = (1e-12, 1e-11) if precision == "fp64" else (1e-5, 1e-6)
    assert force_rel_err(f, ref.get("f"), 0.1).max() <= tol
    assert ia.computeEnergy() == pytest.approx(ref.energy(ria), rel=etol)
    w, t = ref.virial(ria)
    assert ia.computeVirial() == pytest.approx(w, rel=10 * etol)
    if precision == "fp64":
        assert np.allclose(ia.computeVirialTensor(), t, rtol=1e-9, atol=1e-8)


@pytest.mark.parametrize("itype", [1, 3])
def test_fixed_pair_list_tabulated(gpu, tmp_path, itype):
    # FixedPairListTabulated (a FENE table) plus a WCA table on a chain melt, against the oracle
    from oracle import oracle
    pos, box, bonds = random_walk_chains(20, 40, 0.85, seed=31)
    n = len(pos)
    rc, skin = 2 ** (1 / 6), 0.3
    cells = cells_for(box, rc, skin)
    rb = np.linspace(0.0001, 1.49, 1491)
    eb = np.array([oracle.fene_energy(30.0, 0.0, 1.5, x) for x in rb])
    fb = np.array([oracle.fene_force(30.0, 0.0, 1.5, [x, 0.0, 0.0])[0] for x in rb])
    rw, ew, fw = _lj_rows(257, 0.5, rc, 1.0, 1.0, rc)
    ew = ew + 1.0                                     # WCA: LJ shifted by +eps at 2^(1/6)
    fnb, fnw = str(tmp_path / "fene.tab"), str(tmp_path / "wca.tab")
    _write_table(fnb, rb, eb, fb)
    _write_table(fnw, rw, ew, fw, junk=False)
    out = {}
    for precision in ("fp64", "mixed"):
        system = gpu_system(box, cells, skin, precision)
        system.storage.addParticlesArray(np.arange(n), pos)
        system.storage.decompose()
        vl = espressopp.VerletList(system, cutoff=rc)
        wca = espressopp.interaction.VerletListTabulated(vl)
        wca.setPotential(0, 0, espressopp.interaction.Tabulated(itype, fnw, rc))
        system.addInteraction(wca)
        fpl = espressopp.FixedPairList(system.storage)
        fpl.addBonds(bonds.tolist())
        bond = espressopp.interaction.FixedPairListTabulated(system, fpl, espressopp.interaction.Tabulated(itype, fnb))
        system.addInteraction(bond)
        espressopp.integrator.VelocityVerlet(system).run(0)
        out[precision] = (system.storage.getParticleArrays()["f"], wca.computeEnergy(), bond.computeEnergy())
    ref = OracleSim(box, cells, skin)
    ref.add_particles(np.arange(n), pos)
    ref.decompose()
    rw_ia = ref.tab_create(ref.vl_create(rc))
    ref.tab_set(rw_ia, 0, 0, itype, rw, ew, fw, cutoff=rc)
    rfpl = ref.fpl_create()
    ref.fpl_add(rfpl, bonds)
    rb_ia = ref.tabbond_create(rfpl, itype, rb, eb, fb)
    ref.calc_forces()
    for precision, tol, etol in (("fp64", 1e-12, 1e-11), ("mixed", 1e-5, 1e-6)):
        f, ew_gpu, eb_gpu = out[precision]
        assert force_rel_err(f, ref.get("f"), 0.1).max() <= tol, precision
        assert ew_gpu == pytest.approx(ref.energy(rw_ia), rel=etol)
        assert eb_gpu == pytest.approx(ref.energy(rb_ia), rel=etol)


def test_parallel_tempering_replicas(gpu):
    # replicas of uneven size advance concurrently, each on its own stream; every replica ends
    # bit-identical to the same replica advanced alone, and exchange() follows the rule on
    # the replicas' own observables
    sizes, temps = [600, 1100, 1700, 2300], [0.7, 0.9, 1.1, 1.3]
    LJ = espressopp.interaction.LennardJones

    def make(i):
        pos, box = jittered_lattice(sizes[i], 0.8442, seed=40 + i)
        system = gpu_system(box, cells_for(box, 2.5, 0.3), 0.3, "mixed", seed=100 + i)
        system.storage.addParticlesArray(np.arange(sizes[i]), pos)
        system.storage.decompose()
        vl = espressopp.VerletList(system, cutoff=2.5)
        lj = espressopp.interaction.VerletListLennardJones(vl)
        lj.setPotential(0, 0, LJ(1.0, 1.0, 2.5, "auto"))
        system.addInteraction(lj)
        integ = espressopp.integrator.VelocityVerlet(system)
        integ.dt = 0.005
        lang = espressopp.integrator.LangevinThermostat(system)
        lang.gamma = 1.0
        lang.temperature = temps[i]
        integ.addExtension(lang)
        return system, lj, integ, lang

    pt = espressopp.ParallelTempering(NumberOfSystems=4, RNG=np.random.default_rng(7))
    reps = []
    for i in range(4):
        pt.startDefiningSystem(i)
        system, lj, integ, lang = make(i)
        pt.setIntegrator(integ, lang)
        pt.setAnalysisE(lj)
        pt.setAnalysisT(espressopp.analysis.Temperature(system))
        pt.setAnalysisNPart(espressopp.analysis.NPart(system))
        pt.endDefiningSystem(i)
        reps.append((system, lj, lang))
    pt.run(40)
    for i in range(4):
        alone, _, integ, _ = make(i)
        integ.run(40)
        assert np.array_equal(alone.storage.getParticleArrays()["pos"], reps[i][0].storage.getParticleArrays()["pos"])
    E = [r[1].computeEnergy() for r in reps]
    T = [espressopp.analysis.Temperature(r[0]).compute() for r in reps]
    swaps = pt.exchange()
    for m, n, d, u, yes in pt.log:
        assert d == pytest.approx((1 / T[n] - 1 / T[m]) * (E[m] / sizes[m] - E[n] / sizes[n]), rel=1e-12)
        assert yes == (d <= 0 or math.exp(-d) > u)
        if yes:
            assert reps[n][2].temperature == pytest.approx(T[m]) and reps[m][2].temperature == pytest.approx(T[n])
    assert [(m, n) for m, n, _, _, y in pt.log if y] == swaps
    pt.run(10)


@pytest.mark.parametrize("nbins", [100, 500, 20000])
def test_radial_distribution_matches_oracle(gpu, nbins):
    # RadialDistrF on the GPU equals the oracle bit for bit: the same FP64 pair
    # distances fall into the same bins (20000 bins: the global-histogram kernel)
    pid, pos, _, box = read_fixture()
    rc, skin = 2.5, 0.3
    system, vl, ia, ref, rvl, ria = build_lj(pos, box, cells_for(box, rc, skin), rc, skin, "fp64", ids=pid)
    # the same small moves on both sides, some particles leaving the box (no refold before a rebuild)
    moved = pos + np.random.default_rng(3).uniform(-0.1, 0.1, pos.shape)
    order = np.argsort(pid)
    system.storage.setParticleArrays("pos", pid[order], moved[order])
    ref.set("pos", moved[order])
    g = np.asarray(espressopp.analysis.RadialDistrF(system).compute(nbins))
    want = np.asarray(ref.rdf(nbins))
    assert np.abs(g - want).max() <= 1e-12 * np.abs(want).max()
    assert g[: int(0.5 / (box[1] / 2 / nbins))].max() == 0.0       # no pair closer than 0.5 sigma


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_checkpoint_resume_continues_the_run(gpu, tmp_path, precision):
    # a thermostatted run split by a checkpoint and a restore into a fresh system continues
    # the same Langevin noise and rebuild schedule: it ends where the unbroken run ends (up to
    # summation order, the restored system's cells being sorted afresh)
    pos, box = jittered_lattice(3000, 0.8442, seed=51)
    vel = mb_velocities(3000, 1.0, seed=52)
    cells = cells_for(box, 2.5, 0.3)

    def make():
        return gpu_system(box, cells, 0.3, precision, seed=77)

    def attach(system):
        vl = espressopp.VerletList(system, cutoff=2.5)
        lj = espressopp.interaction.VerletListLennardJones(vl)
        lj.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, 2.5, "auto"))
        system.addInteraction(lj)
        integ = espressopp.integrator.VelocityVerlet(system)
        integ.dt = 0.005
        lang = espressopp.integrator.LangevinThermostat(system)
        lang.gamma = 1.0
        lang.temperature = 1.0
        integ.addExtension(lang)
        return integ

    a = make()
    a.storage.addParticlesArray(np.arange(3000), pos, None, vel)
    a.storage.decompose()
    ia = attach(a)
    ia.run(60)
    fn = str(tmp_path / "chk.npz")
    espressopp.tools.checkpoint.save(fn, a, ia)
    ia.run(60)
    b = make()
    ib = attach(b)
    espressopp.tools.checkpoint.restore(fn, b, ib)
    assert ib.step == 60
    ib.run(60)
    pa, pb = a.storage.getParticleArrays(), b.storage.getParticleArrays()
    ua = pa["pos"] + pa["image"] * np.asarray(box)
    ub = pb["pos"] + pb["image"] * np.asarray(box)
    assert np.array_equal(pa["id"], pb["id"])
    assert np.abs(ua - ub).max() < 1e-8
    assert np.abs(pa["v"] - pb["v"]).max() < 1e-7
    assert ia.step == ib.step == 120
