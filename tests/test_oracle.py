"""Pins of the FP64 CPU oracle against the reference's own known answers.

Each check uses a value the reference's test suite or the mathematics fixes;
none restates the oracle's own formula.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from oracle.oracle import OracleSim

from _systems import cells_for, jittered_lattice, read_fixture


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32-10 (Salmon et al., SC'11)
    assert [hex(x) for x in oracle.philox4x32_10([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(x) for x in oracle.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(x) for x in oracle.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                                 [0xA4093822, 0x299F31D0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


def test_langevin_uniforms_moments():
    u = np.array([oracle.langevin_uniforms(7, c, p) for c in range(40) for p in range(250)]).ravel()
    assert np.all((u > 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1.0 / 12.0) < 0.002


def test_fene_known_answers():
    # testsuite/interaction_potentials/unittest/TestFENE.cpp:35-37 and :64-68 (tolerance 0.01 %)
    for r, e in ((0.97, 18.279), (0.80, 11.296), (1.40, 69.147)):
        assert oracle.fene_energy(30.0, 0.0, 1.5, r) == pytest.approx(e, rel=1e-4)
    f = oracle.fene_force(30.0, 0.0, 1.5, [0.9, 0.1, 0.1])
    assert f == pytest.approx([-42.782, -4.7535, -4.7535], rel=1e-4)


def test_lj_known_answers():
    # testsuite/interaction_potentials/unittest/TestLennardJones.cpp:30-49
    inf = float("inf")
    assert oracle.lj_energy(1.0, 1.0, inf, 0.0, 2 ** (1 / 6)) == pytest.approx(-1.0, rel=1e-4)
    assert abs(oracle.lj_energy(1.0, 1.0, inf, 0.0, 2.5)) < 5.0
    assert abs(oracle.lj_energy(1.0, 1.0, 2.0, "auto", 1.9999)) < 0.01
    assert abs(oracle.lj_energy(1.0, 1.0, 2.0, "auto", 2.0)) < 0.01
    assert oracle.lj_energy(1.0, 1.0, 2.0, "auto", 10.0) == 0.0
    # testsuite/interaction_potentials/unittest/testLennardJones.py: eps=sigma=2 -> root at 2,
    # minimum -2 at 2*2^(1/6) with zero force there
    assert oracle.lj_energy(2.0, 2.0, inf, 0.0, 2.0) == pytest.approx(0.0, abs=1e-12)
    rmin = 2.0 * 2 ** (1 / 6)
    assert oracle.lj_energy(2.0, 2.0, inf, 0.0, rmin) == pytest.approx(-2.0, rel=1e-12)
    assert np.allclose(oracle.lj_force(2.0, 2.0, inf, [rmin, 0.0, 0.0]), 0.0, atol=1e-12)
    # force is minus the derivative of the energy (central difference)
    r, h = 1.3, 1e-6
    dE = (oracle.lj_energy(1.0, 1.0, inf, 0.0, r + h) - oracle.lj_energy(1.0, 1.0, inf, 0.0, r - h)) / (2 * h)
    assert oracle.lj_force(1.0, 1.0, inf, [r, 0.0, 0.0])[0] == pytest.approx(-dE, rel=1e-7)


def _lattice_sim(cut):
    # testsuite/unittest/testVerletList.py:38-109: 6^3 simple cubic lattice, box 6, skin 0.001
    N, SIZE = 6, 6.0
    s = OracleSim((SIZE,) * 3, (3, 3, 3), 0.001)
    g = (np.arange(N) + 0.5) / N * SIZE
    pos = np.array([[x, y, z] for x in g for y in g for z in g])
    s.add_particles(np.arange(N ** 3), pos)
    s.decompose()
    return s, s.vl_create(cut)


@pytest.mark.parametrize("cut,count", [(0.0, 0), (1.0, 3 * 216), (math.sqrt(2.0), 9 * 216),
                                       (math.sqrt(3.0), 13 * 216)])
def test_verlet_list_lattice_counts(cut, count):
    s, vl = _lattice_sim(cut)
    assert s.vl_size(vl) == count


def _brute_force_pairs(pos, box, cutsq):
    L = np.asarray(box)
    n = len(pos)
    out = set()
    for i in range(n):
        d = pos[i] - pos[i + 1:]
        d -= L * np.round(d / L)
        r2 = np.sum(d * d, axis=1)
        for j in np.nonzero(r2 <= cutsq)[0]:
            out.add((i, i + 1 + int(j)))
    return out


@pytest.mark.parametrize("seed,box,rc,skin", [(1, (9.0, 9.0, 9.0), 2.5, 0.3),
                                              (2, (8.7, 11.3, 14.9), 2.5, 0.4),
                                              (3, (6.1, 6.1, 9.3), 1.12, 0.3)])
def test_verlet_list_brute_force(seed, box, rc, skin):
    # PTestVerletList.cpp:152-212 pattern: random particles vs O(N^2) pair count
    rng = np.random.default_rng(seed)
    n = 500
    pos = rng.uniform(0, 1, (n, 3)) * np.asarray(box)
    s = OracleSim(box, cells_for(box, rc, skin), skin)
    s.add_particles(np.arange(n), pos)
    s.decompose()
    vl = s.vl_create(rc)
    got = set(map(tuple, np.sort(s.vl_pairs(vl), axis=1).tolist()))
    assert got == _brute_force_pairs(pos, box, (rc + skin) ** 2)


def test_exclusions_remove_pairs():
    rng = np.random.default_rng(5)
    box = (8.0, 8.0, 8.0)
    pos = rng.uniform(0, 8, (300, 3))
    s = OracleSim(box, (2, 2, 2), 0.3)
    s.add_particles(np.arange(300), pos)
    s.decompose()
    vl = s.vl_create(2.5)
    pairs = s.vl_pairs(vl)
    excl = pairs[:25]
    s.vl_exclude(vl, excl[:, ::-1])          # either orientation excludes
    after = set(map(tuple, np.sort(s.vl_pairs(vl), axis=1).tolist()))
    assert len(after) == len(pairs) - 25
    assert not (set(map(tuple, np.sort(excl, axis=1).tolist())) & after)


def test_golden_nve_temperature():
    # testsuite/potential_array/potential_array_test.py:37-163: T = 0.160074388303 +- 1e-8
    pid, pos, _, box = read_fixture()
    rc, skin = 2.5, 0.4
    s = OracleSim(box, [int(b / (rc + skin)) for b in box], skin)
    s.add_particles(pid, pos, types=pid % 2)
    s.decompose()
    ia = s.lj_create(s.vl_create(rc))
    s.lj_set(ia, 0, 0, 0.1, 1.0, rc, 0.0)
    s.lj_set(ia, 0, 1, (0.1 * 0.3) ** 0.5, 0.8, rc, 0.0)
    s.lj_set(ia, 1, 1, 0.9, 0.6, rc, 0.0)
    s.run(100, dt=0.005)
    n = len(pid)
    s.add_particles(np.arange(n, n + 10), pos[:10], types=np.full(10, 6))
    s.decompose()
    s.run(100)
    assert abs(s.temperature() - 0.160074388303) < 1.0e-8


def test_free_flight_returns_to_lattice():
    # testsuite/integrator/PTestVelocityVerlet.cpp:222-256: free particles cross the box and return
    pos, box = jittered_lattice(125, 0.5, seed=3, jitter=0.0)
    s = OracleSim(box, (2, 2, 2), 0.3)
    v = np.tile([1.0, -0.5, 0.25], (125, 1))
    s.add_particles(np.arange(125), pos, vel=v)
    s.decompose()
    L = box[0]
    dt = L / 400.0
    s.run(1600, dt=dt)                      # x moves 4 L, y -2 L, z 1 L
    p = s.get("pos") + s.get("image") * np.asarray(box)
    assert np.allclose(p, pos + np.array([4 * L, -2 * L, L]), atol=1e-6)


def test_harmonic_like_fene_energy_conservation():
    # two bonded particles oscillating in a FENE well: NVE energy is conserved to O(dt^2)
    box = (10.0, 10.0, 10.0)
    s = OracleSim(box, (2, 2, 2), 0.3)
    s.add_particles([0, 1], [[5.0, 5.0, 5.0], [6.1, 5.0, 5.0]])
    s.decompose()
    fpl = s.fpl_create()
    s.fpl_add(fpl, [[0, 1]])
    ia = s.fene_create(fpl, 30.0, 0.0, 1.5)
    s.run(0, dt=0.001)
    e0 = s.energy(ia)
    for _ in range(20):
        s.run(50)
        v = s.get("v")
        etot = s.energy(ia) + 0.5 * np.sum(v * v)
        assert abs(etot - e0) < 1e-4 * abs(e0)


def test_thread_count_does_not_change_results():
    pid, pos, _, box = read_fixture()
    out = []
    for threads in (1, 4):
        s = OracleSim(box, (4, 4, 4), 0.3, threads=threads)
        s.add_particles(pid, pos)
        s.decompose()
        ia = s.lj_create(s.vl_create(2.5))
        s.lj_set(ia, 0, 0, 1.0, 1.0, 2.5, "auto")
        s.run(20, dt=0.005)
        out.append(s.get("pos"))
    assert np.array_equal(out[0], out[1])


# ---------------------------------------------------------------- Tabulated tables
def test_tab_linear_reproduces_linear_functions():
    # InterpolationLinear: a r + b per interval is exact for a linear table
    r = np.linspace(0.3, 2.7, 33)
    e, f = 2.0 - 0.75 * r, 0.5 + 1.25 * r
    at = np.concatenate([np.linspace(0.3, 2.69999, 97), [0.1, 2.7, 2.9]])   # both ends extrapolate
    ev, fv = oracle.tab_eval(1, r, e, f, at)
    assert np.allclose(ev, 2.0 - 0.75 * at, rtol=0, atol=1e-13)
    assert np.allclose(fv, 0.5 + 1.25 * at, rtol=0, atol=1e-13)


def test_tab_cubic_exact_for_cubic_energy_and_linear_force():
    # clamped cubic spline (end slopes -f) reproduces a cubic energy whose force is -dE/dr;
    # a linear force table (secant end slopes) is reproduced exactly as well
    r = np.linspace(0.5, 2.5, 41)
    E = lambda x: 3 * x ** 3 - 2 * x ** 2 + x - 5  # noqa: E731
    Fc = lambda x: -(9 * x ** 2 - 4 * x + 1)       # noqa: E731
    at = np.linspace(0.5, 2.4999, 301)
    ev, _ = oracle.tab_eval(3, r, E(r), Fc(r), at)
    assert np.allclose(ev, E(at), rtol=0, atol=1e-11)
    _, fv = oracle.tab_eval(3, r, E(r), 4.0 - 2.0 * r, at)
    assert np.allclose(fv, 4.0 - 2.0 * at, rtol=0, atol=1e-12)
    # grid points are interpolated exactly for any data
    rng = np.random.default_rng(3)
    y = rng.normal(size=len(r))
    ev, fv = oracle.tab_eval(3, r, y, y, r[:-1])
    assert np.allclose(ev, y[:-1], atol=1e-12) and np.allclose(fv, y[:-1], atol=1e-12)


@pytest.mark.parametrize("delta", [0.05, 1.5])
def test_tab_akima_quadratic_with_integer_square(delta):
    # Akima slopes are exact for quadratic data (equal weights, parabolic end extension), so
    # on interval i: y = y_i + y'(x_i) z + c * int(z*z) + 0 * ..., the reference keeping
    # z*z in an int (src/interaction/InterpolationAkima.hpp:105)
    c, b0, a0 = 0.7, -1.1, 2.0
    r = 0.4 + delta * np.arange(20)
    y = c * r ** 2 + b0 * r + a0
    i = np.arange(len(r) - 1)
    z = delta * np.array([0.1, 0.45, 0.8, 0.97])
    at = (r[i][:, None] + z[None, :]).ravel()
    xi = np.repeat(r[i], len(z))
    zz = np.tile(z, len(i))
    want = (c * xi ** 2 + b0 * xi + a0) + (2 * c * xi + b0) * zz + c * np.floor(zz * zz)
    ev, fv = oracle.tab_eval(2, r, y, y, at)
    assert np.allclose(ev, want, rtol=1e-12, atol=1e-12)
    assert np.allclose(fv, want, rtol=1e-12, atol=1e-12)


def _lj_table(N, lo, hi, eps=1.0, sig=1.0, rc=2.5):
    r = np.linspace(lo, hi, N)
    e = np.array([oracle.lj_energy(eps, sig, rc, 0.0, x) for x in r])
    f = np.array([oracle.lj_force(eps, sig, float("inf"), [x, 0.0, 0.0])[0] for x in r])
    return r, e, f


def test_tabulated_pair_interaction_matches_lennard_jones():
    # testsuite/tabulated/polymer_melt_tabulated.py pattern: a fine LJ table run through
    # VerletListTabulated gives the analytic LJ forces and energy to interpolation accuracy
    pid, pos, _, box = read_fixture()
    rc, skin = 2.5, 0.3
    r, e, f = _lj_table(4001, 0.7, rc)
    out = {}
    for kind in ("lj", "tab"):
        s = OracleSim(box, cells_for(box, rc, skin), skin)
        s.add_particles(pid, pos)
        s.decompose()
        vl = s.vl_create(rc)
        if kind == "lj":
            ia = s.lj_create(vl)
            s.lj_set(ia, 0, 0, 1.0, 1.0, rc, 0.0)
        else:
            ia = s.tab_create(vl)
            s.tab_set(ia, 0, 0, 3, r, e, f, cutoff=rc)
        s.calc_forces()
        out[kind] = (s.get("f"), s.energy(ia), s.virial(ia)[0])
    fl, el, wl = out["lj"]
    ft, et, wt = out["tab"]
    scale = np.sqrt(np.mean(np.sum(fl ** 2, axis=1)))
    assert np.abs(ft - fl).max() < 1e-6 * scale
    assert et == pytest.approx(el, rel=1e-8)
    assert wt == pytest.approx(wl, rel=1e-7)


def test_tabulated_bond_matches_fene():
    box = (20.0, 20.0, 20.0)
    rng = np.random.default_rng(9)
    n = 200
    pos = rng.uniform(0, 20, (n, 3))
    bonds = []
    for k in range(0, n - 1, 2):   # place partners at random bond lengths 0.6-1.3
        u = rng.normal(size=3)
        pos[k + 1] = pos[k] + rng.uniform(0.6, 1.3) * u / np.linalg.norm(u)
        bonds.append((k, k + 1))
    r = np.linspace(0.0001, 1.49, 2981)
    e = np.array([oracle.fene_energy(30.0, 0.0, 1.5, x) for x in r])
    f = np.array([oracle.fene_force(30.0, 0.0, 1.5, [x, 0.0, 0.0])[0] for x in r])
    out = {}
    for kind in ("fene", "tab"):
        s = OracleSim(box, (3, 3, 3), 0.3)
        s.add_particles(np.arange(n), pos)
        s.decompose()
        fpl = s.fpl_create()
        s.fpl_add(fpl, bonds)
        ia = s.fene_create(fpl) if kind == "fene" else s.tabbond_create(fpl, 3, r, e, f)
        s.calc_forces()
        out[kind] = (s.get("f"), s.energy(ia))
    assert np.abs(out["tab"][0] - out["fene"][0]).max() < 1e-5
    assert out["tab"][1] == pytest.approx(out["fene"][1], rel=1e-8)


# ---------------------------------------------------------------- radial distribution
def _rdf_counts(g, nbins, n, box):
    # back from g(r) to pair counts with the reference's normalisation
    # (src/analysis/RadialDistrF.cpp:143-151)
    L = np.asarray(box, dtype=np.float64)
    dr = L[1] / 2 / nbins
    rho = n / (L[0] * L[1] * L[2])
    factor = 2.0 * math.pi * dr * rho * n
    r = (np.arange(nbins) + 0.5) * dr
    return np.asarray(g) * factor * (r * r + dr * dr / 12.0)


def test_rdf_simple_cubic_shells():
    # pair counts of an M^3 simple-cubic lattice: every lattice vector of the
    # minimum-image cube with |v| < L/2 appears once per particle (each pair once)
    M, a = 8, 1.0
    g = (np.arange(M) + 0.25) * a
    pos = np.array([[x, y, z] for x in g for y in g for z in g])
    box = (M * a,) * 3
    s = OracleSim(box, (2, 2, 2), 0.3)
    s.add_particles(np.arange(len(pos)), pos)
    s.decompose()
    nbins = 40                                    # dr = 0.1: shells sit well inside bins
    counts = _rdf_counts(s.rdf(nbins), nbins, len(pos), box)
    want = np.zeros(nbins)
    h = M // 2
    for vx in range(-h, h + 1):
        for vy in range(-h, h + 1):
            for vz in range(-h, h + 1):
                r = math.sqrt(vx * vx + vy * vy + vz * vz) * a
                if 0 < r < M * a / 2 and max(abs(vx), abs(vy), abs(vz)) < h:
                    want[int(r / (M * a / 2 / nbins))] += 1
    want *= len(pos) / 2
    assert np.allclose(counts, want, rtol=1e-9, atol=1e-6)


def test_rdf_ideal_gas_is_one():
    rng = np.random.default_rng(17)
    n, L = 4000, 20.0
    s = OracleSim((L, L, L), (2, 2, 2), 0.3, threads=4)
    s.add_particles(np.arange(n), rng.uniform(0, L, (n, 3)))
    s.decompose()
    nbins = 20
    g = np.asarray(s.rdf(nbins))
    counts = _rdf_counts(g, nbins, n, (L, L, L))
    sel = counts > 2000
    assert sel.sum() >= 15
    # Poisson error of each bin's count
    assert np.all(np.abs(g[sel] - 1.0) < 5.0 / np.sqrt(counts[sel]))


# ---------------------------------------------------------------- virial and pressure
def _two_particles(d, box=(20.0, 20.0, 20.0)):
    s = OracleSim(box, (3, 3, 3), 0.3)
    x1 = np.array([10.3, 9.7, 10.1])
    s.add_particles([0, 1], [x1, x1 - np.asarray(d)])
    s.decompose()
    return s


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("r", [0.95, 1.2, 2.3])
def test_two_particle_lj_virial_closed_form(r):
    # W = d . F1 = -r U'(r) = 48 eps sig^12 / r^12 - 24 eps sig^6 / r^6 for
    # U = 4 eps ((sig/r)^12 - (sig/r)^6); tensor W_ab = W d_a d_b / r^2
    eps, sig, rc = 1.3, 0.9, 2.5
    d = r * _unit([0.3, -0.5, 0.8])
    s = _two_particles(d)
    ia = s.lj_create(s.vl_create(rc))
    s.lj_set(ia, 0, 0, eps, sig, rc, "auto")
    w_exact = 48.0 * eps * sig ** 12 / r ** 12 - 24.0 * eps * sig ** 6 / r ** 6
    w, t = s.virial(ia)
    assert w == pytest.approx(w_exact, rel=1e-12)
    want = w_exact * np.array([d[0] * d[0], d[1] * d[1], d[2] * d[2], d[0] * d[1], d[0] * d[2], d[1] * d[2]]) / r ** 2
    assert np.allclose(t, want, rtol=1e-11, atol=1e-12 * abs(w_exact))
    # the pressure of the pair at rest: P = W / (3V)
    assert s.pressure() == pytest.approx(w_exact / (3.0 * 20.0 ** 3), rel=1e-12)


def test_lj_virial_is_zero_beyond_cutoff():
    s = _two_particles(2.6 * _unit([1.0, 1.0, 0.0]))
    ia = s.lj_create(s.vl_create(2.5))
    s.lj_set(ia, 0, 0, 1.0, 1.0, 2.5, "auto")
    assert s.virial(ia)[0] == 0.0


@pytest.mark.parametrize("r0,r", [(0.0, 0.97), (0.0, 1.4), (0.2, 0.97), (0.35, 1.6)])
def test_two_particle_fene_closed_form(r0, r):
    # U = -1/2 K R^2 ln(1 - ((r - r0)/R)^2)  =>  W = -r U'(r) = -K r (r - r0) / (1 - ((r - r0)/R)^2)
    K, R = 30.0, 1.5
    d = r * _unit([-0.2, 0.9, 0.4])
    s = _two_particles(d)
    fpl = s.fpl_create()
    s.fpl_add(fpl, [[0, 1]])
    ia = s.fene_create(fpl, K, r0, R)
    x = (r - r0) / R
    e_exact = -0.5 * K * R * R * math.log(1.0 - x * x)
    w_exact = -K * r * (r - r0) / (1.0 - x * x)
    assert s.energy(ia) == pytest.approx(e_exact, rel=1e-12)
    w, t = s.virial(ia)
    assert w == pytest.approx(w_exact, rel=1e-12)
    assert t[3] == pytest.approx(w_exact * d[0] * d[1] / r ** 2, rel=1e-11)
    # forces: F1 = -U'(r) d / r on the first particle, the opposite on the second
    s.calc_forces()
    f = s.get("f")
    f1 = -(K * (r - r0) / (1.0 - x * x)) * d / r
    assert np.allclose(f[0], f1, rtol=1e-12, atol=1e-12)
    assert np.allclose(f[1], -f1, rtol=1e-12, atol=1e-12)


def test_fene_nonzero_r0_energy_and_force():
    # FENE(K, r0 > 0, R): minimum at r = r0 (zero force, zero energy), and the
    # energy is symmetric in r - r0 (src/interaction/FENE.hpp:91-112)
    K, r0, R = 30.0, 0.25, 1.5
    assert oracle.fene_energy(K, r0, R, r0) == 0.0
    assert np.allclose(oracle.fene_force(K, r0, R, [r0, 0.0, 0.0]), 0.0, atol=1e-15)
    for h in (0.1, 0.5, 1.2):
        assert oracle.fene_energy(K, r0, R, r0 + h) == pytest.approx(oracle.fene_energy(K, 0.0, R, h), rel=1e-13)
    r, h = 1.1, 1e-6
    dE = (oracle.fene_energy(K, r0, R, r + h) - oracle.fene_energy(K, r0, R, r - h)) / (2 * h)
    assert oracle.fene_force(K, r0, R, [r, 0.0, 0.0])[0] == pytest.approx(-dE, rel=1e-7)


def test_ideal_gas_pressure_and_temperature():
    # no interactions: P = sum m v^2 / (3V), T = sum m v^2 / (3N) (src/analysis/Pressure.cpp:41-121,
    # src/analysis/Temperature.hpp:112-125)
    rng = np.random.default_rng(8)
    n, box = 300, (7.0, 8.0, 9.0)
    pos = rng.uniform(0, 1, (n, 3)) * np.asarray(box)
    v = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 2.0, n)
    s = OracleSim(box, (2, 2, 2), 0.3)
    s.add_particles(np.arange(n), pos, vel=v, mass=m)
    s.decompose()
    mv2 = float(np.sum(m[:, None] * v * v))
    assert s.pressure() == pytest.approx(mv2 / (3.0 * 7.0 * 8.0 * 9.0), rel=1e-13)
    assert s.temperature() == pytest.approx(mv2 / (3.0 * n), rel=1e-13)


def test_virial_brute_force():
    # O(N^2) minimum-image sum of d (x) F over all pairs within rc (LJ) plus the bonds (FENE)
    rng = np.random.default_rng(21)
    box = (9.5, 10.0, 11.0)
    L = np.asarray(box)
    n = 400
    sparse = []
    while len(sparse) < n:                      # keep pairs apart so forces stay moderate
        p = rng.uniform(0, 1, 3) * L
        if all(np.linalg.norm((p - q) - L * np.round((p - q) / L)) > 0.8 for q in sparse):
            sparse.append(p)
    pos = np.asarray(sparse)
    bonds = np.array([(k, k + 1) for k in range(0, 40, 2)], dtype=np.int32)
    for a, b in bonds:                          # bond partners at 0.9-1.2
        u = _unit(rng.normal(size=3))
        pos[b] = pos[a] + rng.uniform(0.9, 1.2) * u
    rc, eps, sig = 2.5, 1.0, 1.0
    s = OracleSim(box, cells_for(box, rc, 0.3), 0.3)
    s.add_particles(np.arange(n), pos)
    s.decompose()
    ia = s.lj_create(s.vl_create(rc))
    s.lj_set(ia, 0, 0, eps, sig, rc, "auto")
    fpl = s.fpl_create()
    s.fpl_add(fpl, bonds)
    ib = s.fene_create(fpl, 30.0, 0.0, 1.5)
    W = np.zeros((3, 3))
    for i in range(n):
        d = pos[i] - pos[i + 1:]
        d -= L * np.round(d / L)
        r2 = np.sum(d * d, axis=1)
        sel = r2 <= rc * rc
        r2s, ds = r2[sel], d[sel]
        fmag_over_r = (48.0 * eps * sig ** 12 / r2s ** 7 - 24.0 * eps * sig ** 6 / r2s ** 4)
        W += np.einsum("ka,kb->ab", ds, ds * fmag_over_r[:, None])
    Wb = np.zeros((3, 3))
    for a, b in bonds:
        d = pos[a] - pos[b]
        d -= L * np.round(d / L)
        Wb += np.outer(d, d * (-30.0 / (1.0 - (d @ d) / 2.25)))
    for (w, t), ref in ((s.virial(ia), W), (s.virial(ib), Wb)):
        want = np.array([ref[0, 0], ref[1, 1], ref[2, 2], ref[0, 1], ref[0, 2], ref[1, 2]])
        assert w == pytest.approx(np.trace(ref), rel=1e-11)
        assert np.allclose(t, want, rtol=1e-10, atol=1e-10 * abs(np.trace(ref)))
    assert s.pressure() == pytest.approx((np.trace(W) + np.trace(Wb)) / (3.0 * np.prod(L)), rel=1e-11)


def test_langevin_holds_the_target_temperature():
    # the thermostat's fluctuation-dissipation balance: <T> -> T (uniform noise of variance
    # 2 gamma T m / dt per component, src/integrator/LangevinThermostat.cpp:141-189)
    pos, box = jittered_lattice(1000, 0.5, seed=4)
    s = OracleSim(box, cells_for(box, 2.5, 0.3), 0.3, threads=4)
    s.add_particles(np.arange(1000), pos)
    s.decompose()
    ia = s.lj_create(s.vl_create(2.5))
    s.lj_set(ia, 0, 0, 1.0, 1.0, 2.5, "auto")
    s.langevin(1.0, 1.5)
    s.run(500, dt=0.005)
    temps = []
    for _ in range(100):
        s.run(20)
        temps.append(s.temperature())
    temps = np.asarray(temps)
    # 100 samples 20 steps apart (about the velocity decorrelation time 1/gamma = 200 steps / 10)
    assert abs(temps.mean() - 1.5) < 0.03 * 1.5
    assert temps.std() < 0.1 * 1.5


# ---------------------------------------------------------------- the reference's melt golden values
def test_melt_golden_step0_energies_and_pressure(tmp_path):
    # testsuite/tabulated/result.dat lines 14 and 22: step-0 epair, ebond (both runs) and P, Pxy
    # (plain run) of polymer_melt_tabulated.py on polymer_melt.start
    import _melt_golden as mg
    ids, pos, box, bonds, angles = mg.load(tmp_path)
    cells = [int(b / (mg.RC + mg.SKIN)) for b in box]
    V = float(np.prod(box))
    for run, golden in (("plain", mg.GOLDEN_PLAIN), ("tab", mg.GOLDEN_TAB)):
        s = OracleSim(box, cells, mg.SKIN, threads=4)
        s.add_particles(ids, pos)
        s.decompose()
        vl = s.vl_create(mg.RC)
        fpl = s.fpl_create()
        s.fpl_add(fpl, bonds)
        if run == "plain":
            ia = s.lj_create(vl)
            s.lj_set(ia, 0, 0, 1.0, 1.0, mg.RC, 0.0)
            ib = s.fene_create(fpl, 30.0, 0.0, 1.5)
        else:
            ia = s.tab_create(vl)
            s.tab_set(ia, 0, 0, 2, *mg.lj_table(), cutoff=mg.RC)
            ib = s.tabbond_create(fpl, 2, *mg.fene_table())
        assert abs(s.energy(ia) - golden["epair"]) <= mg.HALF_UNIT["epair"], run
        assert abs(s.energy(ib) - golden["ebond"]) <= mg.HALF_UNIT["ebond"], run
        if run == "plain":
            # angle forces are normal to both bonds, so they add nothing to the scalar virial;
            # P at rest is therefore the pair and bond virial over 3V
            w_ang, wxy_ang = mg.cosine_virial(pos, box, angles)
            assert abs(w_ang) < 1e-6
            assert abs(s.pressure() - golden["P"]) <= mg.HALF_UNIT["P"]
            # the printed angle energy and Pxy follow K (1 - cos(theta - theta0)), the opposite
            # sign of today's Cosine.hpp; with that sign the tensor closes as well
            assert abs(2 * 1.5 * len(angles) - mg.cosine_energy(pos, box, angles) - golden["eangle"]) <= \
                mg.HALF_UNIT["eangle"]
            wxy = s.virial(ia)[1][3] + s.virial(ib)[1][3] - wxy_ang
            assert abs(wxy / V - golden["Pxy"]) <= mg.HALF_UNIT["Pxy"]


@pytest.mark.parametrize("hci", [2, 3])
def test_half_cells_give_the_same_pairs(hci):
    # DomainDecomposition(..., halfCellInt): cells of (rc+skin)/h with neighbours within +-h cells
    # (src/storage/DomainDecomposition.cpp:325-374); the pair set is the brute-force set
    rng = np.random.default_rng(40 + hci)
    box = (9.3, 10.1, 11.7)
    rc, skin = 2.5, 0.3
    n = 600
    pos = rng.uniform(0, 1, (n, 3)) * np.asarray(box)
    from espressopp_amd.tools import decomp
    cells = decomp.cellGrid(box, (1, 1, 1), rc, skin, hci)
    s = OracleSim(box, cells, skin, half_cell_int=hci)
    s.add_particles(np.arange(n), pos)
    s.decompose()
    vl = s.vl_create(rc)
    got = set(map(tuple, np.sort(s.vl_pairs(vl), axis=1).tolist()))
    assert got == _brute_force_pairs(pos, box, (rc + skin) ** 2)


def test_half_cell_lattice_counts():
    # testsuite/unittest/testVerletList.py:91-109 lattice with half cells: same counts
    N, SIZE = 6, 6.0
    g = (np.arange(N) + 0.5) / N * SIZE
    pos = np.array([[x, y, z] for x in g for y in g for z in g])
    for cut, count in ((1.0, 3 * 216), (math.sqrt(2.0), 9 * 216), (math.sqrt(3.0), 13 * 216)):
        s = OracleSim((SIZE,) * 3, (6, 6, 6), 0.001, half_cell_int=2)
        s.add_particles(np.arange(N ** 3), pos)
        s.decompose()
        assert s.vl_size(s.vl_create(cut)) == count
