"""Host logic of the coarse-graining tools (no GPU): mapping, distributions, tables."""
import numpy as np
import pytest

from espressopp_amd.interaction import read_table
from espressopp_amd.tools import coarse_grain as cg


def test_blob_map_centre_of_mass_and_force_sum():
    rng = np.random.default_rng(1)
    L = 10.0
    pos = rng.uniform(-5, 15, (2 * 8, 3))       # 2 chains x 8 beads, unwrapped
    f = rng.normal(size=(16, 3))
    p, F = cg.blob_map(pos, f, 2, 8, 4, (L, L, L))
    assert p.shape == (4, 3) and F.shape == (4, 3)
    assert np.allclose(p[1], np.mod(pos[4:8].mean(axis=0), L))
    assert np.allclose(F[3], f[12:16].sum(axis=0))
    assert (p >= 0).all() and (p < L).all()
    # the total force is conserved by the mapping
    assert np.allclose(F.sum(axis=0), f.sum(axis=0))
    with pytest.raises(ValueError):
        cg.blob_map(pos, f, 2, 8, 3, (L, L, L))


def test_cg_bonds_join_consecutive_sites_of_a_chain():
    b = cg.cg_bonds(3, 200, 4)
    assert len(b) == 3 * 49
    assert b[0].tolist() == [0, 1] and b[49].tolist() == [50, 51]
    assert not any(a == 49 and c == 50 for a, c in b)


def test_pair_rdf_of_an_ideal_gas_is_one():
    rng = np.random.default_rng(2)
    L = 12.0
    frames = [rng.uniform(0, L, (3000, 3)) for _ in range(3)]
    r, g = cg.pair_rdf(frames, (L, L, L), 3.0, 30)
    assert len(r) == 30 and r[0] == pytest.approx(0.05)
    assert np.abs(g[5:] - 1.0).max() < 0.08


def test_pair_rdf_exclusions_remove_the_bonded_peak():
    L = 20.0
    a = np.array([[1.0, 1.0, 1.0], [2.0, 1.0, 1.0], [15.0, 15.0, 15.0]])
    r, g = cg.pair_rdf([a], (L, L, L), 3.0, 30)
    assert g[10] > 0                              # the pair at distance 1.0
    r, g = cg.pair_rdf([a], (L, L, L), 3.0, 30, exclude=[(0, 1)])
    assert g.max() == 0.0


def test_tables_round_trip_and_shift(tmp_path):
    r = np.linspace(0.0, 3.0, 61)
    V = (1.0 / np.maximum(r, 0.3)) ** 6
    rr, e, f = cg.table_from_potential(r, V, 2.5)
    assert e[np.searchsorted(rr, 2.5)] == pytest.approx(0.0, abs=1e-12)
    assert (e[rr > 2.5] == 0).all() and (f[rr > 2.5] == 0).all()
    fn = str(tmp_path / "t.tab")
    cg.write_table(fn, rr, e, f)
    a, b, c = read_table(fn)
    assert np.allclose(a, rr, rtol=1e-7) and np.allclose(b, e, rtol=1e-7, atol=1e-12)


def test_bond_table_is_minimal_at_the_most_probable_length():
    rng = np.random.default_rng(3)
    L = 50.0
    n = 4000
    a = rng.uniform(10, 40, (n, 3))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    b = a + u * rng.normal(1.9, 0.2, (n, 1))
    frames = [np.concatenate([a, b])]
    bonds = np.stack([np.arange(n), np.arange(n) + n], axis=1)
    r, V, F = cg.bond_table(frames, bonds, (L, L, L), kT=1.0)
    rmin = r[np.argmin(V)]
    assert 1.6 < rmin < 2.1
    # restoring force on both sides of the minimum
    assert F[np.searchsorted(r, 1.3)] > 0 and F[np.searchsorted(r, 2.6)] < 0


def test_ibi_update_raises_the_potential_where_the_model_overshoots():
    r = np.linspace(0.05, 3.0, 60)
    V = np.zeros_like(r)
    g_t = np.ones_like(r)
    g_m = np.ones_like(r)
    g_m[20:30] = 1.5
    _, V2, _ = cg.ibi_update(r, V, g_m, g_t, 1.0, 3.0, alpha=1.0)
    assert V2[25] > V2[10] and V2[25] > V2[45]
