"""Coarse-graining of Kremer-Grest melts and the plain and learned CG models.

The workflow of the coarse-grained workload, on the GPU engine throughout:

  melt_frames      a bead-spring melt (WCA + FENE, Langevin) run by the plain
                   method; frames of unwrapped positions and conservative forces
  blob_map         m consecutive beads of a chain -> one site at their centre of
                   mass, carrying the sum of their forces (the mapped force of
                   force matching)
  cg_bonds / bond_table    bonded sites and their Boltzmann-inverted bond table
  pair_rdf / ibi   the non-bonded radial distribution and iterative Boltzmann
                   inversion of a pair table (interaction.Tabulated), the plain
                   fallback and judge
  train_pair_mlp   force matching of interaction.VerletListPairMLP on the
                   residual forces (mapped force minus bond force)
  force_error / rdf_error  the judge's two measures

Tables are written in the format interaction.Tabulated reads (rows "r e f",
src/interaction/InterpolationTable.cpp:106-140).
"""
import math
import os

import numpy as np


# ------------------------------------------------------------------ mapping

def blob_map(pos, force, nchains, nbeads, per_blob, box):
    """Sites at the centre of mass of per_blob consecutive beads (equal masses) and
    the sum of their forces.  pos: unwrapped bead positions in id order, chains
    stored consecutively.  Site positions are folded into the box."""
    if nbeads % per_blob:
        raise ValueError("beads per chain must be a multiple of beads per site")
    L = np.asarray(box, dtype=np.float64)
    nb = nbeads // per_blob
    p = np.asarray(pos, dtype=np.float64).reshape(nchains, nb, per_blob, 3).mean(axis=2).reshape(-1, 3)
    f = np.asarray(force, dtype=np.float64).reshape(nchains, nb, per_blob, 3).sum(axis=2).reshape(-1, 3)
    return p - L * np.floor(p / L), f


def cg_bonds(nchains, nbeads, per_blob):
    """Pairs of consecutive sites of every chain."""
    nb = nbeads // per_blob
    a = (np.arange(nchains)[:, None] * nb + np.arange(nb - 1)[None, :]).ravel()
    return np.stack([a, a + 1], axis=1).astype(np.int32)


# ------------------------------------------------------------------ distributions

def _min_image(d, L):
    return d - L * np.rint(d / L)


def pair_rdf(frames, box, rmax, nbins, exclude=None):
    """Radial distribution of the non-excluded pairs, averaged over frames; returns
    bin centres and g(r) normalised by the ideal-gas count of those pairs."""
    from scipy.spatial import cKDTree
    L = float(box[0])
    edges = np.linspace(0.0, rmax, nbins + 1)
    hist = np.zeros(nbins)
    n = len(frames[0])
    ex = None
    if exclude is not None and len(exclude):
        e = np.sort(np.asarray(exclude, dtype=np.int64), axis=1)
        ex = np.unique(e[:, 0] * n + e[:, 1])
    for p in frames:
        t = cKDTree(np.mod(p, L), boxsize=L)
        pr = t.query_pairs(rmax, output_type="ndarray")
        if ex is not None:
            key = np.minimum(pr[:, 0], pr[:, 1]).astype(np.int64) * n + np.maximum(pr[:, 0], pr[:, 1])
            pr = pr[~np.isin(key, ex, assume_unique=False)]
        d = _min_image(p[pr[:, 0]] - p[pr[:, 1]], L)
        hist += np.histogram(np.sqrt(np.sum(d * d, axis=1)), bins=edges)[0]
    npairs = n * (n - 1) / 2 - (len(exclude) if exclude is not None else 0)
    shell = 4.0 / 3.0 * math.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    ideal = npairs * shell / L ** 3 * len(frames)
    return 0.5 * (edges[1:] + edges[:-1]), hist / ideal


def bond_lengths(frames, bonds, box):
    L = float(box[0])
    b = np.asarray(bonds)
    return np.concatenate([np.linalg.norm(_min_image(p[b[:, 0]] - p[b[:, 1]], L), axis=1) for p in frames])


# ------------------------------------------------------------------ tables

def write_table(filename, r, e, f):
    with open(filename, "w") as fh:
        for a, b, c in zip(r, e, f):
            fh.write("%15.8g %15.8g %15.8g\n" % (a, b, c))


def _smooth(y, k=2):
    if k <= 0:
        return y
    w = np.ones(2 * k + 1) / (2 * k + 1)
    yp = np.concatenate([np.full(k, y[0]), y, np.full(k, y[-1])])
    return np.convolve(yp, w, mode="valid")


def _repair(r, V, valid):
    """V on the grid r from its valid points: linear extrapolation (steep repulsion)
    below the first valid point, interpolation over gaps."""
    idx = np.nonzero(valid)[0]
    if len(idx) < 3:
        raise ValueError("too few sampled points to build a table")
    out = np.interp(r, r[idx], V[idx])
    i0 = idx[0]
    slope = min((V[idx[1]] - V[idx[0]]) / (r[idx[1]] - r[idx[0]]), -1.0)
    out[:i0] = V[i0] + slope * (r[:i0] - r[i0])
    return out


def table_from_potential(r, V, rc):
    """(r, e, f) rows: V shifted to zero at rc (and zero beyond), f = -dV/dr."""
    V = np.where(r <= rc, V - np.interp(rc, r, V), 0.0)
    f = -np.gradient(V, r)
    f[r > rc] = 0.0
    return r, V, f


def bond_table(frames, bonds, box, kT, rmin=0.5, rmax=4.5, nbins=160):
    """Boltzmann-inverted bond potential V(b) = -kT ln(P(b) / b^2) on a uniform grid."""
    b = bond_lengths(frames, bonds, box)
    edges = np.linspace(rmin, rmax, nbins + 1)
    h, _ = np.histogram(b, bins=edges, density=True)
    r = 0.5 * (edges[1:] + edges[:-1])
    valid = h > 0
    V = np.zeros_like(r)
    V[valid] = -kT * np.log(h[valid] / r[valid] ** 2)
    V = _repair(r, V, valid)
    # beyond the sampled range the bond pulls back with the last slope
    last = np.nonzero(valid)[0][-1]
    slope = max((V[last] - V[last - 3]) / (r[last] - r[last - 3]), 1.0)
    V[last + 1:] = V[last] + slope * (r[last + 1:] - r[last])
    V = _smooth(V, 1)
    V -= V.min()
    return r, V, -np.gradient(V, r)


def pmf_table(r, g, kT, rc):
    """Potential of mean force -kT ln g(r), the first iterate of IBI."""
    valid = g > 1e-3
    V = np.zeros_like(r)
    V[valid] = -kT * np.log(g[valid])
    V = _repair(r, V, valid)
    return table_from_potential(r, _smooth(V, 1), rc)


def ibi_update(r, V, g_model, g_target, kT, rc, alpha=0.5):
    """V + alpha kT ln(g_model / g_target) where both are sampled; unchanged elsewhere."""
    ok = (g_model > 1e-3) & (g_target > 1e-3)
    dV = np.zeros_like(V)
    dV[ok] = alpha * kT * np.log(g_model[ok] / g_target[ok])
    return table_from_potential(r, V + _smooth(dV, 2), rc)


# ------------------------------------------------------------------ systems

def _system(espressopp, box, cutoffs, skin, pos, seed, mass=None, precision="fp64"):
    from .decomp import cellGrid
    system = espressopp.System(precision=precision)
    system.rng = espressopp.esutil.RNG(seed)
    system.bc = espressopp.bc.OrthorhombicBC(system.rng, box)
    system.skin = skin
    cells = cellGrid(box, (1, 1, 1), max(cutoffs), skin)
    system.storage = espressopp.storage.DomainDecomposition(system, (1, 1, 1), cells)
    n = len(pos)
    system.storage.addParticlesArray(np.arange(n, dtype=np.int32), pos, None, None,
                                     None if mass is None else np.full(n, float(mass)))
    system.storage.decompose()
    return system


def random_walk_melt(nchains, nbeads, rho, seed, bond=0.97):
    """Freely jointed chains of fixed bond length with no immediate reversal, in a
    cubic box of density rho (positions unfolded along each chain)."""
    rng = np.random.default_rng(seed)
    n = nchains * nbeads
    L = (n / rho) ** (1.0 / 3.0)
    u = rng.normal(size=(nchains, nbeads, 3))
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    pos = np.zeros((nchains, nbeads, 3))
    pos[:, 0] = rng.uniform(0, L, (nchains, 3))
    for b in range(1, nbeads):
        step = u[:, b]
        if b >= 2:
            prev = pos[:, b - 1] - pos[:, b - 2]
            back = np.sum(step * prev, axis=1) < -0.8 * bond
            step[back] *= -1.0
        pos[:, b] = pos[:, b - 1] + bond * step
    a = (np.arange(nchains)[:, None] * nbeads + np.arange(nbeads - 1)[None, :]).ravel()
    return pos.reshape(-1, 3), (L, L, L), np.stack([a, a + 1], axis=1).astype(np.int32)


def melt_frames(espressopp, nchains, nbeads, nframes, stride, seed=1, rho=0.85, gamma=0.5, T=1.0, dt=0.01,
                pushoff_steps=2000, equil_steps=4000, tmpdir=None, on_frame=None, precision="fp64"):
    """A Kremer-Grest melt run by the plain method: random-walk chains, a soft
    push-off (tabulated cosine repulsion growing in strength, with FENE bonds),
    then WCA (rc 2^(1/6)) + FENE (K 30, R0 1.5) under Langevin (T, gamma) with time
    step dt.  Returns the box, the bonds and nframes frames taken every stride
    steps, each (unwrapped positions, conservative forces) in id order; with
    on_frame, each frame is handed to on_frame(pos, force) and its result kept
    instead."""
    import tempfile
    pos, box, bonds = random_walk_melt(nchains, nbeads, rho, seed)
    rc = 2.0 ** (1.0 / 6.0)
    skin = 0.4
    tmp = tmpdir or tempfile.mkdtemp()
    system = _system(espressopp, box, (rc,), skin, pos, seed, precision=precision)
    fpl = espressopp.FixedPairList(system.storage)
    fpl.addBonds(bonds.tolist())
    fene = espressopp.interaction.FixedPairListFENE(system, fpl, espressopp.interaction.FENE(30.0, 0.0, 1.5))
    vl = espressopp.VerletList(system, cutoff=rc)
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = dt
    lang = espressopp.integrator.LangevinThermostat(system)
    lang.gamma = gamma
    lang.temperature = T
    integ.addExtension(lang)
    # push-off: A (1 + cos(pi r / rc)) with A raised in stages
    rr = np.linspace(0.0, rc, 201)
    soft = espressopp.interaction.VerletListTabulated(vl)
    system.addInteraction(fene)
    system.addInteraction(soft)
    for stage, A in enumerate((2.0, 8.0, 25.0, 60.0)):
        fn = os.path.join(tmp, "pushoff_%d.tab" % stage)
        write_table(fn, rr, A * (1 + np.cos(np.pi * rr / rc)), A * np.pi / rc * np.sin(np.pi * rr / rc))
        soft.setPotential(0, 0, espressopp.interaction.Tabulated(1, fn, rc))
        integ.run(pushoff_steps // 4)
    system.removeInteraction(1)
    wca = espressopp.interaction.VerletListLennardJones(vl)
    wca.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, rc, "auto"))
    system.addInteraction(wca)
    integ.run(equil_steps)
    frames = []
    L = np.asarray(box)
    for _ in range(nframes):
        integ.run(stride)
        system.storage._h.call("espp_compute_forces")
        h = system.storage._h
        p, f = h.get("pos") + h.get("image") * L, h.get("f")
        frames.append(on_frame(p, f) if on_frame else (p, f))
    return box, bonds, frames


class CGModel:
    """A coarse-grained system (one site per blob, mass per_blob) with a bond table
    and, optionally, a pair table or a learned pair force; bonded sites are
    excluded from the non-bonded list."""

    def __init__(self, espressopp, box, bonds, pos, bond_rows, tmpdir, pair_rows=None, mlp=None, rc=3.0, mass=4.0,
                 seed=7, tag="cg"):
        self.espressopp = espressopp
        self.box = box
        skin = 0.3
        cut = max(rc, 1.0)
        self.system = _system(espressopp, box, (cut,), skin, pos, seed, mass)
        fpl = espressopp.FixedPairList(self.system.storage)
        fpl.addBonds(np.asarray(bonds).tolist())
        fn = os.path.join(tmpdir, tag + "_bond.tab")
        write_table(fn, *bond_rows)
        self.bond = espressopp.interaction.FixedPairListTabulated(
            self.system, fpl, espressopp.interaction.Tabulated(1, fn, float(bond_rows[0][-1])))
        self.system.addInteraction(self.bond)
        self.vl = espressopp.VerletList(self.system, cutoff=cut, exclusionlist=np.asarray(bonds).tolist())
        self.pair = None
        if pair_rows is not None:
            fn = os.path.join(tmpdir, tag + "_pair.tab")
            write_table(fn, *pair_rows)
            self.pair = espressopp.interaction.VerletListTabulated(self.vl)
            self.pair.setPotential(0, 0, espressopp.interaction.Tabulated(1, fn, rc))
            self.system.addInteraction(self.pair)
        if mlp is not None:
            spec, theta = mlp
            self.pair = espressopp.interaction.VerletListPairMLP(self.vl, **spec)
            self.pair.setParameters(theta)
            self.system.addInteraction(self.pair)

    def set_positions(self, pos):
        s = self.system.storage
        s.setParticleArrays("pos", np.arange(len(pos), dtype=np.int32), pos)
        s.decompose()

    def forces(self, pos):
        self.set_positions(pos)
        self.system.storage._h.call("espp_compute_forces")
        return self.system.storage.getParticleArrays()["f"]

    def sample(self, pos, steps, every, dt=0.01, T=1.0, gamma=0.5, equil=1000):
        """Langevin CG MD from pos; positions every `every` steps after equil steps."""
        esp = self.espressopp
        self.set_positions(pos)
        integ = esp.integrator.VelocityVerlet(self.system)
        integ.dt = dt
        lang = esp.integrator.LangevinThermostat(self.system)
        lang.gamma = gamma
        lang.temperature = T
        integ.addExtension(lang)
        integ.run(equil)
        out = []
        L = np.asarray(self.box)
        for _ in range(steps // every):
            integ.run(every)
            a = self.system.storage.getParticleArrays()
            out.append(np.mod(a["pos"], L))
        lang.disconnect()
        return out


def ibi(espressopp, box, bonds, start, bond_rows, r, g_target, tmpdir, kT=1.0, rc=3.0, iterations=5,
        steps=3000, every=100, alpha=0.5):
    """Iterative Boltzmann inversion of the non-bonded pair table; returns the final
    (r, e, f) rows and the RDF error of every iterate."""
    rows = pmf_table(r, g_target, kT, rc)
    history = []
    for it in range(iterations):
        m = CGModel(espressopp, box, bonds, start, bond_rows, tmpdir, pair_rows=rows, rc=rc, tag="ibi%d" % it)
        frames = m.sample(start, steps, every)
        _, g = pair_rdf(frames, box, r[-1] + 0.5 * (r[1] - r[0]), len(r), exclude=bonds)
        history.append(rdf_error(g, g_target))
        rows = ibi_update(r, rows[1], g, g_target, kT, rc, alpha)
    return rows, history


def train_pair_mlp(espressopp, box, bonds, frames, residual, spec, theta0, epochs=40, lr=2e-3, seed=3,
                   tmpdir=None, bond_rows=None, batch_frames=None):
    """Force matching of the learned pair force on (positions, residual force)
    frames: each epoch visits the frames in a shuffled order and takes one Adam
    step per batch_frames frames (all frames when None) on their mean loss.
    Returns the parameters and the mean loss of every epoch."""
    theta = np.asarray(theta0, dtype=np.float32)
    from .decomp import cellGrid
    system = espressopp.System(precision="fp64")
    system.rng = espressopp.esutil.RNG(seed)
    system.bc = espressopp.bc.OrthorhombicBC(system.rng, box)
    system.skin = 0.3
    cut = spec.get("cutoff", 3.0)
    system.storage = espressopp.storage.DomainDecomposition(system, (1, 1, 1),
                                                            cellGrid(box, (1, 1, 1), cut, 0.3))
    n = len(frames[0])
    system.storage.addParticlesArray(np.arange(n, dtype=np.int32), frames[0])
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=cut, exclusionlist=np.asarray(bonds).tolist())
    mlp = espressopp.interaction.VerletListPairMLP(vl, **spec)
    mlp.setParameters(theta)
    system.addInteraction(mlp)
    ids = np.arange(n, dtype=np.int32)
    nf = len(frames)
    bsz = nf if not batch_frames else max(1, min(int(batch_frames), nf))
    rng = np.random.default_rng(seed)
    history = []
    for ep in range(epochs):
        order = rng.permutation(nf) if bsz < nf else np.arange(nf)
        tot = 0.0
        for b0 in range(0, nf, bsz):
            chunk = order[b0:b0 + bsz]
            mlp.zeroGradient()
            for k in chunk:
                system.storage.setParticleArrays("pos", ids, frames[k])
                system.storage.decompose()
                tot += mlp.trainFrame(residual[k], weight=1.0 / len(chunk)) * len(chunk) / nf
            mlp.adamStep(lr=lr)
        history.append(tot)
    return mlp.getParameters(), history


def force_error(f_model, f_ref):
    """|F_model - F_ref| / |F_ref| over all sites and frames (RMS ratio)."""
    a = np.concatenate([np.asarray(x).ravel() for x in f_model])
    b = np.concatenate([np.asarray(x).ravel() for x in f_ref])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rdf_error(g, g_target):
    """RMS difference of two RDFs over the bins."""
    return float(np.sqrt(np.mean((np.asarray(g) - np.asarray(g_target)) ** 2)))


def density_norm(frames, box, bonds, radii):
    """Mean and standard deviation of the two local densities over frames (the
    learned force's input normalisation), from the non-bonded neighbours."""
    from scipy.spatial import cKDTree
    L = float(box[0])
    n = len(frames[0])
    b = np.sort(np.asarray(bonds, dtype=np.int64), axis=1)
    ex = np.unique(b[:, 0] * n + b[:, 1])
    vals = [[], []]
    for p in frames[: max(1, min(len(frames), 5))]:
        t = cKDTree(np.mod(p, L), boxsize=L)
        pr = t.query_pairs(max(radii), output_type="ndarray")
        key = np.minimum(pr[:, 0], pr[:, 1]).astype(np.int64) * n + np.maximum(pr[:, 0], pr[:, 1])
        pr = pr[~np.isin(key, ex)]
        d = np.linalg.norm(_min_image(p[pr[:, 0]] - p[pr[:, 1]], L), axis=1)
        for a, R in enumerate(radii):
            w = np.where(d < R, 0.5 * (1 + np.cos(np.pi * d / R)), 0.0)
            rho = np.bincount(pr[:, 0], weights=w, minlength=n) + np.bincount(pr[:, 1], weights=w, minlength=n)
            vals[a].append(rho)
    return tuple((float(np.mean(v)), float(max(np.std(v), 1e-3))) for v in (np.concatenate(vals[0]),
                                                                               np.concatenate(vals[1])))


def main(argv=None):
    """The coarse-grained workload end to end: melt frames by the plain method,
    4 -> 1 mapping, Boltzmann-inverted bonds, IBI, force matching of the learned
    pair force, and the judge; prints progress and one JSON line of results."""
    import argparse
    import json
    import tempfile
    import time

    import espressopp_amd as espressopp
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=5000)
    ap.add_argument("--beads", type=int, default=200)
    ap.add_argument("--per-blob", type=int, default=4)
    ap.add_argument("--frames", type=int, default=500)
    ap.add_argument("--stride", type=int, default=100)
    ap.add_argument("--test-frames", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--batch-frames", type=int, default=10)
    ap.add_argument("--ibi-iterations", type=int, default=4)
    ap.add_argument("--ibi-steps", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, flush=True)  # noqa: E731
    tmp = tempfile.mkdtemp()
    t0 = time.perf_counter()
    nch, nb, m = args.chains, args.beads, args.per_blob
    total = args.frames + args.test_frames

    mapped = []

    def on_frame(p, f):
        mapped.append(None)
        if len(mapped) % 100 == 0:
            log("melt: frame %d of %d (%.1fs)" % (len(mapped), total, time.perf_counter() - t0))
        return blob_map(p, f, nch, nb, m, _box_of(nch * nb))

    box, bonds, raw = melt_frames(espressopp, nch, nb, total, args.stride, seed=args.seed, tmpdir=tmp,
                                  precision="mixed", on_frame=on_frame)
    pos = [x[0] for x in raw]
    F = [x[1] for x in raw]
    del raw
    cgb = cg_bonds(nch, nb, m)
    train, test = list(range(args.frames)), list(range(args.frames, total))
    log("melt: %d beads, %d frames mapped to %d sites (%.1fs)" % (nch * nb, total, len(pos[0]),
                                                                  time.perf_counter() - t0))
    sample = [pos[k] for k in train[:: max(1, len(train) // 25)]]
    r, g_t = pair_rdf(sample, box, 3.0, 60, exclude=cgb)
    bond_rows = bond_table(sample, cgb, box, kT=1.0)
    t1 = time.perf_counter()
    ibi_rows, ibi_hist = ibi(espressopp, box, cgb, pos[train[-1]], bond_rows, r, g_t, tmp,
                             iterations=args.ibi_iterations, steps=args.ibi_steps, every=100)
    log("IBI: RDF error by iteration %s (%.1fs)" % (["%.4f" % e for e in ibi_hist], time.perf_counter() - t1))
    bonds_only = CGModel(espressopp, box, cgb, pos[0], bond_rows, tmp, tag="bonds")
    resid = [F[k] - bonds_only.forces(pos[k]) for k in range(total)]
    radii = (1.5, 3.0)
    spec = dict(cutoff=3.0, densityRadii=radii, densityNorm=density_norm(pos, box, cgb, radii))
    theta0 = espressopp.interaction.VerletListPairMLP.initialParameters(seed=1, scale=0.5)
    t2 = time.perf_counter()
    theta, hist = train_pair_mlp(espressopp, box, cgb, [pos[k] for k in train], [resid[k] for k in train], spec,
                                 theta0, epochs=args.epochs, lr=args.lr, batch_frames=args.batch_frames)
    t_train = time.perf_counter() - t2
    log("training: %d epochs x %d frames (Adam step per %d frames), loss %.4f -> %.4f (%.1fs)"
        % (args.epochs, len(train), args.batch_frames, hist[0], hist[-1], t_train))
    fb = [F[k] - resid[k] for k in test]
    ibi_model = CGModel(espressopp, box, cgb, pos[0], bond_rows, tmp, pair_rows=ibi_rows, tag="ibi_eval")
    mlp_model = CGModel(espressopp, box, cgb, pos[0], bond_rows, tmp, mlp=(spec, theta), tag="mlp_eval")
    e_bond = force_error(fb, [F[k] for k in test])
    e_ibi = force_error([ibi_model.forces(pos[k]) for k in test], [F[k] for k in test])
    e_mlp = force_error([mlp_model.forces(pos[k]) for k in test], [F[k] for k in test])
    _, g_ibi = pair_rdf(ibi_model.sample(pos[test[0]], args.ibi_steps, 100), box, 3.0, 60, exclude=cgb)
    _, g_mlp = pair_rdf(mlp_model.sample(pos[test[0]], args.ibi_steps, 100), box, 3.0, 60, exclude=cgb)
    out = {"beads": nch * nb, "sites": len(pos[0]), "train_frames": len(train), "test_frames": len(test),
           "epochs": args.epochs, "batch_frames": args.batch_frames, "lr": args.lr, "train_s": t_train, "train_s_per_epoch": t_train / args.epochs,
           "loss_first": hist[0], "loss_last": hist[-1],
           "force_error": {"bonds_only": e_bond, "ibi": e_ibi, "learned": e_mlp},
           "explained_force_variance": {"bonds_only": 1 - e_bond ** 2, "ibi": 1 - e_ibi ** 2, "learned": 1 - e_mlp ** 2},
           "rdf_error": {"ibi": rdf_error(g_ibi, g_t), "learned": rdf_error(g_mlp, g_t)},
           "ibi_rdf_error_by_iteration": ibi_hist, "wall_s": time.perf_counter() - t0}
    log(json.dumps(out))
    return out


def _box_of(n, rho=0.85):
    L = (n / rho) ** (1.0 / 3.0)
    return (L, L, L)


if __name__ == "__main__":
    main()
