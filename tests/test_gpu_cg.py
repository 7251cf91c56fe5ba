"""The coarse-grained workload end to end on the GPU, at a reduced size.

A Kremer-Grest melt (40 chains x 200 beads, density 0.85, WCA + FENE, Langevin
T = 1, gamma = 0.5, dt = 0.01) is run by the plain method; 48 frames of its
conservative forces are mapped 4 beads -> 1 site.  The plain CG model is a
Boltzmann-inverted bond table plus an IBI pair table (interaction.Tabulated,
src/interaction/Tabulated.hpp:43-115); the learned model keeps the bond table and
replaces the pair table by the density-dependent pair MLP, force-matched on the
GPU by the BF16 MFMA kernels on 36 frames.  On the 12 held-out frames the test
reports both models' force error against the mapped forces and the RDF error
of CG runs of each against the mapped melt, and requires the learned model to
reduce the force error.
"""
import numpy as np
import pytest

import espressopp_amd as espressopp
from espressopp_amd.tools import coarse_grain as cg

pytestmark = pytest.mark.gpu


def test_learned_cg_force_matching_against_ibi(gpu, tmp_path):
    nchains, nbeads, m = 40, 200, 4
    tmp = str(tmp_path)
    box, bonds, frames = cg.melt_frames(espressopp, nchains, nbeads, nframes=48, stride=100, seed=5, tmpdir=tmp)
    mapped = [cg.blob_map(p, f, nchains, nbeads, m, box) for p, f in frames]
    pos = [c[0] for c in mapped]
    F = [c[1] for c in mapped]
    cgb = cg.cg_bonds(nchains, nbeads, m)
    train, test = list(range(36)), list(range(36, 48))
    print("melt: %d beads, box %.3f; %d sites, %d CG bonds" % (nchains * nbeads, box[0], len(pos[0]), len(cgb)),
          flush=True)
    r, g_t = cg.pair_rdf([pos[k] for k in train], box, 3.0, 60, exclude=cgb)
    bond_rows = cg.bond_table([pos[k] for k in train], cgb, box, kT=1.0)
    # the plain method: IBI pair table
    ibi_rows, ibi_hist = cg.ibi(espressopp, box, cgb, pos[train[-1]], bond_rows, r, g_t, tmp, iterations=4,
                                steps=2000, every=100)
    print("IBI RDF error by iteration:", ["%.4f" % e for e in ibi_hist], flush=True)
    # residual targets for the learned pair force: mapped force minus the bond table's force
    bonds_only = cg.CGModel(espressopp, box, cgb, pos[0], bond_rows, tmp, tag="bonds")
    fb = [bonds_only.forces(p) for p in pos]
    resid = [F[k] - fb[k] for k in range(len(pos))]
    radii = (1.5, 3.0)
    spec = dict(cutoff=3.0, densityRadii=radii, densityNorm=cg.density_norm(pos, box, cgb, radii))
    theta0 = espressopp.interaction.VerletListPairMLP.initialParameters(seed=1, scale=0.5)
    theta, hist = cg.train_pair_mlp(espressopp, box, cgb, [pos[k] for k in train], [resid[k] for k in train], spec,
                                    theta0, epochs=60, lr=3e-3, batch_frames=6)
    print("training loss: %.4f -> %.4f" % (hist[0], hist[-1]), flush=True)
    assert hist[-1] < hist[0]
    ibi_model = cg.CGModel(espressopp, box, cgb, pos[0], bond_rows, tmp, pair_rows=ibi_rows, tag="ibi_eval")
    mlp_model = cg.CGModel(espressopp, box, cgb, pos[0], bond_rows, tmp, mlp=(spec, theta), tag="mlp_eval")
    e_bond = cg.force_error([fb[k] for k in test], [F[k] for k in test])
    e_ibi = cg.force_error([ibi_model.forces(pos[k]) for k in test], [F[k] for k in test])
    e_mlp = cg.force_error([mlp_model.forces(pos[k]) for k in test], [F[k] for k in test])
    _, g_ibi = cg.pair_rdf(ibi_model.sample(pos[test[0]], 2000, 100), box, 3.0, 60, exclude=cgb)
    _, g_mlp = cg.pair_rdf(mlp_model.sample(pos[test[0]], 2000, 100), box, 3.0, 60, exclude=cgb)
    r_ibi, r_mlp = cg.rdf_error(g_ibi, g_t), cg.rdf_error(g_mlp, g_t)
    # the mapped force is mostly thermal noise of the beads: what a model can explain is
    # the fraction 1 - e^2 of the force variance
    x_bond, x_ibi, x_mlp = 1 - e_bond ** 2, 1 - e_ibi ** 2, 1 - e_mlp ** 2
    print("held-out force error |F_model - F_mapped| / |F_mapped|: bonds only %.4f, IBI %.4f, learned %.4f" %
          (e_bond, e_ibi, e_mlp), flush=True)
    print("explained force variance 1 - e^2: bonds only %.4f, IBI %.4f, learned %.4f" % (x_bond, x_ibi, x_mlp),
          flush=True)
    print("RDF error (RMS over 60 bins to r = 3): IBI %.4f, learned %.4f" % (r_ibi, r_mlp), flush=True)
    assert e_mlp < e_ibi < e_bond
    assert x_mlp > 1.1 * x_ibi
    assert r_ibi < 0.1 and r_mlp < 0.3
