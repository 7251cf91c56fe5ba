"""Learned pair force on the GPU (BF16 MFMA kernels) against the oracle.

Forces and the force-matching gradient of the HIP kernels are compared with
oracle/mlp_oracle.py twice: against its bfloat16 emulation, which rounds at the
same points as the kernels (differences come only from FP32 vs FP64 sums and
rare one-ulp bf16 rounding flips), and against plain FP64.
"""
import numpy as np
import pytest

import espressopp_amd as espressopp
from oracle import mlp_oracle as M
from tests._systems import gpu_system

pytestmark = pytest.mark.gpu

RC, RADII = 3.0, (1.5, 3.0)
NORM = ((1.2, 0.8), (12.0, 3.5))


def cg_cloud(n, rho, seed, min_dist=0.5):
    """n sites at number density rho, no two closer than min_dist (rejection on a grid)."""
    L = (n / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    nc = max(int(L / min_dist), 1)
    grid = {}
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, L, 3)
        c = tuple((p // (L / nc)).astype(int) % nc)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                        d = p - pts[q]
                        d -= L * np.rint(d / L)
                        if d @ d < min_dist ** 2:
                            ok = False
        if ok:
            grid.setdefault(c, []).append(len(pts))
            pts.append(p)
    return np.array(pts), L


def mlp_system(pos, L, theta, exclusions=()):
    box = (L, L, L)
    skin = 0.3
    cells = [max(int(L / (RC + skin)), 3)] * 3
    system = gpu_system(box, cells, skin, "fp64")
    system.storage.addParticlesArray(np.arange(len(pos)), pos)
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=RC, exclusionlist=list(exclusions))
    ia = espressopp.interaction.VerletListPairMLP(vl, cutoff=RC, densityRadii=RADII, densityNorm=NORM)
    ia.setParameters(theta)
    system.addInteraction(ia)
    return system, vl, ia


@pytest.mark.parametrize("n,seed", [(300, 1), (2000, 2)])
def test_mlp_forces_match_oracle(gpu, n, seed):
    pos, L = cg_cloud(n, 0.2125, seed)
    theta = M.init_params(100 + seed, scale=0.8).astype(np.float32).astype(np.float64)
    system, vl, ia = mlp_system(pos, L, theta)
    system.storage._h.call("espp_compute_forces")
    f = system.storage.getParticleArrays()["f"]
    f_bf, _ = M.forces(theta, pos, (L, L, L), rc=RC, radii=RADII, norm=NORM, bf=True)
    f64, _ = M.forces(theta, pos, (L, L, L), rc=RC, radii=RADII, norm=NORM)
    rms = np.sqrt(np.mean(f64 ** 2))
    e_bf = np.linalg.norm(f - f_bf) / np.linalg.norm(f_bf)
    e_64 = np.linalg.norm(f - f64) / np.linalg.norm(f64)
    print("n=%d: |F-F_bf16|/|F| %.2e, max %.2e RMS; |F-F_fp64|/|F| %.2e" %
          (n, e_bf, np.abs(f - f_bf).max() / rms, e_64), flush=True)
    assert e_bf < 5e-3
    assert np.abs(f - f_bf).max() < 3e-2 * rms
    assert e_64 < 3e-2
    # Newton's third law holds to FP32 summation
    assert np.abs(f.sum(axis=0)).max() < 1e-4 * rms * n


def test_mlp_gradient_matches_oracle(gpu):
    pos, L = cg_cloud(1500, 0.2125, 7)
    theta = M.init_params(107, scale=0.8).astype(np.float32).astype(np.float64)
    f_ref = np.random.default_rng(8).normal(0, 3.0, (len(pos), 3))
    system, vl, ia = mlp_system(pos, L, theta)
    ia.zeroGradient()
    loss = ia.trainFrame(f_ref, weight=1.0)
    g = ia.gradient().astype(np.float64)
    lb, gb, _, _ = M.loss_and_grad(theta, pos, (L, L, L), f_ref, rc=RC, radii=RADII, norm=NORM, bf=True)
    l64, g64, _, _ = M.loss_and_grad(theta, pos, (L, L, L), f_ref, rc=RC, radii=RADII, norm=NORM)
    print("loss gpu %.6f bf16-oracle %.6f fp64-oracle %.6f" % (loss, lb, l64), flush=True)
    assert loss == pytest.approx(lb, rel=2e-3)
    assert loss == pytest.approx(l64, rel=2e-2)
    o = 0
    for name, size in (("W1", 34 * 64), ("b1", 64), ("W2", 64 * 64), ("b2", 64), ("W3", 64 * 64), ("b3", 64),
                       ("w4", 64), ("b4", 1)):
        a, b, c = g[o:o + size], gb[o:o + size], g64[o:o + size]
        eb = np.linalg.norm(a - b) / np.linalg.norm(b)
        e6 = np.linalg.norm(a - c) / np.linalg.norm(c)
        cos = a @ c / (np.linalg.norm(a) * np.linalg.norm(c))
        print("%-3s |g-g_bf16|/|g| %.2e  |g-g_fp64|/|g| %.2e  cos %.6f" % (name, eb, e6, cos), flush=True)
        assert eb < 2e-2, name
        assert e6 < 6e-2 and cos > 0.998, name
        o += size
    # a second frame adds to the accumulated gradient
    ia.trainFrame(f_ref, weight=1.0)
    assert np.allclose(ia.gradient(), 2 * g, rtol=1e-4, atol=1e-6 * np.abs(g).max())


def test_mlp_training_reduces_the_loss(gpu):
    # a target force field of the same family: the oracle's forces for other parameters
    pos, L = cg_cloud(1200, 0.2125, 9)
    target = M.init_params(201, scale=0.8)
    f_ref, _ = M.forces(target, pos, (L, L, L), rc=RC, radii=RADII, norm=NORM)
    theta = M.init_params(202, scale=0.8)
    system, vl, ia = mlp_system(pos, L, theta)
    losses = []
    for it in range(150):
        ia.zeroGradient()
        losses.append(ia.trainFrame(f_ref))
        ia.adamStep(lr=3e-3)
    print("loss: %.4f -> %.4f" % (losses[0], losses[-1]), flush=True)
    assert losses[-1] < 0.2 * losses[0]
    # the parameters the GPU learned give the oracle the same lower loss
    th = ia.getParameters().astype(np.float64)
    l64 = M.loss_and_grad(th, pos, (L, L, L), f_ref, rc=RC, radii=RADII, norm=NORM)[0]
    assert l64 < 0.3 * losses[0]


def test_mlp_runs_in_md_with_exclusions(gpu):
    pos, L = cg_cloud(800, 0.2125, 11)
    theta = M.init_params(111, scale=0.3).astype(np.float32).astype(np.float64)
    excl = [(k, k + 1) for k in range(0, 799, 2)]
    system, vl, ia = mlp_system(pos, L, theta, exclusions=excl)
    system.storage._h.call("espp_compute_forces")
    f = system.storage.getParticleArrays()["f"]
    f_bf, _ = M.forces(theta, pos, (L, L, L), rc=RC, radii=RADII, norm=NORM, exclude=excl, bf=True)
    assert np.linalg.norm(f - f_bf) / np.linalg.norm(f_bf) < 5e-3
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    lang = espressopp.integrator.LangevinThermostat(system)
    lang.gamma = 1.0
    lang.temperature = 1.0
    integ.addExtension(lang)
    integ.run(200)
    a = system.storage.getParticleArrays()
    assert np.isfinite(a["pos"]).all() and np.isfinite(a["v"]).all()
    with pytest.raises(RuntimeError):
        ia.computeEnergy()
