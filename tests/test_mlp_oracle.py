"""Pins of the learned pair force's FP64 oracle (oracle/mlp_oracle.py).

The model has no reference counterpart, so the oracle is pinned by what the
mathematics fixes: finite differences of the loss against the analytic
gradient in every parameter block, Newton's third law, invariance under
translation and relabelling, equivariance under rotation, and the closed form
of a network whose weight matrices are zero.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M

RC, RADII = 3.0, (1.5, 3.0)
NORM = ((2.0, 1.0), (8.0, 3.0))


def cloud(n, L, seed, min_dist=0.6):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, L, 3)
        ok = True
        for q in pts:
            d = p - q
            d -= L * np.rint(d / L)
            if d @ d < min_dist ** 2:
                ok = False
                break
        if ok:
            pts.append(p)
    return np.array(pts)


def kw():
    return dict(rc=RC, radii=RADII, norm=NORM)


def test_parameter_count_and_packing():
    assert M.N_PARAMS == 10625
    th = np.arange(M.N_PARAMS, dtype=np.float64)
    p = M.unpack(th)
    assert p["W1"].shape == (34, 64) and p["W1"][0, 1] == 1 and p["b1"][0] == 34 * 64
    assert p["b4"][0] == M.N_PARAMS - 1
    assert np.array_equal(M.pack(p), th)


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between bf16 1 and 1 + 2^-7: ties go to the even mantissa (1)
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159], dtype=np.float64)
    b = M.bf16(x)
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == 1.0 + 2 ** -6
    assert b[3] == -2.5
    assert abs(b[4] - 3.140625) < 1e-12


@pytest.mark.parametrize("block", ["W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4"])
def test_gradient_matches_finite_differences(block):
    L = 7.0
    pos = cloud(40, L, seed=3)
    box = (L, L, L)
    theta = M.init_params(11, scale=0.7)
    f_ref = np.random.default_rng(4).normal(0, 2.0, (len(pos), 3))
    loss, grad, _, _ = M.loss_and_grad(theta, pos, box, f_ref, weight=0.5, **kw())
    offsets = {}
    o = 0
    for name, size in (("W1", 34 * 64), ("b1", 64), ("W2", 64 * 64), ("b2", 64), ("W3", 64 * 64), ("b3", 64),
                       ("w4", 64), ("b4", 1)):
        offsets[name] = (o, size)
        o += size
    start, size = offsets[block]
    rng = np.random.default_rng(sum(map(ord, block)))
    idx = start + rng.choice(size, size=min(6, size), replace=False)
    h = 1e-6
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        lp = M.loss_and_grad(tp, pos, box, f_ref, weight=0.5, **kw())[0]
        lm = M.loss_and_grad(tm, pos, box, f_ref, weight=0.5, **kw())[0]
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(grad[k], rel=1e-5, abs=1e-9 * max(1.0, abs(loss))), (block, k)


def test_newton_third_law_and_invariances():
    L = 8.0
    pos = cloud(60, L, seed=5)
    box = (L, L, L)
    theta = M.init_params(12)
    f, po = M.forces(theta, pos, box, **kw())
    assert np.abs(f.sum(axis=0)).max() < 1e-10 * np.abs(f).max()
    # every ordered pair appears with its reverse and equal g
    key = {(i, j): g for i, j, g in zip(po["I"], po["J"], po["g"])}
    assert all(abs(key[(j, i)] - g) <= 1e-12 * max(1.0, abs(g)) for (i, j), g in key.items())
    # translation (with periodic wrap) and relabelling
    f2, _ = M.forces(theta, (pos + np.array([1.7, -3.2, 0.4])) % L, box, **kw())
    assert np.abs(f2 - f).max() < 1e-9 * np.abs(f).max()
    perm = np.random.default_rng(6).permutation(len(pos))
    f3, _ = M.forces(theta, pos[perm], box, **kw())
    assert np.abs(f3 - f[perm]).max() < 1e-10 * np.abs(f).max()


def test_rotation_equivariance():
    # an isolated cluster in a large box: periodic images never interact
    L = 60.0
    pos = cloud(25, 6.0, seed=7) + 27.0
    box = (L, L, L)
    theta = M.init_params(13)
    f, _ = M.forces(theta, pos, box, **kw())
    a, b, c = 0.3, -1.1, 0.7
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rx = np.array([[1, 0, 0], [0, np.cos(c), -np.sin(c)], [0, np.sin(c), np.cos(c)]])
    R = Rz @ Ry @ Rx
    centre = pos.mean(axis=0)
    pr = (pos - centre) @ R.T + centre
    fr, _ = M.forces(theta, pr, box, **kw())
    assert np.abs(fr - f @ R.T).max() < 1e-9 * np.abs(f).max()


def test_zero_weight_network_closed_form():
    rng = np.random.default_rng(8)
    p = M.unpack(np.zeros(M.N_PARAMS))
    b1, b2, b3, w4 = (rng.normal(size=64) for _ in range(4))
    b4 = 0.37
    p = {"W1": np.zeros((34, 64)), "b1": b1, "W2": np.zeros((64, 64)), "b2": b2, "W3": np.zeros((64, 64)),
         "b3": b3, "w4": w4, "b4": np.array([b4])}
    theta = M.pack(p)
    # with zero matrices each layer sees only its bias: h3 = silu(b3)
    s = np.sum(b3 / (1.0 + np.exp(-b3)) * w4) + b4
    L = 20.0
    x0, x1 = np.array([10.0, 10.0, 10.0]), np.array([11.3, 9.5, 10.8])
    f, _ = M.forces(theta, np.stack([x0, x1]), (L, L, L), **kw())
    d = x0 - x1
    r = np.linalg.norm(d)
    expected = s * 0.5 * (1 + np.cos(np.pi * r / RC)) * d / r
    assert np.allclose(f[0], expected, rtol=1e-12)
    assert np.allclose(f[1], -expected, rtol=1e-12)
    # beyond the cutoff the pair does not interact
    f, _ = M.forces(theta, np.stack([x0, x0 + np.array([3.05, 0, 0])]), (L, L, L), **kw())
    assert np.abs(f).max() == 0.0


def test_bf16_emulation_stays_close_to_fp64():
    L = 8.0
    pos = cloud(60, L, seed=9)
    theta = M.init_params(14)
    f64, _ = M.forces(theta, pos, (L, L, L), **kw())
    fb, _ = M.forces(theta, pos, (L, L, L), bf=True, **kw())
    rel = np.linalg.norm(fb - f64) / np.linalg.norm(f64)
    assert 1e-5 < rel < 3e-2
