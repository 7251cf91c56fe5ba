"""NumPy restatement of the learned coarse-grained pair force (FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and bench.py's checks); the
product package never imports it.

The model has no counterpart in the reference (ESPResSo++ has no learned
potentials); it is the density-dependent pair MLP of the coarse-grained
workload, defined here once so that the GPU kernels can be checked against it:

  pair (i, j), d = r_i - r_j (minimum image), r = |d| < rc
  x[0:32]  = exp(-(r - mu_k)^2 / (2 s^2)),  mu_k = k rc / 31,  s = rc / 31
  x[32+a]  = ((rho_i^a + rho_j^a) / 2 - mean_a) / scale_a,   a = 0, 1
             rho_i^a = sum over listed j != i of w(r_ij; R_a),
             w(r; R) = (1 + cos(pi r / R)) / 2 for r < R, else 0
  h1 = silu(x W1 + b1), h2 = silu(h1 W2 + b2), h3 = silu(h2 W3 + b3)   (64 each)
  s  = h3 . w4 + b4,    g = fc(r) s,    fc(r) = (1 + cos(pi r / rc)) / 2
  F_i = sum_j g_ij d_ij / r_ij                      (Newton's third law holds: g_ij = g_ji)
  loss = sum_i |F_i - F_i^ref|^2 / (3 N)

Parameters are one FP32/FP64 vector: W1 (34 x 64, [in][out]), b1, W2 (64 x 64),
b2, W3, b3, w4 (64), b4 (1): 10625 values.

With bf16=True the forward and backward round to bfloat16 (round to nearest
even) exactly where the GPU kernels do: the 34 inputs, the weight matrices
W1..W3, the hidden activations h1..h3, silu'(z) of each layer and the
back-propagated dz of each layer; everything else (sums, biases, w4, b4, the
force and the loss) stays FP64 here, FP32 on the GPU.

Pinned in tests/test_mlp_oracle.py by: central finite differences of the loss
against the analytic gradient (every parameter block), Newton's third law,
invariance under translation and relabelling of sites, equivariance under a
rotation of the frame, and the closed form of a network whose hidden weights
are zero (then s = silu(silu(silu(b1) W2 + b2) W3 + b3) . w4 + b4 is the same for
every pair and the force of two sites is s fc(r) d / r).
"""
import numpy as np

N_RBF = 32
N_IN = N_RBF + 2
N_HID = 64
N_PARAMS = N_IN * N_HID + N_HID + 2 * (N_HID * N_HID + N_HID) + N_HID + 1


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    assert t.size == N_PARAMS
    o = 0
    out = {}
    for name, shape in (("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
                        ("W3", (N_HID, N_HID)), ("b3", (N_HID,)), ("w4", (N_HID,)), ("b4", (1,))):
        k = int(np.prod(shape))
        out[name] = t[o:o + k].reshape(shape)
        o += k
    return out


def pack(p):
    return np.concatenate([np.asarray(p[k], dtype=np.float64).ravel()
                           for k in ("W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4")])


def init_params(seed, scale=1.0):
    """He-style random initialisation (FP64)."""
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, scale * np.sqrt(2.0 / N_IN), (N_IN, N_HID)),
         "b1": np.zeros(N_HID),
         "W2": rng.normal(0, scale * np.sqrt(2.0 / N_HID), (N_HID, N_HID)),
         "b2": np.zeros(N_HID),
         "W3": rng.normal(0, scale * np.sqrt(2.0 / N_HID), (N_HID, N_HID)),
         "b3": np.zeros(N_HID),
         "w4": rng.normal(0, scale * np.sqrt(1.0 / N_HID), N_HID),
         "b4": np.zeros(1)}
    return pack(p)


def bf16(x):
    """Round FP32 values to bfloat16 (nearest, ties to even), returned as FP64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def pairs(pos, box, rc, exclude=None):
    """All ordered pairs (i, j), i != j, with minimum-image distance < rc (O(N^2))."""
    pos = np.asarray(pos, dtype=np.float64)
    L = np.asarray(box, dtype=np.float64)
    n = len(pos)
    I, J, D = [], [], []
    for i in range(n):
        d = pos[i] - pos
        d -= L * np.rint(d / L)
        r2 = np.sum(d * d, axis=1)
        js = np.nonzero(r2 < rc * rc)[0]
        js = js[js != i]
        I.append(np.full(len(js), i))
        J.append(js)
        D.append(d[js])
    I, J, D = np.concatenate(I), np.concatenate(J), np.concatenate(D)
    if exclude is not None and len(exclude):
        ex = {(int(a), int(b)) for a, b in exclude} | {(int(b), int(a)) for a, b in exclude}
        keep = np.array([(int(a), int(b)) not in ex for a, b in zip(I, J)], dtype=bool)
        I, J, D = I[keep], J[keep], D[keep]
    return I, J, D


def cosine_cut(r, R):
    return np.where(r < R, 0.5 * (1.0 + np.cos(np.pi * r / R)), 0.0)


def densities(n, I, r, radii):
    return np.stack([np.bincount(I, weights=cosine_cut(r, R), minlength=n) for R in radii], axis=1)


def features(I, J, r, rho, rc, norm):
    """34 inputs per pair: 32 Gaussian RBFs and the two normalised pair densities."""
    mu = np.arange(N_RBF) * rc / (N_RBF - 1)
    s = rc / (N_RBF - 1)
    x = np.empty((len(r), N_IN))
    x[:, :N_RBF] = np.exp(-((r[:, None] - mu[None, :]) ** 2) / (2.0 * s * s))
    for a in range(2):
        x[:, N_RBF + a] = (0.5 * (rho[I, a] + rho[J, a]) - norm[a][0]) / norm[a][1]
    return x


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    sg = 1.0 / (1.0 + np.exp(-z))
    return sg * (1.0 + z * (1.0 - sg))


def _dense(a, W, b):
    return a @ W + b


def forward(theta, x, bf=False):
    """Per-pair output s and the saved activations."""
    p = unpack(theta)
    rnd = bf16 if bf else (lambda v: v)
    W1, W2, W3 = (rnd(p[k]) for k in ("W1", "W2", "W3"))
    a0 = rnd(x)
    z1 = _dense(a0, W1, p["b1"])
    h1 = rnd(silu(z1))
    z2 = _dense(h1, W2, p["b2"])
    h2 = rnd(silu(z2))
    z3 = _dense(h2, W3, p["b3"])
    h3 = rnd(silu(z3))
    s = h3 @ p["w4"] + p["b4"][0]
    return s, dict(a0=a0, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, W1=W1, W2=W2, W3=W3, p=p)


def pair_outputs(theta, pos, box, rc=3.0, radii=(1.5, 3.0), norm=((0.0, 1.0), (0.0, 1.0)), exclude=None,
                 bf=False):
    I, J, D = pairs(pos, box, rc, exclude)
    r = np.sqrt(np.sum(D * D, axis=1))
    rho = densities(len(pos), I, r, radii)
    x = features(I, J, r, rho, rc, norm)
    s, cache = forward(theta, x, bf)
    fc = cosine_cut(r, rc)
    return dict(I=I, J=J, D=D, r=r, rho=rho, x=x, s=s, g=fc * s, fc=fc, cache=cache)


def forces(theta, pos, box, **kw):
    po = pair_outputs(theta, pos, box, **kw)
    n = len(pos)
    f = np.zeros((n, 3))
    contrib = po["g"][:, None] * po["D"] / po["r"][:, None]
    for c in range(3):
        f[:, c] = np.bincount(po["I"], weights=contrib[:, c], minlength=n)
    return f, po


def loss_and_grad(theta, pos, box, f_ref, weight=1.0, **kw):
    """Loss of one frame and its gradient with respect to theta (analytic backprop).
    weight scales both (several frames: weight = 1 / number of frames)."""
    bf = kw.get("bf", False)
    f, po = forces(theta, pos, box, **kw)
    n = len(pos)
    c = weight / (3.0 * n)
    dF = f - np.asarray(f_ref, dtype=np.float64)
    loss = c * np.sum(dF * dF)
    rhat = po["D"] / po["r"][:, None]
    # d loss / d s for each ordered pair (owner I)
    ds = 2.0 * c * np.sum(dF[po["I"]] * rhat, axis=1) * po["fc"]
    C = po["cache"]
    p = C["p"]
    rnd = bf16 if bf else (lambda v: v)
    g = {}
    g["w4"] = C["h3"].T @ ds
    g["b4"] = np.array([ds.sum()])
    dz3 = rnd((ds[:, None] * p["w4"][None, :]) * rnd(dsilu(C["z3"])))
    g["W3"] = C["h2"].T @ dz3
    g["b3"] = dz3.sum(axis=0)
    dz2 = rnd((dz3 @ C["W3"].T) * rnd(dsilu(C["z2"])))
    g["W2"] = C["h1"].T @ dz2
    g["b2"] = dz2.sum(axis=0)
    dz1 = rnd((dz2 @ C["W2"].T) * rnd(dsilu(C["z1"])))
    g["W1"] = C["a0"].T @ dz1
    g["b1"] = dz1.sum(axis=0)
    return loss, pack(g), f, po
