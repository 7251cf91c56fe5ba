"""espressopp.interaction: LennardJones, FENE, Tabulated and their list interactions.

Reference: src/interaction/LennardJones.py:721-750 (LennardJones(epsilon=1,
sigma=1, cutoff=inf, shift='auto'), VerletListLennardJones),
src/interaction/FENE.py:98-137 (FENE(K=30, r0=0, rMax=1.5, cutoff=inf,
shift=0), FixedPairListFENE), src/interaction/Interaction.py:105-110.

Potential objects are parameter holders.  Their computeEnergy/computeForce
evaluate one distance with the reference formulas (Potential::computeEnergy,
src/interaction/Potential.hpp:300-400), a convenience for scripts; forces,
energies and virials of a system are always computed by the GPU kernels.
"""
import ctypes
import math
import re

import numpy as np

from ._vec3 import Real3D

infinity = float("inf")


class Potential:
    def __init__(self, cutoff=infinity, shift=0.0):
        self._cutoff = float(cutoff)
        self._shift = 0.0 if shift == "auto" else float(shift)
        self._auto = shift == "auto"
        self._owners = []

    # PotentialTemplate cutoff/shift handling (src/interaction/Potential.hpp:190-235)
    @property
    def cutoff(self):
        return self._cutoff

    @cutoff.setter
    def cutoff(self, c):
        self._cutoff = float(c)
        self._update()

    @property
    def shift(self):
        return self._shift

    @shift.setter
    def shift(self, s):
        self._auto = False
        self._shift = float(s)
        self._update()

    def setAutoShift(self):
        self._auto = True
        self._update()
        return self._shift

    def _update(self):
        if self._auto:
            self._shift = 0.0 if math.isinf(self._cutoff) else self._energy_raw(self._cutoff ** 2)
        for o in self._owners:
            o._potential_changed(self)

    def _dist2(self, *args):
        if len(args) == 1:
            a = args[0]
            if hasattr(a, "__len__") and len(a) == 3:
                d = [float(x) for x in a]
                return d[0] * d[0] + d[1] * d[1] + d[2] * d[2], d
            r = float(a)
            return r * r, None
        if len(args) == 2:
            p1, p2 = args
            d = [p1.pos[k] - p2.pos[k] for k in range(3)]
            return d[0] * d[0] + d[1] * d[1] + d[2] * d[2], d
        raise TypeError("expected a distance, a distance vector or two particles")

    def computeEnergy(self, *args):
        d2, _ = self._dist2(*args)
        if d2 > self._cutoff * self._cutoff:
            return 0.0
        return self._energy_raw(d2) - self._shift

    def computeEnergySqr(self, d2):
        if d2 > self._cutoff * self._cutoff:
            return 0.0
        return self._energy_raw(d2) - self._shift

    def computeForce(self, *args):
        d2, d = self._dist2(*args)
        if d is None:
            raise TypeError("computeForce needs a distance vector")
        if d2 > self._cutoff * self._cutoff:
            return Real3D(0.0)
        ff = self._force_factor(d2)
        return Real3D(d[0] * ff, d[1] * ff, d[2] * ff)


class LennardJones(Potential):
    """LJ: E = 4 eps ((sigma/r)^12 - (sigma/r)^6) - shift (src/interaction/LennardJones.hpp:60-160)."""

    def __init__(self, epsilon=1.0, sigma=1.0, cutoff=infinity, shift="auto"):
        self._epsilon = float(epsilon)
        self._sigma = float(sigma)
        Potential.__init__(self, cutoff, shift)
        self._update()

    @property
    def epsilon(self):
        return self._epsilon

    @epsilon.setter
    def epsilon(self, e):
        self._epsilon = float(e)
        self._update()

    @property
    def sigma(self):
        return self._sigma

    @sigma.setter
    def sigma(self, s):
        self._sigma = float(s)
        self._update()

    def _energy_raw(self, d2):
        frac2 = self._sigma * self._sigma / d2
        frac6 = frac2 * frac2 * frac2
        return 4.0 * self._epsilon * (frac6 * frac6 - frac6)

    def _force_factor(self, d2):
        sig2 = self._sigma * self._sigma
        sig6 = sig2 * sig2 * sig2
        ff1 = 48.0 * self._epsilon * sig6 * sig6
        ff2 = 24.0 * self._epsilon * sig6
        frac2 = 1.0 / d2
        frac6 = frac2 * frac2 * frac2
        return frac6 * (ff1 * frac6 - ff2) * frac2

    def _abi_shift(self):
        return float("nan") if self._auto else self._shift


class FENE(Potential):
    """FENE: E = -1/2 rMax^2 K ln(1 - ((r - r0)/rMax)^2) (src/interaction/FENE.hpp:91-112)."""

    def __init__(self, K=30.0, r0=0.0, rMax=1.5, cutoff=infinity, shift=0.0):
        self._K = float(K)
        self._r0 = float(r0)
        self._rMax = float(rMax)
        Potential.__init__(self, cutoff, shift)
        self._update()

    K = property(lambda s: s._K, lambda s, v: (setattr(s, "_K", float(v)), s._update()))
    r0 = property(lambda s: s._r0, lambda s, v: (setattr(s, "_r0", float(v)), s._update()))
    rMax = property(lambda s: s._rMax, lambda s, v: (setattr(s, "_rMax", float(v)), s._update()))

    def _energy_raw(self, d2):
        return -0.5 * self._rMax ** 2 * self._K * math.log(1 - math.pow((math.sqrt(d2) - self._r0) / self._rMax, 2))

    def _force_factor(self, d2):
        rMaxSqr = self._rMax * self._rMax
        if self._r0 > 2.220446049250313e-16:
            r = math.sqrt(d2)
            return -self._K * (r - self._r0) / (r * (1 - ((r - self._r0) * (r - self._r0) / rMaxSqr)))
        return -self._K / (1.0 - d2 / rMaxSqr)


class Interaction:
    """Interaction::computeEnergy / computeVirial (src/interaction/Interaction.cpp:45-53)."""

    _system_attached = None

    def _attach(self, system):
        if self._system_attached is None:
            self._h.call("espp_add_interaction", self._id)
        self._system_attached = system

    def computeEnergy(self):
        return self._h.out_dbl("espp_ia_energy", self._id)

    def computeVirial(self):
        w = ctypes.c_double(0.0)
        self._h.call("espp_ia_virial", self._id, ctypes.byref(w), None)
        return w.value

    def computeVirialTensor(self):
        """xx, yy, zz, xy, xz, yz (src/interaction/VerletListInteractionTemplate.hpp:306-331)."""
        t = np.zeros(6)
        w = ctypes.c_double(0.0)
        self._h.call("espp_ia_virial", self._id, ctypes.byref(w), t.ctypes.data_as(ctypes.c_void_p))
        return tuple(t)

    def computeEnergyDeriv(self):
        return 0.0

    def bondType(self):
        return self._bond_type


class VerletListLennardJones(Interaction):
    """VerletListInteractionTemplate<LennardJones> (src/interaction/VerletListInteractionTemplate.hpp:116-331)."""

    _bond_type = 1   # Nonbonded

    def __init__(self, vl):
        self._vl = vl
        self._h = vl._h
        self._id = self._h.out_int("espp_lj_create", vl._id)
        self._potentials = {}

    def setPotential(self, type1, type2, potential):
        if not isinstance(potential, LennardJones):
            raise TypeError("VerletListLennardJones needs a LennardJones potential")
        self._potentials[(type1, type2)] = potential
        self._potentials[(type2, type1)] = potential
        if self not in potential._owners:
            potential._owners.append(self)
        self._upload(type1, type2, potential)

    def _upload(self, t1, t2, p):
        self._h.call("espp_lj_set", self._id, int(t1), int(t2), p.epsilon, p.sigma, p.cutoff, p._abi_shift())

    def _potential_changed(self, p):
        for (t1, t2), q in self._potentials.items():
            if q is p and t1 <= t2:
                self._upload(t1, t2, p)

    def getPotential(self, type1, type2):
        return self._potentials.get((type1, type2), LennardJones(0.0, 0.0, infinity, 0.0))

    def getVerletList(self):
        return self._vl

    getVerletListLocal = getVerletList

    def getMaxCutoff(self):
        cs = [p.cutoff for p in self._potentials.values()]
        return max(cs) if cs else 0.0


class FixedPairListFENE(Interaction):
    """FixedPairListInteractionTemplate<FENE> (src/interaction/FixedPairListInteractionTemplate.hpp:112-262)."""

    _bond_type = 2   # Pair

    def __init__(self, system, bondlist, potential):
        if not isinstance(potential, FENE):
            raise TypeError("FixedPairListFENE needs a FENE potential")
        self._fpl = bondlist
        self._h = bondlist._h
        self._potential = potential
        p = potential
        self._id = self._h.out_int("espp_fene_create", bondlist._id, p.K, p.r0, p.rMax, p.cutoff, p.shift)
        potential._owners.append(self)

    def setPotential(self, potential):
        """FixedPairListInteractionTemplate::setPotential (src/interaction/FENE.py:114-116)."""
        if not isinstance(potential, FENE):
            raise TypeError("FixedPairListFENE needs a FENE potential")
        if self in self._potential._owners:
            self._potential._owners.remove(self)
        self._potential = potential
        potential._owners.append(self)
        self._potential_changed(potential)

    def _potential_changed(self, p):
        if p is self._potential:
            self._h.call("espp_fene_set", self._id, p.K, p.r0, p.rMax, p.cutoff, p.shift)

    def getPotential(self):
        return self._potential

    def getFixedPairList(self):
        return self._fpl

    def setFixedPairList(self, fpl):
        """FixedPairListInteractionTemplate::setFixedPairList (src/interaction/FENE.py:122-124)."""
        if fpl._h is not self._h:
            raise RuntimeError("the bond list belongs to another system")
        self._h.call("espp_ia_set_fpl", self._id, fpl._id)
        self._fpl = fpl

    def getMaxCutoff(self):
        return self._potential.cutoff if not math.isinf(self._potential.cutoff) else 0.0


# ---------------------------------------------------------------- Tabulated
_NUM = re.compile(r"[ \t\n\r\f\v]*([+-]?(?:inf(?:inity)?|nan|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?))",
                  re.IGNORECASE)


def read_table(filename):
    """Rows (r, e, f) of a table file, read as the reference reads it
    (InterpolationTable::readFile, src/interaction/InterpolationTable.cpp:106-140):
    a line counts when its first three whitespace-separated fields parse as
    numbers (sscanf "%lg %lg %lg" == 3); every other line is skipped."""
    rows = []
    with open(filename) as fh:
        for line in fh:
            vals, pos = [], 0
            for _ in range(3):
                m = _NUM.match(line, pos)
                if not m:
                    break
                vals.append(float(m.group(1)))
                pos = m.end()
            if len(vals) == 3:
                rows.append(vals)
    a = np.asarray(rows, dtype=np.float64).reshape(-1, 3)
    return np.ascontiguousarray(a[:, 0]), np.ascontiguousarray(a[:, 1]), np.ascontiguousarray(a[:, 2])


class Tabulated(Potential):
    """interaction.Tabulated(itype, filename, cutoff) (src/interaction/Tabulated.hpp:43-115,
    src/interaction/Tabulated.py).  itype 1 linear, 2 Akima, 3 cubic spline
    interpolation of the file's (r, e, f) rows; shift 0.  The table is read when
    the potential is created; energies and forces of a system are computed by
    the GPU kernels."""

    def __init__(self, itype, filename, cutoff=infinity):
        if int(itype) not in (1, 2, 3):
            raise ValueError("itype must be 1 (linear), 2 (Akima) or 3 (cubic)")
        self._itype = int(itype)
        self._filename = str(filename)
        self._r, self._e, self._f = read_table(self._filename)
        if len(self._r) < 2 or (self._itype == 2 and len(self._r) < 4):
            raise RuntimeError("illegal file for tabulated potential")
        Potential.__init__(self, cutoff, 0.0)

    @property
    def filename(self):
        return self._filename

    @property
    def itype(self):
        return self._itype

    def getInterpolationType(self):
        return self._itype

    def _update(self):
        for o in self._owners:
            o._potential_changed(self)

    def table(self):
        return self._r, self._e, self._f

    def _energy_raw(self, d2):
        raise NotImplementedError("single-distance evaluation of a Tabulated potential is not provided; "
                                  "use the interaction's computeEnergy")

    _force_factor = _energy_raw


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class VerletListTabulated(Interaction):
    """VerletListInteractionTemplate<Tabulated> (src/interaction/Tabulated.cpp:73,100-103)."""

    _bond_type = 1   # Nonbonded

    def __init__(self, vl):
        self._vl = vl
        self._h = vl._h
        self._id = self._h.out_int("espp_tab_create", vl._id)
        self._potentials = {}

    def setPotential(self, type1, type2, potential):
        if not isinstance(potential, Tabulated):
            raise TypeError("VerletListTabulated needs a Tabulated potential")
        self._potentials[(type1, type2)] = potential
        self._potentials[(type2, type1)] = potential
        if self not in potential._owners:
            potential._owners.append(self)
        self._upload(type1, type2, potential)

    def _upload(self, t1, t2, p):
        r, e, f = p.table()
        self._h.call("espp_tab_set", self._id, int(t1), int(t2), p.itype, len(r), _ptr(r), _ptr(e), _ptr(f),
                     p.cutoff)

    def _potential_changed(self, p):
        for (t1, t2), q in self._potentials.items():
            if q is p and t1 <= t2:
                self._upload(t1, t2, p)

    def getPotential(self, type1, type2):
        return self._potentials.get((type1, type2))

    def getVerletList(self):
        return self._vl

    def getMaxCutoff(self):
        cs = [p.cutoff for p in self._potentials.values()]
        return max(cs) if cs else 0.0


class FixedPairListTabulated(Interaction):
    """FixedPairListInteractionTemplate<Tabulated> (src/interaction/Tabulated.cpp:85,182-190)."""

    _bond_type = 2   # Pair

    def __init__(self, system, fpl, potential):
        if not isinstance(potential, Tabulated):
            raise TypeError("FixedPairListTabulated needs a Tabulated potential")
        self._fpl = fpl
        self._h = fpl._h
        self._potential = potential
        r, e, f = potential.table()
        self._id = self._h.out_int("espp_tabbond_create", fpl._id, potential.itype, len(r), _ptr(r), _ptr(e),
                                   _ptr(f), potential.cutoff)

    def getPotential(self):
        return self._potential

    def getFixedPairList(self):
        return self._fpl

    def setFixedPairList(self, fpl):
        if fpl._h is not self._h:
            raise RuntimeError("the bond list belongs to another system")
        self._h.call("espp_ia_set_fpl", self._id, fpl._id)
        self._fpl = fpl

    def getMaxCutoff(self):
        return self._potential.cutoff if not math.isinf(self._potential.cutoff) else 0.0


class VerletListPairMLP(Interaction):
    """Learned coarse-grained pair force over a Verlet list (no reference counterpart;
    the learned alternative to VerletListTabulated, src/interaction/Tabulated.hpp:43-115).

    For a listed pair at r < cutoff: inputs are 32 Gaussian RBFs (centres k rc/31,
    width rc/31) and the pair means of two smooth local densities (radii
    densityRadii, cosine weights) normalised by densityNorm = ((mean_a, scale_a),
    (mean_b, scale_b)); three 64-wide SiLU layers and a linear output give s; the
    pair force is s (1 + cos(pi r / rc)) / 2 along d / r.  Evaluation and training
    run on the GPU's BF16 matrix cores with FP32 accumulation.  The force is not
    conservative, so computeEnergy / computeVirial raise."""

    _bond_type = 1   # Nonbonded
    N_PARAMS = 10625

    def __init__(self, vl, cutoff=3.0, densityRadii=(1.5, 3.0), densityNorm=((0.0, 1.0), (0.0, 1.0))):
        self._vl = vl
        self._h = vl._h
        spec = np.array([cutoff, densityRadii[0], densityRadii[1], densityNorm[0][0], densityNorm[0][1],
                         densityNorm[1][0], densityNorm[1][1]], dtype=np.float64)
        self._id = self._h.out_int("espp_mlp_create", vl._id, _ptr(spec))
        self.cutoff = float(cutoff)

    @classmethod
    def initialParameters(cls, seed=0, scale=1.0):
        """He-normal weights (std scale sqrt(2 / fan_in); sqrt(1 / 64) for the output
        row), zero biases."""
        rng = np.random.default_rng(seed)
        parts = [rng.normal(0, scale * math.sqrt(2.0 / 34), 34 * 64), np.zeros(64),
                 rng.normal(0, scale * math.sqrt(2.0 / 64), 64 * 64), np.zeros(64),
                 rng.normal(0, scale * math.sqrt(2.0 / 64), 64 * 64), np.zeros(64),
                 rng.normal(0, scale * math.sqrt(1.0 / 64), 64), np.zeros(1)]
        return np.concatenate(parts).astype(np.float32)

    def setParameters(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        if t.size != self.N_PARAMS:
            raise ValueError("the model has %d parameters" % self.N_PARAMS)
        self._h.call("espp_mlp_set_params", self._id, _ptr(t), t.size)

    def getParameters(self):
        t = np.zeros(self.N_PARAMS, dtype=np.float32)
        self._h.call("espp_mlp_get_params", self._id, _ptr(t), t.size)
        return t

    def trainFrame(self, fref, weight=1.0):
        """Gradient of weight * sum_i |F_i - fref_i|^2 / (3N) on the current configuration
        (fref in ascending id order) added to the accumulated gradient; returns the loss."""
        f = np.ascontiguousarray(fref, dtype=np.float64)
        loss = ctypes.c_double(0.0)
        self._h.call("espp_mlp_train_frame", self._id, _ptr(f), float(weight), ctypes.byref(loss))
        return loss.value

    def gradient(self):
        g = np.zeros(self.N_PARAMS, dtype=np.float32)
        self._h.call("espp_mlp_grad", self._id, _ptr(g), g.size)
        return g

    def zeroGradient(self):
        self._h.call("espp_mlp_zero_grad", self._id)

    def adamStep(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self._h.call("espp_mlp_adam", self._id, float(lr), float(beta1), float(beta2), float(eps))

    def getVerletList(self):
        return self._vl

    def getMaxCutoff(self):
        return self.cutoff
