"""espressopp.integrator: VelocityVerlet and LangevinThermostat.

Reference: src/integrator/MDIntegrator.py:56-90 (run/addExtension, dt and
step properties), src/integrator/VelocityVerlet.py:39-50 (getTimers is
pmiinvoke: one tuple per rank), src/integrator/LangevinThermostat.py:59-76.
The whole step loop of run() executes on the GPU through espp_run.
"""
import numpy as np

from . import _lib


class Extension:
    def connect(self):
        pass

    def disconnect(self):
        pass


class LangevinThermostat(Extension):
    """Langevin friction + uniform noise (src/integrator/LangevinThermostat.cpp:141-214)."""

    def __init__(self, system):
        if system.rng is None:
            raise RuntimeError("system has no RNG")
        self.system = system
        self._gamma = 0.0
        self._temperature = 0.0
        self.adress = False
        self._integrator = None
        self._exclusions = []

    def _push(self):
        if self._integrator is not None:
            h = self.system.storage._h
            h.call("espp_langevin", 1, self._gamma, self._temperature, 0, None)

    @property
    def gamma(self):
        return self._gamma

    @gamma.setter
    def gamma(self, g):
        self._gamma = float(g)
        self._push()

    @property
    def temperature(self):
        return self._temperature

    @temperature.setter
    def temperature(self, t):
        self._temperature = float(t)
        self._push()

    def addExclusions(self, pidlist):
        ids = np.asarray(list(pidlist), dtype=np.int32)
        self._exclusions.extend(int(x) for x in ids)
        if self._integrator is not None and len(ids):
            h = self.system.storage._h
            h.call("espp_langevin", 1, self._gamma, self._temperature, len(ids), _lib.ptr(ids))

    def addExclpid(self, pid):
        self.addExclusions([pid])

    def connect(self):
        if self.adress:
            raise NotImplementedError("AdResS thermalisation is not part of the GPU path")
        h = self.system.storage._h
        ids = np.asarray(self._exclusions, dtype=np.int32)
        h.call("espp_langevin", 1, self._gamma, self._temperature, len(ids), _lib.ptr(ids) if len(ids) else None)

    def disconnect(self):
        h = self.system.storage._h
        h.call("espp_langevin", 0, self._gamma, self._temperature, 0, None)


class VelocityVerlet:
    """Velocity Verlet with skin-triggered rebuilds (src/integrator/VelocityVerlet.cpp:63-371)."""

    def __init__(self, system):
        if system.storage is None:
            raise RuntimeError("system has no storage")
        self.system = system
        self._h = system.storage._h
        self._dt = 0.005
        self._extensions = []
        system.integrator = self

    @property
    def dt(self):
        return self._dt

    @dt.setter
    def dt(self, dt):
        if dt == 0.0:
            raise ValueError("Timestep 'dt' must be non-zero!")
        self._dt = float(dt)

    @property
    def step(self):
        return self._h.out_i64("espp_get_step")

    @step.setter
    def step(self, s):
        self._h.call("espp_set_step", int(s))

    def run(self, niter):
        if not isinstance(niter, int):
            raise ValueError('The provided number of steps have to be an integer not {} with value {}'.format(type(niter), niter))
        self._h.call("espp_run", niter, self._dt)

    def addExtension(self, extension):
        extension._integrator = self
        extension.connect()
        self._extensions.append(extension)
        return True

    def getExtension(self, k):
        return self._extensions[k]

    def getNumberOfExtensions(self):
        return len(self._extensions)

    def getNumResorts(self):
        return self._h.out_i64("espp_num_resorts")

    def getTimers(self):
        t = np.zeros(10)
        self._h.call("espp_timers", _lib.ptr(t))
        return [tuple(float(x) for x in t)]

    def resetTimers(self):
        pass
