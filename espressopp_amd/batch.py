"""Batched replicas of one Lennard-Jones model on one GPU.

ReplicaBatch holds many periodic cubic LJ systems of uneven size (the replicas
of a parallel-tempering run) in one device state (libespp_amd's espp_batch).
A step of every replica is one fixed sequence of launches over all replicas,
with each replica's skin check and rebuild decision kept on the device, captured
once into a HIP graph and replayed; run(n) therefore makes no host round trip
per step.  BatchedParallelTempering drives it with the reference's exchange
rule (src/ParallelTempering.py:164-200).
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import FIELD, check, load, ptr


class ReplicaBatch:
    """nrep replicas: n[r] particles in cubic boxes of edge box[r]; LJ(eps, sigma, rc)
    with the automatic shift (interaction.LennardJones(shift='auto')), Verlet skin,
    Langevin friction gamma (0: NVE) at temperature[r], Philox seed seed[r]."""

    def __init__(self, n, box, rc=2.5, skin=0.3, eps=1.0, sigma=1.0, gamma=1.0, temperature=None, seed=None):
        lib = load()
        n = np.ascontiguousarray(n, dtype=np.int32)
        R = len(n)
        box = np.ascontiguousarray(box, dtype=np.float64)
        if temperature is None:
            temperature = np.ones(R)
        temperature = np.ascontiguousarray(temperature, dtype=np.float64)
        if seed is None:
            seed = np.arange(1, R + 1)
        seed = np.ascontiguousarray(seed, dtype=np.uint64)
        if not (len(box) == len(temperature) == len(seed) == R):
            raise ValueError("n, box, temperature and seed need one entry per replica")
        h = ctypes.c_void_p()
        check(lib.espp_batch_create(R, ptr(n), ptr(box), float(rc), float(skin), float(eps), float(sigma),
                                    float(gamma), ptr(temperature), ptr(seed), ctypes.byref(h)))
        self.h = h
        self.lib = lib
        self.n = n
        self.box = box
        self.R = R

    def close(self):
        if self.h:
            self.lib.espp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        return check(getattr(self.lib, name)(self.h, *args))

    def load(self, rep, ids, pos, vel=None, mass=None):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        if len(ids) != self.n[rep] or pos.shape != (self.n[rep], 3):
            raise ValueError("replica %d holds %d particles" % (rep, self.n[rep]))
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        self._call("espp_batch_load", int(rep), ptr(ids), ptr(pos), ptr(v), ptr(m))

    def run(self, nsteps, dt):
        if not isinstance(nsteps, int):
            raise ValueError("number of integration steps must be int")
        self._call("espp_batch_run", nsteps, float(dt))

    def use_graph(self, on=True):
        """Replay steps from captured HIP graphs (default) or launch them directly."""
        self._call("espp_batch_set_graph", 1 if on else 0)

    def set_temperature(self, rep, T):
        self._call("espp_batch_set_temperature", int(rep), float(T))

    def temperature(self, rep):
        v = ctypes.c_double(0.0)
        self._call("espp_batch_get_temperature", int(rep), ctypes.byref(v))
        return v.value

    def energy(self):
        """(potential energy, virial) per replica."""
        e = np.zeros(self.R)
        w = np.zeros(self.R)
        self._call("espp_batch_energy", ptr(e), ptr(w))
        return e, w

    def kinetic(self):
        """sum m v^2 per replica."""
        k = np.zeros(self.R)
        self._call("espp_batch_kinetic", ptr(k))
        return k

    def measured_temperature(self):
        """analysis.Temperature per replica: sum m v^2 / (3 N) (src/analysis/Temperature.hpp:125)."""
        return self.kinetic() / (3.0 * self.n)

    def get(self, rep, field):
        out = np.zeros((self.n[rep], 3))
        self._call("espp_batch_get", int(rep), FIELD[field], ptr(out))
        return out

    def pairs(self, rep):
        total = ctypes.c_int64(0)
        self._call("espp_batch_pairs", int(rep), None, 0, ctypes.byref(total))
        out = np.zeros((max(total.value, 1), 2), dtype=np.int32)
        self._call("espp_batch_pairs", int(rep), ptr(out), total.value, ctypes.byref(total))
        return out[:total.value]

    @property
    def lpp(self):
        out = np.zeros(self.R, dtype=np.int32)
        self._call("espp_batch_get_lpp", ptr(out))
        return out

    @lpp.setter
    def lpp(self, values):
        v = np.ascontiguousarray(values, dtype=np.int32)
        if len(v) != self.R:
            raise ValueError("one lanes-per-particle value per replica")
        self._call("espp_batch_set_lpp", ptr(v))

    def force_time(self, reps=20):
        ms = ctypes.c_double(0.0)
        self._call("espp_batch_force_time", int(reps), ctypes.byref(ms))
        return ms.value

    def build_time(self, reps=10):
        """Average time (ms) of one list build of all replicas from one binning
        (every replica is rebuilt and binned once first)."""
        ms = ctypes.c_double(0.0)
        self._call("espp_batch_build_time", int(reps), ctypes.byref(ms))
        return ms.value

    def autotune(self, reps=20):
        """Choose lanes per particle per size class; returns a report dict."""
        lpp = np.zeros(self.R, dtype=np.int32)
        table = np.zeros(16)
        before, after = ctypes.c_double(0.0), ctypes.c_double(0.0)
        self._call("espp_batch_autotune", int(reps), ptr(lpp), ptr(table), ctypes.byref(before),
                   ctypes.byref(after))
        return {"lpp": lpp, "table_ms": table.reshape(4, 4), "before_ms": before.value, "after_ms": after.value}

    def stats(self, rep):
        a, b, c = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int32(0)
        self._call("espp_batch_stats", int(rep), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"resorts": a.value, "builds": b.value, "maxcount": c.value}


class BatchedParallelTempering:
    """ParallelTempering over a ReplicaBatch: run() advances every replica in the
    batch's graph-replayed launches; exchange() is the reference's rule
    (src/ParallelTempering.py:164-200): neighbours (m, m+1) with m alternating
    odd/even, delta = (1/t2 - 1/t1)(e1/n1 - e2/n2) from the replicas' potential
    energy, measured temperature and particle count, one uniform draw per pair,
    accepted when delta <= 0 or exp(-delta) > draw; on acceptance the two
    thermostats take each other's measured temperature."""

    def __init__(self, batch, RNG, dt=0.005):
        if batch.R < 2:
            raise ValueError("Parallel Tempering needs at least 2 systems to be able to work")
        self.batch = batch
        self._RNG = RNG
        self.dt = dt
        self._oddeven = 0
        self.log = []

    def getNumberOfSystems(self):
        return self.batch.R

    def run(self, nsteps):
        self.batch.run(nsteps, self.dt)

    def exchange(self):
        from .parallel_tempering import ParallelTempering
        self._oddeven = 1 if self._oddeven == 0 else 0
        energies, _ = self.batch.energy()
        temperatures = self.batch.measured_temperature()
        nparts = self.batch.n
        swaps = []
        for i in range(self.batch.R // 2):
            m = 2 * i + self._oddeven
            n = m + 1
            if n < self.batch.R:
                metro = self._RNG.random()
                t1, t2 = temperatures[m], temperatures[n]
                e1, e2 = energies[m], energies[n]
                d = ParallelTempering.delta(t1, t2, e1, e2, nparts[m], nparts[n])
                yes = d <= 0 or math.exp(-d) > metro
                self.log.append((m, n, d, metro, yes))
                if yes:
                    self.batch.set_temperature(n, t1)
                    self.batch.set_temperature(m, t2)
                    swaps.append((m, n))
        return swaps
