"""ctypes front end of the FP64 CPU oracle (oracle/espp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg.  The product package (espressopp_amd) never
imports it.

The oracle restates ESPResSo++'s short-range MD step in FP64; see the header
of espp_oracle.c for the reference file:line of every part.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_espp.so")
_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32, i64, dbl, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64
        P = ctypes.POINTER
        sigs = {
            "or_create": (vp, [P(dbl), P(i32), dbl, u64]),
            "or_destroy": (None, [vp]),
            "or_last_error": (ctypes.c_char_p, [vp]),
            "or_set_threads": (None, [vp, i32]),
            "or_set_halfcell": (i32, [vp, i32]),
            "or_set_skin": (None, [vp, dbl]),
            "or_set_dt": (None, [vp, dbl]),
            "or_get_step": (i64, [vp]),
            "or_set_step": (None, [vp, i64]),
            "or_num_resorts": (i64, [vp]),
            "or_num_particles": (i64, [vp]),
            "or_add_particles": (i32, [vp, i64, vp, vp, vp, vp, vp]),
            "or_decompose": (i32, [vp]),
            "or_vl_create": (i32, [vp, dbl]),
            "or_vl_exclude": (i32, [vp, i32, i64, vp]),
            "or_vl_size": (i64, [vp, i32]),
            "or_vl_builds": (i64, [vp, i32]),
            "or_vl_pairs": (i64, [vp, i32, vp, i64]),
            "or_fpl_create": (i32, [vp]),
            "or_fpl_add": (i32, [vp, i32, i64, vp]),
            "or_ia_lj_create": (i32, [vp, i32]),
            "or_ia_lj_set": (i32, [vp, i32, i32, i32, dbl, dbl, dbl, dbl]),
            "or_ia_fene_create": (i32, [vp, i32, dbl, dbl, dbl, dbl, dbl]),
            "or_ia_energy": (dbl, [vp, i32]),
            "or_ia_virial": (dbl, [vp, i32, vp]),
            "or_calc_forces": (i32, [vp]),
            "or_langevin": (i32, [vp, i32, dbl, dbl, i64, vp]),
            "or_set_therm_calls": (None, [vp, u64]),
            "or_get_therm_calls": (u64, [vp]),
            "or_run": (i32, [vp, i32]),
            "or_temperature": (dbl, [vp]),
            "or_rdf": (i32, [vp, i32, vp]),
            "or_pressure": (dbl, [vp]),
            "or_get_by_id": (i32, [vp, i32, vp]),
            "or_get_ids": (i32, [vp, vp]),
            "or_set_by_id": (i32, [vp, i32, vp]),
            "or_timers": (None, [vp, vp]),
            "or_lj_energy": (dbl, [dbl, dbl, dbl, dbl, dbl]),
            "or_lj_force": (None, [dbl, dbl, dbl, vp, vp]),
            "or_fene_energy": (dbl, [dbl, dbl, dbl, dbl]),
            "or_fene_force": (None, [dbl, dbl, dbl, vp, vp]),
            "or_tab_new": (vp, [i32, i32, vp, vp, vp, dbl]),
            "or_tab_destroy": (None, [vp]),
            "or_tab_eval": (None, [vp, i32, vp, vp, vp]),
            "or_ia_tab_create": (i32, [vp, i32]),
            "or_ia_tab_set": (i32, [vp, i32, i32, i32, i32, i32, vp, vp, vp, dbl]),
            "or_ia_tabbond_create": (i32, [vp, i32, i32, i32, vp, vp, vp, dbl]),
            "or_philox4x32_10": (None, [vp, vp, vp]),
            "or_langevin_uniforms": (None, [u64, u64, i32, vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def langevin_uniforms(seed, call, pid):
    out = np.zeros(3)
    lib().or_langevin_uniforms(seed, call, pid, _ptr(out))
    return out


def lj_energy(eps, sig, rc, shift, r):
    return lib().or_lj_energy(eps, sig, rc, float("nan") if shift == "auto" else shift, r)


def lj_force(eps, sig, rc, d):
    d = np.ascontiguousarray(d, dtype=np.float64)
    F = np.zeros(3)
    lib().or_lj_force(eps, sig, rc, _ptr(d), _ptr(F))
    return F


def fene_energy(K, r0, rMax, r):
    return lib().or_fene_energy(K, r0, rMax, r)


def fene_force(K, r0, rMax, d):
    d = np.ascontiguousarray(d, dtype=np.float64)
    F = np.zeros(3)
    lib().or_fene_force(K, r0, rMax, _ptr(d), _ptr(F))
    return F


def _table(r, e, f):
    r = np.ascontiguousarray(r, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    if not (len(r) == len(e) == len(f)):
        raise ValueError("table columns differ in length")
    return r, e, f


def tab_eval(itype, r, e, f, at, cutoff=float("inf")):
    """Energy and force of a Tabulated table (rows r, e, f) at distances `at`."""
    r, e, f = _table(r, e, f)
    h = lib().or_tab_new(int(itype), len(r), _ptr(r), _ptr(e), _ptr(f), float(cutoff))
    if not h:
        raise ValueError("illegal table")
    try:
        at = np.ascontiguousarray(at, dtype=np.float64)
        eo, fo = np.zeros(len(at)), np.zeros(len(at))
        lib().or_tab_eval(h, len(at), _ptr(at), _ptr(eo), _ptr(fo))
    finally:
        lib().or_tab_destroy(h)
    return eo, fo


class OracleSim:
    """One FP64 system: box, cell grid, skin, particles, lists, interactions."""

    def __init__(self, box, cells, skin, seed=12345, threads=1, half_cell_int=1):
        L = (ctypes.c_double * 3)(*[float(b) for b in box])
        C = (ctypes.c_int * 3)(*[int(c) for c in cells])
        self._h = lib().or_create(L, C, float(skin), int(seed))
        if not self._h:
            raise MemoryError("or_create failed")
        lib().or_set_threads(self._h, int(threads))
        if half_cell_int != 1 and lib().or_set_halfcell(self._h, int(half_cell_int)) != 0:
            raise ValueError("halfCellInt %r does not fit the cell grid %r" % (half_cell_int, tuple(cells)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.or_destroy(h)
            self._h = None

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(lib().or_last_error(self._h).decode())
        return rc

    def add_particles(self, ids, pos, types=None, vel=None, mass=None):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        t = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        self._check(lib().or_add_particles(self._h, len(ids), _ptr(ids), _ptr(t), _ptr(pos), _ptr(v), _ptr(m)))

    def decompose(self):
        self._check(lib().or_decompose(self._h))

    def set_threads(self, n):
        lib().or_set_threads(self._h, int(n))

    def set_skin(self, skin):
        lib().or_set_skin(self._h, float(skin))

    @property
    def n(self):
        return lib().or_num_particles(self._h)

    def vl_create(self, cutoff):
        return self._check(lib().or_vl_create(self._h, float(cutoff)))

    def vl_exclude(self, vl, pairs):
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        self._check(lib().or_vl_exclude(self._h, vl, len(p), _ptr(p)))

    def vl_size(self, vl):
        return lib().or_vl_size(self._h, vl)

    def vl_builds(self, vl):
        return lib().or_vl_builds(self._h, vl)

    def vl_pairs(self, vl):
        n = self.vl_size(vl)
        out = np.zeros((max(n, 1), 2), dtype=np.int32)
        lib().or_vl_pairs(self._h, vl, _ptr(out), n)
        return out[:n]

    def fpl_create(self):
        return self._check(lib().or_fpl_create(self._h))

    def fpl_add(self, fpl, pairs):
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        self._check(lib().or_fpl_add(self._h, fpl, len(p), _ptr(p)))

    def lj_create(self, vl):
        return self._check(lib().or_ia_lj_create(self._h, vl))

    def lj_set(self, ia, t1, t2, epsilon=1.0, sigma=1.0, cutoff=float("inf"), shift="auto"):
        sh = float("nan") if shift == "auto" else float(shift)
        self._check(lib().or_ia_lj_set(self._h, ia, t1, t2, float(epsilon), float(sigma), float(cutoff), sh))

    def fene_create(self, fpl, K=30.0, r0=0.0, rMax=1.5, cutoff=float("inf"), shift=0.0):
        return self._check(lib().or_ia_fene_create(self._h, fpl, K, r0, rMax, cutoff, shift))

    def tab_create(self, vl):
        return self._check(lib().or_ia_tab_create(self._h, vl))

    def tab_set(self, ia, t1, t2, itype, r, e, f, cutoff=float("inf")):
        r, e, f = _table(r, e, f)
        self._check(lib().or_ia_tab_set(self._h, ia, t1, t2, int(itype), len(r), _ptr(r), _ptr(e), _ptr(f),
                                        float(cutoff)))

    def tabbond_create(self, fpl, itype, r, e, f, cutoff=float("inf")):
        r, e, f = _table(r, e, f)
        return self._check(lib().or_ia_tabbond_create(self._h, fpl, int(itype), len(r), _ptr(r), _ptr(e), _ptr(f),
                                                      float(cutoff)))

    def energy(self, ia):
        return lib().or_ia_energy(self._h, ia)

    def virial(self, ia):
        t = np.zeros(6)
        w = lib().or_ia_virial(self._h, ia, _ptr(t))
        return w, t

    def calc_forces(self):
        self._check(lib().or_calc_forces(self._h))

    def langevin(self, gamma, temperature, exclusions=(), on=True):
        e = np.ascontiguousarray(exclusions, dtype=np.int32)
        self._check(lib().or_langevin(self._h, 1 if on else 0, float(gamma), float(temperature), len(e), _ptr(e) if len(e) else None))

    @property
    def therm_calls(self):
        return lib().or_get_therm_calls(self._h)

    @therm_calls.setter
    def therm_calls(self, c):
        lib().or_set_therm_calls(self._h, int(c))

    def run(self, nsteps, dt=None):
        if dt is not None:
            lib().or_set_dt(self._h, float(dt))
        self._check(lib().or_run(self._h, int(nsteps)))

    @property
    def step(self):
        return lib().or_get_step(self._h)

    @property
    def resorts(self):
        return lib().or_num_resorts(self._h)

    def temperature(self):
        return lib().or_temperature(self._h)

    def rdf(self, nbins):
        out = np.zeros(int(nbins))
        self._check(lib().or_rdf(self._h, int(nbins), _ptr(out)))
        return out

    def pressure(self):
        return lib().or_pressure(self._h)

    def ids(self):
        out = np.zeros(self.n, dtype=np.int32)
        lib().or_get_ids(self._h, _ptr(out))
        return out

    def get(self, field):
        code = {"pos": 0, "v": 1, "f": 2, "image": 3}[field]
        out = np.zeros((self.n, 3))
        lib().or_get_by_id(self._h, code, _ptr(out))
        return out

    def set(self, field, values):
        code = {"pos": 0, "v": 1}[field]
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1, 3)
        lib().or_set_by_id(self._h, code, _ptr(a))

    def timers(self):
        t = np.zeros(5)
        lib().or_timers(self._h, _ptr(t))
        return dict(zip(["run", "force", "int1", "int2", "resort"], t))
