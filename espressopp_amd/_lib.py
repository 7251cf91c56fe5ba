"""ctypes binding of libespp_amd.so, the C ABI declared in include/espp_amd.h.

The library is loaded from this package directory only.  There is no CPU
fallback: a missing library or a machine without a GPU raises an error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libespp_amd.so")

ESPP_FP64 = 0
ESPP_MIXED = 1
FIELD = {"pos": 0, "v": 1, "f": 2, "image": 3, "type": 4, "mass": 5}

_c_int, _c_i64, _c_dbl, _c_u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64
_vp = ctypes.c_void_p
_P = ctypes.POINTER

# name -> argument types (every entry point returns int32 status unless listed in _RESTYPE)
SIGNATURES = {
    "espp_last_error": [],
    "espp_version": [],
    "espp_device_count": [_P(_c_int)],
    "espp_create": [_P(_c_dbl), _P(_c_int), _c_int, _c_dbl, _c_u64, _c_int, _P(_vp)],
    "espp_destroy": [_vp],
    "espp_set_skin": [_vp, _c_dbl],
    "espp_add_particles": [_vp, _c_i64, _vp, _vp, _vp, _vp, _vp],
    "espp_decompose": [_vp],
    "espp_num_particles": [_vp, _P(_c_i64)],
    "espp_get_ids": [_vp, _vp],
    "espp_get_particles": [_vp, _c_int, _vp],
    "espp_set_particles": [_vp, _c_int, _c_i64, _vp, _vp],
    "espp_get_particles_by_id": [_vp, _c_int, _c_i64, _vp, _vp],
    "espp_vl_create": [_vp, _c_dbl, _P(_c_int)],
    "espp_vl_exclude": [_vp, _c_int, _c_i64, _vp],
    "espp_vl_rebuild": [_vp, _c_int],
    "espp_vl_size": [_vp, _c_int, _P(_c_i64)],
    "espp_vl_builds": [_vp, _c_int, _P(_c_i64)],
    "espp_vl_pairs": [_vp, _c_int, _vp, _c_i64, _P(_c_i64)],
    "espp_vl_cutoff": [_vp, _c_int, _P(_c_dbl)],
    "espp_fpl_create": [_vp, _P(_c_int)],
    "espp_fpl_add": [_vp, _c_int, _c_i64, _vp],
    "espp_fpl_size": [_vp, _c_int, _P(_c_i64)],
    "espp_lj_create": [_vp, _c_int, _P(_c_int)],
    "espp_lj_set": [_vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_dbl],
    "espp_fene_create": [_vp, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _P(_c_int)],
    "espp_tab_create": [_vp, _c_int, _P(_c_int)],
    "espp_tab_set": [_vp, _c_int, _c_int, _c_int, _c_int, _c_i64, _vp, _vp, _vp, _c_dbl],
    "espp_tabbond_create": [_vp, _c_int, _c_int, _c_i64, _vp, _vp, _vp, _c_dbl, _P(_c_int)],
    "espp_add_interaction": [_vp, _c_int],
    "espp_remove_interaction": [_vp, _c_int],
    "espp_fene_set": [_vp, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _c_dbl],
    "espp_ia_set_fpl": [_vp, _c_int, _c_int],
    "espp_ia_energy": [_vp, _c_int, _P(_c_dbl)],
    "espp_ia_virial": [_vp, _c_int, _P(_c_dbl), _vp],
    "espp_ia_energy_noalloc": [_vp, _c_int, _P(_c_dbl), _P(_c_dbl), _vp],
    "espp_compute_forces": [_vp],
    "espp_langevin": [_vp, _c_int, _c_dbl, _c_dbl, _c_i64, _vp],
    "espp_run": [_vp, _c_int, _c_dbl],
    "espp_get_step": [_vp, _P(_c_i64)],
    "espp_set_step": [_vp, _c_i64],
    "espp_num_resorts": [_vp, _P(_c_i64)],
    "espp_timers": [_vp, _vp],
    "espp_kinetic": [_vp, _P(_c_dbl)],
    "espp_rdf": [_vp, _c_int, _vp],
    "espp_get_rng_state": [_vp, _P(_c_u64), _P(_c_u64)],
    "espp_set_rng_state": [_vp, _c_u64, _c_u64],
    "espp_get_run_state": [_vp, _P(_c_dbl), _P(_c_int)],
    "espp_set_run_state": [_vp, _c_dbl, _c_int],
    "espp_kinetic_tensor": [_vp, _c_u64, _vp, _vp, _vp],
    # learned pair force
    "espp_mlp_create": [_vp, _c_int, _vp, _P(_c_int)],
    "espp_mlp_set_params": [_vp, _c_int, _vp, _c_i64],
    "espp_mlp_get_params": [_vp, _c_int, _vp, _c_i64],
    "espp_mlp_train_frame": [_vp, _c_int, _vp, _c_dbl, _P(_c_dbl)],
    "espp_mlp_grad": [_vp, _c_int, _vp, _c_i64],
    "espp_mlp_zero_grad": [_vp, _c_int],
    "espp_mlp_adam": [_vp, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_dbl],
    # batched replicas
    "espp_batch_create": [_c_int, _vp, _vp, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _vp, _vp, _P(_vp)],
    "espp_batch_destroy": [_vp],
    "espp_batch_load": [_vp, _c_int, _vp, _vp, _vp, _vp],
    "espp_batch_set_temperature": [_vp, _c_int, _c_dbl],
    "espp_batch_get_temperature": [_vp, _c_int, _P(_c_dbl)],
    "espp_batch_run": [_vp, _c_int, _c_dbl],
    "espp_batch_energy": [_vp, _vp, _vp],
    "espp_batch_kinetic": [_vp, _vp],
    "espp_batch_get": [_vp, _c_int, _c_int, _vp],
    "espp_batch_pairs": [_vp, _c_int, _vp, _c_i64, _P(_c_i64)],
    "espp_batch_set_lpp": [_vp, _vp],
    "espp_batch_get_lpp": [_vp, _vp],
    "espp_batch_force_time": [_vp, _c_int, _P(_c_dbl)],
    "espp_batch_build_time": [_vp, _c_int, _P(_c_dbl)],
    "espp_batch_autotune": [_vp, _c_int, _vp, _vp, _P(_c_dbl), _P(_c_dbl)],
    "espp_batch_set_graph": [_vp, _c_int],
    "espp_batch_stats": [_vp, _c_int, _P(_c_i64), _P(_c_i64), _P(_c_int)],
}

# instrumentation entry points (include/espp_amd_prof.h)
PROF_SIGNATURES = {
    "espp_device_synchronize": [],
    "espp_prof_enable": [_vp, _c_int],
    "espp_prof_reset": [_vp],
    "espp_prof_stats": [_vp, _c_int, _P(_c_i64), _P(_c_dbl)],
    "espp_set_prune_shell": [_vp, _c_dbl],
    "espp_set_step_window": [_vp, _c_int],
    "espp_step_window_stats": [_vp, _P(_c_i64), _P(_c_i64), _P(_c_i64)],
    "espp_prune_stats": [_vp, _c_int, _P(_c_i64), _P(_c_i64), _P(_c_i64)],
    "espp_vl_stats": [_vp, _c_int, _P(_c_i64), _P(_c_int), _P(_c_int)],
    "espp_vl_batch_stats": [_vp, _c_int, _c_int, _P(_c_i64), _P(_c_i64)],
    "espp_tile_stats": [_vp, _P(_c_int), _P(_c_int)],
}
KERNEL = {"pair": 0, "nbr_build": 1, "integrate1": 2, "integrate2": 3, "binning": 4, "bond": 5, "pair_step": 6,
          "pair_step_pruned": 7}
_RESTYPE = {"espp_last_error": ctypes.c_char_p, "espp_version": ctypes.c_char_p}

_lib = None


class EsppError(RuntimeError):
    pass


def load():
    """Load libespp_amd.so (built by espressopp_amd.build) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libespp_amd.so is not built (%s); run `python -m espressopp_amd.build`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in list(SIGNATURES.items()) + list(PROF_SIGNATURES.items()):
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, ctypes.c_int32)
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().espp_last_error().decode()
        if rc == -1:
            raise ValueError(msg)
        raise EsppError(msg)
    return rc


def device_count():
    c = _c_int(0)
    check(load().espp_device_count(ctypes.byref(c)))
    return c.value


def ptr(a):
    return None if a is None else a.ctypes.data_as(_vp)


class Handle:
    """Owner of one espp_sim."""

    def __init__(self, box, cells, skin, seed, precision, half_cell_int=1):
        lib = load()
        b = (_c_dbl * 3)(*[float(x) for x in box])
        c = (_c_int * 3)(*[int(x) for x in cells])
        h = _vp()
        check(lib.espp_create(b, c, int(half_cell_int), float(skin), int(seed) & 0xFFFFFFFFFFFFFFFF, int(precision),
                              ctypes.byref(h)))
        self.h = h
        self.lib = lib

    def close(self):
        if self.h:
            self.lib.espp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # small helpers ---------------------------------------------------------
    def call(self, name, *args):
        return check(getattr(self.lib, name)(self.h, *args))

    def out_int(self, name, *args):
        v = _c_int(0)
        self.call(name, *args, ctypes.byref(v))
        return v.value

    def out_i64(self, name, *args):
        v = _c_i64(0)
        self.call(name, *args, ctypes.byref(v))
        return v.value

    def out_dbl(self, name, *args):
        v = _c_dbl(0.0)
        self.call(name, *args, ctypes.byref(v))
        return v.value

    @property
    def n(self):
        return self.out_i64("espp_num_particles")

    def ids(self):
        out = np.zeros(self.n, dtype=np.int32)
        self.call("espp_get_ids", ptr(out))
        return out

    def get(self, field):
        n = self.n
        width = 1 if field in ("type", "mass") else 3
        out = np.zeros((n, width) if width == 3 else n, dtype=np.float64)
        self.call("espp_get_particles", FIELD[field], ptr(out))
        return out

    def get_by_id(self, field, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        width = 1 if field in ("type", "mass") else 3
        out = np.zeros((len(ids), width) if width == 3 else len(ids), dtype=np.float64)
        self.call("espp_get_particles_by_id", FIELD[field], len(ids), ptr(ids), ptr(out))
        return out

    def set(self, field, ids, values):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        v = np.ascontiguousarray(values, dtype=np.float64)
        self.call("espp_set_particles", FIELD[field], len(ids), ptr(ids), ptr(v))
