"""Checkpoint and resume of a system on the GPU path.

The reference stores configurations with io.DumpH5MD / io.RestoreH5MD
(src/io/DumpH5MD.py:353, src/io/RestoreH5MDParallel.py:28) and the RNG with
esutil::RNG::saveState (src/esutil/RNG.cpp:86-143).  HDF5 is not part of this
package; a checkpoint is one NumPy .npz file holding every particle's id,
type, mass, folded position, image, velocity and force, the step counter, the
Langevin noise state (seed, calls) and the integrator's skin bookkeeping
(accumulated displacement, pending resort).  A run resumed from it continues
the same Langevin noise sequence and rebuild schedule.
"""
import ctypes

import numpy as np

from .. import _lib

_FORMAT = 1


def save(filename, system, integrator=None):
    h = system.storage._h
    a = system.storage.getParticleArrays()
    seed, calls = ctypes.c_uint64(0), ctypes.c_uint64(0)
    h.call("espp_get_rng_state", ctypes.byref(seed), ctypes.byref(calls))
    md, rs = ctypes.c_double(0.0), ctypes.c_int32(0)
    h.call("espp_get_run_state", ctypes.byref(md), ctypes.byref(rs))
    step = h.out_i64("espp_get_step")
    box = np.asarray(system.bc.boxL, dtype=np.float64)
    np.savez(filename, format=_FORMAT, id=a["id"], type=a["type"], mass=a["mass"], pos=a["pos"], image=a["image"],
             v=a["v"], f=a["f"], box=box, skin=float(system.skin), step=step, seed=seed.value, calls=calls.value,
             maxdist=md.value, resort=rs.value, dt=(integrator.dt if integrator is not None else np.nan))


def restore(filename, system, integrator=None):
    """Load a checkpoint into a freshly set-up system (same box and cell grid,
    interactions and extensions attached as before); returns the arrays read."""
    d = np.load(filename, allow_pickle=False)
    if int(d["format"]) != _FORMAT:
        raise ValueError("unknown checkpoint format")
    if not np.allclose(np.asarray(system.bc.boxL, dtype=np.float64), d["box"], rtol=0, atol=0):
        raise ValueError("checkpoint box differs from the system's box")
    h = system.storage._h
    ids = d["id"].astype(np.int32)
    box = d["box"]
    # unfolded positions: the image counts come back through the fold of decompose
    pos = d["pos"] + d["image"] * box
    system.storage.addParticlesArray(ids, pos, d["type"].astype(np.int32), d["v"], d["mass"])
    system.storage.decompose()
    # forces are not loaded: VelocityVerlet.run recomputes them before its first step
    h.call("espp_set_step", int(d["step"]))
    h.call("espp_set_rng_state", int(d["seed"]), int(d["calls"]))
    h.call("espp_set_run_state", float(d["maxdist"]), int(d["resort"]))
    if integrator is not None and np.isfinite(float(d["dt"])):
        integrator.dt = float(d["dt"])
    return d
