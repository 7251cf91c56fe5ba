"""espp_ia_energy_noalloc (static device scratch, no allocation) against
espp_ia_energy / espp_ia_virial through the C ABI: two calls in a row, the
second on a smaller system, each bit-identical to the allocating entry point."""
import ctypes

import numpy as np
import pytest

import espressopp_amd as espressopp
from espressopp_amd import _lib
from tests._systems import cells_for, gpu_system, jittered_lattice, mb_velocities

pytestmark = pytest.mark.gpu


def lj_system(n, precision, seed):
    pos, box = jittered_lattice(n, 0.8442, seed=seed)
    system = gpu_system(box, cells_for(box, 2.5, 0.3), 0.3, precision, seed=seed)
    system.storage.addParticlesArray(np.arange(n), pos, None, mb_velocities(n, 1.0, seed + 1))
    system.storage.decompose()
    vl = espressopp.VerletList(system, cutoff=2.5)
    lj = espressopp.interaction.VerletListLennardJones(vl)
    lj.setPotential(0, 0, espressopp.interaction.LennardJones(1.0, 1.0, 2.5, "auto"))
    system.addInteraction(lj)
    integ = espressopp.integrator.VelocityVerlet(system)
    integ.dt = 0.005
    integ.run(20)
    return system, lj


def noalloc(lib, h, ia):
    e, w = ctypes.c_double(0.0), ctypes.c_double(0.0)
    t = np.zeros(6)
    rc = lib.espp_ia_energy_noalloc(h, ia, ctypes.byref(e), ctypes.byref(w), t.ctypes.data_as(ctypes.c_void_p))
    return rc, e.value, w.value, t


def allocating(lib, h, ia):
    e, w = ctypes.c_double(0.0), ctypes.c_double(0.0)
    t = np.zeros(6)
    assert lib.espp_ia_energy(h, ia, ctypes.byref(e)) == 0
    assert lib.espp_ia_virial(h, ia, ctypes.byref(w), t.ctypes.data_as(ctypes.c_void_p)) == 0
    return e.value, w.value, t


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_energy_matches_allocating_entry_point(gpu, precision):
    lib = _lib.load()
    for n, seed in ((27000, 3), (4000, 5)):          # the second call on a smaller system
        system, lj = lj_system(n, precision, seed)
        h = system.storage._h.h
        rc, e, w, t = noalloc(lib, h, lj._id)
        assert rc == 0, lib.espp_last_error()
        e0, w0, t0 = allocating(lib, h, lj._id)
        print("%s n=%d: energy %.12f (allocating %.12f)" % (precision, n, e, e0), flush=True)
        assert e == e0 and w == w0 and np.array_equal(t, t0)
        # a stale state is refused, not rebuilt (rebuilding allocates)
        system.storage.setParticleArrays("pos", [0], [[1.0, 1.0, 1.0]])
        rc, *_ = noalloc(lib, h, lj._id)
        assert rc == -2                     # ESPP_ERR_STATE
        system.storage._h.call("espp_compute_forces")
        rc, e, w, t = noalloc(lib, h, lj._id)
        assert rc == 0 and e == allocating(lib, h, lj._id)[0]
