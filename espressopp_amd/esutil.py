"""espressopp.esutil.RNG (src/esutil/RNG.cpp:40-66).

The seed keys the Philox4x32-10 stream of the GPU Langevin thermostat.  The
host-side draws (``rng()``, ``rng.normal()``) serve set-up scripts only.
"""
import random


class RNG:
    def __init__(self, seed=12345):
        self._seed = int(seed)
        self._py = random.Random(self._seed)

    def seed(self, s):
        self._seed = int(s)
        self._py.seed(self._seed)

    def get_seed(self):
        return self._seed

    def __call__(self, N=None):
        if N is None:
            return self._py.random()
        return self._py.randrange(int(N))

    def normal(self):
        return self._py.gauss(0.0, 1.0)

    def uniformOnSphere(self):
        while True:
            v = [self._py.gauss(0.0, 1.0) for _ in range(3)]
            n = sum(x * x for x in v) ** 0.5
            if n > 0:
                return [x / n for x in v]
