"""ParallelTempering.exchange: the reference's Metropolis rule (src/ParallelTempering.py:164-200),
checked against its closed form with stand-in observables (no GPU needed)."""
import math

import numpy as np
import pytest

import espressopp_amd as espressopp


class _Obs:
    def __init__(self, v):
        self.v = v

    def compute(self):
        return self.v

    def computeEnergy(self):
        return self.v


class _Thermo:
    def __init__(self, T):
        self.temperature = T


class _Integ:
    def __init__(self):
        self.steps = 0

    def run(self, n):
        self.steps += n


class _QueueRNG:
    def __init__(self, draws):
        self.draws = list(draws)
        self.used = 0

    def random(self):
        self.used += 1
        return self.draws.pop(0)


def _pt(E, T, N, rng):
    pt = espressopp.ParallelTempering(NumberOfSystems=len(E), RNG=rng)
    thermos = []
    for i in range(len(E)):
        pt.startDefiningSystem(i)
        th = _Thermo(T[i])
        thermos.append(th)
        pt.setIntegrator(_Integ(), th)
        pt.setAnalysisE(_Obs(E[i]))
        pt.setAnalysisT(_Obs(T[i]))
        pt.setAnalysisNPart(_Obs(N[i]))
        pt.endDefiningSystem(i)
    return pt, thermos


def test_pairs_alternate_odd_even_and_one_draw_per_pair():
    rng = _QueueRNG([0.99] * 20)
    E, T, N = [-1.0, -2.0, -3.0, -4.0, -5.0], [0.7, 0.8, 0.9, 1.0, 1.1], [100] * 5
    pt, _ = _pt(E, T, N, rng)
    pt.exchange()
    assert [(m, n) for m, n, *_ in pt.log] == [(1, 2), (3, 4)]
    pt.exchange()
    assert [(m, n) for m, n, *_ in pt.log[2:]] == [(0, 1), (2, 3)]
    pt.exchange()
    assert [(m, n) for m, n, *_ in pt.log[4:]] == [(1, 2), (3, 4)]
    assert rng.used == 6


def test_metropolis_rule_closed_form():
    # delta = (1/t2 - 1/t1)(e1/n1 - e2/n2); accept iff delta <= 0 or exp(-delta) > U
    E, T, N = [-500.0, -480.0, -700.0, -650.0], [0.7, 0.9, 1.1, 1.3], [200, 200, 300, 300]
    d01 = (1 / T[1] - 1 / T[0]) * (E[0] / N[0] - E[1] / N[1])
    d23 = (1 / T[3] - 1 / T[2]) * (E[2] / N[2] - E[3] / N[3])
    assert d01 > 0 and d23 > 0
    pt, th = _pt(E, T, N, _QueueRNG([]))
    pt._oddeven = 1                          # next exchange pairs (0,1), (2,3)
    pt._RNG = _QueueRNG([math.exp(-d01) - 1e-9, math.exp(-d23) + 1e-9])
    swaps = pt.exchange()
    assert swaps == [(0, 1)]
    # on acceptance each thermostat takes the other's measured temperature
    assert th[0].temperature == T[1] and th[1].temperature == T[0]
    assert th[2].temperature == T[2] and th[3].temperature == T[3]
    assert pt.log[0][2] == pytest.approx(d01, rel=1e-15)


def test_downhill_swaps_always_accepted():
    E, T, N = [-400.0, -500.0], [0.7, 1.0], [100, 100]   # colder replica has the higher energy
    assert espressopp.ParallelTempering.delta(T[0], T[1], E[0], E[1], N[0], N[1]) < 0
    rng = _QueueRNG([0.999999])
    pt, th = _pt(E, T, N, rng)
    pt._oddeven = 1
    assert pt.exchange() == [(0, 1)]
    assert rng.used == 1


def test_acceptance_rate_is_min_one_exp_minus_delta():
    E, T, N = [-500.0, -470.0], [0.8, 1.0], [100, 100]
    d = espressopp.ParallelTempering.delta(T[0], T[1], E[0], E[1], N[0], N[1])
    rng = np.random.default_rng(5)
    acc = 0
    trials = 20000
    for _ in range(trials):
        pt, _ = _pt(E, T, N, rng)
        pt._oddeven = 1
        acc += len(pt.exchange())
    assert abs(acc / trials - min(1.0, math.exp(-d))) < 4 * math.sqrt(0.25 / trials)


def test_run_advances_every_replica_and_checks_arguments():
    pt, _ = _pt([0.0] * 3, [1.0] * 3, [1] * 3, _QueueRNG([]))
    pt.run(7)
    assert all(i.steps == 7 for i in pt._multisystem._integrators)
    with pytest.raises(ValueError):
        espressopp.ParallelTempering(NumberOfSystems=1, RNG=np.random.default_rng(0))
    with pytest.raises(ValueError):
        pt.startDefiningSystem(3)
