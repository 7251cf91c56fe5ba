"""espressopp.ParallelTempering and espressopp.MultiSystem for replicas on one GPU.

Reference: src/ParallelTempering.py:90-200 and src/MultiSystem.py:75-140.  The
reference places each replica on its own group of MPI ranks.  Here every
replica is its own System (its own device buffers and HIP stream) on the one
GPU, and run() advances the replicas from concurrent host threads, so their
kernels overlap on the device.  exchange() restates the reference's Metropolis
rule: neighbours (m, m+1) with m alternating between odd and even,
delta = (1/t2 - 1/t1) * (e1/n1 - e2/n2) from the replicas' energy, temperature
and particle-count observables, one uniform draw per pair, accept when
delta <= 0 or exp(-delta) > draw, and on acceptance the two thermostats take
each other's measured temperature.
"""
import math
from concurrent.futures import ThreadPoolExecutor


class MultiSystem:
    """Integrators and observables of several systems, evaluated together
    (src/MultiSystem.py:75-140)."""

    def __init__(self):
        self._integrators = []
        self._potential = []
        self._temperature = []
        self._npart = []
        self._dump = []
        self._current = None

    def beginSystemDefinition(self, n=None):
        self._current = n

    def _slot(self, lst):
        n = self._current if self._current is not None else len(lst)
        while len(lst) <= n:
            lst.append(None)
        return n

    def setIntegrator(self, integrator):
        self._integrators[self._slot(self._integrators)] = integrator

    def setAnalysisPotential(self, potential):
        self._potential[self._slot(self._potential)] = potential

    def setAnalysisTemperature(self, temperature):
        self._temperature[self._slot(self._temperature)] = temperature

    def setAnalysisNPart(self, npart):
        self._npart[self._slot(self._npart)] = npart

    def setDumpConfXYZ(self, dumpconf):
        self._dump[self._slot(self._dump)] = dumpconf

    def runIntegrator(self, niter):
        integ = [i for i in self._integrators if i is not None]
        if len(integ) <= 1:
            for i in integ:
                i.run(niter)
            return
        # each replica owns its stream; the C ABI releases the interpreter lock, so the
        # replicas' kernels run concurrently
        with ThreadPoolExecutor(max_workers=min(len(integ), 16)) as pool:
            for f in [pool.submit(i.run, niter) for i in integ]:
                f.result()

    def runAnalysisPotential(self):
        return [p.computeEnergy() for p in self._potential]

    def runAnalysisTemperature(self):
        return [t.compute() for t in self._temperature]

    def runAnalysisNPart(self):
        return [n.compute() for n in self._npart]

    def runDumpConfXYZ(self):
        for d in self._dump:
            if d is not None:
                d.dump()


class ParallelTempering:
    """ParallelTempering(NumberOfSystems, RNG) (src/ParallelTempering.py:90-200).

    RNG is any object with a random() method returning a uniform in [0, 1)
    (the reference's example passes Python's random module)."""

    def __init__(self, NumberOfSystems=4, RNG=None):
        if RNG is None:
            raise ValueError("ParallelTempering needs a random number generator")
        if NumberOfSystems < 2:
            raise ValueError("Parallel Tempering needs at least 2 systems to be able to work")
        self._RNG = RNG
        self._multisystem = MultiSystem()
        self._nsystems = int(NumberOfSystems)
        self._thermostat = [None] * self._nsystems
        self._oddeven = 0
        self._defining = None
        self.log = []            # (m, n, delta, draw, accepted) of every attempted pair

    def startDefiningSystem(self, n):
        if n not in range(self._nsystems):
            raise ValueError("System number must be between 0 and %d" % self._nsystems)
        self._defining = n
        self._multisystem.beginSystemDefinition(n)

    def endDefiningSystem(self, n):
        if n not in range(self._nsystems):
            raise ValueError("System number must be between 0 and %d" % self._nsystems)
        self._defining = None

    def getNumberOfSystems(self):
        return self._nsystems

    def getNumberOfCPUsPerSystem(self):
        return 1

    def setIntegrator(self, integrator, thermostat):
        self._multisystem.setIntegrator(integrator)
        n = self._defining if self._defining is not None else sum(t is not None for t in self._thermostat)
        self._thermostat[n] = thermostat

    def setAnalysisE(self, analysisE):
        self._multisystem.setAnalysisPotential(analysisE)

    def setAnalysisT(self, analysisT):
        self._multisystem.setAnalysisTemperature(analysisT)

    def setAnalysisNPart(self, analysisNPart):
        self._multisystem.setAnalysisNPart(analysisNPart)

    def setDumpConfXYZ(self, dumpconf):
        self._multisystem.setDumpConfXYZ(dumpconf)

    def runDumpConfXYZ(self):
        self._multisystem.runDumpConfXYZ()

    def run(self, nsteps):
        self._multisystem.runIntegrator(nsteps)

    @staticmethod
    def delta(t1, t2, e1, e2, n1, n2):
        """Exponent of the exchange rule (src/ParallelTempering.py:182)."""
        return (1 / t2 - 1 / t1) * (e1 / n1 - e2 / n2)

    def exchange(self):
        """One round of neighbour swaps (src/ParallelTempering.py:164-200)."""
        self._oddeven = 1 if self._oddeven == 0 else 0
        energies = self._multisystem.runAnalysisPotential()
        temperatures = self._multisystem.runAnalysisTemperature()
        nparts = self._multisystem.runAnalysisNPart()
        swaps = []
        for i in range(len(energies) // 2):
            m = 2 * i + self._oddeven
            n = m + 1
            if n < len(energies):
                metro = self._RNG.random()
                t1, t2 = temperatures[m], temperatures[n]
                e1, e2 = energies[m], energies[n]
                n1, n2 = nparts[m], nparts[n]
                d = self.delta(t1, t2, e1, e2, n1, n2)
                yes = d <= 0 or math.exp(-d) > metro
                self.log.append((m, n, d, metro, yes))
                if yes:
                    self._thermostat[n].temperature = t1
                    self._thermostat[m].temperature = t2
                    swaps.append((m, n))
        return swaps
