"""espressopp.bc.OrthorhombicBC (src/bc/OrthorhombicBC.cpp:35-161)."""
import math

from ._vec3 import Real3D, toReal3DFromVector


class OrthorhombicBC:
    def __init__(self, rng, boxL):
        self.rng = rng
        self.boxL = toReal3DFromVector(boxL)

    def getBoxL(self):
        return Real3D(self.boxL)

    def getRandomPos(self):
        """A uniform position in the box: boxL * (rng(), rng(), rng())
        (OrthorhombicBC::getRandomPos, src/bc/OrthorhombicBC.cpp:171-178)."""
        return Real3D([self.boxL[k] * self.rng() for k in range(3)])

    def getFoldedPosition(self, pos, imageBox=None):
        """foldCoordinate for x, y, z (src/bc/OrthorhombicBC.cpp:144-161)."""
        p = [float(x) for x in pos]
        img = [0, 0, 0] if imageBox is None else [int(x) for x in imageBox]
        for d in range(3):
            tmp = int(math.floor(p[d] * (1.0 / self.boxL[d])))
            img[d] += tmp
            p[d] -= tmp * self.boxL[d]
        return Real3D(p), img

    def getUnfoldedPosition(self, pos, imageBox):
        return Real3D([pos[d] + imageBox[d] * self.boxL[d] for d in range(3)])

    def getMinimumImageVector(self, pos1, pos2):
        """getMinimumImageVectorBox (src/bc/OrthorhombicBC.cpp:86-105)."""
        d = [pos1[k] - pos2[k] for k in range(3)]
        for k in range(3):
            L = self.boxL[k]
            if d[k] < -0.5 * L:
                d[k] += L
            elif d[k] > 0.5 * L:
                d[k] -= L
        return Real3D(d)
