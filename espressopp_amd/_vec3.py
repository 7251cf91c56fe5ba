"""Real3D and Int3D value types with the reference's Python behaviour
(src/Real3D.py:79-160, src/Int3D.py:60-140)."""
import math
import numbers


class Real3D:
    __slots__ = ("_v",)

    def __init__(self, *args):
        if len(args) == 0:
            v = (0.0, 0.0, 0.0)
        elif len(args) == 1:
            a = args[0]
            if isinstance(a, (Real3D, Int3D)):
                v = tuple(float(x) for x in a)
            elif hasattr(a, "__iter__"):
                v = tuple(float(x) for x in a)
                if len(v) != 3:
                    raise TypeError("Cannot initialize Real3D from %s" % (args,))
            elif isinstance(a, numbers.Number):
                v = (float(a),) * 3
            else:
                raise TypeError("Cannot initialize Real3D from %s" % (args,))
        elif len(args) == 3:
            v = tuple(float(x) for x in args)
        else:
            raise TypeError("Cannot initialize Real3D from %s" % (args,))
        self._v = list(v)

    def __getitem__(self, i):
        return self._v[i]

    def __setitem__(self, i, val):
        self._v[i] = float(val)

    def __len__(self):
        return 3

    def __iter__(self):
        return iter(self._v)

    x = property(lambda s: s._v[0], lambda s, v: s.__setitem__(0, v))
    y = property(lambda s: s._v[1], lambda s, v: s.__setitem__(1, v))
    z = property(lambda s: s._v[2], lambda s, v: s.__setitem__(2, v))

    def __add__(self, o):
        return Real3D(self._v[0] + o[0], self._v[1] + o[1], self._v[2] + o[2])

    def __sub__(self, o):
        return Real3D(self._v[0] - o[0], self._v[1] - o[1], self._v[2] - o[2])

    def __mul__(self, o):
        if isinstance(o, numbers.Number):
            return Real3D(self._v[0] * o, self._v[1] * o, self._v[2] * o)
        return self._v[0] * o[0] + self._v[1] * o[1] + self._v[2] * o[2]

    __rmul__ = __mul__

    def __truediv__(self, o):
        return Real3D(self._v[0] / o, self._v[1] / o, self._v[2] / o)

    def __neg__(self):
        return Real3D(-self._v[0], -self._v[1], -self._v[2])

    def __eq__(self, o):
        if o is None:
            return False
        if isinstance(o, numbers.Number):
            return all(x == o for x in self._v)
        return all(self._v[i] == o[i] for i in range(3))

    def __ne__(self, o):
        return not self.__eq__(o)

    def __hash__(self):
        return hash(tuple(self._v))

    def sqr(self):
        return self._v[0] * self._v[0] + self._v[1] * self._v[1] + self._v[2] * self._v[2]

    def abs(self):
        return math.sqrt(self.sqr())

    def __repr__(self):
        return "Real3D(%r, %r, %r)" % tuple(self._v)

    __str__ = __repr__


class Int3D:
    __slots__ = ("_v",)

    def __init__(self, *args):
        if len(args) == 0:
            v = (0, 0, 0)
        elif len(args) == 1:
            a = args[0]
            if isinstance(a, Int3D):
                v = tuple(a)
            elif hasattr(a, "__iter__"):
                v = tuple(int(x) for x in a)
                if len(v) != 3:
                    raise TypeError("Cannot initialize Int3D from %s" % (args,))
            elif isinstance(a, numbers.Number):
                v = (int(a),) * 3
            else:
                raise TypeError("Cannot initialize Int3D from %s" % (args,))
        elif len(args) == 3:
            v = tuple(int(x) for x in args)   # float arguments truncate
        else:
            raise TypeError("Cannot initialize Int3D from %s" % (args,))
        self._v = list(v)

    def __getitem__(self, i):
        return self._v[i]

    def __setitem__(self, i, val):
        self._v[i] = int(val)

    def __len__(self):
        return 3

    def __iter__(self):
        return iter(self._v)

    x = property(lambda s: s._v[0], lambda s, v: s.__setitem__(0, v))
    y = property(lambda s: s._v[1], lambda s, v: s.__setitem__(1, v))
    z = property(lambda s: s._v[2], lambda s, v: s.__setitem__(2, v))

    def __eq__(self, o):
        if o is None:
            return False
        if isinstance(o, numbers.Number):
            return all(x == o for x in self._v)
        return all(self._v[i] == o[i] for i in range(3))

    def __ne__(self, o):
        return not self.__eq__(o)

    def __hash__(self):
        return hash(tuple(self._v))

    def __repr__(self):
        return "Int3D(%d, %d, %d)" % tuple(self._v)

    __str__ = __repr__


def toReal3DFromVector(*args):
    if len(args) == 1 and isinstance(args[0], Real3D):
        return args[0]
    return Real3D(*args)


def toInt3DFromVector(*args):
    if len(args) == 1 and isinstance(args[0], Int3D):
        return args[0]
    return Int3D(*args)
