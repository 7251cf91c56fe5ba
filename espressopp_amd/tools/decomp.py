"""Node and cell grid sizing (src/tools/decomp.py:122,322-359)."""
from .._vec3 import Int3D


class Error(Exception):
    pass


def nodeGridSimple(n):
    best, ijkmax = (1, 1, 1), 3 * n * n + 1
    for i in range(1, n + 1):
        for j in range(i, n + 1):
            for k in range(j, n + 1):
                if i * j * k == n and i * i + j * j + k * k < ijkmax:
                    best, ijkmax = (k, j, i), i * i + j * j + k * k
    return Int3D(*best)


def nodeGrid(n=None, box_size=None, rc=None, skin=None, eh_size=0, ratioMS=0, idealGas=0, slabMSDims=[0, 0, 0]):
    """One GPU owns the box: the node grid of one rank is (1, 1, 1)."""
    return nodeGridSimple(int(n) if n else 1)


def cellGrid(box_size, node_grid, rc, skin, halfCellInt=1):
    rc_skin = rc + skin
    if rc_skin == 0:
        raise Error("interaction range (cutoff + skin) must be larger than 0")
    if node_grid[0] <= 0 or node_grid[1] <= 0 or node_grid[2] <= 0:
        raise Error("invalid node grid %s" % str(node_grid))
    ixyz = []
    for d in range(3):
        v = (box_size[d] * halfCellInt) / (rc_skin * node_grid[d])
        if v < 1:
            raise Error("local box size in direction %d (=%6f) is smaller than interaction range "
                        "(cutoff + skin = %6f)" % (d, v, rc_skin))
        ixyz.append(v)
    return Int3D(*ixyz)
