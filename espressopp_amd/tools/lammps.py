"""tools.lammps.read: particles, bonds and box of a LAMMPS data file
(src/tools/lammps.py:33-156), with the reference's return tuples:

one atom type:   (x, y, z, Lx, Ly, Lz)                     no bonds/angles/dihedrals
                 (x, y, z, Lx, Ly, Lz, vx, vy, vz)          with 'velocities'
                 (bonds, x, y, z, Lx, Ly, Lz)               bonds only
                 (bonds, angles, x, y, z, Lx, Ly, Lz)       bonds and angles
                 (bonds, angles, dihedrals, x, y, z, Lx, Ly, Lz)
several types:   (p_type, bonds, angles, q, x, y, z, Lx, Ly, Lz)
"""


def _next(f, word):
    line = f.readline()
    while word not in line:
        line = f.readline()
        if not line:
            raise ValueError("LAMMPS data file has no '%s' section" % word)
    return line


def read(fin):
    with open(fin) as f:
        line = _next(f, "atoms")
        num_particles = int(line.split()[0])
        num_bonds = int(f.readline().split()[0])
        num_angles = int(f.readline().split()[0])
        num_dihedrals = int(f.readline().split()[0])
        f.readline()                                   # impropers
        f.readline()                                   # blank line
        line = f.readline()                            # atom types (and maybe "velocities")
        num_types = int(line.split()[0])
        velocities = "velocities" in line

        line = _next(f, "xlo")
        xmin, xmax = map(float, line.split()[0:2])
        ymin, ymax = map(float, f.readline().split()[0:2])
        zmin, zmax = map(float, f.readline().split()[0:2])
        Lx, Ly, Lz = xmax - xmin, ymax - ymin, zmax - zmin

        _next(f, "Atoms")
        f.readline()
        x, y, z, p_type, q = [], [], [], [], []
        if num_types == 1:
            rstart = 3 if num_bonds != 0 else 2
            for _ in range(num_particles):
                rx, ry, rz = map(float, f.readline().split()[rstart:])
                x.append(rx); y.append(ry); z.append(rz)
        else:
            for _ in range(num_particles):
                k, rq, rx, ry, rz = map(float, f.readline().split()[2:])
                p_type.append(int(k)); q.append(rq)
                x.append(rx); y.append(ry); z.append(rz)

        bonds, angles, dihedrals = [], [], []
        if num_bonds != 0:
            _next(f, "Bonds")
            f.readline()
            for _ in range(num_bonds):
                _, _, p1, p2 = map(int, f.readline().split())
                bonds.append((p1, p2))
        if num_angles != 0:
            _next(f, "Angles")
            f.readline()
            for _ in range(num_angles):
                _, _, p1, p2, p3 = map(int, f.readline().split())
                angles.append((p1, p2, p3))
        if num_dihedrals != 0:
            _next(f, "Dihedrals")
            f.readline()
            for _ in range(num_dihedrals):
                _, _, p1, p2, p3, p4 = map(int, f.readline().split())
                dihedrals.append((p1, p2, p3, p4))
        vx, vy, vz = [], [], []
        if velocities:
            _next(f, "Velocities")
            f.readline()
            for _ in range(num_particles):
                a, b, c = map(float, f.readline().split()[1:])
                vx.append(a); vy.append(b); vz.append(c)

    if num_types != 1:
        return p_type, bonds, angles, q, x, y, z, Lx, Ly, Lz
    if num_bonds == 0 and num_angles == 0 and num_dihedrals == 0 and not velocities:
        return x, y, z, Lx, Ly, Lz
    if num_bonds == 0 and num_angles == 0 and num_dihedrals == 0 and velocities:
        return x, y, z, Lx, Ly, Lz, vx, vy, vz
    if num_bonds != 0 and num_angles == 0 and num_dihedrals == 0:
        return bonds, x, y, z, Lx, Ly, Lz
    if num_bonds != 0 and num_angles != 0 and num_dihedrals == 0:
        return bonds, angles, x, y, z, Lx, Ly, Lz
    return bonds, angles, dihedrals, x, y, z, Lx, Ly, Lz
