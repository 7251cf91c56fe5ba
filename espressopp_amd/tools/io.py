"""XYZ reader/writer (src/tools/DumpConfigurations.py:103-145)."""


def readxyz(filename):
    with open(filename) as f:
        num_particles = int(f.readline().split()[0])
        line = f.readline().split()
        if len(line) == 3:
            Lx, Ly, Lz = float(line[0]), float(line[1]), float(line[2])
        else:
            Lx, Ly, Lz = float(line[0]), float(line[4]), float(line[8])
        pid, type_, xpos, ypos, zpos, xvel, yvel, zvel = [], [], [], [], [], [], [], []
        for _ in range(num_particles):
            line = f.readline().split()
            if len(line) == 7 or len(line) == 4:
                line.insert(1, "0")
            pid.append(int(line[0]))
            type_.append(int(line[1]))
            xpos.append(float(line[2]))
            ypos.append(float(line[3]))
            zpos.append(float(line[4]))
            if len(line) > 5:
                xvel.append(float(line[5]))
                yvel.append(float(line[6]))
                zvel.append(float(line[7]))
            else:
                xvel.append(0.0)
                yvel.append(0.0)
                zvel.append(0.0)
    return pid, type_, xpos, ypos, zpos, xvel, yvel, zvel, Lx, Ly, Lz


def writexyz(filename, system, velocities=True, unfolded=False, append=False):
    st = system.storage.getParticleArrays()
    L = system.bc.boxL
    with open(filename, "a" if append else "w") as f:
        f.write("%d\n" % len(st["id"]))
        f.write("%15.10f %15.10f %15.10f\n" % (L[0], L[1], L[2]))
        for k in range(len(st["id"])):
            p = st["pos"][k]
            if unfolded:
                p = [p[d] + st["image"][k][d] * L[d] for d in range(3)]
            row = "%d %d %15.10f %15.10f %15.10f" % (st["id"][k], st["type"][k], p[0], p[1], p[2])
            if velocities:
                v = st["v"][k]
                row += " %15.10f %15.10f %15.10f" % (v[0], v[1], v[2])
            f.write(row + "\n")
