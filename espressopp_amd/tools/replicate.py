"""Replicate a configuration xdim x ydim x zdim times (src/tools/replicate.py:26-80)."""


def replicate(bonds, angles, x, y, z, Lx, Ly, Lz, xdim=1, ydim=1, zdim=1):
    x_replicated, y_replicated, z_replicated = x[:], y[:], z[:]
    bonds_replicated, angles_replicated = bonds[:], angles[:]
    num_particles_original = len(x)
    ct = 0
    for i in range(xdim):
        for j in range(ydim):
            for k in range(zdim):
                if i + j + k != 0:
                    ct += 1
                    for x_, y_, z_ in zip(x, y, z):
                        x_replicated.append(x_ + i * Lx)
                        y_replicated.append(y_ + j * Ly)
                        z_replicated.append(z_ + k * Lz)
                    for p1, p2 in bonds:
                        bonds_replicated.append((p1 + ct * num_particles_original, p2 + ct * num_particles_original))
                    for p1, p2, p3 in angles:
                        angles_replicated.append((p1 + ct * num_particles_original, p2 + ct * num_particles_original,
                                                  p3 + ct * num_particles_original))
    return (bonds_replicated, angles_replicated, x_replicated, y_replicated, z_replicated,
            xdim * Lx, ydim * Ly, zdim * Lz)
