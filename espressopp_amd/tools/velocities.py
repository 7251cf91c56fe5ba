"""Maxwell-Boltzmann velocities rescaled to exactly T (src/tools/velocities.py:27-110)."""
import random


def gaussian(T, N, particle_mass=None, zero_momentum=True, seed=7654321, kb=1.0):
    random.seed(seed)
    boltzmann_factor = kb * T
    E_kin = 0.0
    vx, vy, vz = [], [], []
    for i in range(N):
        mass = particle_mass[i] if particle_mass else 1.0
        sd = (boltzmann_factor / mass) ** 0.5
        v_x = sd * random.gauss(0.0, 1.0)
        E_kin += 0.5 * mass * v_x * v_x
        vx.append(v_x)
        v_y = sd * random.gauss(0.0, 1.0)
        E_kin += 0.5 * mass * v_y * v_y
        vy.append(v_y)
        v_z = sd * random.gauss(0.0, 1.0)
        E_kin += 0.5 * mass * v_z * v_z
        vz.append(v_z)
    local_temp = (2.0 * E_kin) / (3 * N * kb)
    if local_temp > 0:
        temp_scale = (T / local_temp) ** 0.5
        for i in range(N):
            vx[i] *= temp_scale
            vy[i] *= temp_scale
            vz[i] *= temp_scale
    if zero_momentum:
        sumvx = sum(vx) / N
        sumvy = sum(vy) / N
        sumvz = sum(vz) / N
        for i in range(N):
            vx[i] -= sumvx
            vy[i] -= sumvy
            vz[i] -= sumvz
    return vx, vy, vz
