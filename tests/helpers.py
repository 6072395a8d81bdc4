"""Shared set-ups for the parity tests: seeded particles and fields fed
identically to the oracle and the HIP library."""
import numpy as np

EL_CHARGE = 1.6e-19
EL_MASS = 9.1e-31
C = 3e8


def geometry(nr=16, nz=32, dr=1e-3, dz=1e-3):
    return (nr * dr, nz * dz), (nr, nz)


def species_list(beam=False):
    sp = [(-1, 1, False), (1, 1836, False)]
    if beam:
        sp.append((-1, 1, True))
    return sp


def make_particles(rng, size, grid, n, mass_me, fast_fraction=0.3, vmax=2.5e8, temp_ev=1.0, special=True):
    """Positions inside the area (r >= dr/2), a thermal core and a fast tail.
    With `special`, a few particles have v_r = 0 exactly (strict axial motion)
    or v_z = 0 (strict radial motion), and some sit next to the walls."""
    R, L = size
    dr, dz = R / grid[0], L / grid[1]
    r = np.sqrt(rng.uniform((0.5 * dr) ** 2, (R - 0.5 * dr) ** 2, n))
    z = rng.uniform(0.5 * dz, L - 0.5 * dz, n)
    vth = np.sqrt(temp_ev * EL_CHARGE / (mass_me * EL_MASS))
    v = rng.normal(0.0, vth, (3, n))
    nf = int(fast_fraction * n)
    if nf:
        dirs = rng.normal(size=(3, nf))
        dirs /= np.linalg.norm(dirs, axis=0)
        speed = rng.uniform(1e6, vmax, nf)
        v[:, :nf] = dirs * speed
    if special and n >= 40:
        k = n // 20
        v[0, nf:nf + k] = 0.0            # strict axial motion (no radial displacement:
        v[1, nf:nf + k] = 0.0            # the rz rotation moves r unless v_phi = 0 too)
        v[2, nf + k:nf + 2 * k] = 0.0    # strict radial motion
        # next to the walls, heading out
        m = nf + 2 * k
        r[m:m + k] = R - 0.5 * dr - rng.uniform(0, 2e-5, k)
        v[0, m:m + k] = np.abs(v[0, m:m + k]) + 1e7
        z[m + k:m + 2 * k] = 0.5 * dz + rng.uniform(0, 2e-5, k)
        v[2, m + k:m + 2 * k] = -np.abs(v[2, m + k:m + 2 * k]) - 1e7
    w = rng.uniform(0.5, 1.5, n) * 1e7
    return r, z, v[0], v[1], v[2], w


def random_fields(rng, D, nr, nz, scale_e=2e4, scale_h=30.0):
    shape = (D, 3, nr + 4, nz + 4)
    E = rng.normal(0, scale_e, shape)
    H = rng.normal(0, scale_h, shape)
    Hat = rng.normal(0, scale_h, shape)
    return E, H, Hat


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale
