"""FP64 parity over many steps at the uniform-plasma configuration's full size.

Electron-proton plasma, n = 1e17 m^-3, 128 (r) x 512 (z) cells of 1 mm, 16
macro-particles per cell and species (2,097,152 particles), positions uniform
in z and r^2, Maxwellian Te = 1 eV, Ti = 0.1 eV, relativistic Boris +
Villasenor-Buneman + Yee, dt = 0.5 dz / c, conducting walls, 200 steps from
zero fields (src/PiCoPiC.cpp:299-345).  The block is split into 4 x 8 domains
so that the overlay and the migration run every step.

The device and the oracle start from the same seeded particles.  Every 10
steps the largest difference of every particle's r, z, v_r, v_phi, v_z
(matched by id, relative to the largest value of that quantity) and of the
fields is printed; that error curve is the record of how the reordered
atomic sums of the deposit grow through the dynamics.  The bound is 1e-12
relative for positions and momenta at every sample.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as Hh

pytestmark = pytest.mark.gpu

EL_CHARGE = 1.6e-19
EL_MASS = 9.1e-31


def uniform_plasma(rng, size, grid, ppc, density, temp_ev, mass_me):
    R, L = size
    n = ppc * grid[0] * grid[1]
    r = np.sqrt(rng.uniform(0.0, R * R, n))
    dr = R / grid[0]
    dz = L / grid[1]
    r = np.clip(r, 0.5 * dr, R - 0.5 * dr)
    z = rng.uniform(0.5 * dz, L - 0.5 * dz, n)
    vth = np.sqrt(temp_ev * EL_CHARGE / (mass_me * EL_MASS))
    v = rng.normal(0.0, vth, (3, n))
    w = np.full(n, density * np.pi * R * R * L / n)
    return r, z, v[0], v[1], v[2], w


def test_uniform_plasma_200_steps_fp64():
    from picopic_amd import capi
    import os
    grid = (128, 512)
    size = (grid[0] * 1e-3, grid[1] * 1e-3)
    dt = 0.5 * 1e-3 / 3e8
    domains = (4, 8)
    species = [(-1, 1, False), (1, 1836, False)]
    ppc = 16
    n = ppc * grid[0] * grid[1]
    orc = O.SMB(size, grid, domains, dt, species, pusher="boris-adaptive", ccs="vb")
    gpu = capi.Context(size, grid, domains, dt, [s + (n + n // 8,) for s in species], pusher="boris-adaptive",
                       ccs="vb")
    rng = np.random.default_rng(2024)
    for sp, (q, m, b) in enumerate(species):
        r, z, vr, vp, vz, w = uniform_plasma(rng, size, grid, ppc, 1e17, 1.0 if m == 1 else 0.1, m)
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 8)
        orc.add_particles(sp, r, z, vr, vp, vz, w, ids)
        gpu.upload(sp, r, z, vr, vp, vz, w, ids)
        del r, z, vr, vp, vz, w
    O.set_threads(min(16, os.cpu_count() or 1))
    curve = []
    worst = 0.0
    for block in range(20):
        orc.step(10)
        gpu.step(10)
        row = {"step": 10 * (block + 1)}
        for sp in range(len(species)):
            a = orc.particles(sp)
            b = gpu.particles(sp)
            ia, ib = np.argsort(a["id"]), np.argsort(b["id"])
            assert np.array_equal(a["id"][ia], b["id"][ib]), "particle sets differ at step %d" % row["step"]
            for k in ("r", "z", "vr", "vphi", "vz"):
                e = Hh.relerr(b[k][ib], a[k][ia])
                row["%d.%s" % (sp, k)] = e
                worst = max(worst, e)
            del a, b
        for gname in ("E", "H", "J"):
            ga = orc.get_grid(gname)[:, :, 2:-2, 2:-2]
            gb = gpu.get_grid(gname)[:, :, 2:-2, 2:-2]
            row[gname] = Hh.relerr(gb, ga)
        curve.append(row)
        print("step %3d  particles max rel err %.3e  E %.3e  H %.3e  J %.3e" % (
            row["step"], max(v for k, v in row.items() if "." in k), row["E"], row["H"], row["J"]), flush=True)
    assert worst <= 1e-12, "positions/momenta diverge: max rel err %.3e" % worst
