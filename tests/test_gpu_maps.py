"""Species maps on the device (pcp_species_map) against the oracle's
restatement (src/specieP.cpp:665-839, include/algo/weighter.hpp:27-132,
src/outController.cpp:409-491), through the C ABI: density, temperature
(before any energy map, reading the zero momentum_abs map as the reference
does, and after one), energy and momentum of the reference's initial plasma
after a few steps over 2 x 4 domains, on every interior node, within 1e-12
of each map's largest value (the atomics reorder the sums).

An FP64 context is compared with the oracle run alongside it.  An FP32
context is compared with an oracle that holds the device's own particles
(positions rebuilt from cell + offset in FP64 as the map kernel rebuilds
them): the ratio maps (temperature, energy, momentum) at nodes that only the
edges of a few clouds reach depend on overlaps of a small fraction of a cell,
whose relative error from the offsets' 2^-24-cell resolution is then far
above 2^-24, so the FP32 run's own evolution is checked on the density map
alone (1e-5).  A map probe window equals the downloaded map's."""
import numpy as np
import pytest

from oracle import oracle as O
from picopic_amd import capi

pytestmark = pytest.mark.gpu

SIZE, GRID, DOMAINS = (0.0375 / 4, 0.15 / 4), (32, 64), (2, 4)
SPECIES = [(-1, 1, False), (1, 1836, False)]
MACRO = 40000


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_maps_match_the_oracle(precision):
    orc = O.SMB(SIZE, GRID, DOMAINS, 1e-13, SPECIES)
    dev = capi.Context(SIZE, GRID, DOMAINS, 1e-13, [s + (MACRO + 64,) for s in SPECIES], precision=precision)
    for s in (orc, dev):
        s.init_reference(0, MACRO, 1e17, 1e17, 1.0, seed=3)
        s.init_reference(1, MACRO, 1e17, 1e17, 0.1, seed=4)
    O.set_threads(8)
    orc.step(3)
    dev.step(3)
    if precision == "fp32":
        for sp in range(2):
            a = orc.species_map(sp, "density")[:, :, 2:-2, 2:-2]
            b = dev.species_map(sp, "density")[:, :, 2:-2, 2:-2]
            err = np.abs(a - b).max() / np.abs(a).max()
            print("fp32 evolution", sp, "density max rel err %.2e" % err)
            assert err <= 1e-5
        orc = O.SMB(SIZE, GRID, DOMAINS, 1e-13, SPECIES)
        for sp in range(2):
            p = dev.particles(sp)
            orc.add_particles(sp, p["r"], p["z"], p["vr"], p["vphi"], p["vz"], p["w"], p["id"])
    tol = 1e-12
    for sp in range(2):
        for kind in ("density", "temperature", "energy", "temperature", "momentum"):
            a = orc.species_map(sp, kind)[:, :, 2:-2, 2:-2]
            b = dev.species_map(sp, kind)[:, :, 2:-2, 2:-2]
            for c in range(a.shape[1]):
                scale = np.abs(a[:, c]).max()
                err = np.abs(a[:, c] - b[:, c]).max() / scale
                print(precision, sp, kind, c, "max rel err %.2e" % err)
                assert err <= tol, (sp, kind, c)
    full = dev.species_map(0, "momentum")
    win = dev.probe("map", 2, 3, 5, 29, 61)
    nrd, nzd = GRID[0] // DOMAINS[0], GRID[1] // DOMAINS[1]
    for i in range(3, 29):
        for j in range(5, 61):
            d = (i // nrd) * DOMAINS[1] + j // nzd
            assert win[i - 3, j - 5] == full[d, 2, 2 + i % nrd, 2 + j % nzd]
