"""Species maps on the device (pcp_species_map) against the oracle's
restatement (src/specieP.cpp:665-839, include/algo/weighter.hpp:27-132,
src/outController.cpp:409-491), through the C ABI: density, temperature
(before any energy map, reading the zero momentum_abs map as the reference
does, and after one), energy and momentum of the reference's initial plasma
after a few steps over 2 x 4 domains, on every interior node -- within 1e-12
of each map's largest value in FP64 (the atomics reorder the sums), 1e-5 in
FP32 (the particles' positions and velocities are FP32) -- and a map probe
window equal to the downloaded map's."""
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
    tol = 1e-12 if precision == "fp64" else 1e-5
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
