"""The reference's initial plasma on the device (pcp_init_reference) against
the oracle's restatement (src/specieP.cpp:79-189, 280-331, 346-384;
src/math/maxwellJuettner.cpp:182-255), through the C ABI.

- FP64: every particle's position and weight equal the oracle's bit for bit
  (the same radical-inverse arithmetic; digit powers from the host's libm),
  velocities within 1e-12 (the bisections call the device's erf/exp/log1p);
- the device's storage order (domain, order key) is the oracle's vector
  order after the collector moved the particles placed past inner domain
  edges (the migration pass gives them their keys);
- FP32 split: the same particles to FP32 precision;
- five steps of the functional test's physics (Boris, Villasenor-Buneman,
  PML) from this plasma stay within 1e-12 of the oracle on every particle
  and 1e-10 on the grids (FP64).
"""
import numpy as np
import pytest

from oracle import oracle as O
from picopic_amd import capi
from tests import helpers as Hh

pytestmark = pytest.mark.gpu

GRID, DOMAINS = (32, 64), (2, 4)
SIZE = (0.0375 / 4, 0.15 / 4)
DT = 1e-13
SPECIES = [(-1, 1, False), (1, 1836, False)]
MACRO = 60001
PML = (0.0, 0.05, 0.2, 1e-5, 7e-2)


def contexts(precision, pml=None):
    orc = O.SMB(SIZE, GRID, DOMAINS, DT, SPECIES, pml=pml)
    dev = capi.Context(SIZE, GRID, DOMAINS, DT, [s + (MACRO + 64,) for s in SPECIES], pml=pml, precision=precision)
    for sp, (temp, dens) in enumerate(((1.0, (1e17, 1e17)), (0.1, (1e17, 1e17)))):
        orc.init_reference(sp, MACRO, dens[0], dens[1], temp, seed=40 + sp)
        dev.init_reference(sp, MACRO, dens[0], dens[1], temp, seed=40 + sp)
    return orc, dev


def by_id(p):
    o = np.argsort(p["id"], kind="stable")
    return {k: v[o] for k, v in p.items()}


def test_fp64_matches_the_oracle_bitwise_and_in_order():
    orc, dev = contexts("fp64")
    for sp in range(2):
        a, b = orc.particles(sp), dev.particles(sp)
        assert len(a["id"]) == len(b["id"]) == (MACRO // DOMAINS[0] // DOMAINS[1]) * DOMAINS[0] * DOMAINS[1]
        # storage order: device (domain, key) order = oracle vector order
        assert np.array_equal(a["id"], b["id"][np.argsort(b["order"], kind="stable")])
        a, b = by_id(a), by_id(b)
        for k in ("r", "z", "w"):
            assert np.array_equal(a[k], b[k]), k
        for k in ("vr", "vphi", "vz"):
            assert Hh.relerr(b[k], a[k]) <= 1e-12, k


def test_fp32_split_matches_the_oracle():
    orc, dev = contexts("fp32")
    for sp in range(2):
        a, b = by_id(orc.particles(sp)), by_id(dev.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        assert np.abs(b["r"] - a["r"]).max() <= 1e-6 * SIZE[0] / GRID[0]
        assert np.abs(b["z"] - a["z"]).max() <= 1e-6 * SIZE[1] / GRID[1]
        for k in ("vr", "vphi", "vz", "w"):
            assert Hh.relerr(b[k], a[k]) <= 1e-6, k


def test_steps_from_the_reference_plasma_match_the_oracle():
    orc, dev = contexts("fp64", pml=PML)
    O.set_threads(8)
    orc.step(5)
    dev.step(5)
    for sp in range(2):
        a, b = by_id(orc.particles(sp)), by_id(dev.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        for k in ("r", "z", "vr", "vphi", "vz"):
            assert Hh.relerr(b[k], a[k]) <= 1e-12, k
    # the grids: current sums reordered by the atomics, and J cancels between
    # electrons and ions (1e-10 of its largest value)
    for g in ("E", "H", "J"):
        assert Hh.relerr(dev.get_grid(g)[:, :, 2:-2, 2:-2], orc.get_grid(g)[:, :, 2:-2, 2:-2]) <= 1e-10, g
