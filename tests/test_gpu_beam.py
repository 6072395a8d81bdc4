"""Beam injection on the device (pcp_set_beam / pcp_inject / pcp_step) against
the oracle's BeamP restatement (src/beamP.cpp:24-157, src/SMB.cpp:1308-1322),
through the C ABI.

- the bunch windows and the domain (0, 0) counter (a radius reaching into the
  second radial domain, exact arithmetic): the same particles injected at
  the same times, positions and weights bit for bit;
- the reference's functional configuration (test/functional/picopic.json.tmpl:
  128 x 256 cells in 2 x 4 domains, PML, Boris + Villasenor-Buneman, the
  reference's initial plasma and the 2.8e8 m/s beam): 12 whole steps in FP64
  agree with the oracle within 1e-12 on every particle (the beam's strict
  axial motions exercise the order-dependent J_r reset) and 1e-11 on the
  fields; the FP32 split path injects the same particles.
"""
import numpy as np
import pytest

from oracle import oracle as O
from picopic_amd import capi
from tests import helpers as Hh

pytestmark = pytest.mark.gpu

SPECIES = [(-1, 1, False), (1, 1836, False), (-1, 1, True)]
SIZE, GRID, DOMAINS = (0.0375, 0.15), (128, 256), (2, 4)
DT = 1e-13
PML = (0.0, 0.05, 0.2, 1e-5, 7e-2)
BEAM = (2070, 0.0, 0.005, 1e16, 2, 0.0056, 0.0914, 2.8e8)
PLASMA_MACRO = 103964


def by_id(p):
    o = np.argsort(p["id"], kind="stable")
    return {k: v[o] for k, v in p.items()}


def test_windows_and_counter_match_the_oracle():
    size, grid, domains = (8.0, 8.0), (8, 8), (2, 2)
    orc = O.SMB(size, grid, domains, 1.0, SPECIES)
    dev = capi.Context(size, grid, domains, 1.0, [s + (256,) for s in SPECIES])
    for s in (orc, dev):
        s.set_beam(2, 4, 0.0, 6.0, 1.0, 2, 3.0, 2.0, 1.0, seed=1)
    for t in range(12):
        orc.inject()
        dev.inject()
        assert orc.count(2) == dev.count(2), t
        orc.set_time(orc.time() + 1.0)
        dev.set_time(dev.time() + 1.0)
    a, b = by_id(orc.particles(2)), by_id(dev.particles(2))
    assert np.array_equal(a["id"], b["id"])
    for k in ("r", "z", "vr", "vphi", "vz", "w"):
        assert np.array_equal(a[k], b[k]), k


def functional(precision):
    orc = O.SMB(SIZE, GRID, DOMAINS, DT, SPECIES, pml=PML)
    dev = capi.Context(SIZE, GRID, DOMAINS, DT, [(-1, 1, False, PLASMA_MACRO + 64), (1, 1836, False, PLASMA_MACRO + 64),
                                                 (-1, 1, True, 4096)], pml=PML, precision=precision)
    for s in (orc, dev):
        s.init_reference(0, PLASMA_MACRO, 1e17, 1e17, 1.0, seed=11)
        s.init_reference(1, PLASMA_MACRO, 1e17, 1e17, 0.1, seed=12)
        s.set_beam(2, *BEAM, seed=13)
    return orc, dev


def test_functional_configuration_steps_match_the_oracle_fp64():
    orc, dev = functional("fp64")
    O.set_threads(8)
    orc.step(12)
    dev.step(12)
    assert dev.time() == orc.time()
    assert orc.count(2) == dev.count(2) == 12 * 6
    for sp in range(3):
        a, b = by_id(orc.particles(sp)), by_id(dev.particles(sp))
        assert np.array_equal(a["id"], b["id"]), sp
        for k in ("r", "z", "vr", "vphi", "vz", "w"):
            assert Hh.relerr(b[k], a[k]) <= 1e-12, (sp, k)
    for g in ("E", "H", "J"):
        e = Hh.relerr(dev.get_grid(g)[:, :, 2:-2, 2:-2], orc.get_grid(g)[:, :, 2:-2, 2:-2])
        print(g, "rel err", e)
        assert e <= 1e-11, g


def test_fp32_split_injects_the_same_particles():
    orc, dev = functional("fp32")
    for s in (orc, dev):
        for k in range(3):
            s.inject()
            s.set_time(s.time() + DT)
    a, b = by_id(orc.particles(2)), by_id(dev.particles(2))
    assert np.array_equal(a["id"], b["id"])
    dr, dz = SIZE[0] / GRID[0], SIZE[1] / GRID[1]
    assert np.abs(a["r"] - b["r"]).max() <= 1e-6 * dr
    assert np.abs(a["z"] - b["z"]).max() <= 1e-6 * dz
    for k in ("vz", "w"):
        assert Hh.relerr(b[k], a[k]) <= 1e-6, k
