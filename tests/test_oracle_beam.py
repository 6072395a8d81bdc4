"""The oracle's restatement of BeamP (src/beamP.cpp:24-157) and
SMB::inject_beam (src/SMB.cpp:1308-1322), pinned by what the reference's
arithmetic fixes:

- the functional test's beam (test/functional/picopic.json.tmpl: 128 x 256
  cells in 2 x 4 domains, two bunches of length 5.6 mm, radius 5 mm,
  density 1e16 m^-3 at 2.8e8 m/s, dt = 1e-13 s): the macro amount that
  Cfg::weight_macro_amount gives it (src/cfg.cpp:488-570), its per-step
  rate macro / bunches (integer division) x v dt / length, and ceil of that
  rate injected per step, by domain (0, 0) only;
- the injected particles: r uniform on [dr/2, dr/2 + radius), z uniform on
  [dz/2, dz/2 + v dt), v = (0, 0, velocity), weights 2 r / radius times the
  mean macro weight (the total of a bunch's window equals its real
  particle count within the sampling error);
- the bunch windows and the counter: injection while start + k (L + D) / v
  <= t < that + L / v; the counter advances once t passes the window's end,
  and only in domain (0, 0), so a second radial domain reached by a wide
  bunch injects its first bunch only.
"""
import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O

SIZE, GRID, DOMAINS = (0.0375, 0.15), (128, 256), (2, 4)
DT = 1e-13
BEAM = dict(start_time=0.0, radius=0.005, density=1e16, bunches=2, length=0.0056, distance=0.0914, velocity=2.8e8)
SPECIES = [(-1, 1, False), (1, 1836, False), (-1, 1, True)]


def functional_macro_amounts(macro_amount=2.1e5, ratio=2):
    """Cfg::weight_macro_amount for the functional test's configuration."""
    dr, dz = SIZE[0] / GRID[0], SIZE[1] / GRID[1]
    beam_cells = (BEAM["radius"] / dr) * (BEAM["length"] * BEAM["bunches"] / dz)
    norm = macro_amount / (GRID[0] * GRID[1] * 2 + beam_cells * ratio)
    return int(GRID[0] * GRID[1] * norm), int(beam_cells * norm * ratio)


def beam_smb(domains=DOMAINS, size=SIZE, grid=GRID, dt=DT, **kw):
    b = dict(BEAM, **kw)
    s = O.SMB(size, grid, domains, dt, SPECIES)
    plasma, macro = functional_macro_amounts()
    s.set_beam(2, macro, b["start_time"], b["radius"], b["density"], b["bunches"], b["length"], b["distance"],
               b["velocity"], seed=3)
    return s, macro, b


def test_functional_beam_rate_and_placement():
    s, macro, b = beam_smb()
    rate = (macro // b["bunches"]) * b["velocity"] * DT / b["length"]
    per_step = int(np.ceil(rate))
    assert per_step == 6 and macro == 2070
    for k in range(40):
        s.inject()
        s.set_time(s.time() + DT)
        assert s.count(2) == per_step * (k + 1)
    p = s.particles(2)
    dr, dz = SIZE[0] / GRID[0], SIZE[1] / GRID[1]
    assert np.all(p["dom"] == 0)
    assert np.all(p["vr"] == 0) and np.all(p["vphi"] == 0) and np.all(p["vz"] == b["velocity"])
    u = (p["r"] - dr / 2) / b["radius"]
    assert np.all((u >= 0) & (u < 1)) and stats.kstest(u, "uniform").pvalue > 1e-3
    uz = (p["z"] - dz / 2) / (b["velocity"] * DT)
    assert np.all((uz >= 0) & (uz < 1)) and stats.kstest(uz, "uniform").pvalue > 1e-3
    n_total = np.pi * b["radius"] ** 2 * b["length"] * b["bunches"] * b["density"]
    w_avg = n_total / macro
    assert np.allclose(p["w"], w_avg * 2 * p["r"] / b["radius"], rtol=1e-13, atol=0)


def test_bunch_weights_add_up_to_its_real_particles():
    # a short single bunch, injected over its whole window
    s, macro, b = beam_smb(length=40 * 2.8e8 * DT, bunches=1)
    rate = macro * b["velocity"] * DT / b["length"]
    t, steps = 0.0, 0
    for k in range(45):
        s.inject()
        steps += 0.0 <= t < b["length"] / b["velocity"]
        t += DT
        s.set_time(t)
    p = s.particles(2)
    assert len(p["w"]) == steps * int(np.ceil(rate))
    # sum w = real particles x (injected / macro) x E[2 r / radius], r uniform
    # on [dr/2, dr/2 + radius): E = 1 + dr / radius
    real = np.pi * b["radius"] ** 2 * b["length"] * b["density"]
    dr = SIZE[0] / GRID[0]
    expect = real * len(p["w"]) / macro * (1 + dr / b["radius"])
    assert p["w"].sum() == pytest.approx(expect, rel=0.04)


def test_windows_and_the_domain_zero_counter():
    # exact arithmetic: dt = 1, velocity 1, bunch length 3, distance 2; the
    # radius reaches into the second radial domain
    size, grid, domains = (8.0, 8.0), (8, 8), (2, 2)
    s = O.SMB(size, grid, domains, 1.0, SPECIES)
    s.set_beam(2, 4, 0.0, 6.0, 1.0, 2, 3.0, 2.0, 1.0, seed=1)
    counts = []
    for t in range(12):
        n0 = s.count(2)
        s.inject()
        counts.append(s.count(2) - n0)
        s.set_time(s.time() + 1.0)
    p = s.particles(2)
    gen = (p["id"].astype(np.uint64) >> np.uint64(40)) & np.uint64(0xFFFF)
    # rate: macro / bunches = 2; domain 0 covers r 0..4 (share 4/6), domain 2
    # (second radial domain) 2/6: 2 x (4/6) x 1 / 3 and 2 x (2/6) / 3 per step
    per0, per2 = int(np.ceil(2 * (4 / 6) / 3)), int(np.ceil(2 * (2 / 6) / 3))
    # bunch 0 at t = 0, 1, 2 (both domains); bunch 1 at t = 5, 6, 7 (domain 0:
    # domain 2's counter never advances)
    assert counts == [per0 + per2] * 3 + [0, 0] + [per0] * 3 + [0] * 4
    assert (gen == 0).sum() == 6 * per0 and (gen == 2).sum() == 3 * per2
