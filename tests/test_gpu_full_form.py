"""The full-form configuration at its real size, through properties that do
not depend on it: 1024 x 8192 cells in 8 x 8 domains of 128 x 1024 cells,
an electron-proton plasma with 32 macro-particles per cell and species at
the peak and a 0.1 n -> n ramp along z (about 3.0e8 particles), the
Gaussian electron bunch of the wakefield case, Higuera-Cary pusher, ZigZag
deposition.

- FP32 split and FP64 plain contexts built from the same seeded device
  generators hold the same particles, and after three steps their total
  field and kinetic energies agree to FP32 accuracy, with no device error;
- a collision pass with the surrogate (cells outside the network's range,
  here all of them at n = 1e17, take the P12 fallback) keeps the total
  momentum and the particle counts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0
GRID = (1024, 8192)
DOMAINS = (8, 8)
PPC = 32
DENSITY = 1e17
BUNCH = dict(n=1 << 20, n_peak=1e16, sigma_r=2e-3, sigma_z=5e-3, z0=0.1, gamma=5.0)


def _context(precision):
    from picopic_amd import capi
    dr = dz = 1e-3
    dt = 0.5 * dz / C_LIGHT
    n = PPC * GRID[0] * GRID[1]
    species = [(-1, 1, False, n + n // 50), (1, 1836, False, n + n // 50), (-1, 1, True, BUNCH["n"])]
    ctx = capi.Context((GRID[0] * dr, GRID[1] * dz), GRID, DOMAINS, dt, species, pusher="hc", ccs="zigzag",
                       precision=precision)
    ctx.init_plasma(0, PPC, 0.1 * DENSITY, DENSITY, 1.0, 21)
    ctx.init_plasma(1, PPC, 0.1 * DENSITY, DENSITY, 0.1, 22)
    ctx.init_bunch(2, BUNCH["n"], BUNCH["n_peak"], BUNCH["sigma_r"], BUNCH["sigma_z"], BUNCH["z0"], BUNCH["gamma"],
                   23)
    return ctx


def test_full_form_fp32_and_fp64_energies_agree():
    res = {}
    for precision in ("fp64", "fp32"):
        print("full form", precision, flush=True)
        ctx = _context(precision)
        counts = [ctx.count(s) for s in range(3)]
        ctx.step(3)
        ctx.sync()
        res[precision] = dict(counts=counts, counts_after=[ctx.count(s) for s in range(3)],
                              field=ctx.field_energy(), kin=[ctx.kinetic_energy(s) for s in range(3)])
        print(precision, res[precision], flush=True)
        ctx.close()
    a, b = res["fp64"], res["fp32"]
    assert a["counts"] == b["counts"]
    assert sum(a["counts"]) > 2.5e8
    assert a["counts_after"] == b["counts_after"]
    assert a["field"] > 0
    assert abs(b["field"] - a["field"]) <= 1e-5 * a["field"]
    for s in range(3):
        assert abs(b["kin"][s] - a["kin"][s]) <= 1e-5 * a["kin"][s]


def test_full_form_collision_pass_keeps_momentum_and_counts():
    from tests.test_gpu_surrogate import trained
    params, norm, _, _ = trained()
    ctx = _context("fp32")
    ctx.mlp_set_params(params)
    ctx.surrogate_set_norm(norm)
    m0 = [ctx.moments(s) for s in range(2)]
    n0 = [ctx.count(s) for s in range(2)]
    ctx.collide("surrogate", 0, 1, 9, 0)
    fb = ctx.surrogate_fallback_cells()
    m1 = [ctx.moments(s) for s in range(2)]
    print("fallback cells", fb, flush=True)
    assert fb > 0
    assert [ctx.count(s) for s in range(2)] == n0
    # electron-ion collisions move momentum between the species; their sum
    # stays, against the thermal momentum scale sum_s m_s sqrt(sum w v^2 sum w)
    mass = (9.1e-31, 1836 * 9.1e-31)
    dp = sum(mass[s] * (m1[s][1:4] - m0[s][1:4]) for s in range(2))
    scale = sum(mass[s] * np.sqrt(m0[s][4] + m0[s][5]) * np.sqrt(m0[s][0]) for s in range(2))
    moved = max(np.abs(mass[s] * (m1[s][1:4] - m0[s][1:4])).max() for s in range(2))
    print("total momentum change", dp, "moved between species", moved, "scale", scale)
    assert moved > 0
    assert np.abs(dp).max() <= 1e-7 * scale
