"""The full-form configuration at its real size, through properties that do
not depend on it: 1024 x 8192 cells in 8 x 8 domains of 128 x 1024 cells, a
uniform electron-proton plasma (n = 1e17 m^-3) with 32 macro-particles per
cell and species (536,870,912 particles), the Gaussian electron bunch of the
wakefield case, Higuera-Cary pusher, ZigZag deposition, electron-ion
collisions in every step -- the collision surrogate (trained for the plasma's
regime) fused into the FP32 split step, P12 in the FP64 step.

- FP32 split and FP64 plain contexts built from the same seeded device
  generators hold the same particles; over three steps the field energy and
  the total kinetic energy logged on the device in every step agree to
  1e-5, with no particle lost and no device error.  Each species' kinetic
  energy agrees to 1e-4: the two contexts collide differently (the
  surrogate's per-cell rescale against P12 pairs), which moves about 1e-5
  of the ions' energy to or from the electrons;
- the surrogate covers every cell that holds at least 8 macro-particles of
  each species (the network's minimum); the plasma generator puts
  floor(n_col (i+1)^2 / Nr^2) - floor(n_col i^2 / Nr^2) particles into radial
  cell i (uniform in r^2, equal weights, n_col = 32 x 1024 per column), so
  the cells of the innermost ~120 rows hold 1..7 and take the P12 fallback.
  The fallback count after three steps equals the count of such cells in
  the initial placement within 1 % (thermal motion moves a few particles
  across cell boundaries).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C_LIGHT = 3e8
GRID = (1024, 8192)
DOMAINS = (8, 8)
PPC = 32
DENSITY = 1e17
BUNCH = dict(n=1 << 20, n_peak=1e16, sigma_r=2e-3, sigma_z=5e-3, z0=0.1, gamma=5.0)
REGIME = dict(dens=(2e16, 5e17), tepar=(0.3, 3.0), teperp=(0.3, 3.0), ti=(0.03, 0.3))


def _context(precision, net):
    from picopic_amd import capi, surrogate
    dr = dz = 1e-3
    dt = 0.5 * dz / C_LIGHT
    n = PPC * GRID[0] * GRID[1]
    species = [(-1, 1, False, n + n // 50), (1, 1836, False, n + n // 50), (-1, 1, True, BUNCH["n"])]
    ctx = capi.Context((GRID[0] * dr, GRID[1] * dz), GRID, DOMAINS, dt, species, pusher="hc", ccs="zigzag",
                       precision=precision)
    ctx.init_plasma(0, PPC, DENSITY, DENSITY, 1.0, 21)
    ctx.init_plasma(1, PPC, DENSITY, DENSITY, 0.1, 22)
    ctx.init_bunch(2, BUNCH["n"], BUNCH["n_peak"], BUNCH["sigma_r"], BUNCH["sigma_z"], BUNCH["z0"], BUNCH["gamma"],
                   23)
    if precision == "fp32":
        surrogate.install(ctx, *net)
        ctx.set_collisions("surrogate", 0, 1, seed=21)
    else:
        ctx.set_collisions("p12", 0, 1, seed=21)
    ctx.set_energy_log(True)
    return ctx


def test_full_form_fp32_and_fp64_energies_agree():
    from picopic_amd import surrogate
    params, norm, loss = surrogate.train(0.5e-3 / C_LIGHT, 1e-3, 1e-3, grid=(128, 256), ppc=32, passes=20, epochs=30,
                                         seed=1, **REGIME)
    print("surrogate fit loss", loss, flush=True)
    res = {}
    for precision in ("fp64", "fp32"):
        print("full form", precision, flush=True)
        ctx = _context(precision, (params, norm))
        counts = [ctx.count(s) for s in range(3)]
        ctx.step(3)
        ctx.sync()
        log = ctx.energy_log()
        res[precision] = dict(counts=counts, counts_after=[ctx.count(s) for s in range(3)], log=log,
                              fallback=ctx.surrogate_fallback_cells() if precision == "fp32" else None)
        print(precision, counts, log["field"], log["kinetic"], flush=True)
        ctx.close()
    a, b = res["fp64"], res["fp32"]
    assert a["counts"] == b["counts"]
    assert sum(a["counts"]) > 5.37e8
    assert a["counts_after"] == b["counts_after"]
    i = np.arange(GRID[0], dtype=np.int64)
    ncol = PPC * GRID[0]
    per_cell = (ncol * (i + 1) ** 2) // (GRID[0] ** 2) - (ncol * i ** 2) // (GRID[0] ** 2)
    sparse = int(((per_cell >= 1) & (per_cell < 8)).sum()) * GRID[1]
    print("fallback cells", b["fallback"], "sparse cells of the placement", sparse, flush=True)
    assert sparse > 0 and abs(b["fallback"] - sparse) <= 0.01 * sparse
    la, lb = a["log"], b["log"]
    assert len(la["field"]) == len(lb["field"]) == 3
    assert la["field"][-1] > 0
    for t in range(3):
        assert abs(lb["field"][t] - la["field"][t]) <= 1e-5 * la["field"][t], t
        ka, kb = la["kinetic"][t].sum(), lb["kinetic"][t].sum()
        assert abs(kb - ka) <= 1e-5 * ka, t
        for s in range(3):
            assert abs(lb["kinetic"][t, s] - la["kinetic"][t, s]) <= 1e-4 * la["kinetic"][t, s], (t, s)
