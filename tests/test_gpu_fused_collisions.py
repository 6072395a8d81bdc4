"""The fused surrogate collision phase of FP32 steps and the in-step energy
log, through the C ABI.

Fused phase (no reference counterpart): the push kernel sums each cell's
moments of the pushed velocities, the network evaluates every cell, and the
move kernel rescales each colliding particle's velocity about its cell's
mean as it reads it.  Checked against the bucketed surrogate pass of an FP64
context (the same network and rescale, evaluated by staging each cell) on
the same plasma:

- the cells that fall back to P12 are the same (thinned rows add fallback
  cells);
- per-particle velocities after one step agree within 1e-4 of the thermal
  speed: the network's inputs are rounded to BF16 at the MFMA, so features
  computed from FP32 and from FP64 moments can round to neighbouring BF16
  values, which moves a cell's predicted change by about 2^-8 of itself
  (about 1e-5 of v_th at these rates);
- the rescale conserves each cell's momentum and the pair's energy, so with
  zero fields one step changes each species' axial momentum and the total
  kinetic energy only through the move, as in the FP64 context: within
  1e-5 of sum w m |v_z| and 2e-6 of the kinetic energy.  (At this time step
  fast particles cross several cells; the move reflects them at the walls,
  which flips v_z, and removes about 1.5 % of them, which takes 5e-4 of the
  kinetic energy -- the oracle's step without collisions loses the same.)

Energy log: the rows recorded on the device match pcp_field_energy /
pcp_kinetic_energy evaluated after the same steps (FP64 contexts to 1e-12;
FP32 contexts sum w m v^2 / (s (1 + s)) per particle in FP32 before the FP64
wavefront sums, 1e-6).
"""
import numpy as np
import pytest

from picopic_amd import capi, surrogate

pytestmark = pytest.mark.gpu

DR = DZ = 1e-4
DT = 2e-10
GRID = (32, 64)
REGIME = dict(dens=(5e19, 2e20), tepar=(2.0, 16.0), teperp=(0.5, 8.0), ti=(0.5, 2.5))

_NET = {}
_FALLBACK = {}


def network():
    if "net" not in _NET:
        params, norm, loss = surrogate.train(DT, DR, DZ, grid=(64, 128), ppc=32, passes=20, epochs=25, seed=3,
                                             **REGIME)
        print("surrogate fit loss", loss, flush=True)
        _NET["net"] = (params, norm)
    return _NET["net"]


def pair_of_contexts(seed=5, sparse_rows=0):
    rng = np.random.default_rng(seed)
    tepar, teperp, ti, dens = surrogate.regime_cells(rng, GRID, **REGIME)
    data = surrogate.cell_plasma(rng, GRID, DR, DZ, 24, tepar, teperp, ti, dens)
    if sparse_rows:
        # the first rows hold too few macro-particles for the network
        keep = [np.floor(d[0] / DR) >= sparse_rows for d in data]
        for s in range(2):
            r = data[s][0]
            thin = ~keep[s] & (np.arange(len(r)) % 6 != 0)
            data[s] = tuple(a[~thin] for a in data[s])
    ctxs = []
    params, norm = network()
    for prec in ("fp64", "fp32"):
        cap = max(len(d[0]) for d in data) + 64
        c = capi.Context((GRID[0] * DR, GRID[1] * DZ), GRID, (2, 2), DT,
                         [(-1, 1, False, cap), (1, 1836, False, cap)], precision=prec)
        for sp, d in enumerate(data):
            c.upload(sp, *d, np.arange(len(d[0]), dtype=np.uint64) + np.uint64(sp << 32))
        surrogate.install(c, params, norm)
        c.set_collisions("surrogate", 0, 1, seed=17)
        ctxs.append(c)
    return ctxs, data


def by_id(p):
    o = np.argsort(p["id"])
    return {k: v[o] for k, v in p.items()}


@pytest.mark.parametrize("sparse_rows", [0, 3])
def test_fused_phase_matches_bucketed_surrogate(sparse_rows):
    (c64, c32), data = pair_of_contexts(sparse_rows=sparse_rows)
    c64.step(1)
    c32.step(1)
    fb = c32.surrogate_fallback_cells()
    print("fallback cells", fb, flush=True)
    assert fb == c64.surrogate_fallback_cells()
    _FALLBACK[sparse_rows] = fb
    if sparse_rows:
        assert fb > _FALLBACK.get(0, 0)
    for sp in range(2):
        a, b = by_id(c64.particles(sp)), by_id(c32.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        vth = np.sqrt(np.mean(a["vr"] ** 2 + a["vphi"] ** 2 + a["vz"] ** 2))
        for k in ("vr", "vphi", "vz"):
            err = np.abs(a[k] - b[k]).max() / vth
            print(sp, k, "max |dv| / v_th", err)
            assert err < 1e-4
    # and the rescale did change the velocities (the bi-Maxwellian relaxes)
    fresh = {k: v for k, v in zip(("r", "z", "vr", "vphi", "vz", "w"), data[0])}
    b = c32.particles(0)
    assert np.abs(np.sort(b["vz"] ** 2).sum() / np.sort(fresh["vz"] ** 2).sum() - 1) > 1e-6


def test_fused_step_conserves_axial_momentum_and_energy():
    (c64, c32), _ = pair_of_contexts(seed=9)
    m = [9.1e-31, 1836 * 9.1e-31]

    def totals(c):
        pz, pabs, ke = [], [], 0.0
        for sp in range(2):
            p = c.particles(sp)
            pz.append((p["w"] * m[sp] * p["vz"]).sum())
            pabs.append((p["w"] * m[sp] * np.abs(p["vz"])).sum())
            ke += (p["w"] * m[sp] * (p["vr"] ** 2 + p["vphi"] ** 2 + p["vz"] ** 2)).sum() / 2
        return np.array(pz), np.array(pabs), ke

    pz0, pabs, ke0 = totals(c32)
    c32.step(1)
    c64.step(1)
    pz32, _, ke32 = totals(c32)
    pz64, _, ke64 = totals(c64)
    print("kinetic energy change fused %.3e, bucketed %.3e" % (ke32 / ke0 - 1, ke64 / ke0 - 1))
    for s in range(2):
        print("species %d: axial momentum change fused %.3e, bucketed %.3e (of sum w m |v_z| %.3e)" % (
            s, pz32[s] - pz0[s], pz64[s] - pz0[s], pabs[s]))
        assert abs((pz32[s] - pz0[s]) - (pz64[s] - pz0[s])) <= 1e-5 * pabs[s], s
    assert [c32.count(s) for s in range(2)] == [c64.count(s) for s in range(2)]
    assert abs(ke32 / ke0 - ke64 / ke0) < 2e-6


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_energy_log_matches_the_diagnostics(prec):
    size, grid = (16e-3, 64e-3), (16, 64)
    dt = 0.5e-3 / 3e8
    ctx = capi.Context(size, grid, (2, 2), dt, [(-1, 1, False, 1 << 16), (1, 1836, False, 1 << 16),
                                              (-1, 1, True, 1 << 14)], ccs="zigzag", pusher="hc", precision=prec)
    ctx.init_plasma(0, 16, 1e17, 1e17, 1.0, 1)
    ctx.init_plasma(1, 16, 1e17, 1e17, 0.1, 2)
    ctx.init_bunch(2, 8192, 1e16, 2e-3, 5e-3, 0.02, 5.0, 3)
    ctx.set_energy_log(True)
    ctx.step(4)
    log = ctx.energy_log()
    assert len(log["step"]) == 4 and np.all(np.diff(log["step"]) == 1)
    tol = 1e-12 if prec == "fp64" else 1e-6
    fe = ctx.field_energy()
    assert fe > 0 and abs(log["field"][-1] / fe - 1) < 1e-12
    for sp in range(3):
        ke = ctx.kinetic_energy(sp)
        assert abs(log["kinetic"][-1, sp] / ke - 1) < tol, (sp, log["kinetic"][-1, sp], ke)
    # energy grows from zero fields as the bunch drives the wake
    assert log["field"][0] < log["field"][-1]
