"""Device collisions (P12 / TA77S / SK98) against the oracle's keyed pass.

Both sides shuffle each cell by the same Philox key and give pair q of
global cell c the same four uniforms, so one pass is compared particle by
particle.  Tolerances, relative to the largest speed:

- TA77S: 1e-12 (libm and the device math library may differ in the last ulp
  of pow/log/sin/cos).  Cells holding a single macro-particle of a species
  are left out of the TA77S data: correct_velocities then divides a
  round-off difference by a round-off difference (thermal energy of one
  particle), which amplifies any last-ulp difference without bound.
- P12 and SK98: 1e-8.  Their relative speed is the difference of the two
  centre-of-mass speeds (collisionsP12.cpp:117-118), which for a pair of
  equal masses is zero up to round-off; those pairs deflect by an amount
  proportional to that round-off (P12 caps s12 by s12max ~ v_rel), so a
  last-ulp difference upstream shows as ~1e-9 of the speed.
- FP32 split contexts collide in FP64 from FP32 velocities and store FP32:
  1e-6.
"""
import numpy as np
import pytest

from oracle import oracle as O
from picopic_amd import capi

pytestmark = pytest.mark.gpu

M_E = 9.1e-31
Q_E = 1.6e-19


def plasma(rng, grid, size, counts_e, counts_i, temp_e=5.0, temp_i=1.0, fp32=False, weights="equal",
           wall_safe=False):
    """Per-cell counts given as arrays [nr, nz] for electrons and ions.  With
    `wall_safe`, particles of the wall cells start on the inner half of the
    cell (r >= dr/2, r <= R - dr/2, dz/2 <= z <= L - dz/2), where the movers
    of the reference keep them."""
    nr, nz = grid
    dr, dz = size[0] / nr, size[1] / nz
    out = []
    for sp, (cnt, m, T) in enumerate(((counts_e, M_E, temp_e), (counts_i, 1836 * M_E, temp_i))):
        ci, cj = np.meshgrid(np.arange(nr), np.arange(nz), indexing="ij")
        ci = np.repeat(ci.ravel(), cnt.ravel())
        cj = np.repeat(cj.ravel(), cnt.ravel())
        n = len(ci)
        fr = rng.uniform(0.05, 0.95, n)
        fz = rng.uniform(0.05, 0.95, n)
        if wall_safe:
            fr = np.where(ci == 0, 0.55 + 0.4 * fr, np.where(ci == nr - 1, 0.45 * fr, fr))
            fz = np.where(cj == 0, 0.55 + 0.4 * fz, np.where(cj == nz - 1, 0.45 * fz, fz))
        if fp32:
            fr = fr.astype(np.float32).astype(np.float64)
            fz = fz.astype(np.float32).astype(np.float64)
        r = (ci + fr) * dr
        z = (cj + fz) * dz
        vth = np.sqrt(T * Q_E / m)
        v = rng.normal(0, vth, (3, n))
        if weights == "equal":
            w = np.full(n, 1e8)
        else:
            w = rng.uniform(0.5, 1.5, n) * 1e8
        if fp32:
            v = v.astype(np.float32).astype(np.float64)
            w = w.astype(np.float32).astype(np.float64)
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 6)
        order = rng.permutation(n)
        out.append(tuple(a[order] for a in (r, z, v[0], v[1], v[2], w, ids)))
    return out


def counts(rng, grid, lo=1, hi=12, hole=None):
    ce = rng.integers(lo, hi, grid)
    ci = rng.integers(lo, hi, grid)
    ci[0, 0] = ce[0, 0]          # nion == nel
    ci[0, 1] = 2 * ce[0, 1] + 1  # nion > nel, non-integer ratio
    ce[1, 0] = 3 * ci[1, 0]      # nel = 3 nion
    if hole is not None:
        ce[hole] = 0             # an empty electron bucket ends the row
    return ce, ci


def build(grid, domains, data, precision, dt=1e-13):
    size = (grid[0] * 1e-4, grid[1] * 1e-4)
    species = [(-1, 1, False), (1, 1836, False)]
    orc = O.SMB(size, grid, domains, dt, species)
    cap = max(len(d[0]) for d in data) + 16
    gpu = capi.Context(size, grid, domains, dt, [s + (cap,) for s in species], precision=precision)
    for sp, d in enumerate(data):
        orc.add_particles(sp, *d)
        gpu.upload(sp, *d)
    return orc, gpu


def by_id(p):
    o = np.argsort(p["id"])
    return {k: v[o] for k, v in p.items()}


def compare(orc, gpu, tol):
    worst = 0.0
    for sp in range(2):
        a, b = by_id(orc.particles(sp)), by_id(gpu.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        scale = max(np.abs(a[k]).max() for k in ("vr", "vphi", "vz"))
        for k in ("vr", "vphi", "vz"):
            e = np.abs(a[k] - b[k]).max() / scale
            worst = max(worst, e)
    assert worst <= tol, worst
    return worst


@pytest.mark.parametrize("scheme", ["p12", "ta77s", "sk98"])
@pytest.mark.parametrize("domains", [(1, 1), (2, 2)])
def test_collision_pass_fp64(scheme, domains):
    rng = np.random.default_rng(31)
    grid = (6, 8)
    size = (grid[0] * 1e-4, grid[1] * 1e-4)
    ce, ci = counts(rng, grid, lo=2 if scheme == "ta77s" else 1, hole=(2, 5))
    data = plasma(rng, grid, size, ce, ci, weights="equal" if scheme != "ta77s" else "mixed")
    orc, gpu = build(grid, domains, data, "fp64")
    print("collide", scheme, domains, flush=True)
    orc.collide(scheme, 0, 1, seed=77, step=3)
    gpu.collide(scheme, 0, 1, 77, 3)
    err = compare(orc, gpu, 1e-12 if scheme == "ta77s" else 1e-8)
    print(scheme, domains, err)
    # and something happened
    before = by_id({k: v for k, v in zip(("r", "z", "vr", "vphi", "vz", "w", "id"), data[0])})
    after = by_id(gpu.particles(0))
    assert np.abs(after["vr"] - before["vr"]).max() > 0


@pytest.mark.parametrize("scheme", ["p12", "ta77s"])
def test_collision_pass_fp32(scheme):
    rng = np.random.default_rng(32)
    grid = (6, 8)
    size = (grid[0] * 1e-4, grid[1] * 1e-4)
    ce, ci = counts(rng, grid)
    data = plasma(rng, grid, size, ce, ci, fp32=True)
    orc, gpu = build(grid, (2, 2), data, "fp32")
    print("collide fp32", scheme, flush=True)
    orc.collide(scheme, 0, 1, seed=5, step=0)
    gpu.collide(scheme, 0, 1, 5, 0)
    print(scheme, compare(orc, gpu, 1e-6))


@pytest.mark.parametrize("scheme", ["p12", "ta77s", "sk98"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_crowded_cells_match_the_oracle(scheme, precision):
    """Cells of more than 256 macro-particles of a species (the reference's
    buckets have no size limit, src/collisions.cpp:87-108) take the crowded
    path -- a workgroup per cell, keys sorted and velocities staged in
    global memory -- next to cells of every other size class, and match the
    oracle's pass at the tolerances above."""
    rng = np.random.default_rng(35)
    grid = (3, 4)
    size = (grid[0] * 1e-4, grid[1] * 1e-4)
    ce = rng.integers(2, 12, grid)
    ci = rng.integers(2, 12, grid)
    ce[0, 1], ci[0, 1] = 300, 257        # both crowded, nion < nel
    ce[1, 2], ci[1, 2] = 40, 700         # ions crowded, nion > nel
    ce[2, 3], ci[2, 3] = 1000, 1000      # nion == nel
    ce[2, 0], ci[2, 0] = 100, 30         # a wavefront cell of the 256 class
    fp32 = precision == "fp32"
    data = plasma(rng, grid, size, ce, ci, weights="equal" if scheme != "ta77s" else "mixed", fp32=fp32)
    orc, gpu = build(grid, (1, 1), data, precision)
    print("collide crowded", scheme, precision, flush=True)
    orc.collide(scheme, 0, 1, seed=91, step=2)
    gpu.collide(scheme, 0, 1, 91, 2)
    tol = 1e-6 if fp32 else (1e-12 if scheme == "ta77s" else 1e-8)
    err = compare(orc, gpu, tol)
    print(scheme, precision, "crowded max rel err", err)


def test_crowded_stage_overflow_is_reported():
    """A pass whose crowded cells hold more than 2^22 macro-particles in all
    exceeds the global stage: the device reports it and nothing hangs."""
    rng = np.random.default_rng(33)
    grid = (2, 2)
    size = (grid[0] * 1e-4, grid[1] * 1e-4)
    ce = np.array([[2 ** 22 + 1, 4], [4, 4]])
    ci = np.array([[4, 4], [4, 4]])
    data = plasma(rng, grid, size, ce, ci)
    cap = max(len(d[0]) for d in data) + 16
    species = [(-1, 1, False), (1, 1836, False)]
    gpu = capi.Context(size, grid, (1, 1), 1e-13, [s + (cap,) for s in species])
    for sp, d in enumerate(data):
        gpu.upload(sp, *d)
    print("collide overflow", flush=True)
    with pytest.raises(capi.PcpError, match="crowded collision cells"):
        gpu.collide("p12", 0, 1, 1, 0)


@pytest.mark.parametrize("precision,tol,tol_fields", [("fp64", 1e-6, 1e-6), ("fp32", 2e-5, 1e-4)])
def test_steps_with_collision_phase(precision, tol, tol_fields):
    """pcp_step with the collision phase between push and move, against the
    oracle's step with its collision phase (3 steps, fields on).  The FP64
    bound is set by the round-off-driven like-particle P12 deflections (see
    the module docstring) carried through three steps."""
    rng = np.random.default_rng(34)
    grid = (8, 16)
    size = (grid[0] * 1e-4, grid[1] * 1e-4)
    ce, ci = counts(rng, grid, lo=3, hi=10)
    data = plasma(rng, grid, size, ce, ci, temp_e=20.0, temp_i=2.0, fp32=precision == "fp32", wall_safe=True)
    dt = 0.5 * size[1] / grid[1] / 3e8
    orc, gpu = build(grid, (2, 2), data, precision, dt=dt)
    orc.set_collisions("p12", 0, 1, seed=9)
    gpu.set_collisions("p12", 0, 1, 9)
    print("steps with collisions", precision, flush=True)
    orc.step(3)
    gpu.step(3)
    worst = 0.0
    for sp in range(2):
        a, b = by_id(orc.particles(sp)), by_id(gpu.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        for k in ("r", "z", "vr", "vphi", "vz"):
            scale = np.abs(a[k]).max()
            worst = max(worst, np.abs(a[k] - b[k]).max() / scale)
    print(precision, "particles", worst)
    assert worst <= tol
    for nm in ("E", "J"):
        A, B = orc.get_grid(nm)[:, :, 2:-2, 2:-2], gpu.get_grid(nm)[:, :, 2:-2, 2:-2]
        e = np.abs(A - B).max() / np.abs(A).max()
        print(nm, e)
        assert e <= tol_fields
