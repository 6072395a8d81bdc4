"""FP64 parity of the HIP path against the CPU oracle, phase by phase and over
whole steps, through the C ABI.

Tolerances (FP64 mode): single phases 1e-13 relative to the largest value of
the compared quantity (the only differences are last-bit ones from pow/log in
the device math library and the summation order of atomics); several whole
steps 1e-12.  Field copies (overlay of E/H) and the Yee update are compared
bit for bit.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as Hh

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cap():
    from picopic_amd import capi
    return capi


def build_pair(grid=(16, 32), domains=(2, 2), pusher="boris-adaptive", ccs="vb", n=3000, seed=1, beam=False,
               dt_factor=0.5, pml=None, fields=True):
    capi = _cap()
    size, grid = Hh.geometry(*grid)
    dt = dt_factor * (size[1] / grid[1]) / 3e8
    species = Hh.species_list(beam)
    orc = O.SMB(size, grid, domains, dt, species, pusher=pusher, ccs=ccs, pml=pml)
    gpu = capi.Context(size, grid, domains, dt, [s + (4 * n,) for s in species], pusher=pusher, ccs=ccs, pml=pml)
    rng = np.random.default_rng(seed)
    for sp, (q, m, b) in enumerate(species):
        r, z, vr, vp, vz, w = Hh.make_particles(rng, size, grid, n, m, temp_ev=1.0 if m == 1 else 0.1)
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 6)
        orc.add_particles(sp, r, z, vr, vp, vz, w, ids)
        gpu.upload(sp, r, z, vr, vp, vz, w, ids)
    if fields:
        E, H, Hat = Hh.random_fields(rng, orc.n_domains, orc.nr_d, orc.nz_d)
        for name, arr in (("E", E), ("H", H), ("H_at_et", Hat)):
            orc.set_grid(name, arr)
            gpu.set_grid(name, arr)
    return orc, gpu, species


def by_id(p):
    o = np.argsort(p["id"])
    return {k: v[o] for k, v in p.items()}


def compare_particles(orc, gpu, nsp, rtol, label=""):
    for sp in range(nsp):
        a = by_id(orc.particles(sp))
        b = by_id(gpu.particles(sp))
        assert np.array_equal(a["id"], b["id"]), "%s species %d: particle sets differ (%d vs %d)" % (
            label, sp, len(a["id"]), len(b["id"]))
        for k in ("r", "z", "vr", "vphi", "vz", "w"):
            e = Hh.relerr(b[k], a[k])
            assert e <= rtol, "%s species %d %s rel err %.3e" % (label, sp, k, e)


def interior(G):
    return G[:, :, 2:-2, 2:-2]


def test_gather_known_answer_on_device():
    capi = _cap()
    k = json.load(open(os.path.join(GOLD, "kat_reference.json")))
    d = k["domain"]
    n = d["grid"]
    ctx = capi.Context((n[0] * d["dr"], n[1] * d["dz"]), tuple(n), (1, 1), d["dt"], [(-1, 1, False, 16)])
    E = np.zeros((1, 3, n[0] + 4, n[1] + 4))
    H = np.zeros_like(E)
    for c in range(3):
        E[:, c] = k["fields"]["E"][c]
        H[:, c] = k["fields"]["H"][c]
    ctx.set_grid("E", E)
    ctx.set_grid("H", H)
    p = k["particle"]
    e = ctx.get_field_e(0, p["r"], p["z"])
    h = ctx.get_field_h(0, p["r"], p["z"])
    print("gather e", list(e), "h", list(h))
    assert Hh.relerr(e, k["gather_e"]) <= 1e-15
    assert Hh.relerr(h, k["gather_h"]) <= 1e-15
    # Boris with those fields, through a one-particle push
    ctx.upload(0, [p["r"]], [p["z"]], [p["v"][0]], [p["v"][1]], [p["v"][2]], [1.0])
    ctx.phase("push")
    q = ctx.particles(0)
    v = [q["vr"][0], q["vphi"][0], q["vz"][0]]
    print("boris v", v)
    assert Hh.relerr(v, k["boris_relativistic_v"]) <= 1e-15


@pytest.mark.parametrize("domains", [(2, 2), (2, 4), (1, 1)])
def test_field_phases_bitwise(domains):
    orc, gpu, species = build_pair(domains=domains, n=200)
    J = np.random.default_rng(5).normal(0, 1e4, (orc.n_domains, 3, orc.nr_d + 4, orc.nz_d + 4))
    orc.set_grid("J", J)
    gpu.set_grid("J", J)
    for ph, grids in (("yee_e", ["E"]), ("overlay_e", ["E"]), ("yee_h", ["H", "H_at_et"]),
                      ("overlay_h", ["H", "H_at_et"])):
        orc.phase(ph)
        gpu.phase(ph)
        for gname in grids:
            a, b = orc.get_grid(gname), gpu.get_grid(gname)
            diff = np.abs(a - b).max()
            print(ph, gname, "max abs diff", diff)
            assert np.array_equal(a, b), "%s %s differs by %g" % (ph, gname, diff)


def test_pml_sigma_matches():
    orc, gpu, _ = build_pair(grid=(32, 64), domains=(2, 4), n=10, pml=(0.1, 0.1, 0.2, 1e-5, 1.5))
    a, b = orc.get_grid("sigma"), gpu.get_grid("sigma")
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b)


@pytest.mark.parametrize("pusher", ["boris-adaptive", "boris", "vay", "hc"])
def test_push_and_move_parity(pusher):
    orc, gpu, species = build_pair(pusher=pusher, domains=(2, 2), n=3000)
    orc.phase("push")
    gpu.phase("push")
    compare_particles(orc, gpu, len(species), 1e-13, "push")
    orc.phase("move")
    gpu.phase("move")
    compare_particles(orc, gpu, len(species), 1e-13, "move")


@pytest.mark.parametrize("ccs", ["vb", "zigzag"])
@pytest.mark.parametrize("domains", [(1, 1), (2, 2), (2, 4)])
def test_deposit_and_stitch_parity(ccs, domains):
    orc, gpu, species = build_pair(ccs=ccs, domains=domains, n=3000, beam=True)
    for ph in ("push", "move", "migrate"):
        orc.phase(ph)
        gpu.phase(ph)
    compare_particles(orc, gpu, len(species), 1e-13, "migrate")
    # the GPU keys reproduce the reference's per-domain vector order
    for sp in range(len(species)):
        a = orc.particles(sp)
        b = gpu.particles(sp)
        ob = b["id"][np.argsort(b["order"])]
        assert np.array_equal(a["id"], ob), "species %d vector order differs" % sp
    orc.phase("reset_j")
    orc.phase("deposit")
    gpu.phase("deposit")
    a, b = interior(orc.get_grid("J")), interior(gpu.get_grid("J"))
    e = Hh.relerr(b, a)
    print("deposit J rel err", e)
    assert e <= 1e-13
    orc.phase("overlay_j")
    gpu.phase("overlay_j")
    a, b = interior(orc.get_grid("J")), interior(gpu.get_grid("J"))
    e = Hh.relerr(b, a)
    print("stitched J rel err", e)
    assert e <= 1e-13


def test_crowded_migration_order():
    """Thousands of electrons cross each interior z boundary of 1 x 4 domains
    in one step, from both sides, so every destination group holds several
    sorted tiles of migrants from two source domains: the device's order keys
    reproduce the reference's push_back order (src/SMB.cpp:211-340) and the
    particle set is the oracle's."""
    capi = _cap()
    size, grid = Hh.geometry(16, 64)
    dr, dz = size[0] / grid[0], size[1] / grid[1]
    dt = 0.5 * dz / 3e8
    domains = (1, 4)
    species = [(-1, 1, False)]
    orc = O.SMB(size, grid, domains, dt, species)
    rng = np.random.default_rng(11)
    parts = []
    nz_d = grid[1] // domains[1]
    for J in range(1, domains[1]):
        zb = J * nz_d * dz
        for sign in (1, -1):
            n = 3000 + 500 * J
            r = rng.uniform(0.5 * dr, size[0] - 0.5 * dr, n)
            z = zb - sign * rng.uniform(0.02, 0.4, n) * dz
            vz = np.full(n, sign * 0.9 * 3e8)
            vr = rng.normal(0, 1e5, n)
            vp = rng.normal(0, 1e5, n)
            parts.append((r, z, vr, vp, vz, np.full(n, 1e6)))
    r, z, vr, vp, vz, w = (np.concatenate(a) for a in zip(*parts))
    o = rng.permutation(len(r))
    r, z, vr, vp, vz, w = (a[o] for a in (r, z, vr, vp, vz, w))
    ids = np.arange(len(r), dtype=np.uint64)
    gpu = capi.Context(size, grid, domains, dt, [species[0] + (len(r) + 64,)])
    orc.add_particles(0, r, z, vr, vp, vz, w, ids)
    gpu.upload(0, r, z, vr, vp, vz, w, ids)
    for ph in ("push", "move", "migrate"):
        orc.phase(ph)
        gpu.phase(ph)
    compare_particles(orc, gpu, 1, 1e-13, "migrate")
    a = orc.particles(0)
    b = gpu.particles(0)
    ob = b["id"][np.argsort(b["order"])]
    z0 = np.empty_like(z)
    z0[ids.astype(np.int64)] = z
    crossed = np.floor(a["z"] / (nz_d * dz)) != np.floor(z0[a["id"].astype(np.int64)] / (nz_d * dz))
    print("particles", len(r), "changed domain", int(crossed.sum()))
    assert crossed.sum() > 0.9 * len(r)
    assert np.array_equal(a["id"], ob)


def test_strict_axial_reset_order():
    """Particles with v_r = 0 reset J_r nodes (currentVB.cpp:548-549); the
    result depends on the reference's particle order, which the exact path
    reproduces."""
    capi = _cap()
    size, grid = Hh.geometry(8, 16)
    dt = 0.5e-3 / 3e8
    species = [(-1, 1, False), (1, 1836, False)]
    orc = O.SMB(size, grid, (1, 1), dt, species)
    gpu = capi.Context(size, grid, (1, 1), dt, [s + (4000,) for s in species])
    rng = np.random.default_rng(3)
    for sp in range(2):
        n = 1500
        r, z, vr, vp, vz, w = Hh.make_particles(rng, size, grid, n, 1 if sp == 0 else 1836, special=False)
        # a third of the particles move strictly along z, crowded in few cells
        k = n // 3
        vr[::3] = 0.0
        vp[::3] = 0.0
        r[::3] = rng.choice([2.5e-3, 3.5e-3, 4.5e-3], size=len(r[::3])) + rng.uniform(-4e-4, 4e-4, len(r[::3]))
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 6)
        orc.add_particles(sp, r, z, vr, vp, vz, w, ids)
        gpu.upload(sp, r, z, vr, vp, vz, w, ids)
    for ph in ("push", "move", "migrate"):
        orc.phase(ph)
        gpu.phase(ph)
    orc.phase("reset_j")
    orc.phase("deposit")
    gpu.phase("deposit")
    a, b = interior(orc.get_grid("J")), interior(gpu.get_grid("J"))
    e = Hh.relerr(b, a)
    print("strict-reset J rel err", e)
    assert e <= 1e-13


@pytest.mark.parametrize("domains", [(1, 1), (2, 2)])
@pytest.mark.parametrize("ccs", ["vb", "zigzag"])
def test_whole_steps_parity(domains, ccs):
    orc, gpu, species = build_pair(domains=domains, ccs=ccs, n=2000, beam=True, fields=False)
    nsteps = 5
    orc.step(nsteps)
    gpu.step(nsteps)
    compare_particles(orc, gpu, len(species), 1e-12, "steps")
    for gname in ("E", "H", "H_at_et"):
        e = Hh.relerr(interior(gpu.get_grid(gname)), interior(orc.get_grid(gname)))
        print(gname, "rel err", e)
        assert e <= 1e-12
    e = Hh.relerr(interior(gpu.get_grid("J")), interior(orc.get_grid("J")))
    print("J rel err", e)
    assert e <= 1e-12


def test_init_plasma_counts_and_moments():
    capi = _cap()
    size, grid = Hh.geometry(32, 64)
    ctx = capi.Context(size, grid, (2, 2), 1e-12, [(-1, 1, False, 1 << 16)])
    ctx.init_plasma(0, 16, 1e17, 1e17, 1.0, 1234)
    p = ctx.particles(0)
    assert len(p["r"]) == 16 * 32 * 64
    vth = np.sqrt(1.6e-19 / 9.1e-31)
    for k in ("vr", "vphi", "vz"):
        assert abs(p[k].std() / vth - 1) < 0.02
        assert abs(p[k].mean()) < 0.02 * vth
    # uniform in r^2 and z
    R, L = size
    assert abs(np.mean(p["r"] ** 2) / (R * R / 2) - 1) < 0.02
    assert abs(np.mean(p["z"]) / (L / 2) - 1) < 0.02
    # total real particles = n * volume
    assert abs(p["w"].sum() / (1e17 * np.pi * R * R * L) - 1) < 1e-9


@pytest.mark.parametrize("domains", [(1, 1), (2, 4)])
def test_sorted_storage_gives_same_results(domains):
    """Sorting by cell every step (and dropping removed particles) changes
    only the storage order: states match the oracle and the order keys still
    reproduce the reference's vector order."""
    orc, gpu, species = build_pair(domains=domains, n=2000, beam=True, fields=False)
    for sp in range(len(species)):
        gpu.set_sort_interval(sp, 1)
    orc.step(4)
    gpu.step(4)
    compare_particles(orc, gpu, len(species), 1e-12, "sorted steps")
    for sp in range(len(species)):
        a = orc.particles(sp)
        b = gpu.particles(sp)
        assert np.array_equal(a["id"], b["id"][np.argsort(b["order"])])
    e = Hh.relerr(interior(gpu.get_grid("J")), interior(orc.get_grid("J")))
    assert e <= 1e-12
    # storage is cell-ordered after an explicit sort
    gpu.sort(0)
    p = gpu.particles(0)
    dr = orc.dr
    dz = orc.dz
    cells = np.floor(p["r"] / dr).astype(np.int64) * 10 ** 6 + np.floor(p["z"] / dz).astype(np.int64)
    dom = (np.floor(p["r"] / dr).astype(np.int64) // orc.nr_d) * domains[1] + np.floor(p["z"] / dz).astype(np.int64) // orc.nz_d
    key = dom * 10 ** 12 + cells
    assert np.all(np.diff(key) >= 0)
