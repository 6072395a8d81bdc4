"""FP32 split-precision mode against the FP64 oracle.

The FP32 path runs the reference algorithm in single precision with
positions split into a cell index and an FP32 in-cell offset.  Tolerances
written here are FP32-level: single-call results within 1e-6 relative to the
largest value, particle states after a few steps within 1e-5, and the total
field energy after 20 steps within 1e-5 relative (the target stated for this
mode).
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


def pair32(grid=(16, 32), domains=(2, 2), ccs="zigzag", pusher="boris-adaptive", n=3000, seed=2, beam=False,
           fields=False, fast=0.3, special=True):
    capi = _cap()
    size, grid = Hh.geometry(*grid)
    dt = 0.5 * (size[1] / grid[1]) / 3e8
    species = Hh.species_list(beam)
    orc = O.SMB(size, grid, domains, dt, species, pusher=pusher, ccs=ccs)
    gpu = capi.Context(size, grid, domains, dt, [s + (4 * n,) for s in species], pusher=pusher, ccs=ccs,
                       precision="fp32")
    rng = np.random.default_rng(seed)
    for sp, (q, m, b) in enumerate(species):
        r, z, vr, vp, vz, w = Hh.make_particles(rng, size, grid, n, m, fast_fraction=fast, special=special,
                                                temp_ev=1.0 if m == 1 else 0.1)
        # start both from the FP32-representable state so that the comparison
        # measures the arithmetic, not the input rounding
        dr, dz = size[0] / grid[0], size[1] / grid[1]
        ir, iz = np.floor(r / dr), np.floor(z / dz)
        r = (ir + (r / dr - ir).astype(np.float32).astype(np.float64)) * dr
        z = (iz + (z / dz - iz).astype(np.float32).astype(np.float64)) * dz
        vr, vp, vz, w = (a.astype(np.float32).astype(np.float64) for a in (vr, vp, vz, w))
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 6)
        orc.add_particles(sp, r, z, vr, vp, vz, w, ids)
        gpu.upload(sp, r, z, vr, vp, vz, w, ids)
    if fields:
        E, H, Hat = Hh.random_fields(rng, orc.n_domains, orc.nr_d, orc.nz_d)
        E, H, Hat = (a.astype(np.float32).astype(np.float64) for a in (E, H, Hat))
        for name, arr in (("E", E), ("H", H), ("H_at_et", Hat)):
            orc.set_grid(name, arr)
            gpu.set_grid(name, arr)
    return orc, gpu, species


def by_id(p):
    o = np.argsort(p["id"])
    return {k: v[o] for k, v in p.items()}


def interior(G):
    return G[:, :, 2:-2, 2:-2]


def test_fp32_gather_known_answer():
    capi = _cap()
    k = json.load(open(os.path.join(GOLD, "kat_reference.json")))
    d = k["domain"]
    n = d["grid"]
    ctx = capi.Context((n[0] * d["dr"], n[1] * d["dz"]), tuple(n), (1, 1), d["dt"], [(-1, 1, False, 16)],
                       precision="fp32")
    E = np.zeros((1, 3, n[0] + 4, n[1] + 4))
    H = np.zeros_like(E)
    for c in range(3):
        E[:, c] = k["fields"]["E"][c]
        H[:, c] = k["fields"]["H"][c]
    ctx.set_grid("E", E)
    ctx.set_grid("H", H)
    p = k["particle"]
    assert Hh.relerr(ctx.get_field_e(0, p["r"], p["z"]), k["gather_e"]) <= 1e-6
    assert Hh.relerr(ctx.get_field_h(0, p["r"], p["z"]), k["gather_h"]) <= 1e-6


@pytest.mark.parametrize("domains", [(1, 1), (2, 2)])
def test_fp32_fields_phases(domains):
    orc, gpu, _ = pair32(domains=domains, n=100, fields=True)
    for ph, names in (("yee_e", ["E"]), ("overlay_e", ["E"]), ("yee_h", ["H", "H_at_et"]),
                      ("overlay_h", ["H", "H_at_et"])):
        orc.phase(ph)
        gpu.phase(ph)
        for nm in names:
            e = Hh.relerr(gpu.get_grid(nm), orc.get_grid(nm))
            print(ph, nm, e)
            assert e <= 1e-6


@pytest.mark.parametrize("ccs", ["zigzag", "vb"])
@pytest.mark.parametrize("pusher", ["boris-adaptive", "vay", "hc"])
def test_fp32_steps_against_fp64_oracle(ccs, pusher):
    orc, gpu, species = pair32(ccs=ccs, pusher=pusher, domains=(2, 2), n=3000, beam=True, fields=True)
    nsteps = 3
    orc.step(nsteps)
    gpu.step(nsteps)
    for sp in range(len(species)):
        a, b = by_id(orc.particles(sp)), by_id(gpu.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        for k in ("r", "z", "vr", "vphi", "vz"):
            e = Hh.relerr(b[k], a[k])
            print(ccs, pusher, sp, k, e)
            assert e <= 1e-5, (sp, k, e)
    for nm in ("E", "H", "J"):
        e = Hh.relerr(interior(gpu.get_grid(nm)), interior(orc.get_grid(nm)))
        print(nm, e)
        assert e <= 1e-4, (nm, e)


@pytest.mark.parametrize("ccs", ["zigzag", "vb"])
def test_fp32_field_energy_error(ccs):
    """Total field energy after 20 steps from a zero field: FP32 split against
    the FP64 oracle (relative error <= 1e-5)."""
    orc, gpu, species = pair32(grid=(32, 64), ccs=ccs, domains=(2, 2), n=20000, fields=False, fast=0.0,
                               special=False)
    nsteps = 20
    orc.step(nsteps)
    gpu.step(nsteps)
    E, H = orc.get_grid("E"), orc.get_grid("H")
    # FP64 energy of the oracle's fields with the same formula as pcp_field_energy
    dr, dz = orc.dr, orc.dz
    w_ref = 0.0
    for d in range(orc.n_domains):
        g = orc.geometry(d)
        i = np.arange(orc.nr_d) + g.cell_dims[0]
        vol = np.pi * dz * dr * dr * (2.0 * i + 1.0)
        e2 = (E[d, :, 2:-2, 2:-2] ** 2).sum(0)
        h2 = (H[d, :, 2:-2, 2:-2] ** 2).sum(0)
        w_ref += (0.5 * (8.85e-12 * e2 + 1.26e-6 * h2) * vol[:, None]).sum()
    w_gpu = gpu.field_energy()
    rel = abs(w_gpu - w_ref) / w_ref
    print("field energy", ccs, w_ref, w_gpu, rel)
    assert w_ref > 0
    assert rel <= 1e-5


def test_fp32_vb_sub_resolution_radial_step():
    """Villasenor-Buneman with radial steps far below the FP32 offset's
    resolution (|v_r| dt ~ 1e-14 m, a 1e-11 fraction of a cell): the FP64
    reference treats them as ordinary moves (they are above its 1e-15 m
    strict-motion threshold), so the FP32 path must deposit them from the
    unrounded end point and must not take the strict branch that resets J_r
    nodes."""
    capi = _cap()
    size, grid = Hh.geometry(16, 32)
    dt = 0.5 * (size[1] / grid[1]) / 3e8
    species = [(-1, 1, False), (1, 1836, False)]
    orc = O.SMB(size, grid, (2, 2), dt, species, ccs="vb")
    gpu = capi.Context(size, grid, (2, 2), dt, [s + (8000,) for s in species], ccs="vb", precision="fp32")
    rng = np.random.default_rng(9)
    dr, dz = size[0] / grid[0], size[1] / grid[1]
    for sp, (q, m, b) in enumerate(species):
        n = 2000
        r, z, vr, vp, vz, w = Hh.make_particles(rng, size, grid, n, m, fast_fraction=0.0, special=False,
                                                temp_ev=1.0 if m == 1 else 0.1)
        if sp == 1:
            # a quarter of the ions: radial speed of a few mm/s, no azimuthal speed
            k = n // 4
            vr[:k] = rng.choice([-1.0, 1.0], k) * rng.uniform(2e-3, 5e-3, k)
            vp[:k] = 0.0
        ir, iz = np.floor(r / dr), np.floor(z / dz)
        r = (ir + (r / dr - ir).astype(np.float32).astype(np.float64)) * dr
        z = (iz + (z / dz - iz).astype(np.float32).astype(np.float64)) * dz
        vr, vp, vz, w = (a.astype(np.float32).astype(np.float64) for a in (vr, vp, vz, w))
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 6)
        orc.add_particles(sp, r, z, vr, vp, vz, w, ids)
        gpu.upload(sp, r, z, vr, vp, vz, w, ids)
    orc.step(1)
    gpu.step(1)
    e = Hh.relerr(interior(gpu.get_grid("J")), interior(orc.get_grid("J")))
    print("J", e)
    assert e <= 1e-5


@pytest.mark.parametrize("shape", [(256, 0, 0), (128, 4, 8), (512, 2, 16)])
def test_fp32_batched_uneven_domains(shape):
    """The batched-domain shape at small scale: 8 x 8 non-square domains
    (8 x 32 cells each), density ramping 0.1 -> 1 along z so that the
    domains hold about 10x different particle counts, Vay pusher, VB, FP32,
    with the persistent launch at several workgroup sizes and sort bins,
    against the FP64 oracle after 3 steps."""
    block, br, bz = shape
    capi = _cap()
    size, grid = Hh.geometry(64, 256)
    domains = (8, 8)
    dt = 0.5 * (size[1] / grid[1]) / 3e8
    species = Hh.species_list(True)
    n = 24000
    orc = O.SMB(size, grid, domains, dt, species, pusher="vay", ccs="vb")
    gpu = capi.Context(size, grid, domains, dt, [s + (2 * n,) for s in species], pusher="vay", ccs="vb",
                       precision="fp32")
    gpu.set_launch(block)
    if br:
        gpu.set_sort_bin(br, bz)
    rng = np.random.default_rng(31)
    dr, dz = size[0] / grid[0], size[1] / grid[1]
    for sp, (q, m, b) in enumerate(species):
        r, z, vr, vp, vz, w = Hh.make_particles(rng, size, grid, n, m, fast_fraction=0.2, special=False,
                                                temp_ev=1.0 if m == 1 else 0.1)
        # linear ramp 0.1 -> 1 along z by inverse transform sampling
        u = rng.uniform(0, 1, n)
        x = (-0.1 + np.sqrt(0.01 + 4 * 0.45 * 0.55 * u)) / 0.9
        z = np.clip(x * size[1], 0.5 * dz, size[1] - 0.5 * dz)
        ir, iz = np.floor(r / dr), np.floor(z / dz)
        r = (ir + (r / dr - ir).astype(np.float32).astype(np.float64)) * dr
        z = (iz + (z / dz - iz).astype(np.float32).astype(np.float64)) * dz
        vr, vp, vz, w = (a.astype(np.float32).astype(np.float64) for a in (vr, vp, vz, w))
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 6)
        orc.add_particles(sp, r, z, vr, vp, vz, w, ids)
        gpu.upload(sp, r, z, vr, vp, vz, w, ids)
        if sp == 0:
            dz_dom = size[1] / domains[1]
            counts = np.bincount((z // dz_dom).astype(int), minlength=domains[1])
            print("particles per domain column", counts, "ratio", counts.max() / counts.min())
            assert counts.max() / counts.min() > 5
    for sp in range(len(species)):
        gpu.sort(sp)
    nsteps = 3
    orc.step(nsteps)
    gpu.step(nsteps)
    for sp in range(len(species)):
        a, b = by_id(orc.particles(sp)), by_id(gpu.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        for k in ("r", "z", "vr", "vphi", "vz"):
            e = Hh.relerr(b[k], a[k])
            assert e <= 1e-5, (sp, k, e)
    for nm in ("E", "H", "J"):
        e = Hh.relerr(interior(gpu.get_grid(nm)), interior(orc.get_grid(nm)))
        print(shape, nm, e)
        assert e <= 1e-4, (nm, e)


@pytest.mark.parametrize("ccs", ["zigzag", "vb"])
def test_fp32_error_per_component(ccs):
    """Error of the FP32 split path against the FP64 oracle after 20 steps
    from zero fields, for each field component (max |difference| over the
    interior nodes / max |oracle value|) and each velocity moment of each
    species (weighted mean and mean square of v_r, v_phi, v_z, relative to
    the oracle's value scale).  The printed table is the per-component
    error statement of this mode; the bounds are FP32-level."""
    orc, gpu, species = pair32(grid=(32, 64), ccs=ccs, domains=(2, 2), n=20000, fields=False, fast=0.1,
                               special=False)
    nsteps = 20
    orc.step(nsteps)
    gpu.step(nsteps)
    names = {"E": ("E_r", "E_phi", "E_z"), "H": ("H_r", "H_phi", "H_z"), "J": ("J_r", "J_phi", "J_z")}
    for g, comps in names.items():
        A, B = interior(orc.get_grid(g)), interior(gpu.get_grid(g))
        for c, nm in enumerate(comps):
            a, b = A[:, c], B[:, c]
            e = np.abs(b - a).max() / max(np.abs(a).max(), 1e-300)
            print("error %s %s %.3e" % (ccs, nm, e))
            assert e <= 1e-3, (nm, e)
    for sp in range(len(species)):
        a, b = by_id(orc.particles(sp)), by_id(gpu.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        wa, wb = a["w"], b["w"]
        for k in ("vr", "vphi", "vz"):
            scale = np.sqrt((wa * a[k] ** 2).sum() / wa.sum())
            m1a, m1b = (wa * a[k]).sum() / wa.sum(), (wb * b[k]).sum() / wb.sum()
            m2a, m2b = (wa * a[k] ** 2).sum() / wa.sum(), (wb * b[k] ** 2).sum() / wb.sum()
            e1 = abs(m1b - m1a) / scale
            e2 = abs(m2b - m2a) / m2a
            print("error %s species %d <%s> %.3e <%s^2> %.3e" % (ccs, sp, k, e1, k, e2))
            assert e1 <= 1e-5 and e2 <= 1e-5, (sp, k, e1, e2)


def _drifting_pair(ccs, n, seed=5, compensated=False, domains=(2, 2)):
    """One FP64 oracle and one FP32 context holding the same FP32-representable
    plasma: ~n / 512 particles per cell and species on a 16 x 32 grid, both
    species drifting in phi so that their azimuthal currents add with one
    sign (every J_phi node sums many contributions of like sign, which is
    where rounding in the sum shows)."""
    capi = _cap()
    size, grid = Hh.geometry(16, 32)
    dt = 0.5 * (size[1] / grid[1]) / 3e8
    species = Hh.species_list(False)
    orc = O.SMB(size, grid, domains, dt, species, pusher="boris-adaptive", ccs=ccs)
    gpu = capi.Context(size, grid, domains, dt, [s + (n + 64,) for s in species], pusher="boris-adaptive", ccs=ccs,
                       precision="fp32")
    if compensated:
        gpu.set_compensated_current(True)
    rng = np.random.default_rng(seed)
    dr, dz = size[0] / grid[0], size[1] / grid[1]
    for sp, (q, m, b) in enumerate(species):
        r = np.sqrt(rng.uniform((0.5 * dr) ** 2, (size[0] - 0.5 * dr) ** 2, n))
        z = rng.uniform(0.5 * dz, size[1] - 0.5 * dz, n)
        vth = np.sqrt(1.6e-19 / (m * 9.1e-31))
        v = rng.normal(0.0, vth, (3, n))
        v[1] -= 2e6 * q / np.sqrt(m)
        w = rng.uniform(0.5, 1.5, n) * 1e7
        ir, iz = np.floor(r / dr), np.floor(z / dz)
        r = (ir + (r / dr - ir).astype(np.float32).astype(np.float64)) * dr
        z = (iz + (z / dz - iz).astype(np.float32).astype(np.float64)) * dz
        v, w = v.astype(np.float32).astype(np.float64), w.astype(np.float32).astype(np.float64)
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 7)
        orc.add_particles(sp, r, z, v[0], v[1], v[2], w, ids)
        gpu.upload(sp, r, z, v[0], v[1], v[2], w, ids)
    return orc, gpu


@pytest.mark.parametrize("ccs", ["zigzag", "vb"])
def test_fp32_compensated_current_error(ccs):
    """Compensated deposition against the plain FP32 one, both against the
    FP64 oracle's J after one step of the same particles (~800 per cell and
    species, like-signed azimuthal currents).  J_phi's terms carry no
    position differences, so its error is the summation's: the plain deposit
    rounds at every scan step and atomic, the compensated one keeps those
    errors and rounds once per node, so its J_phi error is at most half the
    plain one and within 2e-7 of J_phi's magnitude.  J_r and J_z are printed
    only: their FP32 error comes from the displacement of FP32 offsets
    (about 1e-3 of a cell per step), which no summation changes."""
    n = 400000
    errs = {}
    for comp in (False, True):
        orc, gpu = _drifting_pair(ccs, n, compensated=comp)
        orc.step(1)
        gpu.step(1)
        A, B = interior(orc.get_grid("J")), interior(gpu.get_grid("J"))
        errs[comp] = [np.abs(B[:, c] - A[:, c]).max() / np.abs(A[:, c]).max() for c in range(3)]
        gpu.close()
    print(ccs, "J (r, phi, z) error plain", errs[False], "compensated", errs[True])
    assert errs[True][1] <= 0.5 * errs[False][1], errs
    assert errs[True][1] <= 2e-7, errs


def test_fp32_compensated_steps_agree_with_plain():
    """Three VB steps with strict axial and radial motions, fields and a
    beam (the reset marks and the fix-up pass run): compensated and plain FP32
    contexts hold the same particles within FP32 tolerance and the same J
    within 1e-5 of its magnitude, graph replays included."""
    orc, plain, species = pair32(ccs="vb", pusher="vay", domains=(2, 2), n=3000, beam=True, fields=True)
    _, comp, _ = pair32(ccs="vb", pusher="vay", domains=(2, 2), n=3000, beam=True, fields=True)
    comp.set_compensated_current(True)
    plain.step(2)
    comp.step(2)
    comp.set_graph(True)
    plain.step(1)
    comp.step(1)
    for sp in range(len(species)):
        a, b = by_id(plain.particles(sp)), by_id(comp.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        for k in ("r", "z", "vr", "vphi", "vz"):
            assert Hh.relerr(b[k], a[k]) <= 1e-5, (sp, k)
    e = Hh.relerr(interior(comp.get_grid("J")), interior(plain.get_grid("J")))
    print("compensated vs plain J", e)
    assert e <= 1e-5
    orc.step(3)
    e = Hh.relerr(interior(comp.get_grid("J")), interior(orc.get_grid("J")))
    print("compensated vs oracle J", e)
    assert e <= 1e-4


def test_compensated_current_is_fp32_only():
    capi = _cap()
    size, grid = Hh.geometry(16, 32)
    ctx = capi.Context(size, grid, (1, 1), 1e-12, [(-1, 1, False, 64)])
    with pytest.raises(RuntimeError):
        ctx.set_compensated_current(True)
    ctx.close()


def test_fp32_fix_list_overflow_falls_back_to_full_screen():
    """Every particle in strict axial motion (v_r = v_phi = 0 exactly, quirk
    Q8 resets in every cell): the fix-up's list of flagged particles (2^20
    entries) overflows with 2 x 700,000 particles, and the redo screens every
    particle itself.  Particles and J match the FP64 oracle at the FP32
    tolerances of the steps test."""
    capi = _cap()
    size, grid = Hh.geometry(16, 32)
    dt = 0.5 * (size[1] / grid[1]) / 3e8
    species = Hh.species_list(False)
    n = 700000
    orc = O.SMB(size, grid, (2, 2), dt, species, pusher="boris-adaptive", ccs="vb")
    gpu = capi.Context(size, grid, (2, 2), dt, [s + (n + 64,) for s in species], pusher="boris-adaptive", ccs="vb",
                       precision="fp32")
    rng = np.random.default_rng(12)
    dr, dz = size[0] / grid[0], size[1] / grid[1]
    for sp, (q, m, b) in enumerate(species):
        r = np.sqrt(rng.uniform((0.5 * dr) ** 2, (size[0] - 0.5 * dr) ** 2, n))
        z = rng.uniform(0.5 * dz, size[1] - 0.5 * dz, n)
        ir, iz = np.floor(r / dr), np.floor(z / dz)
        r = (ir + (r / dr - ir).astype(np.float32).astype(np.float64)) * dr
        z = (iz + (z / dz - iz).astype(np.float32).astype(np.float64)) * dz
        vz = rng.normal(0.0, np.sqrt(1.6e-19 / (m * 9.1e-31)), n).astype(np.float32).astype(np.float64)
        zero = np.zeros(n)
        w = (rng.uniform(0.5, 1.5, n) * 1e7).astype(np.float32).astype(np.float64)
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 7)
        orc.add_particles(sp, r, z, zero, zero, vz, w, ids)
        gpu.upload(sp, r, z, zero, zero, vz, w, ids)
    orc.step(1)
    gpu.step(1)
    for sp in range(len(species)):
        a, b = by_id(orc.particles(sp)), by_id(gpu.particles(sp))
        assert np.array_equal(a["id"], b["id"])
        for k in ("r", "z", "vz"):
            assert Hh.relerr(b[k], a[k]) <= 1e-5, (sp, k)
    e = Hh.relerr(interior(gpu.get_grid("J")), interior(orc.get_grid("J")))
    print("fix-up fallback J error", e)
    assert e <= 1e-4
    gpu.close()
