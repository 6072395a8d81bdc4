"""Edge cases of the step on the device against the oracle: a species with
no particles, species whose counts are not multiples of a wavefront (1, 63,
65, 129 particles: partial trips, partial 4-particle fix-up groups), in one
and in four domains (the single-domain specialisation and the general
kernel), both precisions and both deposition schemes.  Three steps from
random fields; particles are compared by id, fields over the interior."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as Hh

pytestmark = pytest.mark.gpu


def _pair(precision, ccs, domains, counts, seed=5):
    from picopic_amd import capi
    size, grid = Hh.geometry(16, 32)
    dt = 0.5 * (size[1] / grid[1]) / 3e8
    species = Hh.species_list(True)
    orc = O.SMB(size, grid, domains, dt, species, pusher="boris-adaptive", ccs=ccs)
    gpu = capi.Context(size, grid, domains, dt, [s + (max(4 * n, 16),) for s, n in zip(species, counts)],
                       pusher="boris-adaptive", ccs=ccs, precision=precision)
    rng = np.random.default_rng(seed)
    dr, dz = size[0] / grid[0], size[1] / grid[1]
    for sp, ((q, m, b), n) in enumerate(zip(species, counts)):
        if n == 0:
            continue
        r, z, vr, vp, vz, w = Hh.make_particles(rng, size, grid, n, m, fast_fraction=0.3, special=False,
                                                temp_ev=1.0 if m == 1 else 0.1)
        if precision == "fp32":
            # both sides start from the FP32-representable state
            ir, iz = np.floor(r / dr), np.floor(z / dz)
            r = (ir + (r / dr - ir).astype(np.float32).astype(np.float64)) * dr
            z = (iz + (z / dz - iz).astype(np.float32).astype(np.float64)) * dz
            vr, vp, vz, w = (a.astype(np.float32).astype(np.float64) for a in (vr, vp, vz, w))
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 6)
        orc.add_particles(sp, r, z, vr, vp, vz, w, ids)
        gpu.upload(sp, r, z, vr, vp, vz, w, ids)
    E, H, Hat = Hh.random_fields(rng, orc.n_domains, orc.nr_d, orc.nz_d)
    if precision == "fp32":
        E, H, Hat = (a.astype(np.float32).astype(np.float64) for a in (E, H, Hat))
    for name, arr in (("E", E), ("H", H), ("H_at_et", Hat)):
        orc.set_grid(name, arr)
        gpu.set_grid(name, arr)
    return orc, gpu, species


def _by_id(p):
    o = np.argsort(p["id"])
    return {k: v[o] for k, v in p.items()}


@pytest.mark.parametrize("counts", [(0, 65, 1), (63, 0, 129)])
@pytest.mark.parametrize("domains", [(1, 1), (2, 2)])
@pytest.mark.parametrize("ccs", ["zigzag", "vb"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_empty_and_ragged_species(precision, ccs, domains, counts):
    print("case", precision, ccs, domains, counts, flush=True)
    orc, gpu, species = _pair(precision, ccs, domains, counts)
    tol_p, tol_f = (1e-12, 1e-12) if precision == "fp64" else (1e-5, 1e-4)
    orc.step(3)
    gpu.step(3)
    for sp in range(len(species)):
        a, b = _by_id(orc.particles(sp)), _by_id(gpu.particles(sp))
        assert len(a["id"]) == len(b["id"]) == gpu.count(sp)
        if counts[sp] == 0 or len(a["id"]) == 0:
            assert len(b["id"]) == len(a["id"])
            continue
        assert np.array_equal(a["id"], b["id"])
        for k in ("r", "z", "vr", "vphi", "vz"):
            e = Hh.relerr(b[k], a[k])
            assert e <= tol_p, (sp, k, e)
    for nm in ("E", "H", "J"):
        e = Hh.relerr(gpu.get_grid(nm)[:, :, 2:-2, 2:-2], orc.get_grid(nm)[:, :, 2:-2, 2:-2])
        print(nm, e)
        assert e <= tol_f, (nm, e)
