"""The counting sort on the device.  A workgroup counts chunks of 1024
consecutive particles per key in a 256-entry LDS table and falls back to the
global counters for keys that find no entry; the cases drive both paths:
random order over 4096 cells (most chunks hold more distinct keys than the
table has entries), two crowded cells holding 20,000 particles (one or two
keys per chunk, ranges of thousands of slots), and the crowded cells mixed
with a random sprinkle.  Counts are not multiples of the chunk.  Both
precisions, one and four domains.  After the sort the storage is in
(domain, cell) order and every particle keeps its data bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRID = (64, 64)
DR = DZ = 1e-3


def _positions(rng, case):
    def random(n):
        return rng.uniform(0.5 * DR, (GRID[0] - 0.5) * DR, n), rng.uniform(0.5 * DZ, (GRID[1] - 0.5) * DZ, n)

    def crowded(n):
        r = (10 + rng.uniform(0.01, 0.99, n)) * DR
        z = (20 + rng.integers(0, 2, n) + rng.uniform(0.01, 0.99, n)) * DZ
        return r, z

    if case == "random":
        return random(50_123)
    if case == "crowded":
        return crowded(20_000)
    a, b = crowded(20_000), random(3_001)
    r, z = np.concatenate([a[0], b[0]]), np.concatenate([a[1], b[1]])
    o = rng.permutation(len(r))
    return r[o], z[o]


def _by_id(p):
    o = np.argsort(p["id"])
    return {k: v[o] for k, v in p.items() if k not in ("order", "r_old", "z_old")}


@pytest.mark.parametrize("case", ["random", "crowded", "mixed"])
@pytest.mark.parametrize("domains", [(1, 1), (2, 2)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sort_orders_by_cell_and_keeps_every_particle(precision, domains, case):
    from picopic_amd import capi
    print("sort", precision, domains, case, flush=True)
    rng = np.random.default_rng(3)
    r, z = _positions(rng, case)
    n = len(r)
    ctx = capi.Context((GRID[0] * DR, GRID[1] * DZ), GRID, domains, 0.5 * DZ / 3e8, [(-1, 1, False, n + 64)],
                       pusher="boris-adaptive", ccs="zigzag", precision=precision)
    v = rng.normal(0, 1e5, (3, n))
    ctx.upload(0, r, z, v[0], v[1], v[2], rng.uniform(1.0, 2.0, n), np.arange(n, dtype=np.uint64) + 7)
    before = _by_id(ctx.particles(0))
    ctx.sort(0)
    after = ctx.particles(0)
    ctx.close()
    assert len(after["id"]) == n
    a = _by_id(after)
    assert np.array_equal(a["id"], np.arange(n, dtype=np.uint64) + 7)
    for k in ("r", "z", "vr", "vphi", "vz", "w"):
        assert np.array_equal(a[k], before[k]), k
    cr, cz = np.floor(after["r"] / DR).astype(np.int64), np.floor(after["z"] / DZ).astype(np.int64)
    nr_d, nz_d = GRID[0] // domains[0], GRID[1] // domains[1]
    dom = (cr // nr_d) * domains[1] + cz // nz_d
    key = dom * 10 ** 12 + cr * 10 ** 6 + cz
    assert np.all(np.diff(key) >= 0)
