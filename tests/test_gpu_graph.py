"""Graph replay of the step: with graph mode on, the kernels of a step (after
the sorts) are captured once per buffer parity and replayed.  Every replay
must agree with the eager launches, including after sorts (which rebuild the
graphs), with the library's own eager work (grid reads, a kinetic energy
reduction) between the replays.

The deposits accumulate with atomics, whose order differs from run to run,
so two runs (eager or not) agree to rounding, not bit for bit: FP64 to 1e-10
and FP32 to 1e-4 of each array's largest value after 13 steps.  A replay of
the wrong buffer parity or of stale arguments differs by O(1)."""
import numpy as np
import pytest

from picopic_amd import capi
from tests import helpers as Hh

pytestmark = pytest.mark.gpu


def make(precision, pusher, ccs, graph):
    size, grid = Hh.geometry(32, 64)
    dt = 0.5 * (size[1] / grid[1]) / 3e8
    species = Hh.species_list(True)
    n = 6000
    ctx = capi.Context(size, grid, (2, 2), dt, [s + (2 * n,) for s in species], pusher=pusher, ccs=ccs,
                       precision=precision)
    rng = np.random.default_rng(17)
    for sp, (q, m, b) in enumerate(species):
        r, z, vr, vp, vz, w = Hh.make_particles(rng, size, grid, n, m, fast_fraction=0.3,
                                                temp_ev=1.0 if m == 1 else 0.1)
        ids = np.arange(n, dtype=np.uint64) + np.uint64(sp * 10 ** 6)
        ctx.upload(sp, r, z, vr, vp, vz, w, ids)
    E, H, Hat = Hh.random_fields(rng, 4, grid[0] // 2, grid[1] // 2)
    for name, arr in (("E", E), ("H", H), ("H_at_et", Hat)):
        ctx.set_grid(name, arr)
    for sp in range(len(species)):
        ctx.set_sort_interval(sp, 3)
    ctx.set_graph(graph)
    return ctx, len(species)


def state(ctx, nsp):
    out = [ctx.get_grid(g) for g in ("E", "H", "J")]
    for sp in range(nsp):
        p = ctx.particles(sp)
        o = np.argsort(p["id"])
        out += [p[k][o] for k in sorted(p) if k != "order"]
    return out


def close(x, y, tol):
    assert x.shape == y.shape
    if x.dtype == np.uint64:
        return np.array_equal(x, y)
    scale = max(np.abs(x).max(initial=0.0), 1e-300)
    return np.abs(x - y).max(initial=0.0) <= tol * scale


@pytest.mark.parametrize("precision,pusher,ccs", [("fp64", "boris-adaptive", "vb"), ("fp32", "vay", "vb"),
                                                  ("fp32", "hc", "zigzag")])
def test_graph_replay_matches_eager(precision, pusher, ccs):
    tol = 1e-10 if precision == "fp64" else 1e-4
    eager, nsp = make(precision, pusher, ccs, False)
    graph, _ = make(precision, pusher, ccs, True)
    for t in range(8):
        eager.step(1)
        graph.step(1)
        # eager work of the library between replays
        ke = [graph.kinetic_energy(s) for s in range(nsp)]
        assert all(np.isfinite(ke))
        a, b = state(eager, nsp), state(graph, nsp)
        for x, y in zip(a, b):
            assert close(x, y, tol), "step %d differs" % t
    # several steps in one call (replays back to back)
    eager.step(5)
    graph.step(5)
    for x, y in zip(state(eager, nsp), state(graph, nsp)):
        assert close(x, y, tol)
