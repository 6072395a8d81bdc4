"""Graph replay of the step: with graph mode on, the kernels of a step (after
the sorts) are captured once per buffer state and replayed.  Every replay
must agree with the eager launches, including after sorts (which exchange
the buffers the graphs are keyed by), with the library's own eager work
(grid reads, a kinetic energy reduction) between the replays.  Steps with a
collision phase (P12 in FP64 and FP32, the fused surrogate in FP32) replay
too: their pass number is a device counter that the graph advances, so a
replay that reused a captured number would draw the same pairs again and
differ by O(1); the beam is sorted every step, and the cache of captured
graphs stays small.

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


def diff(x, y):
    if x.dtype == np.uint64:
        return 0.0 if np.array_equal(x, y) else np.inf
    return np.abs(x - y).max(initial=0.0) / max(np.abs(x).max(initial=0.0), 1e-300)


@pytest.mark.parametrize("precision,scheme", [("fp64", "p12"), ("fp32", "p12"), ("fp32", "surrogate")])
def test_graph_replay_with_collisions_matches_eager(precision, scheme):
    """Two eager contexts and a graph context on the same plasma (n =
    5e19-2e20 m^-3, time step 1e-13 s: plasma frequency x time step below 0.1).  The fields
    and the like-particle pairs of P12 amplify rounding (the atomics' order
    alone separates two eager runs), so the runs are compared after the
    first step -- the graph run within ten times the eager runs' own
    difference, or the rounding level of the precision (1e-12 FP64, 1e-6
    FP32: the atomics' order differs between launch modes) -- and then
    stepped on: the device's collision
    pass counter, which keys every pass's draws, advances once per replayed
    step as in the eager runs, and the graph cache stays small."""
    from picopic_amd import surrogate
    from tests.test_gpu_fused_collisions import DR, DZ, GRID, REGIME, network
    DT = 1e-13
    rng = np.random.default_rng(23)
    tepar, teperp, ti, dens = surrogate.regime_cells(rng, GRID, **REGIME)
    data = surrogate.cell_plasma(rng, GRID, DR, DZ, 24, tepar, teperp, ti, dens)
    nb = 400
    beam = (rng.uniform(2 * DR, 10 * DR, nb), rng.uniform(2 * DZ, 30 * DZ, nb), rng.normal(0, 1e5, nb),
            np.zeros(nb), np.full(nb, 1.5e7), np.full(nb, 1e6))
    data = list(data) + [beam]
    ctxs = []
    for graph in (False, False, True):
        cap = max(len(d[0]) for d in data) + 64
        c = capi.Context((GRID[0] * DR, GRID[1] * DZ), GRID, (2, 2), DT,
                         [(-1, 1, False, cap), (1, 1836, False, cap), (-1, 1, True, 2 * nb)], precision=precision)
        for sp, d in enumerate(data):
            c.upload(sp, *d, np.arange(len(d[0]), dtype=np.uint64) + np.uint64(sp << 32))
        if scheme == "surrogate":
            params, norm = network()
            surrogate.install(c, params, norm)
        c.set_collisions(scheme, 0, 1, seed=29)
        c.set_sort_interval(0, 3)
        c.set_sort_interval(1, 3)
        c.set_sort_interval(2, 1)
        c.set_graph(graph)
        ctxs.append(c)
    eager, eager2, graph = ctxs
    print("graph collisions", precision, scheme, flush=True)
    for c in ctxs:
        c.step(1)
    a, a2, b = state(eager, 3), state(eager2, 3), state(graph, 3)
    floor = 1e-12 if precision == "fp64" else 1e-6
    for k, (x, x2, y) in enumerate(zip(a, a2, b)):
        d_eager, d_graph = diff(x, x2), diff(x, y)
        assert d_graph <= max(floor, 10 * d_eager), (k, d_graph, d_eager)
    for c in ctxs:
        c.step(11)
    assert [c.collision_passes() for c in ctxs] == [12, 12, 12]
    launches, captures = graph.graph_stats()
    print("graph launches", launches, "captures", captures)
    assert launches == 12 and captures <= 8


@pytest.mark.parametrize("collisions", [False, True])
def test_step_phase_timing(collisions):
    """pcp_set_timing: eager steps split into sort, fields, push, collide,
    particles, fixup and stitch_j phases (HIP events on the context stream)
    that add up to the step; the push and collision phases are empty without
    collisions and not with them; with graph mode only the step is timed."""
    ctx, nsp = make("fp32", "boris-adaptive", "vb", False)
    if collisions:
        ctx.set_collisions("p12", 0, 1, seed=3)
    L = capi.load()
    L.pcp_set_timing(ctx.h, 1)
    ctx.step(4)
    names = ("sort", "fields", "push", "collide", "particles", "fixup", "stitch_j")
    ph = {k: L.pcp_kernel_ms(ctx.h, k.encode()) for k in names + ("step",)}
    print("phases", collisions, ph)
    assert all(ph[k] >= 0 for k in ph)
    assert abs(sum(ph[k] for k in names) - ph["step"]) <= 1e-3 * ph["step"] + 1e-4
    assert ph["particles"] > 0 and ph["fields"] > 0 and ph["sort"] > 0
    if collisions:
        assert ph["push"] > 0 and ph["collide"] > 0
    else:
        assert ph["push"] < 0.01 and ph["collide"] < 0.01
    ctx.set_graph(True)
    ctx.step(4)
    assert L.pcp_kernel_ms(ctx.h, b"step") > 0
    assert L.pcp_kernel_ms(ctx.h, b"particles") == 0
    ctx.close()
