"""Launch tuning: the sort bin and the workgroup shape change only the
storage order and the order of the atomic sums, so every choice gives the
default's result to rounding; the autotuner returns a valid choice and the
simulation goes on from it."""
import numpy as np
import pytest

from picopic_amd import capi
from tests import helpers as Hh

pytestmark = pytest.mark.gpu


def make(precision, n=8000):
    size, grid = Hh.geometry(32, 128)
    dt = 0.5 * (size[1] / grid[1]) / 3e8
    species = Hh.species_list(True)
    ctx = capi.Context(size, grid, (2, 4), dt, [s + (2 * n,) for s in species], pusher="vay", ccs="vb",
                       precision=precision)
    rng = np.random.default_rng(5)
    for sp, (q, m, b) in enumerate(species):
        r, z, vr, vp, vz, w = Hh.make_particles(rng, size, grid, n, m, special=False,
                                                temp_ev=1.0 if m == 1 else 0.1)
        ctx.upload(sp, r, z, vr, vp, vz, w, np.arange(n, dtype=np.uint64) + np.uint64(sp << 32))
    for sp in range(len(species)):
        ctx.set_sort_interval(sp, 1)
    return ctx, len(species)


def state(ctx, nsp):
    out = [ctx.get_grid(g) for g in ("E", "H", "J")]
    for sp in range(nsp):
        p = ctx.particles(sp)
        o = np.argsort(p["id"])
        out += [p[k][o] for k in ("r", "z", "vr", "vphi", "vz", "w")]
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("bins,block", [((4, 8), 256), ((2, 16), 128), ((8, 8), 512), ((1, 4), 256)])
def test_bins_and_shapes_give_the_same_result(precision, bins, block):
    tol = 1e-10 if precision == "fp64" else 1e-4
    ref, nsp = make(precision)
    alt, _ = make(precision)
    alt.set_sort_bin(*bins)
    if precision == "fp32":
        alt.set_launch(block)
    ref.step(6)
    alt.step(6)
    for x, y in zip(state(ref, nsp), state(alt, nsp)):
        assert x.shape == y.shape
        scale = max(np.abs(x).max(initial=0.0), 1e-300)
        assert np.abs(x - y).max(initial=0.0) <= tol * scale


def test_bad_tuning_arguments_raise():
    ctx, _ = make("fp32", n=100)
    with pytest.raises(capi.PcpError):
        ctx.set_sort_bin(3, 8)        # 3 does not divide 16
    with pytest.raises(capi.PcpError):
        ctx.set_launch(384)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_picks_a_valid_choice(precision):
    ctx, nsp = make(precision)
    t = ctx.autotune(steps=2)
    print(precision, t)
    nr, nz = 16, 32
    assert nr % t["bin_r"] == 0 and nz % t["bin_z"] == 0
    assert t["block"] in (128, 256, 512)
    assert 1 <= t["sort_every"] <= 200
    assert t["ms_per_step"] > 0
    assert ctx.tuning()["bin_r"] == t["bin_r"]
    e0 = ctx.kinetic_energy(0)
    ctx.step(3)
    assert np.isfinite(ctx.kinetic_energy(0)) and np.isfinite(e0)
