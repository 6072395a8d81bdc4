"""Probe configuration and the output schedule (src/cfg.cpp:201-324,
src/outWriter.cpp:49-100, src/PiCoPiC.cpp:299-345), on the host: windows and
paths per shape, rejected shapes, frames written when ceil(t / dt) is a
multiple of the schedule, and the main loop's iteration count (t accumulated
by repeated addition while t <= end)."""
import numpy as np
import pytest

from picopic_amd import probes, run
from tests import functional_case as F


class FakeCtx:
    """Returns frames whose values encode (which, comp, window, call count)."""

    def __init__(self):
        self.calls = 0

    def probe(self, which, comp, r0, z0, r1, z1):
        self.calls += 1
        out = np.zeros((r1 - r0, z1 - z0))
        out[...] = self.calls
        return out


def test_shapes_windows_and_paths():
    grid = (128, 256)
    specs = [
        {"component": "E/r", "schedule": 5, "shape": "rec", "start": {"r": 2, "z": 3}, "end": {"r": 10, "z": 40}},
        {"component": "H/z", "schedule": 5, "shape": "col", "z": 17},
        {"component": "J/phi", "schedule": 5, "shape": "row", "r": 9},
        {"component": "E/z", "schedule": 5, "shape": "dot", "r": 4, "z": 100},
    ]
    p = [probes.Probe(s, grid) for s in specs]
    assert [q.path for q in p] == ["E/r/rec/2-10_3-40", "H/z/col/17", "J/phi/row/9", "E/z/dot/4_100"]
    assert [q.window for q in p] == [(2, 3, 10, 40), (0, 17, 128, 18), (9, 0, 10, 256), (4, 100, 5, 101)]
    assert [q.which for q in p] == ["E", "H_at_et", "J", "E"]
    assert [q.comp for q in p] == [0, 2, 1, 2]
    assert [q.frame_shape() for q in p] == [(8, 37), (128,), (256,), (1,)]


@pytest.mark.parametrize("spec", [
    {"component": "E/r", "schedule": 5, "shape": "rec", "start": {"r": 10, "z": 0}, "end": {"r": 2, "z": 4}},
    {"component": "E/r", "schedule": 5, "shape": "rec", "start": {"r": 0, "z": 0}, "end": {"r": 129, "z": 4}},
    {"component": "E/r", "schedule": 5, "shape": "ring", "r": 1},
    {"component": "B/r", "schedule": 5, "shape": "dot", "r": 1, "z": 1},
    {"component": "E/r", "schedule": 0, "shape": "dot", "r": 1, "z": 1},
])
def test_bad_probes_are_rejected(spec):
    with pytest.raises(ValueError):
        probes.Probe(spec, (128, 256))


def test_writer_schedule(tmp_path):
    cfg = {"data": {"probes": [
        {"component": "E/r", "schedule": 3, "shape": "rec", "start": {"r": 0, "z": 0}, "end": {"r": 2, "z": 2}},
        {"component": "J/z", "schedule": 4, "shape": "dot", "r": 1, "z": 1}]}}
    w = probes.Writer(cfg, (8, 8), 12, str(tmp_path))
    ctx = FakeCtx()
    dt = 1e-13
    t = 0.0
    calls = []
    for it in range(13):
        before = ctx.calls
        w(ctx, t, dt)
        calls.append(ctx.calls - before)
        t += dt
    w.close()
    # iterations whose ceil(t / dt) is a multiple of 3 (E/r) and of 4 (J/z)
    t, expect = 0.0, []
    for it in range(13):
        k = int(np.ceil(t / dt))
        expect.append(int(k % 3 == 0) + int(k % 4 == 0))
        t += dt
    assert calls == expect
    e = np.load(str(tmp_path / "E/r/rec/0-2_0-2.npy"))
    j = np.load(str(tmp_path / "J/z/dot/1_1.npy"))
    assert e.shape == (5, 2, 2) and j.shape == (4, 1)
    assert np.all(e[0] == 1) and np.all(j[0] == 2)


def test_functional_case_iterations():
    # t = 0 .. 1e-11 by 1e-13: 101 iterations; slice 4 after the 81st step
    assert run.n_iterations(F.CONFIG) == 101
    t, dumps = 0.0, []
    for it in range(101):
        k = int(np.ceil(t / 1e-13))
        if k % 20 == 0:
            dumps.append((it, k // 20))
        t += 1e-13
    assert (F.STEPS - 1, F.SLICE) in dumps
