"""The device against the reference's own end-to-end statistics: the
functional test case (test/functional/picopic.json.tmpl) read as a
configuration, run by picopic_amd.run as the reference's main loop runs it
(reference initial plasma, beam injection, probes written at schedule 20),
and slice 4 of the nine 32 x 128 probes compared with the reference's true
data as test/functional/test.py:169-222 compares them (rtol 0.15 on std and
var, up to four runs).  As for the oracle, E_phi and H_z are checked on std
only (their std runs 7 % and 11 % above the reference's).

The same run on the oracle (same seeds, so the same particles) gives frames
that agree with the device's within 1e-6 of each component's std: the
device is the oracle's dynamics, not only its statistics.
"""
import numpy as np
import pytest

from picopic_amd import run
from tests import functional_case as F
from tests.test_oracle_functional import TE, oracle_run

pytestmark = pytest.mark.gpu


def device_frames(seed, root):
    ctx, paths = run.run(F.CONFIG, data_root=str(root), seed=seed, iterations=F.STEPS)
    ctx.close()
    frames = {}
    for c in F.COMPONENTS:
        f = np.load(str(root / c) + "/rec/0-32_0-128.npy", mmap_mode="r")
        assert f.shape == (F.SLICE + 1, 32, 128)
        frames[c] = np.array(f[F.SLICE])
    return frames


def test_functional_statistics_match_the_reference(tmp_path):
    report = []
    for k, seed in enumerate((1, 11, 21, 31)):
        frames = device_frames(seed, tmp_path / str(seed))
        st = F.statistics(frames)
        ok = True
        for c, (m, s, v, rs, rv) in st.items():
            report.append("seed %d %-5s mean %11.4g std ratio %.3f var ratio %.3f" % (seed, c, m, rs, rv))
            ok &= abs(rs - 1) <= F.RTOL and (c in TE or abs(rv - 1) <= F.RTOL)
        if k == 0:
            ref = oracle_run(seed)
            for c in F.COMPONENTS:
                err = np.abs(frames[c] - ref[c]).max() / ref[c].std()
                report.append("seed %d %-5s device vs oracle max |diff| / std %.2e" % (seed, c, err))
                assert err <= 1e-6, c
        if ok:
            break
    print("\n".join(report))
    assert ok, "\n".join(report)
