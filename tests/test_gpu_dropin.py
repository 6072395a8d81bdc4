"""The drop-in solver objects of include/picopic_solvers.hpp driven the way
the reference's SMB drives its domains (per-domain calls from OpenMP loops,
then the SMB phase boundary): every reference phase mapped to its own entry
point must give the same end state as one fused step per time step."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_phase_by_phase_equals_fused_step():
    drv = os.path.join(ROOT, "picopic_amd", "_lib", "smb_driver")
    env = dict(os.environ, OMP_NUM_THREADS="4")
    print("smb_driver 5 steps, 2x2 domains", flush=True)
    res = subprocess.run([drv, "5"], capture_output=True, text=True, timeout=100, env=env)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    print(out)
    assert out["launched_phases"] == 5 and out["launched_fused"] == 5
    # atomics reorder the deposit sums; everything else is the same arithmetic
    assert out["particles"] <= 1e-12
    for k in ("E", "H", "J"):
        assert out[k] <= 1e-11, (k, out[k])
