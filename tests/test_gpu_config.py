"""A device context built from a JSON configuration in the reference's schema
runs the same steps as one built by hand with the same parameters."""
import copy
import json
import os

import numpy as np
import pytest

from picopic_amd import capi, config

pytestmark = pytest.mark.gpu

PATH = os.path.join(os.path.dirname(__file__), "golden", "config_uniform.json")


def small():
    cfg = copy.deepcopy(config.read(PATH))
    cfg["geometry"]["size"] = {"radius": 0.032, "longitude": 0.064}
    cfg["geometry"]["grid"] = {"radius": 32, "longitude": 64}
    cfg["macro_amount"] = 32 * 64 * 2 * 8
    cfg["beams"] = []
    cfg["device"]["init"] = "cells"
    return cfg


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config_context_equals_manual(precision):
    cfg = small()
    cfg["device"]["precision"] = precision
    ctx, sp = config.context(cfg, seed=5)
    assert config.per_cell(cfg) == 8
    dt = config.time(cfg)[2]
    man = capi.Context((0.032, 0.064), (32, 64), (2, 4), dt,
                       [(-1, 1, False, ctx.cfg.species[0].capacity), (1, 1836, False, ctx.cfg.species[1].capacity)],
                       precision=precision)
    man.init_plasma(0, 8, 1e17, 1e17, 1.0, 5)
    man.init_plasma(1, 8, 1e17, 1e17, 0.1, 6)
    for c in (ctx, man):
        c.step(3)
    # the deposits' atomics sum in a different order in the two contexts
    tol = 1e-12 if precision == "fp64" else 1e-5
    for s in range(2):
        assert ctx.count(s) == man.count(s) == 32 * 64 * 8
        a, b = ctx.particles(s), man.particles(s)
        oa, ob = np.argsort(a["id"]), np.argsort(b["id"])
        assert np.array_equal(a["id"][oa], b["id"][ob])
        for k in ("r", "z", "vr", "vz"):
            scale = np.abs(a[k]).max()
            assert np.abs(a[k][oa] - b[k][ob]).max() <= tol * scale


def test_config_context_reference_init_and_beam():
    cfg = small()
    cfg["device"]["init"] = "reference"
    cfg["beams"] = copy.deepcopy(config.read(PATH)["beams"])
    cfg["beams"][0]["bunch"]["radius"] = 0.008
    ctx, sp = config.context(cfg, seed=5)
    _, amounts = config.macro_amounts(cfg)
    dt = config.time(cfg)[2]
    man = capi.Context((0.032, 0.064), (32, 64), (2, 4), dt,
                       [(-1, 1, False, amounts[0]), (1, 1836, False, amounts[1]),
                        (-1, 1, True, ctx.cfg.species[2].capacity)])
    man.init_reference(0, amounts[0], 1e17, 1e17, 1.0, 5)
    man.init_reference(1, amounts[1], 1e17, 1e17, 0.1, 6)
    b = cfg["beams"][0]
    man.set_beam(2, amounts[2], 0.0, b["bunch"]["radius"], b["bunch"]["density"], b["bunch"]["amount"],
                 b["bunch"]["length"], b["bunch"]["distance"], b["velocity"], 7)
    for c in (ctx, man):
        c.step(4)
    assert ctx.count(2) == man.count(2) > 0
    for s in range(3):
        a, b = ctx.particles(s), man.particles(s)
        oa, ob = np.argsort(a["id"]), np.argsort(b["id"])
        assert np.array_equal(a["id"][oa], b["id"][ob])
        for k in ("r", "z", "vr", "vz"):
            scale = np.abs(a[k]).max()
            assert np.abs(a[k][oa] - b[k][ob]).max() <= 1e-12 * scale
