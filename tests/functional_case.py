"""The reference's functional test case (test/functional/picopic.json.tmpl with
macro_amount 2.1e5, as test/functional/test.py:118-122 renders it) and its
statistics check (test/functional/test.py:169-222).

The reference compares [mean, std, var] of the 32 x 128 probes (r 0..32,
z 0..128) of E, H (at integer time) and J at slice 4 of schedule 20 -- the
iteration at t = 80 dt, i.e. after 81 steps -- with the values its
``collect_true_data`` averaged over many runs, passing when std and var are
within rtol 0.15 (the mean is computed but not checked).
"""
import json
import os

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "functional_true_data.json")
COMPONENTS = ["E/r", "E/phi", "E/z", "H/r", "H/phi", "H/z", "J/r", "J/phi", "J/z"]
RTOL = 0.15
SLICE, SCHEDULE = 4, 20
STEPS = SLICE * SCHEDULE + 1

CONFIG = {
    "macro_amount": 2.1e5,
    "beam_plasma_macro_size_ratio": 2,
    "geometry": {
        "size": {"radius": 0.0375, "longitude": 0.15},
        "grid": {"radius": 128, "longitude": 256},
        "domains_amount": {"radius": 2, "longitude": 4},
        "pml": {"left_wall": 0, "right_wall": 0.05, "outer_wall": 0.2, "sigma_1": 1e-5, "sigma_2": 7e-2},
    },
    "time": {"start": 0, "end": 1e-11, "step": 1e-13},
    "particles": [
        {"name": "electrons", "charge": -1, "mass": 1, "density": 1e17, "temperature": 1},
        {"name": "ions", "charge": 1, "mass": 1836, "density": 1e17, "temperature": 0.1},
    ],
    "beams": [
        {"name": "electrons", "charge": -1, "mass": 1, "velocity": 2.8e8, "start_time": 0,
         "bunch": {"amount": 2, "distance": 0.0914, "length": 0.0056, "radius": 0.005, "density": 1e16}},
    ],
    "data": {
        "data_root": ".",
        "probes": [{"component": c, "schedule": SCHEDULE, "shape": "rec", "start": {"r": 0, "z": 0},
                    "end": {"r": 32, "z": 128}} for c in COMPONENTS],
    },
}


def true_data():
    with open(GOLD) as f:
        return json.load(f)["values"]


def statistics(frames):
    """{component: (mean, std, var, std ratio, var ratio)} of 32 x 128 frames."""
    import numpy as np
    ref = true_data()
    out = {}
    for c in COMPONENTS:
        x = np.asarray(frames[c], dtype=np.float64).ravel()
        m, s, v = np.nanmean(x), np.nanstd(x), np.nanvar(x)
        out[c] = (m, s, v, s / ref[c][1], v / ref[c][2])
    return out
