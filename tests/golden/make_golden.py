"""Writes the known-answer fixtures held in this folder.

Sources (data only, no reference code is run here):
  * ``kat_reference.json`` -- the known-answer vector printed by the reference
    compiled with IEEE flags (g++ 11.4 / glibc 2.35) for one electron in a
    uniform field on a 16x16 domain: gathered E and H and the Boris velocity.
    The numbers are transcribed from the survey of the reference.
  * the overlay known answers of test/unit/gridTest.cpp:111-215 (o_s = 3,
    interior/halo 8.3 and 20.3 -> 28.6).
  * ``functional_true_data.json`` -- the [mean, std, var] triples the
    reference's functional test compares against
    (test/functional/true_data/<E|H|J>/<r|phi|z>/rec/0-32_0-128/4.dat), read
    as numbers from a checkout of the reference when one is present.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

kat = {
    "source": "reference compiled with -O3 -ffp-contract=off (IEEE), g++ 11.4, glibc 2.35",
    "domain": {"grid": [16, 16], "dr": 1e-3, "dz": 1e-3, "walls": [1, 1, 1, 1],
               "dt": 0.5e-3 / 3e8},
    "fields": {"E": [2e3, 2e3, 1e4], "H": [10.0, 10.0, 50.0],
               "note": "uniform on every node including the halo"},
    "particle": {"species": "electron", "charge_e": -1, "mass_me": 1,
                 "r": 5.3e-3, "z": 7.7e-3, "v": [1e5, -2e5, 3e5]},
    "gather_e": [1831.9999999999986, 1467.3706293706277, 9527.3146853146864],
    "gather_h": [9.162237762237762, 9.1600000000000001, 45.799999999999969],
    "boris_relativistic_v": [99467.547189765683, -200429.32376443502, 297207.10309333273],
}

grid_kat = {
    "source": "test/unit/gridTest.cpp:111-215",
    "o_s": 3, "size": [10, 10], "value_a": 8.3, "value_b": 20.3, "sum": 28.6,
    "overlay_x": {"a_eq_b": [[[15, 4], [0, 4]], [[14, 5], [1, 5]], [[13, 6], [2, 6]]],
                  "a_is_sum": [[15, 4], [14, 5], [13, 6]],
                  "b_not_sum": [[15, 4], [14, 5], [13, 6]]},
    "overlay_y": {"a_eq_b": [[[4, 15], [4, 0]], [[5, 14], [5, 1]], [[6, 13], [6, 2]]],
                  "a_is_sum": [[4, 15], [5, 14], [6, 13]],
                  "b_not_sum": [[4, 15], [5, 14], [6, 13]]},
    "overlay_xy": {"a_corner_eq_b_corner": True, "corner": 13},
}


def functional_true_data(root="/root/reference"):
    base = os.path.join(root, "test", "functional", "true_data")
    out = {"source": "test/functional/true_data/<E|H|J>/<r|phi|z>/rec/0-32_0-128/4.dat of cosmonaut-ok/PiCoPiC: "
                     "[mean, std, var] of the 32 x 128 probe (r 0..32, z 0..128) at slice 4 (schedule 20) of "
                     "test/functional/picopic.json.tmpl with macro_amount 2.1e5; compared by "
                     "test/functional/test.py:169-222 (rtol 0.15 on std and var)",
           "values": {}}
    for f in ("E", "H", "J"):
        for c in ("r", "phi", "z"):
            with open(os.path.join(base, f, c, "rec", "0-32_0-128", "4.dat")) as fh:
                out["values"][f + "/" + c] = [float(x) for x in fh.read().split()]
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "kat_reference.json"), "w") as f:
        json.dump(kat, f, indent=1)
    with open(os.path.join(HERE, "kat_grid.json"), "w") as f:
        json.dump(grid_kat, f, indent=1)
    if os.path.isdir("/root/reference"):
        with open(os.path.join(HERE, "functional_true_data.json"), "w") as f:
            json.dump(functional_true_data(), f, indent=1)
