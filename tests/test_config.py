"""The JSON run configuration in the reference's schema (src/cfg.cpp): the
geometry, time and species blocks, and the macro-particle split of
Cfg::weight_macro_amount (src/cfg.cpp:488-570), checked by hand on a
configuration file written for these tests."""
import json
import os

import pytest

from picopic_amd import config

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PATH = os.path.join(GOLD, "config_uniform.json")


def test_geometry_time_and_species():
    cfg = config.read(PATH)
    size, grid, domains = config.geometry(cfg)
    assert size == (0.128, 0.512) and grid == (128, 512) and domains == (2, 4)
    assert config.pml(cfg) is None                 # zero-width walls: PML off
    assert config.steps(cfg) == 200
    sp = config.species(cfg)
    assert [s[0] for s in sp] == ["electrons", "ions", "beam_electrons"]
    assert [(s[1], s[2], s[3]) for s in sp] == [(-1, 1, False), (1, 1836, False), (-1, 1, True)]


def test_macro_amount_split():
    cfg = config.read(PATH)
    norm, amounts = config.macro_amounts(cfg)
    # beam: 4 mm / 1 mm radial cells x 2 bunches x 10 mm / 1 mm axial cells
    beam_cells = 4.0 * 20.0
    expect_norm = 2097152 / (128 * 512 * 2 + beam_cells * 2)
    assert norm == pytest.approx(expect_norm, rel=1e-15)
    assert amounts[0] == amounts[1] == int(128 * 512 * expect_norm)
    assert amounts[2] == int(beam_cells * expect_norm * 2)
    assert config.per_cell(cfg) == 16


def test_pml_fractions_and_bad_names(tmp_path):
    cfg = config.read(PATH)
    cfg = json.loads(json.dumps(cfg))
    cfg["geometry"]["pml"].update({"left_wall": 0.1, "right_wall": 0.05, "outer_wall": 0.2})
    assert config.pml(cfg) == (0.1, 0.05, 0.2, 1e-5, 1.5)
    cfg["device"]["pml"] = False
    assert config.pml(cfg) is None
    cfg["particles"].append(dict(cfg["particles"][0]))
    with pytest.raises(ValueError):
        config.species(cfg)
    # a JSON string reads like a path
    assert config.read(json.dumps({"a": 1})) == {"a": 1}
