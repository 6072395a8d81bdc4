"""The reference's JSON run configuration (PiCoPiC.json) read into a device
context.

Keys read as the reference reads them (src/cfg.cpp):

- ``geometry.size.radius/longitude`` [m], ``geometry.grid.radius/longitude``
  [cells], ``geometry.domains_amount.radius/longitude``
  (``Cfg::init_geometry``, src/cfg.cpp:377-426);
- ``geometry.pml.left_wall/right_wall/outer_wall`` (fractions of the grid)
  and ``sigma_1/sigma_2`` when PML is on (src/cfg.cpp:394-408);
- ``time.start/end/step`` [s] (``Cfg::init_time``, src/cfg.cpp:428-437);
- ``particles[]``: ``name``, ``charge`` [e], ``mass`` [electron masses]
  (both truncated to integers as the reference does, src/cfg.cpp:188-189),
  ``density`` [m^-3], ``temperature`` [eV] (``Cfg::init_particles``,
  src/cfg.cpp:162-199);
- ``beams[]``: ``name`` (the species is ``beam_<name>``), ``charge``,
  ``mass``, ``velocity`` [m/s] and ``bunch.amount/distance/length/radius/
  density`` (``Cfg::init_beam``, src/cfg.cpp:326-375);
- ``macro_amount`` and ``beam_plasma_macro_size_ratio``: split over the
  species as ``Cfg::weight_macro_amount`` does (src/cfg.cpp:488-570): a
  plasma species gets norm macro-particles per cell and a beam
  r_cells x z_cells x norm x ratio, with norm = macro_amount over the sum of
  those sizes.

The switches the reference takes from ``configure`` become additive keys of
an optional ``"device"`` object (none of them changes a reference key):

- ``pusher`` (``boris``, ``boris-relativistic``, ``boris-adaptive``, ``vay``,
  ``hc``; default ``boris-adaptive``, the reference default), ``ccs``
  (``vb`` or ``zigzag``), ``pml`` (bool, ENABLE_PML; default true),
  ``precision`` (``fp64`` or ``fp32``), ``collisions`` (``p12``, ``ta77s``,
  ``sk98``, ``surrogate`` or absent), ``sort_bin`` ([bin_r, bin_z]),
  ``graph`` (bool), ``capacity_factor`` (particle slots per initial
  particle, default 1.1);
- per plasma species, ``right_density``: the z ramp of SMB::SMB
  (src/SMB.cpp:134-136), which the reference's reader sets equal to
  ``density``.

The initial plasma is the reference's own (``device.init`` = ``"reference"``,
the default): ``pcp_init_reference`` places each species' macro amount as
SMB::SMB + Domain::distribute do (flat placement, r-proportional weights,
Maxwell-Juettner thermal velocities; seeded Philox draws in place of the
reference's time seed).  ``device.init`` = ``"cells"`` draws the synthetic
per-cell plasma of ``pcp_init_plasma`` instead (uniform in r^2 inside each
cell, equal weights, the rounded norm per cell).  ``beams[]`` become beam
species injected over time (``pcp_set_beam``, BeamP::inject).  Collisions
pair the plasma species named ``ions`` with the first other plasma species,
as ``Collisions`` sorts them (src/collisions.cpp:87-108).
"""
import json
import math

from . import capi

PUSHERS = tuple(capi.PUSHERS)


def read(src):
    """The configuration as a dict, from a path, a JSON string or a dict."""
    if isinstance(src, dict):
        return src
    text = src
    if not src.lstrip().startswith("{"):
        with open(src) as f:
            text = f.read()
    return json.loads(text)


def _dev(cfg):
    return cfg.get("device", {})


def geometry(cfg):
    """(size, grid, domains) as the reference's Geometry gets them."""
    g = cfg["geometry"]
    size = (float(g["size"]["radius"]), float(g["size"]["longitude"]))
    grid = (int(g["grid"]["radius"]), int(g["grid"]["longitude"]))
    domains = (int(g["domains_amount"]["radius"]), int(g["domains_amount"]["longitude"]))
    return size, grid, domains


def pml(cfg):
    """(left, right, outer, sigma_1, sigma_2) or None when PML is off."""
    if not _dev(cfg).get("pml", True) or "pml" not in cfg["geometry"]:
        return None
    p = cfg["geometry"]["pml"]
    vals = (float(p["left_wall"]), float(p["right_wall"]), float(p["outer_wall"]), float(p["sigma_1"]),
            float(p["sigma_2"]))
    return None if vals[:3] == (0.0, 0.0, 0.0) else vals


def time(cfg):
    t = cfg["time"]
    return float(t["start"]), float(t["end"]), float(t["step"])


def steps(cfg):
    """Number of time steps from time.start to time.end."""
    start, end, dt = time(cfg)
    return int(math.floor((end - start) / dt + 0.5))


def species(cfg):
    """[(name, charge_e, mass_me, is_beam, spec)] -- plasma species first,
    then beams (named beam_<name>), in file order."""
    out = []
    for p in cfg.get("particles", []):
        out.append((p["name"], int(p["charge"]), int(p["mass"]), False, p))
    for b in cfg.get("beams", []):
        out.append(("beam_" + b["name"], int(b["charge"]), int(b["mass"]), True, b))
    names = [o[0] for o in out]
    if len(set(names)) != len(names):
        raise ValueError("species names must be unique (src/cfg.cpp:176-179, 344-347)")
    return out


def macro_amounts(cfg):
    """(norm, [macro-particles of each species]) as Cfg::weight_macro_amount
    computes them (src/cfg.cpp:488-570)."""
    size, grid, _ = geometry(cfg)
    dr, dz = size[0] / grid[0], size[1] / grid[1]
    sp = species(cfg)
    ratio = float(cfg.get("beam_plasma_macro_size_ratio", 1))
    n_plasma = sum(1 for s in sp if not s[3])
    norm_sum = float(grid[0] * grid[1] * n_plasma)
    beam_cells = []
    for s in sp:
        if s[3]:
            b = s[4]["bunch"]
            cells = (float(b["radius"]) / dr) * (float(b["length"]) * float(b["amount"]) / dz)
            beam_cells.append(cells)
            norm_sum += cells * ratio
    norm = float(cfg["macro_amount"]) / norm_sum
    out, k = [], 0
    for s in sp:
        if s[3]:
            out.append(int(beam_cells[k] * norm * ratio))
            k += 1
        else:
            out.append(int(grid[0] * grid[1] * norm))
    return norm, out


def per_cell(cfg):
    """Macro-particles per cell of a plasma species (the rounded norm)."""
    return max(1, int(round(macro_amounts(cfg)[0])))


def beam_capacity(cfg, spec, amount):
    """Slots a beam needs for every particle BeamP::inject can place: per
    domain at z = 0 reached by the radius, ceil of its per-step rate
    (src/beamP.cpp:40-66) for each step of the bunch windows (only domain
    (0, 0) passes to later bunches)."""
    size, grid, domains = geometry(cfg)
    _, _, dt = time(cfg)
    dr = size[0] / grid[0]
    nrd = grid[0] // domains[0]
    b = spec["bunch"]
    radius, length, bunches = float(b["radius"]), float(b["length"]), int(b["amount"])
    velocity = float(spec["velocity"])
    steps = int(math.ceil(length / velocity / dt)) + 2
    total = 0
    for i in range(domains[0]):
        bot, top = i * nrd * dr, (i + 1) * nrd * dr
        if radius > top:
            share = size[0] / domains[0]
        elif radius > bot:
            share = radius - bot
        else:
            share = 0.0
        rate = (amount // bunches) * (share / radius) * velocity * dt / length
        total += int(math.ceil(rate)) * steps * (bunches if i == 0 else 1)
    return max(1, total)


def context(src, seed=1):
    """A device context for the configuration, with the plasma species
    initialised.  Returns (ctx, species)."""
    cfg = read(src)
    dev = _dev(cfg)
    size, grid, domains = geometry(cfg)
    _, _, dt = time(cfg)
    pusher = dev.get("pusher", "boris-adaptive")
    if pusher not in PUSHERS:
        raise ValueError("unknown pusher %r" % pusher)
    ccs = dev.get("ccs", "vb")
    if ccs not in capi.CCS:
        raise ValueError("unknown current scheme %r" % ccs)
    sp = species(cfg)
    if len(sp) > 8:
        raise ValueError("at most 8 species")
    ppc = per_cell(cfg)
    _, amounts = macro_amounts(cfg)
    init = dev.get("init", "reference")
    if init not in ("reference", "cells"):
        raise ValueError("unknown device.init %r" % init)
    factor = float(dev.get("capacity_factor", 1.1))
    cells = grid[0] * grid[1]
    descs = []
    for (name, q, m, beam, spec), amount in zip(sp, amounts):
        if beam:
            cap = beam_capacity(cfg, spec, amount)
        elif init == "reference":
            cap = max(1, amount)
        else:
            cap = int(math.ceil(ppc * cells * factor))
        descs.append((q, m, beam, cap))
    ctx = capi.Context(size, grid, domains, dt, descs, pusher=pusher, ccs=ccs, pml=pml(cfg),
                       precision=dev.get("precision", "fp64"))
    for k, ((name, q, m, beam, spec), amount) in enumerate(zip(sp, amounts)):
        if beam:
            b = spec["bunch"]
            ctx.set_beam(k, amount, float(spec["start_time"]), float(b["radius"]), float(b["density"]),
                         int(b["amount"]), float(b["length"]), float(b["distance"]), float(spec["velocity"]),
                         seed + k)
            continue
        left = float(spec["density"])
        right = float(spec.get("right_density", left))
        if init == "reference":
            ctx.init_reference(k, amount, left, right, float(spec["temperature"]), seed + k)
        else:
            ctx.init_plasma(k, ppc, left, right, float(spec["temperature"]), seed + k)
    if "sort_bin" in dev:
        ctx.set_sort_bin(*dev["sort_bin"])
    if dev.get("collisions"):
        plasma = [k for k, s in enumerate(sp) if not s[3]]
        ions = [k for k in plasma if sp[k][0] == "ions"]
        others = [k for k in plasma if sp[k][0] != "ions"]
        if not ions or not others:
            raise ValueError("collisions need a plasma species named 'ions' and one other plasma species")
        ctx.set_collisions(dev["collisions"], others[0], ions[0], int(dev.get("collision_seed", 0)))
    if dev.get("graph"):
        ctx.set_graph(True)
    return ctx, sp
