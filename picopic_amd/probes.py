"""Probe output of a run: the reference's ``data.probes`` (src/cfg.cpp:201-324)
extracted on the device and written as frames.

A probe names a component (``E/r``, ``E/phi``, ``E/z``, ``H/..``, ``J/..``),
a shape and a schedule.  Shapes map onto global cell windows the way
OutController maps them onto the domains (src/outController.cpp:49-309):

- ``rec`` with ``start`` {r, z} and ``end`` {r, z}: [r0, r1) x [z0, z1);
- ``col`` with ``z``: every radial cell at z, [0, Nr) x [z, z + 1);
- ``row`` with ``r``: [r, r + 1) x [0, Nz);
- ``dot`` with ``r`` and ``z``: one cell.

Its path is the reference's (``component/shape/r0-r1_z0-z1``, ``col/z``,
``row/r``, ``dot/r_z``).  A frame is written when ceil(t / dt) is a multiple
of the schedule, t being the time of the step just taken (OutWriter,
src/outWriter.cpp:49-57), as slice ceil(t / dt) / schedule.  Where the
reference's HDF5 engine keeps one dataset per path shaped [frames, ...]
(src/outEngine/outEngineHDF5.cpp:52-117), this writer keeps one ``.npy``
file per path (``<data_root>/<path>.npy``, float64, [frames, r, z] for
``rec``, [frames, n] for ``col`` / ``row``, [frames, 1] for ``dot``).

H probes read the field at integer time (``field_h_at_et``), as the
reference's writers do.  Species components (``density``, ``temperature``,
``energy``, ``momentum/r|phi|z`` with a ``specie`` name; path
``component/specie/shape/...``) read the species map that the dump computes
on the device (``pcp_species_map``: weight_cylindrical, SpecieP::calc_*,
the pre-dump overlay).
"""
import math
import os

import numpy as np

GRID_OF = {"E": "E", "H": "H_at_et", "J": "J"}
COMP_OF = {"r": 0, "phi": 1, "z": 2}


SPECIES_MAPS = ("density", "temperature", "energy", "momentum")


class Probe:
    """One entry of ``data.probes``; `species` maps specie names to indices."""

    def __init__(self, spec, grid, species=None):
        comp = spec["component"]
        head, _, c = comp.partition("/")
        self.map = None
        if head in SPECIES_MAPS:
            if (head == "momentum") != (c in COMP_OF) or (head != "momentum" and c):
                raise ValueError("unsupported probe component %r" % comp)
            name = spec.get("specie", "unknown")
            if species is not None and name not in species:
                raise ValueError("probe %r names an unknown specie %r" % (comp, name))
            self.map, self.specie = head, name
            self.sp = species[name] if species is not None else None
            self.which = "map"
            self.comp = COMP_OF[c] if head == "momentum" else 0
        elif head in GRID_OF and c in COMP_OF:
            self.which = GRID_OF[head]
            self.comp = COMP_OF[c]
        else:
            raise ValueError("unsupported probe component %r" % comp)
        self.schedule = int(spec["schedule"])
        if self.schedule <= 0:
            raise ValueError("probe schedule must be positive")
        shape = spec["shape"]
        Nr, Nz = grid
        if shape == "rec":
            r0, z0 = int(spec["start"]["r"]), int(spec["start"]["z"])
            r1, z1 = int(spec["end"]["r"]), int(spec["end"]["z"])
            tail = "rec/%d-%d_%d-%d" % (r0, r1, z0, z1)
        elif shape == "col":
            z = int(spec["z"])
            r0, z0, r1, z1 = 0, z, Nr, z + 1
            tail = "col/%d" % z
        elif shape == "row":
            r = int(spec["r"])
            r0, z0, r1, z1 = r, 0, r + 1, Nz
            tail = "row/%d" % r
        elif shape == "dot":
            r, z = int(spec["r"]), int(spec["z"])
            r0, z0, r1, z1 = r, z, r + 1, z + 1
            tail = "dot/%d_%d" % (r, z)
        else:
            raise ValueError("unknown probe shape %r" % shape)
        # src/cfg.cpp:299-317: reversed or out-of-area windows are fatal
        if r0 > r1 or z0 > z1 or r1 > Nr or z1 > Nz:
            raise ValueError("probe %s outside the simulation area" % tail)
        self.shape = shape
        self.window = (r0, z0, r1, z1)
        self.path = comp + "/" + (self.specie + "/" if self.map else "") + tail

    def frame_shape(self):
        r0, z0, r1, z1 = self.window
        if self.shape == "rec":
            return (r1 - r0, z1 - z0)
        if self.shape == "col":
            return (r1 - r0,)
        if self.shape == "row":
            return (z1 - z0,)
        return (1,)

    def read(self, ctx):
        r0, z0, r1, z1 = self.window
        if self.map:
            ctx.species_map(self.sp, self.map, download=False)
        return ctx.probe(self.which, self.comp, r0, z0, r1, z1).reshape(self.frame_shape())


def probes(cfg, grid, species=None):
    return [Probe(p, grid, species) for p in cfg.get("data", {}).get("probes", [])]


class Writer:
    """Frames of every probe of a run, one ``.npy`` per probe path."""

    def __init__(self, cfg, grid, n_steps, data_root, species=None):
        self.probes = probes(cfg, grid, species)
        self.root = data_root
        self.files = {}
        for p in self.probes:
            frames = n_steps // p.schedule + 1
            path = os.path.join(data_root, p.path + ".npy")
            os.makedirs(os.path.dirname(path), exist_ok=True)
            self.files[p.path] = np.lib.format.open_memmap(path, mode="w+", dtype=np.float64,
                                                          shape=(frames,) + p.frame_shape())

    def __call__(self, ctx, t, dt):
        """The OutController pass after a step taken at time t."""
        current_time_step = int(math.ceil(t / dt))
        for p in self.probes:
            if current_time_step % p.schedule == 0:
                f = self.files[p.path]
                s = current_time_step // p.schedule
                if s < f.shape[0]:
                    f[s] = p.read(ctx)

    def close(self):
        for f in self.files.values():
            f.flush()
        self.files = {}
