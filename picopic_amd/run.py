"""A run of a configuration in the reference's JSON schema, as the reference's
main loop drives it (src/PiCoPiC.cpp:293-345): the plasma is distributed,
then while the time t <= time.end each iteration injects the beams, solves
the fields, advances the particles and deposits the current (one
``pcp_step``), the probes of ``data.probes`` are written for t, and t
advances by dt (the context keeps t; both add dt the same way).

Usage: ``python -m picopic_amd.run PiCoPiC.json [data_root]``.
"""
import json
import os
import sys
import time as _time

from . import config, probes


def n_iterations(cfg):
    """Iterations of the reference's loop: t = 0, dt, 2 dt, ... while t <= end
    (t accumulated by repeated addition, src/PiCoPiC.cpp:299,343)."""
    _, end, dt = config.time(cfg)
    t, n = 0.0, 0
    while t <= end:
        n += 1
        t += dt
    return n


def run(src, data_root=None, seed=1, iterations=None, progress=None):
    """Runs the configuration; returns (context, probe writer paths).  With
    `iterations`, stops after that many iterations of the loop."""
    cfg = config.read(src)
    ctx, _ = config.context(cfg, seed=seed)
    _, _, dt = config.time(cfg)
    _, grid, _ = config.geometry(cfg)
    total = n_iterations(cfg)
    if iterations is not None:
        total = min(total, int(iterations))
    root = data_root if data_root is not None else cfg.get("data", {}).get("data_root", ".")
    names = {name: k for k, (name, _, _, _, _) in enumerate(config.species(cfg))}
    writer = probes.Writer(cfg, grid, total, root, names)
    for it in range(total):
        t = ctx.time()
        ctx.step(1)
        writer(ctx, t, dt)
        if progress and (it + 1) % progress == 0:
            print("iteration %d / %d, t = %.4e s" % (it + 1, total, t), flush=True)
    paths = sorted(os.path.join(root, p.path + ".npy") for p in writer.probes)
    writer.close()
    return ctx, paths


def main(argv):
    if len(argv) < 2:
        print(__doc__)
        return 2
    t0 = _time.time()
    ctx, paths = run(argv[1], argv[2] if len(argv) > 2 else None, progress=100)
    ctx.close()
    print(json.dumps({"seconds": _time.time() - t0, "probes": paths}))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
