"""Benchmark of the per-step particle loop on one MI355X.

Workloads (inputs generated on the device from fixed seeds, resident in HBM
before timing starts):

``full`` (default) -- the full form: electron-proton plasma n = 1e17 m^-3,
uniform, 32 macro-particles per cell and species on a 1024 (r) x 8192 (z)
grid of 1 mm cells (536,870,912 plasma particles) in 8 x 8 domains of
128 x 1024, Te = 1 eV, Ti = 0.1 eV, plus a Gaussian electron bunch (gamma 5,
sigma_z 5 mm, sigma_r 2 mm, peak density 0.1 n, 1,048,576 macro-particles),
Higuera-Cary pusher, ZigZag deposition, dt = 0.5 dz / c, conducting walls,
binary collisions between the plasma electrons and ions in every step.  The
FP32 split path runs the collision surrogate (trained beforehand, untimed, on
P12 passes over a scratch plasma spanning n = 2e16..5e17 m^-3, Te 0.3..3 eV,
Ti 0.03..0.3 eV) fused into the step, with P12 on the cells it does not
cover; the plain FP64 path runs P12.  Both log the FP64 field and kinetic
energies on the device in every step.

``wakefield`` is the 512 x 4096 plasma with 64 per cell and the bunch
(ZigZag, Boris, no collisions); ``batched`` 8 x 8 uneven domains (1024 x
4096 cells, domains of 128 x 512, density ramping 0.1 n -> n along z, 32 per
cell at the peak, Vay, Villasenor-Buneman); ``uniform`` the 128 x 512 plasma
with 16 per cell (VB, Boris).  The FP32 leg autotunes the sort bin, the launch
shape and the plasma sort interval before its warm-up (``--no-tune`` keeps
the defaults).

One step = Yee E + stitch, Yee H + stitch, push (+ collision phase), move +
deposit of every particle, migration keys, J stitch (+ energy log).

Prints one JSON line (see the repository documentation for the fields).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
C_LIGHT = 3e8
BUNCH = dict(n=1 << 20, n_peak=1e16, sigma_r=2e-3, sigma_z=5e-3, z0=0.1, gamma=5.0)
# plasma regime the full form's collision surrogate is trained on
FULL_REGIME = dict(dens=(2e16, 5e17), tepar=(0.3, 3.0), teperp=(0.3, 3.0), ti=(0.03, 0.3))


def workload(name):
    if name == "full":
        return dict(key="full",
                    name="full form 1024x8192, 8x8 domains of 128x1024, uniform 32 ppc/species + Gaussian bunch "
                         "(gamma 5), HC, ZigZag, e-i collisions every step",
                    grid=(1024, 8192), domains=(8, 8), ppc=32, density=1e17, ccs="zigzag", pusher="hc",
                    bunch=BUNCH, collisions=True)
    if name == "wakefield":
        return dict(key="wakefield",
                    name="beam-plasma wakefield 512x4096, 64 ppc/species + Gaussian bunch (gamma 5)",
                    grid=(512, 4096), domains=(1, 1), ppc=64, density=1e17, ccs="zigzag", pusher="boris-adaptive",
                    bunch=BUNCH, collisions=False)
    if name == "batched":
        return dict(key="batched",
                    name="batched uneven domains 1024x4096, 8x8 domains of 128x512, ramp 0.1n->n, 32 ppc at peak, Vay",
                    grid=(1024, 4096), domains=(8, 8), ppc=32, density=1e17, density_left=1e16, ccs="vb",
                    pusher="vay", bunch=None, collisions=False)
    if name == "uniform":
        return dict(key="uniform", name="uniform e-p plasma 128x512, 16 ppc/species, VB", grid=(128, 512),
                    domains=(1, 1), ppc=16, density=1e17, ccs="vb", pusher="boris-adaptive", bunch=None,
                    collisions=False)
    raise SystemExit("unknown workload " + name)


def train_network(w):
    """The collision surrogate for the workload's plasma regime (untimed)."""
    from picopic_amd import surrogate
    t0 = time.perf_counter()
    params, norm, loss = surrogate.train(0.5e-3 / C_LIGHT, 1e-3, 1e-3, grid=(128, 256), ppc=32, passes=30, epochs=40,
                                         seed=1, **FULL_REGIME)
    return dict(params=params, norm=norm, loss=loss, seconds=time.perf_counter() - t0)


def run_gpu(w, steps, warmup, prec, tune, net=None, compensated=False):
    from picopic_amd import capi, surrogate
    dr = dz = 1e-3
    grid = w["grid"]
    size = (grid[0] * dr, grid[1] * dz)
    dt = 0.5 * dz / C_LIGHT
    n_plasma = w["ppc"] * grid[0] * grid[1]
    n_left = w.get("density_left", w["density"])
    species = [(-1, 1, False, n_plasma + n_plasma // 50), (1, 1836, False, n_plasma + n_plasma // 50)]
    if w["bunch"]:
        species.append((-1, 1, True, w["bunch"]["n"]))
    ctx = capi.Context(size, grid, w["domains"], dt, species, pusher=w["pusher"], ccs=w["ccs"], precision=prec)
    ctx.init_plasma(0, w["ppc"], n_left, w["density"], 1.0, 11)
    ctx.init_plasma(1, w["ppc"], n_left, w["density"], 0.1, 12)
    if w["bunch"]:
        b = w["bunch"]
        ctx.init_bunch(2, b["n"], b["n_peak"], b["sigma_r"], b["sigma_z"], b["z0"], b["gamma"], 13)
    n_part = sum(ctx.count(s) for s in range(len(species)))
    scheme = None
    if w["collisions"]:
        if prec == "fp32":
            surrogate.install(ctx, net["params"], net["norm"])
            scheme = "surrogate"
        else:
            scheme = "p12"
        ctx.set_collisions(scheme, 0, 1, seed=21)
    ctx.set_energy_log(True)
    if compensated:
        ctx.set_compensated_current(True)
    # the autotuner times sort bins and launch shapes on the live state (its
    # steps come before the warm-up and are not timed)
    if isinstance(tune, dict):
        ctx.set_sort_bin(tune["bin_r"], tune["bin_z"])
        if prec == "fp32":
            ctx.set_launch(tune["block"], tune["blocks_per_cu"])
        for s in range(len(species)):
            if not species[s][2]:
                ctx.set_sort_interval(s, tune["sort_every"])
        tuning = ctx.tuning()
    else:
        tuning = ctx.autotune(steps=2) if tune else ctx.tuning()
    ctx.step(warmup)
    L = capi.load()
    # the timed steps replay captured step graphs (one graph launch per step
    # after the sorts); the warm-up above ran eagerly, and two plasma sort
    # intervals with the graph on capture the graphs of every buffer state
    # the sorts cycle through, so that the timed window replays only
    ctx.set_graph(True)
    every = max(1, int(tuning.get("sort_every", 1)))
    ctx.step(min(2 * every + 2, 300))

    def to_sort_boundary():
        # the plasma sorts at steps_done % every == 0; the timed window and
        # the eager window below both start there and have the same length,
        # so that their particle launches see the same sort-cycle phases
        done = int(round(ctx.time() / dt))
        ctx.step((-done) % every)
        return (-done) % every

    to_sort_boundary()
    launches0, captures0 = ctx.graph_stats()
    L.pcp_set_timing(ctx.h, 1)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    launches1, captures1 = ctx.graph_stats()
    step_ms_graph = L.pcp_kernel_ms(ctx.h, b"step")
    # the phases (and the particle launch of the roofline figure) from HIP
    # events around eager launches of the same kernels, over as many steps
    # from the same point of the sort cycle
    gap = to_sort_boundary()
    ctx.set_graph(False)
    eager_steps = steps
    ctx.step(eager_steps)
    phases = {k: L.pcp_kernel_ms(ctx.h, k.encode())
              for k in ("sort", "fields", "push", "collide", "particles", "fixup", "stitch_j", "step")}
    graph = {"launches_per_step": (launches1 - launches0) / steps, "captures_in_window": captures1 - captures0,
             "step_ms_events": step_ms_graph, "eager_steps_for_phases": eager_steps,
             "steps_between_windows": gap,
             "eager_ms_per_step": phases["step"]}
    n_end = sum(ctx.count(s) for s in range(len(species)))
    log = ctx.energy_log(steps)
    energies = {"step": int(log["step"][-1]), "field_J": float(log["field"][-1]),
                "kinetic_J": [float(x) for x in log["kinetic"][-1]]}
    fallback = ctx.surrogate_fallback_cells() if scheme == "surrogate" else None
    copy_gbs = ctx.copy_bandwidth(1 << 30, 20)
    ctx.close()
    return dict(n_particles=n_part, n_particles_end=n_end, wall=wall, phases=phases, cells=grid[0] * grid[1],
                copy_gbs=copy_gbs, n_species=len(species), tuning=tuning, collisions=scheme, energies=energies,
                fallback_cells=fallback, graph=graph)


def host_cpu():
    """lscpu model and physical core count of this host."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                info[k.strip()] = v.strip()
    except (OSError, subprocess.CalledProcessError):
        pass
    model = info.get("Model name", "unknown")
    try:
        physical = int(info["Core(s) per socket"]) * int(info["Socket(s)"])
    except (KeyError, ValueError):
        physical = os.cpu_count() or 1
    return model, physical


def cpu_baseline(w):
    """The oracle (a C restatement of the reference CPU path: AoS particles,
    OpenMP over domains, IEEE FP64) built -O3 -march=native on this host, on
    all the host cores this process may use, timed on a slab of the same
    workload: every radial cell and the 128 axial cells around the bunch,
    split into 8 x 2 domains, the same plasma and bunch, the same pusher,
    deposition and P12 collisions; one warm-up step, then enough steps for
    about 15 s.  A single-thread run of the uniform plasma (Boris + VB,
    128 x 512, 16 per cell, 3 steps) gives the per-core speed to compare with
    the reference compiled in the same way (1.82e6 particle-steps/s on the
    container's Xeon)."""
    import numpy as np

    from oracle import oracle as O
    O.use_native_build()
    model, physical = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or physical
    try:
        threads = min(threads, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    dr = dz = 1e-3
    dt = 0.5 * dz / C_LIGHT
    rng = np.random.default_rng(5)

    def plasma(smb, grid, ppc, species, z_lo):
        size_r = grid[0] * dr
        n = ppc * grid[0] * grid[1]
        for sp, (q, m, b) in enumerate(species[:2]):
            r = np.sqrt(rng.uniform(0.0, size_r ** 2, n))
            r = np.clip(r, 0.5 * dr, size_r - 0.5 * dr)
            z = rng.uniform(0.5 * dz, grid[1] * dz - 0.5 * dz, n)
            vth = np.sqrt((1.0 if m == 1 else 0.1) * 1.6e-19 / (m * 9.1e-31))
            v = rng.normal(0, vth, (3, n))
            wgt = np.full(n, w["density"] * np.pi * size_r ** 2 * grid[1] * dz / n)
            smb.add_particles(sp, r, z, v[0], v[1], v[2], wgt)
        return 2 * n

    # the slab of the workload
    sgrid = (w["grid"][0], 128)
    z_lo = 0.1 - 0.064 if w["bunch"] else 0.0
    species = [(-1, 1, False), (1, 1836, False)] + ([(-1, 1, True)] if w["bunch"] else [])
    smb = O.SMB((sgrid[0] * dr, sgrid[1] * dz), sgrid, (8, 2), dt, species, pusher=w["pusher"], ccs=w["ccs"])
    total = plasma(smb, sgrid, w["ppc"], species, z_lo)
    if w["bunch"]:
        b = w["bunch"]
        # the bunch's macro-particles inside the slab (the same macro weight
        # as the device's bunch)
        nb = b["n"]
        real = b["n_peak"] * (2 * np.pi) ** 1.5 * b["sigma_r"] ** 2 * b["sigma_z"]
        r = np.abs(rng.normal(0, b["sigma_r"], (2, nb)))
        r = np.hypot(r[0], r[1])
        z = rng.normal(b["z0"], b["sigma_z"], nb) - z_lo
        keep = (r > 0.5 * dr) & (r < sgrid[0] * dr - 0.5 * dr) & (z > 0.5 * dz) & (z < sgrid[1] * dz - 0.5 * dz)
        r, z = r[keep], z[keep]
        vz = np.full(len(r), C_LIGHT * np.sqrt(1 - 1 / b["gamma"] ** 2))
        zero = np.zeros(len(r))
        smb.add_particles(2, r, z, zero, zero, vz, np.full(len(r), real / nb))
        total += len(r)
    if w["collisions"]:
        smb.set_collisions("p12", 0, 1, seed=21)
    O.set_threads(threads)
    t0 = time.perf_counter()
    smb.step(1)
    t1 = time.perf_counter() - t0
    nsteps = int(max(3, min(20, round(15.0 / max(t1, 1e-3)))))
    t0 = time.perf_counter()
    smb.step(nsteps)
    dt_wall = time.perf_counter() - t0
    value = total * nsteps / dt_wall
    del smb
    # per-core calibration: the uniform plasma, Boris + VB, one thread
    cal = O.SMB((0.128, 0.512), (128, 512), (1, 1), dt, [(-1, 1, False), (1, 1836, False)],
                pusher="boris-adaptive", ccs="vb")
    ncal = plasma(cal, (128, 512), 16, [(-1, 1, False), (1, 1836, False)], 0.0)
    O.set_threads(1)
    cal.step(1)
    t0 = time.perf_counter()
    cal.step(3)
    per_core = ncal * 3 / (time.perf_counter() - t0)
    return dict(value=value, unit="particle-steps/s", cores=threads, kind="port",
                cpu_model=model, physical_cores=physical,
                sample="oracle -O3 -march=native (IEEE FP64, OpenMP over domains) on a %dx%d-cell slab of the "
                       "workload around the bunch (%d particles, 8x2 domains, %s collisions), %d timed steps after "
                       "1 warm-up, %.1f s" % (sgrid[0], sgrid[1], total, "P12" if w["collisions"] else "no",
                                             nsteps, dt_wall),
                per_core=value / threads,
                calibration={"case": "uniform plasma 128x512, 16 ppc/species, Boris + VB, 1 thread, 3 steps",
                             "per_core": per_core, "reference_per_core": 1.82e6,
                             "ratio": per_core / 1.82e6})


def traffic_from_profile(w, prec, n_particles, n_launches_per_step):
    """HBM bytes per launch of the fused particle kernel, from the committed
    rocprofv3 PMC summary of this workload (FETCH_SIZE and WRITE_SIZE per
    dispatch, each scaled by the same counter's reading for a copy kernel of
    known size); None when no summary for this workload and precision exists."""
    path = os.path.join(ROOT, "profiles", "%s_%s_traffic.txt" % (KERNEL[prec], w["key"]))
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        per_particle = float(vals["hbm_bytes_per_particle"])
    except (KeyError, ValueError):
        return None
    return per_particle * n_particles / n_launches_per_step


KERNEL = {"fp32": "k32_step_batch", "fp64": "k_step_fused"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 50 timed steps span about two plasma sort intervals, so that the figure
    # does not depend on where the window falls in the sort cycle
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="full")
    ap.add_argument("--prec", default="both", choices=["fp64", "fp32", "both"],
                    help="'both' times the FP32 split path (the reported value) and the plain FP64 path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="FP32: keep the default sort bin and launch shape")
    ap.add_argument("--tuning", default=None,
                    help="FP32: bin_r,bin_z,block,blocks_per_cu,sort_every instead of autotuning (e.g. to profile "
                         "the configuration an earlier run chose)")
    ap.add_argument("--compensated", action="store_true",
                    help="FP32: deposit the current with compensated run sums and atomics")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: picopic_amd runs the particle loop on a single MI355X "
              "(the multi-GPU sharding plan is described in the documentation)", file=sys.stderr)
        sys.exit(2)
    w = workload(args.workload)
    runs = {}
    net = None
    precs = ["fp32", "fp64"] if args.prec == "both" else [args.prec]
    if w["collisions"] and "fp32" in precs:
        net = train_network(w)
    for prec in precs:
        tune = prec == "fp32" and not args.no_tune
        if prec == "fp32" and args.tuning:
            v = [int(x) for x in args.tuning.split(",")]
            tune = dict(zip(("bin_r", "bin_z", "block", "blocks_per_cu", "sort_every"), v))
        if prec == "fp64" and "fp32" in runs:
            # the FP64 leg sorts with the FP32 leg's bin and plasma interval,
            # so that both timed windows hold the same sorts
            t = runs["fp32"]["tuning"]
            tune = {"bin_r": t["bin_r"], "bin_z": t["bin_z"], "block": 0, "blocks_per_cu": 0,
                    "sort_every": t["sort_every"]}
        runs[prec] = run_gpu(w, args.steps, args.warmup, prec, tune=tune, net=net,
                             compensated=args.compensated and prec == "fp32")
    main_prec = "fp32" if "fp32" in runs else "fp64"
    g = runs[main_prec]
    n = g["n_particles"]
    value = n * args.steps / g["wall"]
    ms_step = g["wall"] * 1e3 / args.steps
    # the dominant kernel: the fused move + deposit kernel (with the push too
    # in steps without collisions), one batched launch per step in FP32, one
    # launch per species and step in FP64; the "particles" phase is exactly
    # those launches, timed with HIP events on the context stream
    t_part = g["phases"]["particles"]
    launches = 1 if main_prec == "fp32" else g["n_species"]
    bytes_pp = 88 if main_prec == "fp64" else 52
    alg_bytes = bytes_pp * n
    achieved = alg_bytes / (t_part * 1e-3) / 1e9 if t_part > 0 else None
    traffic = traffic_from_profile(w, main_prec, n, launches)
    config = {"workload": w["name"], "particles": n, "cells": g["cells"], "precision": main_prec,
              "deposition": w["ccs"], "pusher": w["pusher"], "domains": list(w["domains"]),
              "collisions": g["collisions"], "phases_ms": g["phases"], "tuning": g["tuning"],
              "energies": g["energies"], "graph": g["graph"]}
    if args.compensated:
        config["compensated_current"] = True
    if g["fallback_cells"] is not None:
        config["surrogate"] = {"fallback_cells_last_step": g["fallback_cells"], "fit_loss": net["loss"],
                               "train_seconds": net["seconds"]}
    if "fp64" in runs and main_prec != "fp64":
        g64 = runs["fp64"]
        v64 = g64["n_particles"] * args.steps / g64["wall"]
        config["fp64_plain"] = {"value": v64, "ms_per_step": g64["wall"] * 1e3 / args.steps,
                                "collisions": g64["collisions"], "phases_ms": g64["phases"],
                                "energies": g64["energies"], "graph": g64["graph"],
                                "achieved_gbs": 88 * g64["n_particles"] / (g64["phases"]["particles"] * 1e-3) / 1e9}
        config["speedup_fp32_vs_fp64"] = value / v64
    line = {
        "metric": "particle-steps/s per GPU",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32",
        "data": "synthetic (seeded, generated on device)",
        "config": config,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "kernel": "%s (%s)" % (KERNEL[main_prec],
                                            "move+deposit" if g["collisions"] else "push+move+deposit"),
                     "launches_per_step": launches, "avg_launch_ms": t_part / launches,
                     "alg_bytes_per_launch": alg_bytes / launches,
                     "alg_bytes_per_particle": bytes_pp, "copy_kernel_gbs": g["copy_gbs"]},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
