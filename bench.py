"""Benchmark of the per-step particle loop on one MI355X.

Workloads: ``wakefield`` (default) is the beam-plasma wakefield case -- electron-proton plasma
n = 1e17 m^-3 on a 512 (r) x 4096 (z) grid, dr = dz = 1 mm, 64 macro-particles
per cell per species (268,435,456 plasma particles), Te = 1 eV, Ti = 0.1 eV,
plus a Gaussian electron bunch (gamma = 5, sigma_z = 5 mm, sigma_r = 2 mm,
peak density 0.1 n, 1,048,576 macro-particles), ZigZag deposition, Boris
pusher, dt = 0.5 dz / c, conducting walls.  Inputs are generated on the
device from a fixed seed and are resident in HBM before timing starts.
``batched`` is 8 x 8 uneven domains (1024 x 4096 cells, domains of 128 x 512,
density ramping 0.1 n -> n along z, 32 per cell at the peak, about 1.5e8
particles, Vay, Villasenor-Buneman); ``full`` is the full form (1024 x 8192
cells in 8 x 8 domains of 128 x 1024, the same ramp, 32 per cell at the peak,
the bunch, Higuera-Cary, ZigZag, about 3e8 particles); ``uniform`` is the
128 x 512 plasma with 16 per cell.  The FP32 leg autotunes the sort bin, the
launch shape and the plasma sort interval before its warm-up (``--no-tune``
keeps the defaults).

One step = Yee E + stitch, Yee H + stitch, push + move + deposit of every
particle, migration keys, J stitch.

Prints one JSON line (see the repository documentation for the fields).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
C_LIGHT = 3e8


def workload(name):
    if name == "wakefield":
        return dict(key="wakefield",
                    name="beam-plasma wakefield 512x4096, 64 ppc/species + Gaussian bunch (gamma 5)",
                    grid=(512, 4096), domains=(1, 1), ppc=64, density=1e17, ccs="zigzag", pusher="boris-adaptive",
                    bunch=dict(n=1 << 20, n_peak=1e16, sigma_r=2e-3, sigma_z=5e-3, z0=0.1, gamma=5.0))
    if name == "batched":
        return dict(key="batched",
                    name="batched uneven domains 1024x4096, 8x8 domains of 128x512, ramp 0.1n->n, 32 ppc at peak, Vay",
                    grid=(1024, 4096), domains=(8, 8), ppc=32, density=1e17, density_left=1e16, ccs="vb",
                    pusher="vay", bunch=None)
    if name == "full":
        return dict(key="full",
                    name="full form 1024x8192, 8x8 domains of 128x1024, ramp 0.1n->n, 32 ppc at peak + Gaussian bunch, HC",
                    grid=(1024, 8192), domains=(8, 8), ppc=32, density=1e17, density_left=1e16, ccs="zigzag",
                    pusher="hc", bunch=dict(n=1 << 20, n_peak=1e16, sigma_r=2e-3, sigma_z=5e-3, z0=0.1, gamma=5.0))
    if name == "uniform":
        return dict(key="uniform", name="uniform e-p plasma 128x512, 16 ppc/species, VB", grid=(128, 512),
                    domains=(1, 1), ppc=16, density=1e17, ccs="vb", pusher="boris-adaptive", bunch=None)
    raise SystemExit("unknown workload " + name)


def run_gpu(w, steps, warmup, prec, tune):
    from picopic_amd import capi
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
    L.pcp_set_timing(ctx.h, 1)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    phases = {k: L.pcp_kernel_ms(ctx.h, k.encode()) for k in ("fields", "particles", "fixup", "stitch_j", "step")}
    n_end = sum(ctx.count(s) for s in range(len(species)))
    copy_gbs = ctx.copy_bandwidth(1 << 30, 20)
    ctx.close()
    return dict(n_particles=n_part, n_particles_end=n_end, wall=wall, phases=phases, cells=grid[0] * grid[1],
                copy_gbs=copy_gbs, n_species=len(species), tuning=tuning)


def cpu_baseline(w, threads):
    """The oracle (a C restatement of the reference CPU path, AoS particles,
    OpenMP over domains) on a slab of the same plasma: all 512 radial cells x
    512 axial cells, 32 domains, same ppc, 1 warm-up + 20 timed steps (about
    13 s on 16 cores)."""
    import numpy as np

    from oracle import oracle as O
    dr = dz = 1e-3
    grid = (w["grid"][0], 512)
    size = (grid[0] * dr, grid[1] * dz)
    dt = 0.5 * dz / C_LIGHT
    species = [(-1, 1, False), (1, 1836, False)]
    smb = O.SMB(size, grid, (8, 4), dt, species, pusher=w["pusher"], ccs=w["ccs"])
    rng = np.random.default_rng(5)
    n = w["ppc"] * grid[0] * grid[1]
    for sp, (q, m, b) in enumerate(species):
        r = np.sqrt(rng.uniform(0.0, size[0] ** 2, n))
        r = np.clip(r, 0.5 * dr, size[0] - 0.5 * dr)
        z = rng.uniform(0.5 * dz, size[1] - 0.5 * dz, n)
        vth = np.sqrt((1.0 if m == 1 else 0.1) * 1.6e-19 / (m * 9.1e-31))
        v = rng.normal(0, vth, (3, n))
        wgt = np.full(n, w["density"] * np.pi * size[0] ** 2 * size[1] / n)
        smb.add_particles(sp, r, z, v[0], v[1], v[2], wgt)
        del r, z, v, wgt
    O.set_threads(threads)
    smb.step(1)
    t0 = time.perf_counter()
    nsteps = 20
    smb.step(nsteps)
    dt_wall = time.perf_counter() - t0
    total = 2 * n
    return dict(value=total * nsteps / dt_wall, unit="particle-steps/s", cores=threads, kind="port",
                sample="oracle on a 512x512-cell slab of the same plasma (%d particles, 32 domains), "
                       "%d timed steps after 1 warm-up, %.1f s" % (total, nsteps, dt_wall))


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
    ap.add_argument("--workload", default="wakefield")
    ap.add_argument("--prec", default="both", choices=["fp64", "fp32", "both"],
                    help="'both' times the FP32 split path (the reported value) and the plain FP64 path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true", help="FP32: keep the default sort bin and launch shape")
    ap.add_argument("--tuning", default=None,
                    help="FP32: bin_r,bin_z,block,blocks_per_cu,sort_every instead of autotuning (e.g. to profile "
                         "the configuration an earlier run chose)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: picopic_amd runs the particle loop on a single MI355X "
              "(the multi-GPU sharding plan is described in the documentation)", file=sys.stderr)
        sys.exit(2)
    w = workload(args.workload)
    runs = {}
    for prec in (["fp32", "fp64"] if args.prec == "both" else [args.prec]):
        tune = prec == "fp32" and not args.no_tune
        if prec == "fp32" and args.tuning:
            v = [int(x) for x in args.tuning.split(",")]
            tune = dict(zip(("bin_r", "bin_z", "block", "blocks_per_cu", "sort_every"), v))
        runs[prec] = run_gpu(w, args.steps, args.warmup, prec, tune=tune)
    main_prec = "fp32" if "fp32" in runs else "fp64"
    g = runs[main_prec]
    n = g["n_particles"]
    value = n * args.steps / g["wall"]
    ms_step = g["wall"] * 1e3 / args.steps
    # the fused push+move+deposit kernel: one batched launch per step in
    # FP32, one launch per species and step in FP64; the "particles" phase is
    # exactly those launches, timed with HIP events on the context stream
    t_part = g["phases"]["particles"]
    launches = 1 if main_prec == "fp32" else g["n_species"]
    bytes_pp = 88 if main_prec == "fp64" else 52
    alg_bytes = bytes_pp * n
    achieved = alg_bytes / (t_part * 1e-3) / 1e9 if t_part > 0 else None
    traffic = traffic_from_profile(w, main_prec, n, launches)
    config = {"workload": w["name"], "particles": n, "cells": g["cells"], "precision": main_prec,
              "deposition": w["ccs"], "pusher": w["pusher"], "domains": list(w["domains"]), "phases_ms": g["phases"],
              "tuning": g["tuning"]}
    if "fp64" in runs and main_prec != "fp64":
        g64 = runs["fp64"]
        v64 = g64["n_particles"] * args.steps / g64["wall"]
        config["fp64_plain"] = {"value": v64, "ms_per_step": g64["wall"] * 1e3 / args.steps,
                                "phases_ms": g64["phases"],
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
                     "kernel": "%s (push+move+deposit)" % KERNEL[main_prec],
                     "launches_per_step": launches, "avg_launch_ms": t_part / launches,
                     "alg_bytes_per_launch": alg_bytes / launches,
                     "alg_bytes_per_particle": bytes_pp, "copy_kernel_gbs": g["copy_gbs"]},
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(w, threads)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
