"""Training of the per-cell collision surrogate for a plasma regime.

The network (6 moment features -> 2 x 64 tanh -> 4 relative temperature
changes per pass, BF16 MFMA kernels) is fitted on rows that the device P12
pass generates on a scratch context whose cells sample the regime: each cell
draws its density and temperatures log-uniformly from the given ranges, so
that the cells of the simulation that uses the network fall inside the
training distribution (the surrogate sends cells whose features lie beyond
4 standard deviations of the training set to the P12 fallback).

No reference counterpart: the reference has only the binary collision
operators (src/collisions/collisionsP12.cpp).
"""
import numpy as np

from . import capi

EL_CHARGE = 1.6e-19
EL_MASS = 9.1e-31


def cell_plasma(rng, grid, dr, dz, ppc, tepar, teperp, ti, dens, ion_mass_me=1836.0):
    """Electron and ion particles, ppc per cell and species, uniform inside
    each cell, with per-cell (scalars or [nr, nz] arrays) electron
    temperatures T_par (z) and T_perp, isotropic ion temperature [eV] and
    density [m^-3]; weights density * CELL_VOLUME(i) / ppc.  Returns
    [(r, z, vr, vphi, vz, w)] for the two species."""
    nr, nz = grid
    ci, cj = np.meshgrid(np.arange(nr), np.arange(nz), indexing="ij")
    ci = np.repeat(ci.ravel(), ppc)
    cj = np.repeat(cj.ravel(), ppc)
    n = len(ci)

    def at(v):
        v = np.asarray(v, dtype=np.float64)
        return np.repeat(v.ravel(), ppc) if v.ndim else np.full(n, float(v))

    vol = np.pi * dz * dr * dr * (2.0 * ci + 1)
    out = []
    for m, tpar, tperp in ((EL_MASS, at(tepar), at(teperp)), (ion_mass_me * EL_MASS, at(ti), at(ti))):
        r = (ci + rng.uniform(0.02, 0.98, n)) * dr
        z = (cj + rng.uniform(0.02, 0.98, n)) * dz
        sp = np.sqrt(tperp * EL_CHARGE / m)
        sz = np.sqrt(tpar * EL_CHARGE / m)
        vr = rng.normal(0, 1, n) * sp
        vp = rng.normal(0, 1, n) * sp
        vz = rng.normal(0, 1, n) * sz
        w = at(dens) * vol / ppc
        out.append((r, z, vr, vp, vz, w))
    return out


def regime_cells(rng, grid, dens, tepar, teperp, ti):
    """Per-cell parameters drawn log-uniformly from (lo, hi) ranges."""
    u = lambda ab: np.exp(rng.uniform(np.log(ab[0]), np.log(ab[1]), grid))
    return u(tepar), u(teperp), u(ti), u(dens)


def train(dt, dr, dz, dens, tepar, teperp, ti, grid=(128, 256), ppc=32, passes=30, epochs=40, batch=4096, lr=2e-3,
          seed=0, precision="fp32"):
    """Fit the surrogate on `passes` P12 passes over a scratch plasma of
    `grid` cells whose cells sample the regime (dens, tepar, teperp, ti:
    (lo, hi) ranges); dt and the cell size are those of the simulation.
    Returns (params, norm, loss) for Context.mlp_set_params /
    surrogate_set_norm."""
    rng = np.random.default_rng(seed)
    params = regime_cells(rng, grid, dens, tepar, teperp, ti)
    data = cell_plasma(rng, grid, dr, dz, ppc, *params)
    cap = max(len(d[0]) for d in data) + 16
    ctx = capi.Context((grid[0] * dr, grid[1] * dz), grid, (1, 1), dt, [(-1, 1, False, cap), (1, 1836, False, cap)],
                       precision=precision)
    try:
        for sp, d in enumerate(data):
            ids = np.arange(len(d[0]), dtype=np.uint64) + np.uint64(sp << 32)
            ctx.upload(sp, *d, ids)
        del data
        for k in range(passes):
            ctx.surrogate_collect(0, 1, seed=1000 + seed, step=k)
        loss = ctx.surrogate_fit(epochs=epochs, batch=batch, lr=lr, seed=seed)
        return ctx.mlp_params(), ctx.surrogate_norm(), loss
    finally:
        ctx.close()


def install(ctx, params, norm):
    """Give a context the fitted network."""
    ctx.mlp_set_params(params)
    ctx.surrogate_set_norm(norm)
