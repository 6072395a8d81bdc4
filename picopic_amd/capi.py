"""ctypes binding of the picopic_amd C ABI (include/picopic.h).

The HIP library must be built (``python -m picopic_amd.build``); importing
this module on a machine without it raises instead of falling back to any CPU
path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libpicopic_hip.so")

PUSHERS = {"boris": 0, "boris-relativistic": 1, "boris-adaptive": 2, "vay": 3, "hc": 4}
CCS = {"vb": 0, "zigzag": 1}
GRIDS = {"E": 0, "H": 1, "H_at_et": 2, "J": 3, "eps": 4, "sigma": 5, "map": 6}
MAPS = {"density": 0, "temperature": 1, "energy": 2, "momentum": 3}
PREC = {"fp64": 0, "fp32": 1}

EL_CHARGE = 1.6e-19
EL_MASS = 9.1e-31

EXPORTS = [
    "pcp_last_error", "pcp_create", "pcp_destroy", "pcp_sync", "pcp_upload_species", "pcp_count",
    "pcp_download_species", "pcp_init_plasma", "pcp_init_bunch", "pcp_download_grid", "pcp_upload_grid",
    "pcp_grid_points", "pcp_yee_e", "pcp_yee_h", "pcp_stitch", "pcp_push", "pcp_move", "pcp_migrate",
    "pcp_deposit", "pcp_step", "pcp_get_field_e", "pcp_get_field_h", "pcp_stream", "pcp_set_timing",
    "pcp_kernel_ms", "pcp_sort", "pcp_set_sort_interval", "pcp_copy_bandwidth", "pcp_field_energy", "pcp_kinetic_energy",
    "pcp_collide", "pcp_set_collisions", "pcp_mlp_nparams", "pcp_mlp_init", "pcp_mlp_get_params",
    "pcp_mlp_set_params", "pcp_mlp_forward", "pcp_mlp_grad", "pcp_mlp_adam", "pcp_surrogate_collect",
    "pcp_surrogate_fit", "pcp_surrogate_get_norm", "pcp_surrogate_set_norm", "pcp_surrogate_fallback_cells",
    "pcp_species_moments", "pcp_continuity_residual", "pcp_set_graph", "pcp_set_compensated_current", "pcp_graph_stats", "pcp_collision_passes", "pcp_set_sort_bin", "pcp_set_launch",
    "pcp_get_tuning", "pcp_autotune", "pcp_surrogate_calibrate", "pcp_init_reference",
    "pcp_set_beam", "pcp_inject", "pcp_set_time", "pcp_time", "pcp_probe", "pcp_species_map",
]
COLL = {"p12": 0, "ta77s": 1, "sk98": 2, "surrogate": 3}


class SpeciesDesc(C.Structure):
    _fields_ = [("charge", C.c_double), ("mass", C.c_double), ("is_beam", C.c_int), ("capacity", C.c_long)]


class Config(C.Structure):
    _fields_ = [("size_r", C.c_double), ("size_z", C.c_double), ("grid_r", C.c_long), ("grid_z", C.c_long),
                ("domains_r", C.c_long), ("domains_z", C.c_long), ("dt", C.c_double), ("pusher", C.c_int),
                ("ccs", C.c_int), ("pml_on", C.c_int), ("pml_left", C.c_double), ("pml_right", C.c_double),
                ("pml_outer", C.c_double), ("pml_sigma1", C.c_double), ("pml_sigma2", C.c_double),
                ("precision", C.c_int), ("n_species", C.c_int), ("species", SpeciesDesc * 8),
                ("max_migrants", C.c_long)]


class Tuning(C.Structure):
    _fields_ = [("bin_r", C.c_int), ("bin_z", C.c_int), ("block", C.c_int), ("blocks_per_cu", C.c_int),
                ("sort_every", C.c_int), ("ms_per_step", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PcpError(RuntimeError):
    pass


_lib = None


def load(path=LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PcpError("picopic_amd HIP library not built: %s (run python -m picopic_amd.build)" % path)
    L = C.CDLL(path)
    vp, dp, u64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)
    L.pcp_last_error.restype = C.c_char_p
    L.pcp_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.pcp_destroy.argtypes = [vp]
    L.pcp_sync.argtypes = [vp]
    L.pcp_upload_species.argtypes = [vp, C.c_int, C.c_long] + [dp] * 6 + [u64p]
    L.pcp_count.argtypes = [vp, C.c_int, C.POINTER(C.c_long)]
    L.pcp_download_species.argtypes = [vp, C.c_int] + [dp] * 6 + [u64p, u64p, dp, dp]
    L.pcp_init_plasma.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64]
    L.pcp_init_bunch.argtypes = [vp, C.c_int, C.c_long] + [C.c_double] * 5 + [C.c_uint64]
    L.pcp_init_reference.argtypes = [vp, C.c_int, C.c_long, C.c_double, C.c_double, C.c_double, C.c_uint64]
    L.pcp_set_beam.argtypes = [vp, C.c_int, C.c_long, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                               C.c_double, C.c_double, C.c_uint64]
    L.pcp_inject.argtypes = [vp]
    L.pcp_probe.argtypes = [vp, C.c_int, C.c_int, C.c_long, C.c_long, C.c_long, C.c_long, dp]
    L.pcp_species_map.argtypes = [vp, C.c_int, C.c_int, dp]
    L.pcp_set_time.argtypes = [vp, C.c_double]
    L.pcp_time.argtypes = [vp]
    L.pcp_time.restype = C.c_double
    L.pcp_download_grid.argtypes = [vp, C.c_int, dp]
    L.pcp_upload_grid.argtypes = [vp, C.c_int, dp]
    L.pcp_grid_points.argtypes = [vp]
    L.pcp_grid_points.restype = C.c_long
    for f in ("pcp_yee_e", "pcp_yee_h", "pcp_push", "pcp_move", "pcp_migrate", "pcp_deposit"):
        getattr(L, f).argtypes = [vp]
    L.pcp_stitch.argtypes = [vp, C.c_int]
    L.pcp_step.argtypes = [vp, C.c_int]
    L.pcp_get_field_e.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp]
    L.pcp_get_field_h.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp]
    L.pcp_stream.argtypes = [vp]
    L.pcp_stream.restype = vp
    L.pcp_set_timing.argtypes = [vp, C.c_int]
    L.pcp_kernel_ms.argtypes = [vp, C.c_char_p]
    L.pcp_kernel_ms.restype = C.c_double
    L.pcp_sort.argtypes = [vp, C.c_int]
    L.pcp_set_sort_interval.argtypes = [vp, C.c_int, C.c_int]
    L.pcp_copy_bandwidth.argtypes = [vp, C.c_long, C.c_int, C.POINTER(C.c_double)]
    L.pcp_field_energy.argtypes = [vp, C.POINTER(C.c_double)]
    L.pcp_kinetic_energy.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.pcp_collide.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64]
    L.pcp_set_collisions.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_uint64]
    fp = C.POINTER(C.c_float)
    L.pcp_mlp_nparams.argtypes = []
    L.pcp_mlp_init.argtypes = [vp, C.c_uint64]
    L.pcp_mlp_get_params.argtypes = [vp, fp]
    L.pcp_mlp_set_params.argtypes = [vp, fp]
    L.pcp_mlp_forward.argtypes = [vp, fp, C.c_long, fp]
    L.pcp_mlp_grad.argtypes = [vp, fp, fp, C.c_long, fp, fp]
    L.pcp_mlp_adam.argtypes = [vp, C.c_float]
    L.pcp_surrogate_collect.argtypes = [vp, C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.POINTER(C.c_long)]
    L.pcp_surrogate_fit.argtypes = [vp, C.c_int, C.c_long, C.c_float, C.c_uint64, fp]
    L.pcp_surrogate_get_norm.argtypes = [vp, fp]
    L.pcp_surrogate_set_norm.argtypes = [vp, fp]
    L.pcp_surrogate_fallback_cells.argtypes = [vp, C.POINTER(C.c_long)]
    L.pcp_surrogate_calibrate.argtypes = [vp, C.c_int, C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_double)]
    L.pcp_species_moments.argtypes = [vp, C.c_int, dp]
    L.pcp_continuity_residual.argtypes = [vp, dp]
    L.pcp_set_graph.argtypes = [vp, C.c_int]
    L.pcp_set_compensated_current.argtypes = [vp, C.c_int]
    L.pcp_graph_stats.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]
    L.pcp_collision_passes.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.pcp_set_sort_bin.argtypes = [vp, C.c_int, C.c_int]
    L.pcp_set_launch.argtypes = [vp, C.c_int, C.c_int]
    L.pcp_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.pcp_autotune.argtypes = [vp, C.c_int, C.POINTER(Tuning)]
    L.pcp_set_energy_log.argtypes = [vp, C.c_int]
    L.pcp_energy_log.argtypes = [vp, C.c_long, dp, C.POINTER(C.c_long)]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise PcpError("picopic_amd error %d: %s" % (rc, _lib.pcp_last_error().decode()))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _rows8(x):
    x = np.asarray(x, dtype=np.float32)
    out = np.zeros((len(x), 8), dtype=np.float32)
    out[:, :x.shape[1]] = x
    return out


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


class Context:
    """One shared-memory block of domains resident on the GPU."""

    def __init__(self, size, grid, domains, dt, species, pusher="boris-adaptive", ccs="vb", pml=None,
                 capacity=None, precision="fp64", max_migrants=0):
        """species: list of (charge_in_e, mass_in_me, is_beam[, capacity])."""
        L = load()
        cfg = Config()
        cfg.size_r, cfg.size_z = size
        cfg.grid_r, cfg.grid_z = grid
        cfg.domains_r, cfg.domains_z = domains
        cfg.dt = dt
        cfg.pusher = PUSHERS[pusher]
        cfg.ccs = CCS[ccs]
        if pml is None:
            cfg.pml_on = 0
        else:
            cfg.pml_on = 1
            cfg.pml_left, cfg.pml_right, cfg.pml_outer, cfg.pml_sigma1, cfg.pml_sigma2 = pml
        cfg.precision = PREC[precision]
        cfg.n_species = len(species)
        for i, spc in enumerate(species):
            q, m, beam = spc[:3]
            cap = spc[3] if len(spc) > 3 else (capacity if capacity is not None else 1 << 20)
            cfg.species[i] = SpeciesDesc(q * EL_CHARGE, m * EL_MASS, int(beam), int(cap))
        cfg.max_migrants = max_migrants
        self.cfg = cfg
        h = C.c_void_p()
        _check(L.pcp_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.n_species = len(species)
        self.D = domains[0] * domains[1]
        self.nr_d = grid[0] // domains[0]
        self.nz_d = grid[1] // domains[1]
        self.P = L.pcp_grid_points(h)

    def close(self):
        if getattr(self, "h", None):
            _lib.pcp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, sp, r, z, vr, vphi, vz, w, ids=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (r, z, vr, vphi, vz, w)]
        n = len(arrs[0])
        idp = None
        if ids is not None:
            ids = np.ascontiguousarray(ids, dtype=np.uint64)
            idp = _u64p(ids)
        _check(_lib.pcp_upload_species(self.h, sp, n, *[_dp(a) for a in arrs], idp))

    def count(self, sp):
        n = C.c_long()
        _check(_lib.pcp_count(self.h, sp, C.byref(n)))
        return n.value

    def particles(self, sp):
        n = self.count(sp)
        out = {k: np.zeros(n) for k in ("r", "z", "vr", "vphi", "vz", "w", "r_old", "z_old")}
        ids = np.zeros(n, dtype=np.uint64)
        order = np.zeros(n, dtype=np.uint64)
        _check(_lib.pcp_download_species(self.h, sp, *[_dp(out[k]) for k in ("r", "z", "vr", "vphi", "vz", "w")],
                                         _u64p(ids), _u64p(order), _dp(out["r_old"]), _dp(out["z_old"])))
        out["id"] = ids
        out["order"] = order
        return out

    def init_plasma(self, sp, ppc, density_left, density_right, temperature_ev, seed):
        _check(_lib.pcp_init_plasma(self.h, sp, ppc, density_left, density_right, temperature_ev, seed))

    def init_reference(self, sp, macro_amount, density_left, density_right, temperature, seed):
        """The reference's flat placement, r-proportional weights and
        Maxwell-Juettner thermal velocities (macro_amount over all domains)."""
        _check(_lib.pcp_init_reference(self.h, sp, int(macro_amount), density_left, density_right, temperature, seed))

    def set_beam(self, sp, macro_amount, start_time, radius, density, bunches, length, distance, velocity, seed):
        """BeamP for beam species sp: bunches injected at the start of each step."""
        _check(_lib.pcp_set_beam(self.h, sp, int(macro_amount), start_time, radius, density, int(bunches), length,
                                 distance, velocity, seed))

    def inject(self):
        _check(_lib.pcp_inject(self.h))

    def species_map(self, sp, kind, download=True):
        """density / temperature / energy / momentum map of species sp as the
        reference computes it before a dump: [domain][ncomp][Nr_d+4][Nz_d+4]
        (interior nodes); with download=False it stays on the device for
        probe("map", ...)."""
        nc = 3 if kind == "momentum" else 1
        if not download:
            _check(_lib.pcp_species_map(self.h, sp, MAPS[kind], None))
            return None
        out = np.zeros((self.D, nc, self.nr_d + 4, self.nz_d + 4))
        _check(_lib.pcp_species_map(self.h, sp, MAPS[kind], _dp(out)))
        return out

    def probe(self, which, comp, r0, z0, r1, z1):
        """Component comp of grid `which` over global cells [r0, r1) x [z0, z1)."""
        out = np.zeros((r1 - r0, z1 - z0))
        _check(_lib.pcp_probe(self.h, GRIDS[which], comp, r0, z0, r1, z1, _dp(out)))
        return out

    def time(self):
        return _lib.pcp_time(self.h)

    def set_time(self, t):
        _check(_lib.pcp_set_time(self.h, t))

    def init_bunch(self, sp, n, n_peak, sigma_r, sigma_z, z0, gamma, seed):
        _check(_lib.pcp_init_bunch(self.h, sp, n, n_peak, sigma_r, sigma_z, z0, gamma, seed))

    def _ncomp(self, which):
        return 1 if which in ("eps", "sigma") else 3

    def get_grid(self, which):
        out = np.zeros((self.D, self._ncomp(which), self.nr_d + 4, self.nz_d + 4))
        _check(_lib.pcp_download_grid(self.h, GRIDS[which], _dp(out)))
        return out

    def set_grid(self, which, arr):
        a = np.ascontiguousarray(np.broadcast_to(arr, (self.D, self._ncomp(which), self.nr_d + 4, self.nz_d + 4)),
                                 dtype=np.float64)
        _check(_lib.pcp_upload_grid(self.h, GRIDS[which], _dp(a)))

    def phase(self, name):
        L = _lib
        if name == "yee_e":
            _check(L.pcp_yee_e(self.h))
        elif name == "overlay_e":
            _check(L.pcp_stitch(self.h, GRIDS["E"]))
        elif name == "yee_h":
            _check(L.pcp_yee_h(self.h))
        elif name == "overlay_h":
            _check(L.pcp_stitch(self.h, GRIDS["H"]))
        elif name == "push":
            _check(L.pcp_push(self.h))
        elif name == "move":
            _check(L.pcp_move(self.h))
        elif name == "migrate":
            _check(L.pcp_migrate(self.h))
        elif name == "reset_j":
            pass   # folded into deposit
        elif name == "deposit":
            _check(L.pcp_deposit(self.h))
        elif name == "overlay_j":
            _check(L.pcp_stitch(self.h, GRIDS["J"]))
        else:
            raise ValueError(name)

    def step(self, n=1):
        _check(_lib.pcp_step(self.h, n))

    def sort(self, sp):
        _check(_lib.pcp_sort(self.h, sp))

    def collide(self, scheme, se, si, seed, step):
        """One collision pass (P12 / TA77S / SK98) on the current velocities."""
        _check(_lib.pcp_collide(self.h, COLL[scheme], se, si, seed, step))

    def temperatures(self, sp, mass):
        """(T_par, T_perp) [eV] of species sp over the whole block (z is par)."""
        m = np.zeros(7)
        _check(_lib.pcp_species_moments(self.h, sp, _dp(m)))
        sw = m[0]
        V = m[1:4] / sw
        tpar = mass * (m[4] / sw - V[2] ** 2) / 1.6e-19
        tperp = mass * (m[5] / sw - V[0] ** 2 - V[1] ** 2) / (2 * 1.6e-19)
        return tpar, tperp

    def set_graph(self, on=True):
        _check(_lib.pcp_set_graph(self.h, 1 if on else 0))

    def set_compensated_current(self, on=True):
        """FP32 split contexts: deposit J with compensated run sums and atomics."""
        _check(_lib.pcp_set_compensated_current(self.h, 1 if on else 0))

    def collision_passes(self):
        """Collision passes the step has run since set_collisions (device counter)."""
        n = C.c_uint64()
        _check(_lib.pcp_collision_passes(self.h, C.byref(n)))
        return n.value

    def graph_stats(self):
        """(graph launches, graph captures) since the context was created."""
        a, b = C.c_long(), C.c_long()
        _check(_lib.pcp_graph_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def continuity_residual(self):
        """(max |residual|, max |charge change|, max |q w|) of the last step."""
        out = np.zeros(3)
        _check(_lib.pcp_continuity_residual(self.h, _dp(out)))
        return out

    def moments(self, sp):
        m = np.zeros(7)
        _check(_lib.pcp_species_moments(self.h, sp, _dp(m)))
        return m

    # ---- collision surrogate network ----
    def mlp_init(self, seed=0):
        _check(_lib.pcp_mlp_init(self.h, seed))

    def mlp_params(self):
        out = np.zeros(_lib.pcp_mlp_nparams(), dtype=np.float32)
        _check(_lib.pcp_mlp_get_params(self.h, _fp(out)))
        return out

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib.pcp_mlp_set_params(self.h, _fp(p)))

    def mlp_forward(self, x):
        x = _rows8(x)
        y = np.zeros((len(x), 4), dtype=np.float32)
        _check(_lib.pcp_mlp_forward(self.h, _fp(x), len(x), _fp(y)))
        return y

    def mlp_grad(self, x, t):
        x = _rows8(x)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(len(x), 4)
        g = np.zeros(_lib.pcp_mlp_nparams(), dtype=np.float32)
        loss = C.c_float()
        _check(_lib.pcp_mlp_grad(self.h, _fp(x), _fp(t), len(x), _fp(g), C.byref(loss)))
        return g, loss.value

    def mlp_adam(self, lr):
        _check(_lib.pcp_mlp_adam(self.h, lr))

    def surrogate_collect(self, se, si, seed, step):
        rows = C.c_long()
        _check(_lib.pcp_surrogate_collect(self.h, se, si, seed, step, C.byref(rows)))
        return rows.value

    def surrogate_fit(self, epochs, batch, lr, seed=0):
        loss = C.c_float()
        _check(_lib.pcp_surrogate_fit(self.h, epochs, batch, lr, seed, C.byref(loss)))
        return loss.value

    def surrogate_norm(self):
        out = np.zeros(16, dtype=np.float32)
        _check(_lib.pcp_surrogate_get_norm(self.h, _fp(out)))
        return out

    def surrogate_set_norm(self, v):
        v = np.ascontiguousarray(v, dtype=np.float32)
        _check(_lib.pcp_surrogate_set_norm(self.h, _fp(v)))

    def surrogate_calibrate(self, se, si, seed, passes):
        """Scale the fitted network's electron outputs so that the electron
        anisotropy relaxes at P12's rate over `passes` passes on this
        context's plasma (run on scratch copies); returns the factor."""
        f = C.c_double()
        _check(_lib.pcp_surrogate_calibrate(self.h, se, si, seed, passes, C.byref(f)))
        return f.value

    def surrogate_fallback_cells(self):
        n = C.c_long()
        _check(_lib.pcp_surrogate_fallback_cells(self.h, C.byref(n)))
        return n.value

    def set_collisions(self, scheme, se=0, si=1, seed=0):
        """Collision phase inside step() (scheme None: off)."""
        _check(_lib.pcp_set_collisions(self.h, -1 if scheme is None else COLL[scheme], se, si, seed))

    def set_sort_interval(self, sp, every):
        _check(_lib.pcp_set_sort_interval(self.h, sp, every))

    def set_sort_bin(self, bin_r, bin_z):
        _check(_lib.pcp_set_sort_bin(self.h, bin_r, bin_z))

    def set_launch(self, block, blocks_per_cu=0):
        _check(_lib.pcp_set_launch(self.h, block, blocks_per_cu))

    def tuning(self):
        t = Tuning()
        _check(_lib.pcp_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def autotune(self, steps=4):
        """Time the sort-bin / launch-shape candidates on the live state (the
        simulation advances) and keep the fastest; returns the choice."""
        t = Tuning()
        _check(_lib.pcp_autotune(self.h, steps, C.byref(t)))
        return t.as_dict()

    def copy_bandwidth(self, nbytes=1 << 30, iters=20):
        g = C.c_double()
        _check(_lib.pcp_copy_bandwidth(self.h, nbytes, iters, C.byref(g)))
        return g.value

    def field_energy(self):
        g = C.c_double()
        _check(_lib.pcp_field_energy(self.h, C.byref(g)))
        return g.value

    def kinetic_energy(self, sp):
        g = C.c_double()
        _check(_lib.pcp_kinetic_energy(self.h, sp, C.byref(g)))
        return g.value

    def set_energy_log(self, on=True):
        """Record the field and kinetic energies on the device in every step."""
        _check(_lib.pcp_set_energy_log(self.h, 1 if on else 0))

    def energy_log(self, max_rows=4096):
        """Logged steps, oldest first: dict of arrays 'step', 'field' and
        'kinetic' [rows, n_species] (J)."""
        out = np.zeros((max_rows, 10))
        rows = C.c_long()
        _check(_lib.pcp_energy_log(self.h, max_rows, _dp(out), C.byref(rows)))
        out = out[:rows.value]
        return {"step": out[:, 0].astype(np.int64), "field": out[:, 1], "kinetic": out[:, 2:2 + self.n_species]}

    def sync(self):
        _check(_lib.pcp_sync(self.h))

    def get_field_e(self, d, r, z):
        out = np.zeros(3)
        _check(_lib.pcp_get_field_e(self.h, d, r, z, _dp(out)))
        return out

    def get_field_h(self, d, r, z):
        out = np.zeros(3)
        _check(_lib.pcp_get_field_h(self.h, d, r, z, _dp(out)))
        return out
