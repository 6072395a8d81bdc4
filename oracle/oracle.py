"""ctypes binding of the CPU oracle (test infrastructure only).

Only the test suite, the smoke check and the benchmark's CPU-baseline leg
import this module, to check the device results; it is never the product
path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libpicopic_oracle.so")

PUSHERS = {"boris": 0, "boris-relativistic": 1, "boris-adaptive": 2, "vay": 3, "hc": 4}
CCS = {"vb": 0, "zigzag": 1}
GRIDS = {"E": 0, "H": 1, "H_at_et": 2, "J": 3, "eps": 4, "sigma": 5}
PHASES = ["yee_e", "overlay_e", "yee_h", "overlay_h", "push", "collide", "move", "migrate",
          "reset_j", "deposit", "overlay_j"]
COLL = {"p12": 0, "ta77s": 1, "sk98": 2}

EL_CHARGE = 1.6e-19
EL_MASS = 9.1e-31


class OrcGeometry(C.Structure):
    _fields_ = [("size", C.c_double * 2), ("global_size", C.c_double * 2),
                ("cell_size", C.c_double * 2), ("cell_dims", C.c_long * 4),
                ("cell_amount", C.c_long * 2), ("global_cell_amount", C.c_long * 2),
                ("walls", C.c_int * 4), ("pml_size", C.c_long * 4),
                ("pml_sigma", C.c_double * 2)]


class OrcGrid(C.Structure):
    _fields_ = [("o_s", C.c_int), ("x_real", C.c_int), ("y_real", C.c_int),
                ("x_size", C.c_int), ("y_size", C.c_int), ("g", C.POINTER(C.c_double))]


class OrcConfig(C.Structure):
    _fields_ = [("size", C.c_double * 2), ("grid", C.c_long * 2), ("domains", C.c_long * 2),
                ("dt", C.c_double), ("pusher", C.c_int), ("ccs", C.c_int), ("pml_on", C.c_int),
                ("pml_left", C.c_double), ("pml_right", C.c_double), ("pml_outer", C.c_double),
                ("pml_sigma1", C.c_double), ("pml_sigma2", C.c_double),
                ("n_species", C.c_int), ("charge", C.c_double * 8), ("mass", C.c_double * 8),
                ("is_beam", C.c_int * 8)]


class OrcRng(C.Structure):
    _fields_ = [("u", C.POINTER(C.c_double)), ("n", C.c_long), ("pos", C.c_long)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


NATIVE_LIB_PATH = os.path.join(HERE, "_build", "libpicopic_oracle_native.so")


def use_native_build():
    """Load the -O3 -march=native build (made for this host on first use)
    instead of the portable one; must be called before the first lib()."""
    global LIB_PATH
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    subprocess.run(["make", "-s", "-C", HERE, "native"], check=True)
    LIB_PATH = NATIVE_LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_smb_create.restype = C.c_void_p
        L.orc_smb_create.argtypes = [C.POINTER(OrcConfig)]
        L.orc_smb_destroy.argtypes = [C.c_void_p]
        L.orc_smb_n_domains.argtypes = [C.c_void_p]
        L.orc_smb_n_domains.restype = C.c_int
        L.orc_smb_domain_geometry.argtypes = [C.c_void_p, C.c_int, C.POINTER(OrcGeometry)]
        L.orc_smb_grid_points.argtypes = [C.c_void_p]
        L.orc_smb_grid_points.restype = C.c_long
        L.orc_smb_add_particles.argtypes = [C.c_void_p, C.c_int, C.c_long] + [dp] * 6 + [C.POINTER(C.c_uint64)]
        L.orc_smb_add_particles.restype = C.c_int
        L.orc_smb_count.argtypes = [C.c_void_p, C.c_int]
        L.orc_smb_count.restype = C.c_long
        L.orc_smb_get_particles.argtypes = [C.c_void_p, C.c_int] + [dp] * 6 + [
            C.POINTER(C.c_uint64), C.POINTER(C.c_int), dp, dp]
        L.orc_smb_get_grid.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_smb_set_grid.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_smb_phase.argtypes = [C.c_void_p, C.c_int]
        L.orc_smb_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_threads.argtypes = [C.c_int]
        for f in ("orc_grid_overlay_x", "orc_grid_overlay_y", "orc_grid_overlay_xy"):
            getattr(L, f).argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcGrid)]
        L.orc_grid_overlay_set.argtypes = [C.POINTER(OrcGrid), C.c_double]
        L.orc_smb_get_field_e.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, dp]
        L.orc_smb_get_field_h.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, dp]
        L.orc_push_velocity.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.orc_lorenz_factor.argtypes = [C.c_double]
        L.orc_lorenz_factor.restype = C.c_double
        L.orc_lorenz_factor_inv.argtypes = [C.c_double]
        L.orc_lorenz_factor_inv.restype = C.c_double
        L.orc_p12_collide_single.argtypes = [C.c_double] * 4 + [dp, C.c_double, dp, C.c_double] + [
            C.c_double] * 4 + [C.POINTER(OrcRng)]
        L.orc_smb_collide_p12.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(OrcRng)]
        for f in ("orc_ta77s_collide_single", "orc_sk98_collide_single"):
            getattr(L, f).argtypes = L.orc_p12_collide_single.argtypes
        L.orc_smb_collide.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64]
        L.orc_smb_set_collisions.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64]
        L.orc_coll_uniforms.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint64, dp]
        L.orc_coll_key.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_coll_key.restype = C.c_uint32
        L.orc_philox4x32.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint32)]
        L.orc_philox4x32_raw.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_debye_length.argtypes = [C.c_double] * 4
        L.orc_debye_length.restype = C.c_double
        L.orc_coulomb_logarithm.argtypes = [C.c_double] * 4
        L.orc_coulomb_logarithm.restype = C.c_double
        L.orc_smb_init_reference.argtypes = [C.c_void_p, C.c_int, C.c_long, C.c_double, C.c_double, C.c_double,
                                             C.c_uint64, C.c_uint64]
        L.orc_smb_init_reference.restype = C.c_int
        L.orc_random_reverse.argtypes = [C.c_double, C.c_int]
        L.orc_random_reverse.restype = C.c_double
        L.orc_gamma32_quantile.argtypes = [C.c_double]
        L.orc_gamma32_quantile.restype = C.c_double
        L.orc_mj_g.argtypes = [C.c_double]
        L.orc_mj_g.restype = C.c_double
        L.orc_mj_g_inverse.argtypes = [C.c_double, C.c_double]
        L.orc_mj_g_inverse.restype = C.c_double
        L.orc_mj_energy.argtypes = [C.c_double, C.c_uint64, C.c_uint64]
        L.orc_mj_energy.restype = C.c_double
        L.orc_smb_time.argtypes = [C.c_void_p]
        L.orc_smb_time.restype = C.c_double
        L.orc_smb_set_time.argtypes = [C.c_void_p, C.c_double]
        L.orc_smb_set_beam.argtypes = [C.c_void_p, C.c_int, C.c_long, C.c_double, C.c_double, C.c_double, C.c_long,
                                       C.c_double, C.c_double, C.c_double, C.c_uint64]
        L.orc_smb_set_beam.restype = C.c_int
        L.orc_smb_inject.argtypes = [C.c_void_p]
        L.orc_smb_species_map.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Grid:
    """A reference-layout halo grid backed by a numpy array (tests only)."""

    def __init__(self, xs, ys, os_):
        self.arr = np.zeros((xs + 2 * os_, ys + 2 * os_))
        self.s = OrcGrid(os_, xs + 2 * os_, ys + 2 * os_, xs, ys, _dp(self.arr))


def overlay(kind, a, b):
    getattr(lib(), "orc_grid_overlay_" + kind)(C.byref(a.s), C.byref(b.s))


def overlay_set(a, v):
    lib().orc_grid_overlay_set(C.byref(a.s), v)


MAPS = {"density": 0, "temperature": 1, "energy": 2, "momentum": 3}


class SMB:
    """One shared-memory block of domains (reference src/SMB.cpp)."""

    def __init__(self, size, grid, domains, dt, species, pusher="boris-adaptive", ccs="vb",
                 pml=None):
        """species: list of (charge_in_e, mass_in_me, is_beam)."""
        L = lib()
        cfg = OrcConfig()
        cfg.size[0], cfg.size[1] = size
        cfg.grid[0], cfg.grid[1] = grid
        cfg.domains[0], cfg.domains[1] = domains
        cfg.dt = dt
        cfg.pusher = PUSHERS[pusher]
        cfg.ccs = CCS[ccs]
        if pml is None:
            cfg.pml_on = 0
        else:
            cfg.pml_on = 1
            (cfg.pml_left, cfg.pml_right, cfg.pml_outer, cfg.pml_sigma1, cfg.pml_sigma2) = pml
        cfg.n_species = len(species)
        for i, (q, m, beam) in enumerate(species):
            cfg.charge[i] = q * EL_CHARGE
            cfg.mass[i] = m * EL_MASS
            cfg.is_beam[i] = int(beam)
        self.cfg = cfg
        self.h = L.orc_smb_create(C.byref(cfg))
        if not self.h:
            raise ValueError("oracle: bad configuration")
        self.n_domains = L.orc_smb_n_domains(self.h)
        self.points = L.orc_smb_grid_points(self.h)
        g = OrcGeometry()
        L.orc_smb_domain_geometry(self.h, 0, C.byref(g))
        self.nr_d, self.nz_d = g.cell_amount[0], g.cell_amount[1]
        self.dr, self.dz = g.cell_size[0], g.cell_size[1]
        self.n_species = len(species)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_smb_destroy(self.h)
            self.h = None

    def geometry(self, d):
        g = OrcGeometry()
        lib().orc_smb_domain_geometry(self.h, d, C.byref(g))
        return g

    def add_particles(self, sp, r, z, vr, vphi, vz, w, ids=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (r, z, vr, vphi, vz, w)]
        n = len(arrs[0])
        if ids is None:
            ids = np.arange(n, dtype=np.uint64)
        ids = np.ascontiguousarray(ids, dtype=np.uint64)
        rc = lib().orc_smb_add_particles(self.h, sp, n, *[_dp(a) for a in arrs],
                                         ids.ctypes.data_as(C.POINTER(C.c_uint64)))
        if rc != 0:
            raise ValueError("particle outside the simulation area")

    def count(self, sp):
        return lib().orc_smb_count(self.h, sp)

    def particles(self, sp):
        n = self.count(sp)
        out = {k: np.zeros(n) for k in ("r", "z", "vr", "vphi", "vz", "w", "r_old", "z_old")}
        ids = np.zeros(n, dtype=np.uint64)
        dom = np.zeros(n, dtype=np.int32)
        lib().orc_smb_get_particles(self.h, sp, *[_dp(out[k]) for k in ("r", "z", "vr", "vphi", "vz", "w")],
                                    ids.ctypes.data_as(C.POINTER(C.c_uint64)),
                                    dom.ctypes.data_as(C.POINTER(C.c_int)),
                                    _dp(out["r_old"]), _dp(out["z_old"]))
        out["id"] = ids
        out["dom"] = dom
        return out

    def _ncomp(self, which):
        return 1 if which in ("eps", "sigma") else 3

    def get_grid(self, which):
        nc = self._ncomp(which)
        out = np.zeros((self.n_domains, nc, self.nr_d + 4, self.nz_d + 4))
        lib().orc_smb_get_grid(self.h, GRIDS[which], _dp(out))
        return out

    def set_grid(self, which, arr):
        nc = self._ncomp(which)
        a = np.ascontiguousarray(np.broadcast_to(arr, (self.n_domains, nc, self.nr_d + 4, self.nz_d + 4)),
                                 dtype=np.float64)
        lib().orc_smb_set_grid(self.h, GRIDS[which], _dp(a))

    def phase(self, name):
        lib().orc_smb_phase(self.h, PHASES.index(name))

    def step(self, n=1):
        lib().orc_smb_step(self.h, n)

    def collide(self, scheme, se, si, seed, step):
        """One collision pass over every domain with the counter-based draws."""
        for d in range(self.n_domains):
            lib().orc_smb_collide(self.h, d, COLL[scheme], se, si, seed, step)

    def set_collisions(self, scheme, se=0, si=1, seed=0):
        """Collision phase inside step() (scheme None: off)."""
        lib().orc_smb_set_collisions(self.h, -1 if scheme is None else COLL[scheme], se, si, seed)

    def init_reference(self, sp, macro_amount, density_left, density_right, temperature, seed, id_base=0):
        """The reference's flat placement, weights and thermal velocities for
        plasma species sp (macro_amount over all domains)."""
        rc = lib().orc_smb_init_reference(self.h, sp, int(macro_amount), density_left, density_right, temperature,
                                          seed, id_base)
        if rc:
            raise RuntimeError("orc_smb_init_reference failed (%d)" % rc)

    def set_beam(self, sp, macro_amount, start_time, radius, density, bunches, length, distance, velocity, seed):
        """BeamP for species sp: bunches injected at the start of each step."""
        if lib().orc_smb_set_beam(self.h, sp, int(macro_amount), start_time, radius, density, int(bunches), length,
                                  distance, velocity, seed):
            raise ValueError("bad beam")

    def inject(self):
        lib().orc_smb_inject(self.h)

    def species_map(self, sp, kind):
        """density / temperature / energy / momentum map of species sp after the
        pre-dump overlay: [domain][ncomp][Nr_d+4][Nz_d+4]."""
        k = MAPS[kind]
        out = np.zeros((self.n_domains, 3 if kind == "momentum" else 1, self.nr_d + 4, self.nz_d + 4))
        lib().orc_smb_species_map(self.h, sp, k, _dp(out))
        return out

    def time(self):
        return lib().orc_smb_time(self.h)

    def set_time(self, t):
        lib().orc_smb_set_time(self.h, t)

    def get_field_e(self, d, r, z):
        out = np.zeros(3)
        lib().orc_smb_get_field_e(self.h, d, r, z, _dp(out))
        return out

    def get_field_h(self, d, r, z):
        out = np.zeros(3)
        lib().orc_smb_get_field_h(self.h, d, r, z, _dp(out))
        return out


def set_threads(n):
    lib().orc_set_threads(n)


def push_velocity(pusher, charge, mass, dt, e, h, v):
    v = np.array(v, dtype=np.float64)
    e = np.array(e, dtype=np.float64)
    h = np.array(h, dtype=np.float64)
    lib().orc_push_velocity(PUSHERS[pusher], charge, mass, dt, _dp(e), _dp(h), _dp(v))
    return v


def lorenz_factor(sq):
    return lib().orc_lorenz_factor(sq)


def lorenz_factor_inv(sq):
    return lib().orc_lorenz_factor_inv(sq)


def collide_single(scheme, m_a, m_b, q_a, q_b, va, wa, vb, wb, dens_a, dens_b, debye, dt, uniforms):
    """One pair collision of `scheme` with the given uniform draws."""
    va = np.array(va, dtype=np.float64)
    vb = np.array(vb, dtype=np.float64)
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    rng = OrcRng(_dp(u), len(u), 0)
    getattr(lib(), "orc_%s_collide_single" % scheme)(m_a, m_b, q_a, q_b, _dp(va), wa, _dp(vb), wb, dens_a, dens_b,
                                                       debye, dt, C.byref(rng))
    return va, vb, rng.pos


def p12_collide_single(m_a, m_b, q_a, q_b, va, wa, vb, wb, dens_a, dens_b, debye, dt, uniforms):
    """One P12 pair collision with the given uniform draws (3 per call)."""
    return collide_single("p12", m_a, m_b, q_a, q_b, va, wa, vb, wb, dens_a, dens_b, debye, dt, uniforms)


def philox4x32(ctr, draw, seed):
    out = (C.c_uint32 * 4)()
    lib().orc_philox4x32(ctr, draw, seed, out)
    return list(out)


def philox4x32_raw(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 4)()
    lib().orc_philox4x32_raw(c, k, out)
    return list(out)


def coll_uniforms(seed, cell, q, step):
    u = np.zeros(4)
    lib().orc_coll_uniforms(seed, cell, q, step, _dp(u))
    return u


def random_reverse(vel, power):
    return lib().orc_random_reverse(vel, power)


def gamma32_quantile(u):
    return lib().orc_gamma32_quantile(u)


def mj_g(x):
    return lib().orc_mj_g(x)


def mj_g_inverse(t, x_lo):
    return lib().orc_mj_g_inverse(t, x_lo)


def mj_energy(temperature, seed, ctr):
    return lib().orc_mj_energy(temperature, seed, ctr)
