"""Pins the CPU oracle with what the mathematics fixes (closed forms,
conservation laws, invariants), chosen so that a dropped term, a wrong sign
or index, or a swapped factor fails one of them.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as Hh

Q_E = -1.6e-19
M_E = 9.1e-31
C = 3e8
MU0 = 1.26e-6


def one_particle_smb(r, z, v, nr=10, nz=10, dr=1e-3, dz=1e-3, ccs="vb", pusher="boris-adaptive"):
    dt = 0.5 * dz / C
    smb = O.SMB((nr * dr, nz * dz), (nr, nz), (1, 1), dt, [(-1, 1, False)], ccs=ccs, pusher=pusher)
    smb.add_particles(0, [r], [z], [v[0]], [v[1]], [v[2]], [1.0])
    return smb, dt


# --------------------------------------------------------------------------
# Villasenor-Buneman: discrete charge conservation
# --------------------------------------------------------------------------
def cloud_charge(r, z, q, nr, nz, dr, dz):
    """Charge of the ring cloud [r -+ dr/2] x [z -+ dz/2] (uniform volume
    density) inside the control volumes centred on the nodes (i dr, k dz)."""
    out = np.zeros((nr + 4, nz + 4))
    vc = np.pi * ((r + dr / 2) ** 2 - (r - dr / 2) ** 2) * dz
    for i in range(-2, nr + 2):
        a0, a1 = max(0.0, (i - 0.5) * dr), (i + 0.5) * dr
        lo, hi = max(a0, r - dr / 2), min(a1, r + dr / 2)
        if a1 <= 0 or hi <= lo:
            continue
        vr = np.pi * (hi * hi - lo * lo)
        for k in range(-2, nz + 2):
            hz = max(0.0, min((k + 0.5) * dz, z + dz / 2) - max((k - 0.5) * dz, z - dz / 2))
            if hz > 0:
                out[i + 2, k + 2] = q * vr * hz / vc
    return out


@pytest.mark.parametrize("r_range", [(1.0, 8.5), (0.5, 1.5)])
def test_vb_discrete_continuity(r_range):
    """Q(n+1) - Q(n) + dt * (flux out) = 0 at every node for single particles
    crossing 0, 1 or 2 cell boundaries (including the axis branch): J_z(i,k)
    is the flux through z = (k+1/2) dz over the annulus around r = i dr,
    J_r(i,k) the flux through the cylinder r = (i+1/2) dr over the band
    around z = k dz."""
    rng = np.random.default_rng(11)
    nr = nz = 10
    dr = dz = 1e-3
    worst = 0.0
    for _ in range(120):
        r = rng.uniform(r_range[0] * dr, r_range[1] * dr)
        z = rng.uniform(2 * dz, 8 * dz)
        v = rng.normal(0, 1, 3)
        v[1] = 0.0
        v *= rng.uniform(1e7, 2.5e8) / np.linalg.norm(v)
        smb, dt = one_particle_smb(r, z, v, nr, nz, dr, dz)
        for ph in ("push", "move", "reset_j", "deposit"):
            smb.phase(ph)
        p = smb.particles(0)
        J = smb.get_grid("J")[0]
        dQ = cloud_charge(p["r"][0], p["z"][0], Q_E, nr, nz, dr, dz) - \
            cloud_charge(p["r_old"][0], p["z_old"][0], Q_E, nr, nz, dr, dz)
        i = np.arange(-2, nr + 2)[:, None] * np.ones((1, nz + 4))
        Fr = J[0] * 2 * np.pi * (i + 0.5) * dr * dz * dt
        Fz = J[2] * np.where(i == 0, np.pi * dr * dr / 4, 2 * np.pi * np.abs(i) * dr * dr) * dt
        div = Fr.copy()
        div[1:, :] -= Fr[:-1, :]
        div += Fz
        div[:, 1:] -= Fz[:, :-1]
        worst = max(worst, np.abs((dQ + div)[2:-2, 2:-2]).max() / abs(Q_E))
    assert worst < 1e-13, worst


def test_vb_path_through_a_cell_corner_deposits_nothing():
    """The reference's ten-node case splits a path that crosses both cell
    boundaries at the crossing of r = i dr, z1, and branches on z1 < k dz or
    z1 > k dz (currentVB.cpp:298-309 and siblings); a path whose z1 equals
    k dz exactly deposits no r-z current.  The same move shifted off the
    corner deposits and conserves charge."""
    nr = nz = 10
    dr = dz = 1e-3
    out = {}
    # with these roundings the crossing of r = 4 dr lands exactly on z = 8 dz
    for name, (r, z, vz) in {"corner": (4 * dr - 0.1 * dr, 8 * dz + 0.1 * dz, -1.0e8),
                             "off": (4 * dr - 0.1 * dr, 8 * dz + 0.05 * dz, -0.8e8)}.items():
        smb, dt = one_particle_smb(r, z, np.array([1.0e8, 0.0, vz]), nr, nz, dr, dz)
        for ph in ("push", "move", "reset_j", "deposit"):
            smb.phase(ph)
        p = smb.particles(0)
        assert p["r"][0] > 4 * dr and p["z"][0] < 8 * dz   # both boundaries crossed
        out[name] = np.abs(smb.get_grid("J")[0][[0, 2]]).max()
    assert out["corner"] == 0.0
    assert out["off"] > 0.0


def test_vb_azimuthal_current_matches_charge_times_velocity():
    """The azimuthal deposit weights the charge density of the mid-position
    ring with the cell volumes of CELL_VOLUME(i) and CELL_VOLUME(i+1) (= 2i+2,
    quirk Q1): sum_n J_phi(n) V_n = q v_phi (V_n of those formulas)."""
    rng = np.random.default_rng(5)
    dr = dz = 1e-3
    for _ in range(50):
        r = rng.uniform(2.2 * dr, 7.5 * dr)
        z = rng.uniform(3 * dz, 6 * dz)
        v = np.array([rng.uniform(-1e6, 1e6), rng.uniform(1e6, 1e7), rng.uniform(-1e6, 1e6)])
        smb, dt = one_particle_smb(r, z, v)
        for ph in ("push", "move", "reset_j", "deposit"):
            smb.phase(ph)
        p = smb.particles(0)
        Jp = smb.get_grid("J")[0, 1]
        i_n = int(np.floor(p["r"][0] / dr))
        pos_r = (p["r"][0] + p["r_old"][0]) / 2
        vol = lambda m: np.pi * dz * dr * dr * m
        s = 0.0
        for ii, m in ((i_n, 2 * i_n + 1), (i_n + 1, 2 * i_n + 2)):
            s += Jp[ii + 2, :].sum() * vol(m)
        # the ring of width dr at pos_r: charge density q / (2 pi dz dr pos_r)
        expect = Q_E * p["vphi"][0] * (np.pi * dz * ((pos_r + dr / 2) ** 2 - (pos_r - dr / 2) ** 2)) \
            / (2 * np.pi * dz * dr * pos_r)
        assert abs(s - expect) <= 1e-12 * abs(expect)


# --------------------------------------------------------------------------
# gather: ring-volume weights
# --------------------------------------------------------------------------
def node_weights(smb, which, comp, r, z, i0, k0):
    """weight of node (i0+a, k0+b) in the gathered component `comp` of the
    field `which` when only field component `src` is 1 at that node"""
    out = {}
    for src in range(3):
        for a in (0, 1):
            for b in (0, 1):
                G = np.zeros((1, 3, smb.nr_d + 4, smb.nz_d + 4))
                G[0, src, i0 + a + 2, k0 + b + 2] = 1.0
                smb.set_grid(which, G)
                val = smb.get_field_e(0, r, z) if which == "E" else smb.get_field_h(0, r, z)
                out[(src, a, b)] = val
    return out


@pytest.mark.parametrize("which", ["E", "H"])
def test_gather_ring_volume_identity(which):
    """For each component, sum_n w_n V_n = 2 pi r dr dz (the particle ring's
    volume), with V_n the reference's CELL_VOLUME factors: (2i+2, 2i+4) for
    the components whose radial node is floor(r/dr - 1/2) (E_r, H_z, H_phi;
    quirk Q1) and (2i+1, 2i+3) for floor(r/dr) (E_phi, E_z, H_r).  E_phi's
    fourth node lands in E_z (quirk Q2)."""
    rng = np.random.default_rng(3)
    dr = dz = 1e-3
    smb = O.SMB((10 * dr, 10 * dz), (10, 10), (1, 1), 1e-12, [(-1, 1, False)])
    for _ in range(20):
        r = rng.uniform(1.0 * dr, 8.5 * dr)
        z = rng.uniform(2 * dz, 8 * dz)
        ring = 2 * np.pi * r * dr * dz
        half = {"E": [0], "H": [2, 1]}[which]   # radial node floor(u - 1/2)
        for comp in range(3):
            ih = int(np.floor(r / dr - 0.5)) if comp in half else int(np.floor(r / dr))
            kz = int(np.floor(z / dz - 0.5)) if (which == "E" and comp == 2) or (which == "H" and comp in (0, 1)) \
                else int(np.floor(z / dz))
            W = node_weights(smb, which, comp, r, z, ih, kz)
            if comp in half:
                V = [np.pi * dz * dr * dr * (2 * ih + 2), np.pi * dz * dr * dr * (2 * ih + 4)]
            else:
                V = [np.pi * dz * dr * dr * (2 * ih + 1), np.pi * dz * dr * dr * (2 * ih + 3)]
            s = 0.0
            for a in (0, 1):
                for b in (0, 1):
                    w = W[(comp, a, b)]
                    if which == "E" and comp == 1 and (a, b) == (1, 1):
                        assert w[1] == 0.0 and w[2] > 0.0        # quirk Q2
                        s += w[2] * V[a]
                    else:
                        s += w[comp] * V[a]
            assert abs(s - ring) <= 1e-13 * ring, (which, comp, s / ring)


# --------------------------------------------------------------------------
# pushers
# --------------------------------------------------------------------------
def textbook_boris_relativistic(q, m, dt, E, B, v):
    """Boris (1970) for u = gamma v, written independently: half kick,
    rotation with t = qB dt / (2 m gamma), s = 2t / (1 + t^2), half kick."""
    v = np.asarray(v, float)
    gamma = 1 / np.sqrt(1 - v @ v / C ** 2)
    u = gamma * v
    um = u + q * dt / (2 * m) * np.asarray(E)
    gm = np.sqrt(1 + um @ um / C ** 2)
    t = q * dt / (2 * m * gm) * np.asarray(B)
    up_ = um + np.cross(um, t)
    s = 2 * t / (1 + t @ t)
    upl = um + np.cross(up_, s)
    u2 = upl + q * dt / (2 * m) * np.asarray(E)
    return u2 / np.sqrt(1 + u2 @ u2 / C ** 2)


def test_boris_equals_textbook_relativistic():
    rng = np.random.default_rng(9)
    dt = 1.6e-12
    for _ in range(200):
        v = rng.normal(0, 1, 3)
        v *= rng.uniform(1e3, 2.7e8) / np.linalg.norm(v)
        E = rng.normal(0, 1e5, 3)
        H = rng.normal(0, 1e4, 3)
        got = O.push_velocity("boris-relativistic", Q_E, M_E, dt, E, H, v)
        ref = textbook_boris_relativistic(Q_E, M_E, dt, E, MU0 * H, v)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_boris_nonrelativistic_rotation_conserves_speed():
    rng = np.random.default_rng(4)
    for _ in range(100):
        v = rng.normal(0, 1e6, 3)
        H = rng.normal(0, 1e4, 3)
        got = O.push_velocity("boris", Q_E, M_E, 1.6e-12, [0, 0, 0], H, v)
        assert abs(got @ got - v @ v) <= 1e-13 * (v @ v)


@pytest.mark.parametrize("pusher", ["vay", "hc", "boris-relativistic"])
def test_pushers_closed_form_without_magnetic_field(pusher):
    """With B = 0 every relativistic pusher reduces to
    v' = u' / sqrt(1 + u'^2 / c^2), u' = gamma v + q E dt / m."""
    rng = np.random.default_rng(6)
    dt = 1.6e-12
    for _ in range(100):
        v = rng.normal(0, 1, 3)
        v *= rng.uniform(1e4, 2.5e8) / np.linalg.norm(v)
        E = rng.normal(0, 1e5, 3)
        got = O.push_velocity(pusher, Q_E, M_E, dt, E, [0, 0, 0], v)
        u = v / np.sqrt(1 - v @ v / C ** 2) + Q_E * E * dt / M_E
        ref = u / np.sqrt(1 + u @ u / C ** 2)
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


# --------------------------------------------------------------------------
# kinematics
# --------------------------------------------------------------------------
def test_rz_rotation_conserves_transverse_speed_and_radius():
    rng = np.random.default_rng(8)
    dr = 1e-3
    for _ in range(100):
        r = rng.uniform(2 * dr, 7 * dr)
        v = rng.normal(0, 3e7, 3)
        smb, dt = one_particle_smb(r, 5e-3, v, pusher="boris")
        smb.phase("move")
        p = smb.particles(0)
        assert abs(p["vr"][0] ** 2 + p["vphi"][0] ** 2 - (v[0] ** 2 + v[1] ** 2)) <= 1e-13 * (v[0] ** 2 + v[1] ** 2)
        rr = np.hypot(r + v[0] * dt, v[1] * dt)
        assert abs(p["r"][0] - rr) <= 1e-15 * rr


def test_single_bounce_mirrors_position():
    dr = dz = 1e-3
    nr = 10
    R = nr * dr
    wall = R - dr / 2
    dt = 0.5 * dz / C
    for excess in (1e-6, 3e-5, 2e-4):
        r0 = wall - 1e-6
        vr = (excess + 1e-6) / dt
        smb, _ = one_particle_smb(r0, 5e-3, [vr, 0.0, 1e5], pusher="boris")
        smb.phase("move")
        p = smb.particles(0)
        assert abs(p["r"][0] - (2 * wall - (r0 + vr * dt))) <= 1e-15
        assert p["vr"][0] == -vr
    # axis: positions never fall below dr/2
    smb, _ = one_particle_smb(0.6e-3, 5e-3, [-2.9e8 / 2, 0.0, 0.0], pusher="boris")
    smb.phase("move")
    assert smb.particles(0)["r"][0] >= dr / 2


# --------------------------------------------------------------------------
# Yee
# --------------------------------------------------------------------------
def test_yee_linear_and_zero_fixed_point():
    rng = np.random.default_rng(12)
    size, grid = Hh.geometry(8, 12)
    smb = O.SMB(size, grid, (2, 2), 1e-12, [(-1, 1, False)])
    for nm in ("E", "H", "H_at_et", "J"):
        smb.set_grid(nm, np.zeros((4, 3, 8, 10)))
    smb.phase("yee_e")
    smb.phase("yee_h")
    assert np.abs(smb.get_grid("E")).max() == 0 and np.abs(smb.get_grid("H")).max() == 0
    F = {nm: rng.normal(0, 1, (4, 3, 8, 10)) for nm in ("E", "H_at_et", "J")}
    out = []
    for scale in (1.0, 4.0):
        for nm, a in F.items():
            smb.set_grid(nm, scale * a)
        smb.phase("yee_e")
        smb.phase("yee_h")
        out.append((smb.get_grid("E"), smb.get_grid("H")))
    for a, b in zip(out[0], out[1]):
        assert np.abs(4.0 * a - b).max() <= 1e-14 * np.abs(b).max()


def test_yee_single_node_update():
    """E_z at an interior node with only H_phi_at_et(i, k) = h set:
    E_z -= koef_h * (-h/dr - h / (2 dr i)) with koef_h = dt / eps0."""
    dr = dz = 1e-3
    dt = 1e-12
    smb = O.SMB((8 * dr, 8 * dz), (8, 8), (1, 1), dt, [(-1, 1, False)])
    i, k, h = 3, 4, 7.0
    Hat = np.zeros((1, 3, 12, 12))
    Hat[0, 1, i + 2, k + 2] = h
    smb.set_grid("H_at_et", Hat)
    smb.phase("yee_e")
    E = smb.get_grid("E")[0]
    expect = -(dt * 2 / (2 * 8.85e-12)) * (-h / dr - h / (2 * dr * i))
    assert abs(E[2, i + 2, k + 2] - expect) <= 1e-15 * abs(expect)
    # E_r(i, k) sees +h/dz, E_r(i, k+1) sees -h/dz
    assert abs(E[0, i + 2, k + 2] + dt / 8.85e-12 * h / dz) <= 1e-15 * abs(dt / 8.85e-12 * h / dz)
    assert abs(E[0, i + 2, k + 3] - dt / 8.85e-12 * h / dz) <= 1e-15 * abs(dt / 8.85e-12 * h / dz)


# --------------------------------------------------------------------------
# migration
# --------------------------------------------------------------------------
def test_runaway_collector_places_every_particle_in_its_domain():
    rng = np.random.default_rng(2)
    size, grid = Hh.geometry(16, 32)
    smb = O.SMB(size, grid, (2, 4), 0.5e-3 / C, Hh.species_list(True))
    for sp, (q, m, b) in enumerate(Hh.species_list(True)):
        smb.add_particles(sp, *Hh.make_particles(rng, size, grid, 2000, m, vmax=2.9e8))
    n0 = [smb.count(sp) for sp in range(3)]
    for _ in range(3):
        for ph in ("push", "move", "migrate"):
            smb.phase(ph)
    for sp in range(3):
        p = smb.particles(sp)
        dom = (np.floor(p["r"] / smb.dr).astype(int) // smb.nr_d) * 4 + np.floor(p["z"] / smb.dz).astype(int) // smb.nz_d
        assert np.array_equal(dom, p["dom"])
        assert smb.count(sp) <= n0[sp]


# --------------------------------------------------------------------------
# P12 binary collision
# --------------------------------------------------------------------------
@pytest.mark.parametrize("pair", ["ee", "ei"])
def test_p12_pair_conserves_momentum_and_energy(pair):
    """A P12 collision rotates the relative momentum in the centre-of-mass
    frame: total relativistic momentum and energy are conserved per pair
    (equal weights, collisionsP12.cpp:29-328)."""
    rng = np.random.default_rng(21)
    m_i = 1836 * M_E
    ma, mb = (M_E, M_E) if pair == "ee" else (M_E, m_i)
    qa, qb = (Q_E, Q_E) if pair == "ee" else (Q_E, -Q_E)
    moved = 0
    for _ in range(200):
        va = rng.normal(0, 1, 3) * rng.uniform(1e5, 5e7)
        vb = rng.normal(0, 1, 3) * (rng.uniform(1e5, 5e7) if pair == "ee" else rng.uniform(1e3, 1e5))
        u = rng.uniform(0, 1, 3)
        a, b, used = O.p12_collide_single(ma, mb, qa, qb, va, 1e6, vb, 1e6, 1e20, 1e20, 1e-7, 1e-12, u)
        def pe(m, v):
            g = 1 / np.sqrt(1 - v @ v / C ** 2)
            return g * m * v, g * m * C * C
        p0 = pe(ma, va)[0] + pe(mb, vb)[0]
        p1 = pe(ma, a)[0] + pe(mb, b)[0]
        e0 = pe(ma, va)[1] + pe(mb, vb)[1]
        e1 = pe(ma, a)[1] + pe(mb, b)[1]
        if used:
            moved += int(not np.allclose(a, va))
            assert np.abs(p1 - p0).max() <= 1e-9 * np.abs(p0).max() + 1e-12 * ma * np.abs(va).max()
            # the centre-of-mass transform works with gamma m c, so the total
            # energy is conserved to the rounding of the rest energy
            assert abs(e1 - e0) <= 1e-14 * e0
    if pair == "ei":
        assert moved > 150
