"""The oracle's species maps (src/specieP.cpp:665-839, weight_cylindrical
include/algo/weighter.hpp:27-132, the pre-dump overlay
src/outController.cpp:409-491), pinned independently of their formulas:

- weight_cylindrical: a particle's deposit into each node, times the
  divisor volume (CELL_VOLUME of the lower node -- CYL_VOL(dz, dr) for
  dr/2 < r <= dr, also at r = dr where the lower node is 1 --;
  CELL_VOLUME(r_i + 1), i.e. (2 r_i + 2), for the upper: quirk Q1) and the
  cloud volume, equals
  the volume of its ring cloud [r -+ dr/2] x [z -+ dz/2] inside the node's
  dual cell [(i -+ 1/2) dr] x [(k -+ 1/2) dz], computed by quadrature (the
  cloud folded at the axis below r = dr/2);
- momentum, energy and temperature maps of particles sharing one velocity:
  m gamma v (gamma applied below REL_LIMIT, the reference's inverted test),
  m gamma^2 v^2 / 2e, and 0 (|P_abs| = |P|) after an energy map; a
  temperature map taken before any energy map reads the stale zero
  |P_abs| and is -m gamma^2 v^2 / 2e; particles paired with opposite
  velocities give temperature m gamma^2 v^2 / 2e;
- the overlay: a density map over 2 x 2 domains equals the single-domain map
  everywhere except the 2 x 2 nodes at the domain corner (no diagonal
  overlay), and a derived map is taken per domain (no neighbour's particles
  enter it).
"""
import numpy as np
import pytest

from oracle import oracle as O

EL_CHARGE = 1.6e-19
EL_MASS = 9.1e-31
C = 3e8


def one(size=(0.008, 0.008), grid=(8, 8), domains=(1, 1)):
    return O.SMB(size, grid, domains, 1e-12, [(-1, 1, False)])


def cloud_overlap(r, z, dr, dz, i, k, n=4000):
    """Volume of the ring cloud of a particle at (r, z) inside node (i, k)'s
    dual cell, the cloud folded at the axis; midpoint quadrature in r."""
    def seg(a, b):   # ring volume per unit z of [a, b] inside the node's band
        lo, hi = max(a, (i - 0.5) * dr, 0.0), min(b, (i + 0.5) * dr)
        return np.pi * (hi * hi - lo * lo) if hi > lo else 0.0
    zlo, zhi = max(z - dz / 2, (k - 0.5) * dz), min(z + dz / 2, (k + 0.5) * dz)
    lz = max(0.0, zhi - zlo)
    a, b = r - dr / 2, r + dr / 2
    v = seg(max(a, 0.0), b)
    if a < 0:
        v += seg(0.0, -a)   # the part below the axis folds back
    return v * lz


@pytest.mark.parametrize("pos", [(2.3e-3, 3.7e-3), (5.9e-3, 1.2e-3), (0.8e-3, 4.4e-3), (0.3e-3, 2.6e-3),
                                 (0.5e-3, 6.5e-3), (1.0e-3, 5.0e-3)])
def test_weight_cylindrical_is_the_cloud_overlap(pos):
    r, z = pos
    dr = dz = 1e-3
    s = one()
    w = 7.0
    s.add_particles(0, [r], [z], [0.0], [0.0], [0.0], [w])
    m = s.species_map(0, "density")[0, 0]
    ri, zk = int(np.floor(r / dr)), int(np.floor(z / dz))
    if r <= dr / 2:
        v0 = np.pi * dz * (2 * r * r + dr * dr / 2)
    else:
        v0 = 2 * np.pi * dz * dr * r
    lower = 0 if r <= dr / 2 else ri
    for di in (0, 1):
        for dk in (0, 1):
            i, k = lower + di, zk + dk
            vdiv = np.pi * dz * dr * dr * (2.0 * lower + 1 + di)   # CELL_VOLUME(r_i [+ 1]) as expanded
            if dr / 2 < r <= dr and di == 0:
                vdiv = np.pi * dz * dr * dr   # CYL_VOL(dz, dr) in this branch, whatever r_i (r = dr: r_i = 1)
            got = m[2 + i, 2 + k] * vdiv * v0 / w
            assert got == pytest.approx(cloud_overlap(r, z, dr, dz, i, k), rel=1e-9, abs=1e-22), (i, k)


def uniform_velocity_plasma(s, v, pairs=False, n=4000, seed=2):
    rng = np.random.default_rng(seed)
    R, L = 0.008, 0.008
    r = rng.uniform(0.6e-3, R - 0.6e-3, n)
    z = rng.uniform(0.6e-3, L - 0.6e-3, n)
    vv = np.tile(np.asarray(v, dtype=float)[:, None], (1, n))
    w = rng.uniform(0.5, 1.5, n) * 1e6
    if pairs:
        r, z, w = np.repeat(r[: n // 2], 2), np.repeat(z[: n // 2], 2), np.repeat(w[: n // 2], 2)
        vv[:, 1::2] *= -1
    s.add_particles(0, r, z, vv[0], vv[1], vv[2], w)


def test_momentum_energy_temperature_closed_forms():
    v = np.array([3e5, -2e5, 4e5])
    gamma = 1 / np.sqrt(1 - (v @ v) / C ** 2)
    s = one()
    uniform_velocity_plasma(s, v)
    n = s.species_map(0, "density")[0, 0]
    mask = n[2:-2, 2:-2] != 0
    t_stale = s.species_map(0, "temperature")[0, 0][2:-2, 2:-2]
    assert np.allclose(t_stale[mask], -0.5 * EL_MASS * gamma ** 2 * (v @ v) / EL_CHARGE, rtol=1e-9)
    p = s.species_map(0, "momentum")[0][:, 2:-2, 2:-2]
    for c in range(3):
        assert np.allclose(p[c][mask], EL_MASS * gamma * v[c], rtol=1e-9)
    e = s.species_map(0, "energy")[0, 0][2:-2, 2:-2]
    assert np.allclose(e[mask], 0.5 * EL_MASS * gamma ** 2 * (v @ v) / EL_CHARGE, rtol=1e-9)
    t = s.species_map(0, "temperature")[0, 0][2:-2, 2:-2]
    assert np.all(np.abs(t[mask]) <= 1e-6 * e[mask])
    # opposite velocities in pairs: |P| = 0, |P_abs| = n m gamma v
    s2 = one()
    uniform_velocity_plasma(s2, v, pairs=True)
    s2.species_map(0, "energy")
    t2 = s2.species_map(0, "temperature")[0, 0][2:-2, 2:-2]
    m2 = s2.species_map(0, "density")[0, 0][2:-2, 2:-2] != 0
    assert np.allclose(t2[m2], 0.5 * EL_MASS * gamma ** 2 * (v @ v) / EL_CHARGE, rtol=1e-9)


def test_each_simulation_keeps_its_own_maps():
    """The momentum_abs map an energy map leaves behind belongs to its
    simulation: a simulation created after another one is freed (malloc may
    hand it the same address) starts from zero maps, so its first
    temperature map is the stale -|P|^2 form."""
    v = np.array([3e5, -2e5, 4e5])
    gamma = 1 / np.sqrt(1 - (v @ v) / C ** 2)
    for _ in range(3):
        s = one()
        uniform_velocity_plasma(s, v)
        mask = s.species_map(0, "density")[0, 0][2:-2, 2:-2] != 0
        t = s.species_map(0, "temperature")[0, 0][2:-2, 2:-2]
        assert np.allclose(t[mask], -0.5 * EL_MASS * gamma ** 2 * (v @ v) / EL_CHARGE, rtol=1e-9)
        s.species_map(0, "energy")
        del s


def test_overlay_matches_a_single_domain_except_the_corner():
    rng = np.random.default_rng(4)
    n = 6000
    r = rng.uniform(0.6e-3, 8e-3 - 0.6e-3, n)
    z = rng.uniform(0.6e-3, 8e-3 - 0.6e-3, n)
    w = rng.uniform(0.5, 1.5, n) * 1e6
    v = rng.normal(0, 3e5, (3, n))
    a, b = one(), one(domains=(2, 2))
    for s in (a, b):
        s.add_particles(0, r, z, v[0], v[1], v[2], w)
    ma = a.species_map(0, "density")[0, 0, 2:-2, 2:-2]
    mb = b.species_map(0, "density")[:, 0, 2:-2, 2:-2]
    glued = np.block([[mb[0], mb[1]], [mb[2], mb[3]]])
    diff = np.abs(glued - ma) > 1e-9 * np.abs(ma).max()
    corner = np.zeros_like(diff)
    corner[2:6, 2:6] = True          # nodes (3..4, 3..4) around the corner at (4, 4), with margin
    assert diff.any() and not (diff & ~corner).any()
    # derived maps per domain: at a domain edge the energy map misses the
    # neighbour's particles, so it differs from the single-domain map there
    ea = a.species_map(0, "energy")[0, 0, 2:-2, 2:-2]
    eb = b.species_map(0, "energy")[:, 0, 2:-2, 2:-2]
    glued_e = np.block([[eb[0], eb[1]], [eb[2], eb[3]]])
    assert np.allclose(glued_e[:2, :2], ea[:2, :2], rtol=1e-9)
    assert not np.allclose(glued_e[3:5, 1], ea[3:5, 1], rtol=1e-6)
