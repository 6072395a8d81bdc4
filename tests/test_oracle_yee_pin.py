"""The oracle's Yee update (src/maxwellSolver/maxwellSolverYee.cpp:111-257)
pinned by Maxwell's equations in integral form on the staggered r-z lattice
and by the scheme's discrete energy conservation.

Lattice (global cell indices, axisymmetric): E_r and H_z at (i + 1/2, k),
E_phi at (i, k), E_z and H_r at (i, k + 1/2), H_phi at (i + 1/2, k + 1/2);
node radii r_i = i dr.  Each update follows from Stokes' theorem on the face
dual to the node:

- Faraday, mu0 dH/dt = -curl E.  H_z: the annulus [r_i, r_i+1] at z_k, with
  circulation 2 pi (r_i+1 E_phi(i+1) - r_i E_phi(i)) over the area
  pi (r_i+1^2 - r_i^2).  H_r: the strip of the cylinder r_i between z_k and
  z_k+1 (dE_phi/dz).  H_phi: the rectangle [r_i, r_i+1] x [z_k, z_k+1] in the
  r-z plane.  field_h_at_et holds H at half steps and field_h the time-centred
  average of the old and new values.
- Ampere with a conductivity, eps dE/dt = curl H - J - sigma E, integrated by
  the trapezoidal rule in the sigma term:
  E' = E (2 eps - sigma dt)/(2 eps + sigma dt) + 2 dt (curl H - J)/(2 eps + sigma dt).
  E_z: the annulus [r_i-1/2, r_i+1/2] (the disk of radius dr/2 on the axis,
  where the circulation over the area gives 4 H_phi / dr).

The expected values are evaluated in this conservation form; the reference
writes the same operators expanded ((E1 + E0)/(2 r) + (E1 - E0)/dr), so the
comparison is to round-off, not bitwise.  Which nodes a domain with walls
updates is the reference's (loop bounds maxwellSolverYee.cpp:29-48,
168-257): a conducting wall freezes the tangential E on its last node row,
E_phi stays 0 on the axis, and E_z at the corner (0, 0) is never updated.
The outer wall's H_r is updated twice per step (the wall block
:119-130 and the regular loop :133-140).
"""
import numpy as np
import pytest

from oracle import oracle as O

EPS0 = 8.85e-12
MU0 = 1.26e-6


def _smb(grid, domains, pml=None):
    dr = dz = 1e-3
    dt = 0.5 * dz / 3e8
    return O.SMB((grid[0] * dr, grid[1] * dz), grid, domains, dt, [(-1, 1, False)], pml=pml), dr, dz, dt


def updated_e_nodes(nr, nz, walls):
    """Masks [3][nr][nz] of the E nodes a domain updates (walls: r=0, z=0,
    r=R, z=L), from the loop bounds of calc_field_e."""
    rb, zb = (1 if walls[0] else 0), (1 if walls[1] else 0)
    re, ze = (nr - 1 if walls[2] else nr), (nz - 1 if walls[3] else nz)
    m = np.zeros((3, nr, nz), dtype=bool)
    m[:, rb:re, zb:ze] = True
    if walls[0]:
        m[0, 0, zb:ze] = True
        m[2, 0, zb:ze] = True
    if walls[1]:
        m[2, rb:re, 0] = True
    return m


def curl_e_steps(E, dr, dz, dt, i0):
    """dH over one step for every interior H node, from Stokes on the dual
    faces; E is a local halo grid [3][nr+4][nz+4], i0 the first global row."""
    nr, nz = E.shape[1] - 4, E.shape[2] - 4
    i = np.arange(nr)[:, None] + i0
    ephi = E[1]
    C = lambda a, di, dk: a[2 + di:2 + di + nr, 2 + dk:2 + dk + nz]
    r0, r1 = i * dr, (i + 1) * dr
    circ = 2 * np.pi * (r1 * C(ephi, 1, 0) - r0 * C(ephi, 0, 0))
    d_hz = -dt / MU0 * circ / (np.pi * (r1 * r1 - r0 * r0))
    d_hr = dt / MU0 * (C(ephi, 0, 1) - C(ephi, 0, 0)) / dz
    d_hphi = -dt / MU0 * ((C(E[0], 0, 1) - C(E[0], 0, 0)) / dz - (C(E[2], 1, 0) - C(E[2], 0, 0)) / dr)
    return np.stack([d_hr, d_hphi, d_hz])


def curl_h(H, dr, dz, i0):
    """curl H at every interior E node (same staggering)."""
    nr, nz = H.shape[1] - 4, H.shape[2] - 4
    i = np.arange(nr)[:, None] + i0
    C = lambda a, di, dk: a[2 + di:2 + di + nr, 2 + dk:2 + dk + nz]
    c_r = -(C(H[1], 0, 0) - C(H[1], 0, -1)) / dz
    c_phi = (C(H[0], 0, 0) - C(H[0], 0, -1)) / dz - (C(H[2], 0, 0) - C(H[2], -1, 0)) / dr
    ro, ri = (i + 0.5) * dr, np.maximum(i - 0.5, 0.0) * dr
    inner = np.where(i == 0, 0.0, ri * C(H[1], -1, 0))
    c_z = 2 * np.pi * (ro * C(H[1], 0, 0) - inner) / (np.pi * (ro * ro - ri * ri))
    return np.stack([c_r, c_phi, c_z])


@pytest.mark.parametrize("domains", [(1, 1), (2, 2)])
def test_h_update_is_faraday_on_dual_faces(domains):
    grid = (12, 16)
    smb, dr, dz, dt = _smb(grid, domains)
    rng = np.random.default_rng(3)
    shape = (smb.n_domains, 3, smb.nr_d + 4, smb.nz_d + 4)
    E = rng.normal(0, 1e3, shape)
    Hat0 = rng.normal(0, 1.0, shape)
    smb.set_grid("E", E)
    smb.set_grid("H_at_et", Hat0)
    smb.phase("yee_h")
    H, Hat = smb.get_grid("H"), smb.get_grid("H_at_et")
    for d in range(smb.n_domains):
        g = smb.geometry(d)
        dH = curl_e_steps(E[d], dr, dz, dt, g.cell_dims[0])
        twice = np.zeros_like(dH, dtype=bool)
        if g.walls[2]:
            twice[0, -1, :] = True          # H_r on the outer wall, updated twice
        steps = np.where(twice, 2 * dH, dH)
        old = Hat0[d][:, 2:-2, 2:-2]
        scale = np.abs(dH).max()
        assert np.abs(Hat[d][:, 2:-2, 2:-2] - (old + steps)).max() <= 1e-12 * scale
        # field_h = the at-et value before the (last) increment + half of it
        centred = np.where(twice, old + dH + dH / 2, old + dH / 2)
        assert np.abs(H[d][:, 2:-2, 2:-2] - centred).max() <= 1e-12 * scale
        if g.walls[2]:
            assert np.abs(dH[0, -1]).max() > 0.1 * scale


@pytest.mark.parametrize("domains", [(1, 1), (2, 2)])
@pytest.mark.parametrize("with_sigma", [False, True])
def test_e_update_is_ampere_with_trapezoidal_conductivity(domains, with_sigma):
    grid = (12, 16)
    pml = (0.25, 0.25, 0.25, 1e-5, 0.07) if with_sigma else None
    smb, dr, dz, dt = _smb(grid, domains, pml=pml)
    rng = np.random.default_rng(4)
    shape = (smb.n_domains, 3, smb.nr_d + 4, smb.nz_d + 4)
    E0 = rng.normal(0, 1e3, shape)
    Hat = rng.normal(0, 1.0, shape)
    J = rng.normal(0, 1e3, shape)
    smb.set_grid("E", E0)
    smb.set_grid("H_at_et", Hat)
    smb.set_grid("J", J)
    if with_sigma:
        sig = rng.uniform(0, 0.5, (smb.n_domains, 1, smb.nr_d + 4, smb.nz_d + 4))
        smb.set_grid("sigma", sig)
    smb.phase("yee_e")
    E = smb.get_grid("E")
    for d in range(smb.n_domains):
        g = smb.geometry(d)
        s = sig[d, 0, 2:-2, 2:-2] * dt if with_sigma else 0.0
        ke = (2 * EPS0 - s) / (2 * EPS0 + s)
        kh = 2 * dt / (2 * EPS0 + s)
        want = E0[d][:, 2:-2, 2:-2] * ke + kh * (curl_h(Hat[d], dr, dz, g.cell_dims[0]) - J[d][:, 2:-2, 2:-2])
        mask = updated_e_nodes(smb.nr_d, smb.nz_d, list(g.walls))
        want = np.where(mask, want, E0[d][:, 2:-2, 2:-2])
        scale = np.abs(want).max()
        assert np.abs(E[d][:, 2:-2, 2:-2] - want).max() <= 1e-12 * scale
        # the halos are cleared before the update
        assert not E[d][:, :2].any() and not E[d][:, -2:].any()


def test_discrete_energy_is_conserved():
    """With J = 0 and no conductivity the leapfrog conserves
    W = eps0 sum V E^2 + mu0 sum V H^(n-1/2) . H^(n+1/2) exactly (V: the ring
    volume of each node's control cell), because the discrete curls are
    adjoint in these weights -- the property that makes the scheme stable.
    A wrong radius in one update, or a sign, breaks it at once."""
    nr, nz = 24, 32
    smb, dr, dz, dt = _smb((nr, nz), (1, 1))
    rng = np.random.default_rng(1)
    E = np.zeros((1, 3, nr + 4, nz + 4))
    mask = updated_e_nodes(nr, nz, [1, 1, 1, 1])
    E[0, :, 2:-2, 2:-2] = np.where(mask, rng.normal(0, 1e3, mask.shape), 0.0)
    smb.set_grid("E", E)
    i = np.arange(nr)
    rr = lambda x: np.maximum(x, 0.0) * dr
    v_half = np.pi * (rr(i + 1) ** 2 - rr(i) ** 2) * dz          # control [i, i+1]
    v_int = np.pi * (rr(i + 0.5) ** 2 - rr(i - 0.5) ** 2) * dz    # control [i-1/2, i+1/2]
    VE = [v_half, v_int, v_int]
    VH = [v_int, v_half, v_half]

    def w_e(E):
        return sum(EPS0 * (VE[c][:, None] * E[0, c, 2:-2, 2:-2] ** 2).sum() for c in range(3))

    def w_h(H0, H1):
        return sum(MU0 * (VH[c][:, None] * H0[0, c, 2:-2, 2:-2] * H1[0, c, 2:-2, 2:-2]).sum() for c in range(3))

    h_prev = smb.get_grid("H_at_et")
    W = []
    for _ in range(300):
        smb.phase("yee_e")
        e = smb.get_grid("E")
        smb.phase("yee_h")
        h = smb.get_grid("H_at_et")
        W.append(w_e(e) + w_h(h_prev, h))
        h_prev = h
    W = np.array(W)
    assert np.abs(W / W[0] - 1).max() < 1e-12
    # and the fields did evolve: the electric energy alone oscillates
    assert abs(w_e(e) / w_e(E) - 1) > 1e-3
