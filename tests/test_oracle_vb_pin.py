"""Villasenor-Buneman deposit of the oracle near the axis and in the azimuthal
component, pinned by the cloud geometry (src/current/currentVB.cpp:22-173,
398-498).

- The axis cell (i = 0, with the reference's own density factor,
  currentVB.cpp:110-172): each J_z node receives the charge the ring cloud
  carries across its face, by quadrature of rho v_z (the same independent
  computation as away from the axis), for moves inside the axis cell and
  across r = dr.
- J_phi: the mid-position ring cloud [r -+ dr/2] x [z -+ dz/2] with charge
  density q / (2 pi r dr dz) (uniform in the ring's volume) is split over the
  nodes (i dr, k dz), (i dr, (k+1) dz), ((i+1) dr, ...) by ring-area and
  axial overlaps with each node's control cell, normalised by the volumes
  CELL_VOLUME(r_i) and CELL_VOLUME(r_i + 1) expanded textually (quirk Q1),
  times v_phi.  (i, k) is the cell of the NEW position (P_CELL_R/Z), so for a
  move whose mid-point lies in another cell the weights are the same linear
  forms extrapolated; the axial first moment sum J V z = q v_phi z_mid holds
  for every move.
- Quirk Q6: rz_current_distribution starts with current.overlay_set(0)
  (currentVB.cpp:177), after the azimuthal deposit, so J_phi halo
  contributions are dropped while the J_r / J_z halo deposits stay.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_physics import Q_E, _vb_axial_flux, one_particle_smb

CELL_VOLUME_TEXT = "PI * (dz) * (dr) * (dr) * (2.0 * {i} + 1)"


def cell_volume(arg_text, r_i, dr, dz):
    return eval(CELL_VOLUME_TEXT.format(i=arg_text), {"PI": np.pi, "dr": dr, "dz": dz, "r_i": r_i})


def _deposit(r, z, v, nr=10, nz=10, dr=1e-3, dz=1e-3):
    smb, dt = one_particle_smb(r, z, v, nr, nz, dr, dz)
    for ph in ("push", "move", "reset_j", "deposit"):
        smb.phase(ph)
    return smb, dt, smb.particles(0)


@pytest.mark.parametrize("where", ["inside", "across"])
def test_vb_axis_cell_axial_current_equals_swept_flux(where):
    rng = np.random.default_rng(90 if where == "inside" else 91)
    nr = nz = 10
    dr = dz = 1e-3
    done = 0
    while done < 25:
        r, z = rng.uniform(0.5 * dr, 1.4 * dr), rng.uniform(2.2 * dz, 7.5 * dz)
        v = rng.normal(0, 1, 3)
        v[1] = 0.0
        v *= rng.uniform(2e7, 2.5e8) / np.linalg.norm(v)
        smb, dt, p = _deposit(r, z, v, nr, nz, dr, dz)
        ro, zo, rn, zn = p["r_old"][0], p["z_old"][0], p["r"][0], p["z"][0]
        inside = ro < dr and rn < dr
        across = (ro < dr) != (rn < dr)
        if (where == "inside" and not inside) or (where == "across" and not across):
            continue
        if min(abs(rn - ro) / dr, abs(zn - zo) / dz) < 1e-3:
            continue
        J = smb.get_grid("J")[0]
        i = np.arange(-2, nr + 2)[:, None] * np.ones((1, nz + 4))
        Qz = J[2] * np.where(i == 0, np.pi * dr * dr / 4, 2 * np.pi * np.abs(i) * dr * dr) * dt
        Bz = _vb_axial_flux(ro, zo, rn, zn, Q_E, nr, nz, dr, dz)
        err = np.abs(Qz - Bz).max() / abs(Q_E)
        assert err <= 1e-11, (where, err, ro / dr, rn / dr)
        done += 1


def test_vb_azimuthal_four_node_split():
    rng = np.random.default_rng(12)
    nr = nz = 10
    dr = dz = 1e-3
    checked = 0
    for _ in range(300):
        r = rng.uniform(0.6 * dr, 8.0 * dr)
        z = rng.uniform(2.5 * dz, 7.0 * dz)
        v = np.array([rng.uniform(-3e6, 3e6), rng.uniform(-3e7, 3e7), rng.uniform(-3e6, 3e6)])
        smb, dt, p = _deposit(r, z, v, nr, nz, dr, dz)
        rn, zn = p["r"][0], p["z"][0]
        rm, zm = (rn + p["r_old"][0]) / 2, (zn + p["z_old"][0]) / 2
        ri, zk = int(np.floor(rn / dr)), int(np.floor(zn / dz))
        if int(np.floor(rm / dr)) != ri or int(np.floor(zm / dz)) != zk:
            continue        # the extrapolated case is covered by the moment test
        Jp = smb.get_grid("J")[0, 1]
        rho = Q_E / (2 * np.pi * dz * dr * rm)
        want = np.zeros_like(Jp)
        # node radial control cells [(i - 1/2) dr, (i + 1/2) dr], axial [(k - 1/2) dz, (k + 1/2) dz]
        for di, vtext in ((0, "r_i"), (1, "r_i + 1")):
            lo, hi = max(rm - dr / 2, (ri + di - 0.5) * dr), min(rm + dr / 2, (ri + di + 0.5) * dr)
            area = np.pi * (hi * hi - lo * lo)
            for dk in (0, 1):
                zl, zh = max(zm - dz / 2, (zk + dk - 0.5) * dz), min(zm + dz / 2, (zk + dk + 0.5) * dz)
                want[ri + di + 2, zk + dk + 2] = rho * area * (zh - zl) / cell_volume(vtext, ri, dr, dz) * p["vphi"][0]
        if ri + 1 + 2 >= Jp.shape[0] - 2:
            continue        # the upper node sits in the halo, wiped by quirk Q6
        scale = np.abs(want).max()
        assert np.abs(Jp - want).max() <= 1e-12 * scale, (rm / dr, zm / dz)
        checked += 1
    assert checked > 100


def test_vb_azimuthal_axial_moment_for_any_move():
    rng = np.random.default_rng(13)
    dr = dz = 1e-3
    for _ in range(100):
        r = rng.uniform(1.2 * dr, 7.5 * dr)
        z = rng.uniform(2.5 * dz, 7.0 * dz)
        v = rng.normal(0, 1, 3)
        v *= rng.uniform(1e7, 2e8) / np.linalg.norm(v)
        smb, dt, p = _deposit(r, z, v)
        Jp = smb.get_grid("J")[0, 1]
        rm = (p["r"][0] + p["r_old"][0]) / 2
        zm = (p["z"][0] + p["z_old"][0]) / 2
        ri = int(np.floor(p["r"][0] / dr))
        tot0 = tot1 = 0.0
        for di, vtext in ((0, "r_i"), (1, "r_i + 1")):
            V = cell_volume(vtext, ri, dr, dz)
            row = Jp[ri + di + 2, :]
            k = np.arange(-2, len(row) - 2)
            tot0 += (row * V).sum()
            tot1 += (row * V * k * dz).sum()
        q_vphi = Q_E * p["vphi"][0]
        assert abs(tot0 - q_vphi) <= 1e-12 * abs(q_vphi)
        assert abs(tot1 - q_vphi * zm) <= 1e-12 * abs(q_vphi) * zm


def test_vb_azimuthal_halo_deposits_are_wiped():
    """Two domains along r; a particle at the inner domain's outer edge puts
    part of its ring cloud on the halo nodes: J_phi there is zero after the
    deposit (quirk Q6), J_r / J_z halo nodes keep their share."""
    dr = dz = 1e-3
    smb = O.SMB((8 * dr, 8 * dz), (8, 8), (2, 1), 0.5 * dz / 3e8, [(-1, 1, False)], ccs="vb")
    smb.add_particles(0, [3.8 * dr], [4.3 * dz], [2e7], [3e7], [1e7], [1.0])
    for ph in ("push", "move", "reset_j", "deposit"):
        smb.phase(ph)
    J = smb.get_grid("J")[0]            # inner domain, rows 0..3 interior, 4..5 halo
    assert np.abs(J[1, 2 + 4:]).max() == 0.0
    assert np.abs(J[1, 2 + 3]).max() > 0.0
    assert np.abs(J[0, 2 + 4:]).max() > 0.0 or np.abs(J[2, 2 + 4:]).max() > 0.0
