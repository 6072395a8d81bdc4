"""The oracle's field gather (include/maxwellSolver.hpp:101-316) pinned by the
geometry of cloud-in-cell weighting, for each of the six components.

The gather is area weighting of a particle's ring cloud
[r - dr/2, r + dr/2] x [z - dz/2, z + dz/2] over the control cells of the
component's nodes on the staggered r-z Yee lattice:

    component   node radius    node axial position
    E_r, H_z    (i + 1/2) dr   k dz
    E_z, H_r    i dr           (k + 1/2) dz
    E_phi       i dr           k dz
    H_phi       (i + 1/2) dr   (k + 1/2) dz

Node (i, k) receives  pi (b^2 - a^2) * ovz / V(i),  where [a, b] is the
overlap of the cloud's radial extent with the node's radial control interval
(ring area) and ovz the overlap in z (metres), and V(i) is the reference's
normalising volume.  That volume is the CELL_VOLUME macro of
include/geometry.hpp:186, `PI * (dz) * (dr) * (dr) * (2.0 * i + 1)`, whose
argument is NOT parenthesised, expanded textually at the call sites
(`CELL_VOLUME(i_r+1)` and `CELL_VOLUME(i_r+3)` for the half-integer radial
components, `CELL_VOLUME(i_r)` and `CELL_VOLUME(i_r+2)` for the integer ones).
The test expands the macro text itself rather than restating the volumes.
The fourth E_phi term -- node (i_r + 1, k_z + 1) -- is added to E_z
(maxwellSolver.hpp:313).

Everything here is computed from interval overlaps, independently of the
oracle's split points and index arithmetic: a swapped pair of axial widths,
a wrong staggering or a transposed node moves the result by O(1).
"""
import numpy as np
import pytest

from oracle import oracle as O

# (radial offset of the node in cells, axial offset, the CELL_VOLUME call
# arguments for the lower and upper radial node) -- maxwellSolver.hpp lines
# 131-132 (H_z), 160-161 (H_r), 191-192 (H_phi), 242-243 (E_r), 266-267 (E_z),
# 296-297 (E_phi)
COMPONENTS = {
    ("E", 0): (0.5, 0.0, ("i_r+1", "i_r+3")),
    ("E", 1): (0.0, 0.0, ("i_r", "i_r+2")),
    ("E", 2): (0.0, 0.5, ("i_r", "i_r+2")),
    ("H", 0): (0.0, 0.5, ("i_r", "i_r+2")),
    ("H", 1): (0.5, 0.5, ("i_r+1", "i_r+3")),
    ("H", 2): (0.5, 0.0, ("i_r+1", "i_r+3")),
}

CELL_VOLUME_TEXT = "PI * (dz) * (dr) * (dr) * (2.0 * {i} + 1)"


def cell_volume(arg_text, i_r, dr, dz):
    """Textual expansion of the CELL_VOLUME macro with the call site's
    argument text, evaluated for the given i_r."""
    return eval(CELL_VOLUME_TEXT.format(i=arg_text), {"PI": np.pi, "dr": dr, "dz": dz, "i_r": i_r})


def overlap(lo1, hi1, lo2, hi2):
    """Intersection of two intervals; touching ends (within round-off of the
    1 mm cells) count as no overlap."""
    lo, hi = max(lo1, lo2), min(hi1, hi2)
    return (lo, hi) if hi - lo > 1e-15 else None


def node_weights(r, z, dr, dz, ar, az, vol_args):
    """{(i, k): weight} over the nodes whose control cell the cloud touches."""
    out = {}
    i_r = int(np.floor((r - ar * dr) / dr))     # lower radial node (floor of the node lattice)
    for i in range(i_r - 1, i_r + 3):
        ri = (i + ar) * dr
        ov = overlap(r - dr / 2, r + dr / 2, ri - dr / 2, ri + dr / 2)
        if ov is None:
            continue
        a, b = ov
        for k in range(int(np.floor(z / dz)) - 2, int(np.floor(z / dz)) + 3):
            zk = (k + az) * dz
            oz = overlap(z - dz / 2, z + dz / 2, zk - dz / 2, zk + dz / 2)
            if oz is None:
                continue
            if i == i_r:
                vol = cell_volume(vol_args[0], i_r, dr, dz)
            elif i == i_r + 1:
                vol = cell_volume(vol_args[1], i_r, dr, dz)
            else:
                raise AssertionError("a ring cloud of width dr touches at most two radial nodes")
            out[(i, k)] = np.pi * (b * b - a * a) * (oz[1] - oz[0]) / vol
    return out


def expected_gather(fields, r, z, dr, dz, dims):
    """fields[name][c] are local halo grids; dims = the domain's first global cell."""
    out = {"E": np.zeros(3), "H": np.zeros(3)}
    scale = {"E": np.zeros(3), "H": np.zeros(3)}
    for (name, c), (ar, az, vargs) in COMPONENTS.items():
        W = node_weights(r, z, dr, dz, ar, az, vargs)
        top = None
        if (name, c) == ("E", 1):
            # the E_phi node right of and above the cloud's lower corner
            top = (int(np.floor(r / dr)) + 1, int(np.floor(z / dz)) + 1)
        for (i, k), w in W.items():
            v = fields[name][c][i - dims[0] + 2, k - dims[1] + 2] * w
            dest = 2 if (i, k) == top else c
            out[name][dest] += v
            scale[name][dest] += abs(v)
    return out, scale


def _smb(grid, domains):
    dr = dz = 1e-3
    return O.SMB((grid[0] * dr, grid[1] * dz), grid, domains, 1e-12, [(-1, 1, False)])


@pytest.mark.parametrize("domains,dom", [((1, 1), 0), ((2, 2), 3)])
def test_gather_matches_cloud_overlap_weights(domains, dom):
    grid = (16, 24)
    smb = _smb(grid, domains)
    rng = np.random.default_rng(41)
    E = rng.normal(0, 1e4, (smb.n_domains, 3, smb.nr_d + 4, smb.nz_d + 4))
    H = rng.normal(0, 30, E.shape)
    smb.set_grid("E", E)
    smb.set_grid("H", H)
    g = smb.geometry(dom)
    dims = (g.cell_dims[0], g.cell_dims[1])
    dr, dz = smb.dr, smb.dz
    for t in range(400):
        # inside the domain, r >= dr/2 (the mover's clamp), including cell edges
        r = (dims[0] + rng.uniform(0, smb.nr_d)) * dr
        z = (dims[1] + rng.uniform(0, smb.nz_d)) * dz
        if t % 10 == 0:
            r = (dims[0] + rng.integers(0, smb.nr_d)) * dr + (0.5 * dr if t % 20 == 0 else 0.0)
        r = max(r, 0.5 * dr)
        want, scale = expected_gather({"E": E[dom], "H": H[dom]}, r, z, dr, dz, dims)
        got = {"E": smb.get_field_e(dom, r, z), "H": smb.get_field_h(dom, r, z)}
        for name in ("E", "H"):
            err = np.abs(got[name] - want[name]) / np.maximum(scale[name], 1e-300)
            assert err.max() < 1e-13, (name, r, z, got[name], want[name])


@pytest.mark.parametrize("name,c", list(COMPONENTS.keys()))
def test_gather_reproduces_fields_linear_in_z(name, c):
    """A field linear in the node's axial position, F = a + b z_k, constant in
    r, is gathered as S(r) (a + b z): linear interpolation is exact for linear
    data, where S(r) is the gathered value of the uniform field F = 1 (the sum
    of the radial weights, which carries the normalising-volume quirk)."""
    grid = (12, 20)
    smb = _smb(grid, (1, 1))
    dr, dz = smb.dr, smb.dz
    ar, az, _ = COMPONENTS[(name, c)]
    k = np.arange(-2, grid[1] + 2)
    a, b = 3.0, 2.5e3
    lin = np.zeros((1, 3, grid[0] + 4, grid[1] + 4))
    lin[0, c] = a + b * (k + az) * dz
    one = np.zeros_like(lin)
    one[0, c] = 1.0
    other = "H" if name == "E" else "E"
    smb.set_grid(other, np.zeros_like(lin))
    get = smb.get_field_e if name == "E" else smb.get_field_h
    rng = np.random.default_rng(5 + c)
    for _ in range(200):
        r = rng.uniform(0.5 * dr, (grid[0] - 1) * dr)
        z = rng.uniform(dz, (grid[1] - 1) * dz)
        smb.set_grid(name, one)
        s = get(0, r, z)
        smb.set_grid(name, lin)
        f = get(0, r, z)
        if (name, c) == ("E", 1):
            # the fourth term of E_phi lands in E_z: compare the sums
            assert np.isclose(f[1] + f[2], (s[1] + s[2]) * (a + b * z), rtol=1e-12, atol=0)
        else:
            assert np.isclose(f[c], s[c] * (a + b * z), rtol=1e-12, atol=0), (r, z, f[c], s[c])
