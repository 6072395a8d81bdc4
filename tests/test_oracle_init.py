"""The oracle's restatement of the reference's initial plasma (flat placement,
weights, thermal velocities; src/specieP.cpp:79-189, 280-331, 346-384,
src/math/maxwellJuettner.cpp:182-255), pinned by what the mathematics fixes:

- random_reverse is the radical inverse (digit reversal) of its argument;
- the Gamma(3/2) inverse CDF of the classical branch matches
  scipy.special.gammaincinv, and the relativistic branch's g^-1 inverts
  g(x) = x - log(1 + x + x^2/2) (the log-tail of Gamma(3));
- the sampled energies follow their laws: Kolmogorov-Smirnov against the
  Gamma(3/2) CDF (T < 0.1) and against the Maxwell-Juettner CDF of gamma
  (T >= 0.1, numerically integrated), and the means against the closed forms
  1.5 T and K1(1/T)/K2(1/T) + 3T - 1;
- placement: every particle inside the area, r uniform over the placement
  extent, z uniform (flat density) or following the linear ramp's CDF,
  weights proportional to r, and sum w equal to the species' real particle
  count n pi dr^2 (top^2 - bot^2) z_size per domain;
- thermal velocities: mean kinetic energy half the mean sampled energy
  (three components uniform on [-1, 1) times sqrt(E/m_e));
- domains: M = macro_amount / Dr / Dz per domain, and after the collector
  every particle sits in the domain that owns its cell.
"""
import numpy as np
import pytest
from scipy import integrate, special, stats

from oracle import oracle as O

EL_CHARGE = 1.6e-19
EL_MASS = 9.1e-31


def digits_reversed(n, base):
    r, f = 0.0, 1.0 / base
    while n:
        n, d = divmod(n, base)
        r += d * f
        f /= base
    return r


@pytest.mark.parametrize("n", [0, 1, 12, 13, 14, 168, 169, 2197, 123456, 99999999])
def test_random_reverse_is_the_radical_inverse(n):
    for base in (11, 13):
        assert O.random_reverse(float(n), base) == pytest.approx(digits_reversed(n, base), rel=1e-15, abs=0)


def test_gamma32_quantile_matches_scipy():
    for u in np.concatenate([np.geomspace(6.2e-4, 0.5, 40), 1 - np.geomspace(0.5, 5.7e-9, 40)]):
        x = O.gamma32_quantile(u)
        assert x == pytest.approx(special.gammaincinv(1.5, u), rel=1e-11)


def test_mj_g_inverse_inverts_the_gamma3_log_tail():
    for x_lo in (1.0, 2.5, 10.0):
        for dt in (1e-6, 0.3, 5.0, 30.0):
            t = O.mj_g(x_lo) + dt
            x = O.mj_g_inverse(t, x_lo)
            # g(x) = -log of the Gamma(3) tail at x
            assert -np.log(special.gammaincc(3, x)) == pytest.approx(t, rel=1e-10)


def energies(temperature, n, seed=3):
    return np.array([O.mj_energy(temperature, seed, k) for k in range(n)])


def test_classical_energies_follow_gamma_three_halves():
    T = 0.05
    e = energies(T, 20000)
    assert np.mean(e) == pytest.approx(1.5 * T, rel=0.02)
    assert stats.kstest(e / T, stats.gamma(1.5).cdf).pvalue > 1e-3


@pytest.mark.parametrize("theta", [0.1, 1.0, 3.0])
def test_relativistic_energies_follow_maxwell_juettner(theta):
    e = energies(theta, 20000)
    gamma = e + 1
    mean = special.kn(1, 1 / theta) / special.kn(2, 1 / theta) + 3 * theta
    assert np.mean(gamma) == pytest.approx(mean, rel=0.02)
    pdf = lambda g: g * np.sqrt(g * g - 1) * np.exp(-(g - 1) / theta)
    norm = integrate.quad(pdf, 1, np.inf)[0]
    grid = np.linspace(1, 1 + 60 * theta, 4001)
    cdf_vals = np.concatenate([[0], np.cumsum([integrate.quad(pdf, a, b)[0] for a, b in zip(grid[:-1], grid[1:])])])
    cdf = lambda g: np.interp(g, grid, cdf_vals / norm)
    assert stats.kstest(gamma, cdf).pvalue > 1e-3


def smb(grid=(32, 64), domains=(2, 2), size=(0.0375 / 4, 0.15 / 4)):
    dt = 1e-13
    return O.SMB(size, grid, domains, dt, [(-1, 1, False), (1, 1836, False)])


@pytest.mark.parametrize("ramp", [False, True])
def test_placement_weights_and_domains(ramp):
    grid, domains, size = (32, 64), (2, 2), (0.0375 / 4, 0.15 / 4)
    s = smb(grid, domains, size)
    macro = 40003
    dl, dr_ = (1e16, 1e17) if ramp else (1e17, 1e17)
    s.init_reference(0, macro, dl, dr_, 1.0, seed=5)
    M = macro // domains[0] // domains[1]
    assert s.count(0) == M * domains[0] * domains[1]
    p = s.particles(0)
    dr, dz = size[0] / grid[0], size[1] / grid[1]
    R, L = size
    assert np.all(p["r"] >= 0.5 * dr) and np.all(p["r"] < R - 0.5 * dr)
    assert np.all(p["z"] >= 0.5 * dz) and np.all(p["z"] < L - 0.5 * dz)
    # every particle in the domain owning its cell (the collector ran)
    nrd, nzd = grid[0] // domains[0], grid[1] // domains[1]
    own = (np.floor(p["r"] / dr) // nrd) * domains[1] + np.floor(p["z"] / dz) // nzd
    assert np.array_equal(own.astype(int), p["dom"])
    # ids: id_base + d M + k; per generating domain (id // M) the closed forms
    for d in range(domains[0] * domains[1]):
        i, j = divmod(d, domains[1])
        sel = (p["id"] // M) == d
        assert sel.sum() == M
        r, z, w = p["r"][sel], p["z"][sel], p["w"][sel]
        r_size = nrd * dr - (dr if i == domains[0] - 1 else 0)
        z_size = nzd * dz - (dz if j == domains[1] - 1 else 0)
        r0, z0 = i * nrd * dr + 0.5 * dr, j * nzd * dz + 0.5 * dz
        assert stats.kstest((r - r0) / r_size, "uniform").pvalue > 1e-3
        zz = (z - z0) / z_size
        drho = (dr_ - dl) / L
        n0 = dl + drho * j * nzd * dz
        n1 = dl + drho * (j + 1) * nzd * dz
        if n0 == n1:
            assert stats.kstest(zz, "uniform").pvalue > 1e-3
        else:
            # z CDF of a density growing linearly from n0 to n1 over the extent
            cdf = lambda x: (n0 * x + 0.5 * (n1 - n0) * x * x) / (n0 + 0.5 * (n1 - n0))
            assert stats.kstest(zz, cdf).pvalue > 1e-3
        # weights proportional to r, summing to the domain's real particles
        ratio = w / r
        assert np.ptp(ratio) <= 1e-12 * ratio.mean()
        top = (i + 1) * nrd - (1 if i == domains[0] - 1 else 0)
        bot = i * nrd
        n_total = (n0 + n1) / 2 * np.pi * dr * dr * (top * top - bot * bot) * z_size
        assert w.sum() == pytest.approx(n_total, rel=1e-12)


def test_thermal_velocities_carry_half_the_sampled_energy():
    s = smb()
    s.init_reference(0, 40000, 1e17, 1e17, 1.0, seed=9)
    p = s.particles(0)
    ke_ev = 0.5 * EL_MASS * (p["vr"] ** 2 + p["vphi"] ** 2 + p["vz"] ** 2) / EL_CHARGE
    theta = 1.0
    mean_e = special.kn(1, 1 / theta) / special.kn(2, 1 / theta) + 3 * theta - 1
    assert np.mean(ke_ev) == pytest.approx(mean_e / 2, rel=0.02)
    # each component uniform on [-1, 1) times the particle's speed scale
    for k in ("vr", "vphi", "vz"):
        assert abs(np.mean(p[k])) < 0.02 * np.sqrt(np.mean(p[k] ** 2))
