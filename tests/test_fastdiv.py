"""Division by invariant integers as the kernels do it (cell index -> (r, z)
cell, cell -> domain): q = floor(n m / 2^p) with p = 31 + ceil(log2 d) and
m = ceil(2^p / d), exact for every dividend below 2^31.  Checked here
against integer division on the divisors the grids use and on random and
boundary dividends (multiples of d and their neighbours, 2^31 - 1)."""
import numpy as np


def magic(d):
    ell = 0
    while (1 << ell) < d:
        ell += 1
    p = 31 + ell
    m = ((1 << p) + d - 1) // d
    assert m < (1 << 32)
    return m, p


def fastdiv(n, d):
    m, p = magic(d)
    return (n.astype(object) * m) >> p


def check(d, rng):
    n = np.concatenate([rng.integers(0, 2 ** 31, 4000, dtype=np.int64),
                        np.array([0, 1, d - 1, d, d + 1, 2 ** 31 - 1, 2 ** 31 - 2], dtype=np.int64)])
    k = rng.integers(0, (2 ** 31 - 1) // d, 200, dtype=np.int64)
    n = np.concatenate([n, k * d, k * d + d - 1, np.maximum(k * d - 1, 0)])
    n = n[(n >= 0) & (n < 2 ** 31)]
    q = fastdiv(n, d)
    assert np.array_equal(q.astype(np.int64), n // d), d


def test_fastdiv_small_divisors():
    rng = np.random.default_rng(1)
    for d in range(1, 1100):
        check(d, rng)


def test_fastdiv_grid_divisors():
    rng = np.random.default_rng(2)
    for d in [2048, 4096, 8192, 16384, 3 * 4096, 1000, 1023, 1025, 65535, 65536, 99991, 2 ** 20 + 7, 2 ** 30,
              2 ** 31 - 1]:
        check(d, rng)
