"""The CPU restatement of the surrogate network, pinned by mathematics:
its float64 gradient against central finite differences of its loss, and
its bfloat16 rounding against exactly representable cases."""
import numpy as np

from oracle import mlp as M


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 3.0e-3], dtype=np.float32)
    got = M.bf16(x)
    # 1 + 2^-8 is the midpoint between 1 and 1 + 2^-7: ties go to the even (1)
    assert got[0] == 1.0
    assert got[1] == 1.0
    assert got[2] == 1.0 + 2.0 ** -6   # 1 + 3*2^-8 is the midpoint of 1+2^-7 and 1+2^-6: even mantissa
    assert got[3] == 1.0 + 2.0 ** -7
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8


def test_float64_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init(3).astype(np.float64)
    X = rng.normal(size=(40, 6))
    T = rng.normal(size=(40, 4)) * 0.3
    g, _ = M.grad(p, X, T, bf16_ops=False)
    # every real parameter of a random subset, central differences
    w = M.unpack(np.arange(M.NPARAM, dtype=np.float64))
    real = np.concatenate([w["W1"][:, :6].ravel(), w["b1"], w["W2"].ravel()[::37], w["b2"],
                           w["W3"][:4].ravel(), w["b3"][:4]]).astype(int)
    h = 1e-6
    num = []
    for k in real:
        a, b = p.copy(), p.copy()
        a[k] += h
        b[k] -= h
        num.append((M.loss(a, X, T, False) - M.loss(b, X, T, False)) / (2 * h))
    num = np.array(num)
    assert np.abs(num - g[real]).max() <= 1e-6 * np.abs(num).max()
    # padding receives no gradient
    pad = np.concatenate([w["W1"][:, 6:].ravel(), w["W3"][4:].ravel(), w["b3"][4:]]).astype(int)
    assert np.all(g[pad] == 0)


def test_bf16_forward_close_to_exact():
    rng = np.random.default_rng(2)
    p = M.init(4)
    X = rng.normal(size=(100, 6))
    y16, _ = M.forward(p, X, True)
    y64, _ = M.forward(p, X, False)
    assert np.abs(y16 - y64).max() <= 0.05 * np.abs(y64).max()
