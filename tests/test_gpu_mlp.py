"""The surrogate network's BF16 MFMA kernels against the CPU restatement
(oracle/mlp.py) with bfloat16 rounding at the same places: forward outputs,
loss and the full parameter gradient.  The remaining differences are FP32
accumulation order and the last ulp of tanh (which can flip a bfloat16
rounding), hence the tolerances below."""
import numpy as np
import pytest

from oracle import mlp as M
from picopic_amd import capi

pytestmark = pytest.mark.gpu


def ctx():
    return capi.Context((8e-4, 8e-4), (8, 8), (1, 1), 1e-12, [(-1, 1, False, 64), (1, 1836, False, 64)])


def batch(rng, n):
    X = rng.normal(size=(n, 6)).astype(np.float32)
    T = (np.tanh(X[:, :4] * 0.7) + 0.2 * X[:, 4:5] - 0.1 * X[:, 5:6]).astype(np.float32)
    return X, T


@pytest.mark.parametrize("n", [1, 31, 32, 1000])
def test_forward_matches_cpu(n):
    rng = np.random.default_rng(n)
    c = ctx()
    p = M.init(5)
    c.mlp_set_params(p)
    assert np.array_equal(c.mlp_params(), p)
    X, _ = batch(rng, n)
    y = c.mlp_forward(X)
    y_ref, _ = M.forward(p, X, True)
    err = np.abs(y - y_ref).max() / np.abs(y_ref).max()
    print("forward", n, err)
    assert err <= 5e-3


@pytest.mark.parametrize("n", [32, 777, 4096])
def test_gradient_matches_cpu(n):
    rng = np.random.default_rng(100 + n)
    c = ctx()
    p = M.init(7)
    c.mlp_set_params(p)
    X, T = batch(rng, n)
    g, loss = c.mlp_grad(X, T)
    g_ref, loss_ref = M.grad(p, X, T, True)
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    cos = float(g @ g_ref / (np.linalg.norm(g) * np.linalg.norm(g_ref)))
    print("grad", n, err, cos, loss, loss_ref)
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    assert err <= 2e-2
    assert cos >= 0.9999
    # the padding of the parameter block gets no gradient
    w = M.unpack(np.arange(M.NPARAM))
    pad = np.concatenate([w["W1"][:, 6:].ravel(), w["W3"][4:].ravel(), w["b3"][4:]]).astype(int)
    assert np.all(g[pad] == 0)


def test_training_reduces_loss():
    rng = np.random.default_rng(3)
    c = ctx()
    c.mlp_init(11)
    X, T = batch(rng, 4096)
    _, l0 = c.mlp_grad(X, T)
    for _ in range(300):
        c.mlp_grad(X, T)
        c.mlp_adam(3e-3)
    _, l1 = c.mlp_grad(X, T)
    print("loss", l0, l1)
    assert l1 <= 0.1 * l0
    # the padding stays zero under the optimiser
    p = c.mlp_params()
    w = M.unpack(np.arange(M.NPARAM))
    pad = np.concatenate([w["W1"][:, 6:].ravel(), w["W3"][4:].ravel(), w["b3"][4:]]).astype(int)
    assert np.all(p[pad] == 0)
