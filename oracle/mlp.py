"""CPU restatement of the collision-surrogate network (test infrastructure).

The network has no counterpart in the reference; this numpy version is the
checker of the device kernels' forward and backward passes.  With
``bf16=True`` it rounds to bfloat16 (round to nearest even) exactly where the
device kernels feed an MFMA -- inputs, weights, hidden activations and
back-propagated errors -- and keeps biases, activations and accumulation in
higher precision, so the device result differs from it only by FP32
accumulation order and the last ulp of tanh.  With ``bf16=False`` it is the
exact float64 network, whose gradient is pinned by finite differences.

Parameter block (float32): W1[64][32] b1[64] W2[64][64] b2[64] W3[16][64]
b3[16]; W1 columns >= 6 and W3 rows >= 4 (and b3 entries >= 4) are padding.
"""
import numpy as np

IN, HID, OUT, INP, OUTP = 6, 64, 4, 32, 16
SIZES = [("W1", (HID, INP)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (OUTP, HID)),
         ("b3", (OUTP,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init(seed=0):
    rng = np.random.default_rng(seed)
    d = {name: np.zeros(shape) for name, shape in SIZES}
    a1 = np.sqrt(6.0 / (IN + HID))
    a2 = np.sqrt(6.0 / (2 * HID))
    a3 = np.sqrt(6.0 / (HID + OUT))
    d["W1"][:, :IN] = rng.uniform(-a1, a1, (HID, IN))
    d["W2"][:, :] = rng.uniform(-a2, a2, (HID, HID))
    d["W3"][:OUT, :] = rng.uniform(-a3, a3, (OUT, HID))
    d["b1"][:] = rng.normal(0, 0.1, HID)
    d["b2"][:] = rng.normal(0, 0.1, HID)
    d["b3"][:OUT] = rng.normal(0, 0.1, OUT)
    return pack(d).astype(np.float32)


def _q(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def forward(p, X, bf16_ops=True):
    """Y[n][4] and the activations; X rows carry at least 6 features."""
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)[:, :IN]
    Xp = np.zeros((len(X), INP))
    Xp[:, :IN] = X
    q = bf16_ops
    Z1 = _q(Xp, q) @ _q(w["W1"], q).T + w["b1"]
    H1 = np.tanh(Z1)
    Z2 = _q(H1, q) @ _q(w["W2"], q).T + w["b2"]
    H2 = np.tanh(Z2)
    Y = _q(H2, q) @ _q(w["W3"], q).T + w["b3"]
    return Y[:, :OUT], (Xp, H1, H2, Y)


def loss(p, X, T, bf16_ops=True):
    Y, _ = forward(p, X, bf16_ops)
    return float(((Y - np.asarray(T, dtype=np.float64)) ** 2).sum() / len(X))


def grad(p, X, T, bf16_ops=True):
    """Gradient of sum (y - t)^2 / n in the parameter-block layout."""
    w = unpack(p)
    q = bf16_ops
    n = len(X)
    Y4, (Xp, H1, H2, Y) = forward(p, X, bf16_ops)
    dY = np.zeros((n, OUTP))
    dY[:, :OUT] = 2.0 * (Y4 - np.asarray(T, dtype=np.float64)) / n
    g = {}
    g["W3"] = _q(dY, q).T @ _q(H2, q)
    g["b3"] = dY.sum(0)
    dZ2 = (_q(dY, q) @ _q(w["W3"], q)) * (1 - H2 * H2)
    g["W2"] = _q(dZ2, q).T @ _q(H1, q)
    g["b2"] = dZ2.sum(0)
    dZ1 = (_q(dZ2, q) @ _q(w["W2"], q)) * (1 - H1 * H1)
    g["W1"] = _q(dZ1, q).T @ _q(Xp, q)
    g["W1"][:, IN:] = 0.0
    g["b1"] = dZ1.sum(0)
    return pack(g), loss(p, X, T, bf16_ops)
