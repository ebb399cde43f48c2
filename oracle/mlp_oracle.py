"""CPU restatement of the learned screen's network (test infrastructure only:
tests/ use it to check the GPU forward / backward kernels; the product never
imports it).

16 -> 64 -> 64 -> 1 MLP with ReLU, loss = mean (y - t)^2.  Parameters are one
flat FP32 vector: W1 (16 x 64, row-major [in][out]), b1 (64), W2 (64 x 64),
b2 (64), W3 (64 x 1), b3 (1).  With rnd=True every GEMM operand is rounded to
BF16 (round to nearest even) as the GPU kernels do, and the products are
accumulated in float64 (the GPU accumulates in FP32).  No counterpart in the
reference; its math is pinned by central differences in tests.
"""
import numpy as np

IN, HID = 16, 64
OFF = {"W1": 0, "b1": IN * HID}
OFF["W2"] = OFF["b1"] + HID
OFF["b2"] = OFF["W2"] + HID * HID
OFF["W3"] = OFF["b2"] + HID
OFF["b3"] = OFF["W3"] + HID
NPARAMS = OFF["b3"] + 1


def bf16(x):
    """Round float32 values to the nearest BF16 value (ties to even)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where(special, (u & 0xFFFF0000).astype(np.uint32), r)
    return r.view(np.float32).reshape(a.shape)


def unpack(P):
    P = np.asarray(P)
    return (P[OFF["W1"]:OFF["b1"]].reshape(IN, HID), P[OFF["b1"]:OFF["W2"]],
            P[OFF["W2"]:OFF["b2"]].reshape(HID, HID), P[OFF["b2"]:OFF["W3"]],
            P[OFF["W3"]:OFF["b3"]].reshape(HID, 1), P[OFF["b3"]:])


def _mm(a, b, rnd):
    if rnd:
        a, b = bf16(a), bf16(b)
    return np.asarray(a, np.float64) @ np.asarray(b, np.float64)


def forward(P, X, rnd=True):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    f = np.float32 if rnd else np.float64
    H1 = np.maximum(_mm(X, W1, rnd) + b1, 0.).astype(f)
    H2 = np.maximum(_mm(H1, W2, rnd) + b2, 0.).astype(f)
    y = (_mm(H2, W3, rnd)[:, 0] + b3[0]).astype(f)
    return H1, H2, y


def loss_and_grads(P, X, t, rnd=True):
    """(loss, grads (flat, like P), y)."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    n = len(t)
    f = np.float32 if rnd else np.float64
    H1, H2, y = forward(P, X, rnd)
    d = np.asarray(y, np.float64) - np.asarray(t, np.float64)
    loss = float(np.mean(d * d))
    dy = (2.0 * d / n).astype(f)[:, None]
    g = np.zeros(NPARAMS)
    g[OFF["W3"]:OFF["b3"]] = _mm(H2.T, dy, rnd)[:, 0]
    g[OFF["b3"]] = dy.sum(dtype=np.float64)
    dZ2 = (_mm(dy, W3.T, rnd) * (H2 > 0)).astype(f)
    g[OFF["W2"]:OFF["b2"]] = _mm(H1.T, dZ2, rnd).ravel()
    g[OFF["b2"]:OFF["W3"]] = dZ2.sum(axis=0, dtype=np.float64)
    dZ1 = (_mm(dZ2, W2.T, rnd) * (H1 > 0)).astype(f)
    g[OFF["W1"]:OFF["b1"]] = _mm(np.asarray(X).T, dZ1, rnd).ravel()
    g[OFF["b1"]:OFF["W2"]] = dZ1.sum(axis=0, dtype=np.float64)
    return loss, g, np.asarray(y, np.float64)


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAMS)
    P[OFF["W1"]:OFF["b1"]] = rng.uniform(-1, 1, IN * HID) * np.sqrt(6 / (IN + HID)) * scale
    P[OFF["b1"]:OFF["W2"]] = rng.uniform(-0.1, 0.1, HID)
    P[OFF["W2"]:OFF["b2"]] = rng.uniform(-1, 1, HID * HID) * np.sqrt(6 / (2 * HID)) * scale
    P[OFF["b2"]:OFF["W3"]] = rng.uniform(-0.1, 0.1, HID)
    P[OFF["W3"]:OFF["b3"]] = rng.uniform(-1, 1, HID) * np.sqrt(6 / (HID + 1)) * scale
    P[OFF["b3"]] = 0.3
    return P
