"""The CPU restatement of the screen's MLP, pinned by central differences and
BF16 rounding identities (CPU only)."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_identities():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -2.5, 0.0, np.inf, -np.inf],
                 dtype=np.float32)
    r = M.bf16(x)
    # ties to even: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6 (even mantissa)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2**-7)
    assert r[3] == np.float32(1.0 + 2**-6)
    assert r[4] == -2.5 and r[5] == 0.0 and np.isinf(r[6]) and np.isinf(r[7])
    y = np.random.default_rng(0).standard_normal(10000).astype(np.float32)
    assert np.max(np.abs(M.bf16(y) - y) / np.abs(y)) <= 2.0**-8


def test_gradients_match_central_differences():
    rng = np.random.default_rng(3)
    P = M.random_params(5)
    X = rng.standard_normal((37, M.IN))
    t = rng.standard_normal(37)
    loss, g, _ = M.loss_and_grads(P, X, t, rnd=False)
    idx = np.concatenate([rng.choice(M.NPARAMS, 60, replace=False), [M.OFF["b3"], M.OFF["b1"]]])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        num = (M.loss_and_grads(Pp, X, t, rnd=False)[0] - M.loss_and_grads(Pm, X, t, rnd=False)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, num, g[i])


def test_planted_mistake_fails_the_check():
    """A dropped ReLU mask in the backward pass is caught by the difference check."""
    rng = np.random.default_rng(4)
    P = M.random_params(6)
    X = rng.standard_normal((20, M.IN))
    t = rng.standard_normal(20)
    W1, b1, W2, b2, W3, b3 = M.unpack(P)
    H1, H2, y = M.forward(P, X, rnd=False)
    dy = (2 * (y - t) / len(t))[:, None]
    wrong = (H1.T @ (dy @ W3.T)).ravel()  # mask of layer 2 dropped
    _, g, _ = M.loss_and_grads(P, X, t, rnd=False)
    assert np.abs(wrong - g[M.OFF["W2"]:M.OFF["b2"]]).max() > 1e-6
