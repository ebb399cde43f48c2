"""Pins the oracle's Hermitian eigensolver (restating kokkosZHEEV,
common_kokkos.cpp:11-100) against LAPACK zheev and closed forms."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from oracle import oracle as O
from tests.helpers import projector_diff


def _rand_herm(rng, n):
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return X + X.conj().T


@pytest.mark.parametrize("n", [1, 2, 3, 6, 8, 12, 16, 18, 32])
def test_against_lapack_zheev(n):
    rng = np.random.default_rng(n)
    for _ in range(5):
        H = _rand_herm(rng, n)
        W, V = O.zheev(H)
        wl, vl, info = lapack.zheev(H, compute_v=1, lower=1)
        assert info == 0
        scale = max(1.0, np.abs(wl).max())
        assert np.abs(W - wl).max() <= 1e-12 * scale
        assert np.all(np.diff(W) >= 0)
        assert projector_diff(V, vl, wl) < 1e-10
        assert np.abs(V.conj().T @ V - np.eye(n)).max() < 1e-12


def test_planted_degeneracies():
    rng = np.random.default_rng(3)
    for n, lam in [(6, [1, 1, 2, 3, 3, 3]), (8, [-2, -2, 0, 0, 0, 5, 7, 7]),
                   (16, [0.5] * 3 + list(range(13)))]:
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        H = (Q * np.array(lam, dtype=float)) @ Q.conj().T
        W, V = O.zheev(H)
        assert np.abs(W - np.sort(lam)).max() < 1e-12
        assert np.abs(V @ np.diag(W) @ V.conj().T - H).max() < 1e-12
        assert projector_diff(V, Q[:, np.argsort(lam, kind="stable")], np.sort(lam)) < 1e-10


def test_two_by_two_closed_form():
    a, d, b = 0.3, -1.1, 0.4 - 0.7j
    H = np.array([[a, np.conj(b)], [b, d]])
    W, _ = O.zheev(H)
    m, r = (a + d) / 2, np.sqrt(((a - d) / 2) ** 2 + abs(b) ** 2)
    assert np.allclose(W, [m - r, m + r], atol=1e-15, rtol=0)


def test_diagonal_and_lower_triangle_reference():
    H = np.diag([3.0, -1.0, 2.0]).astype(complex)
    W, V = O.zheev(H)
    assert np.array_equal(W, [-1.0, 2.0, 3.0])
    # uplo = L: the upper triangle is ignored
    rng = np.random.default_rng(1)
    A = _rand_herm(rng, 5)
    B = A.copy()
    B[np.triu_indices(5, 1)] = 99.0
    assert np.abs(O.zheev(A)[0] - O.zheev(B)[0]).max() < 1e-13
