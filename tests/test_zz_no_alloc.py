"""phx_el_bloch_diag_noalloc (the entry point that allocates nothing and works
in the library's fixed device scratch) against phx_el_bloch_diag, both
through the C ABI: the same kernels on the same inputs, so energies and
eigenvectors agree to FP64 rounding (they come out bit for bit alike).  Two
calls in a row, the second smaller, on a tiny model and on a 16-band model
with 343 lattice vectors; a call larger than the scratch is rejected before
any device work."""
import sys

import numpy as np
import pytest

import phoebe_amd as P
from tests.helpers import tiny_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _model16(seed):
    """16 Wannier functions on a 7^3 grid of lattice vectors, Hermitian H(R)."""
    rng = np.random.default_rng(seed)
    n = np.stack(np.meshgrid(*[np.arange(-3, 4)] * 3, indexing="ij"), -1).reshape(-1, 3)
    cell = 5.0 * np.array([[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])
    R = n @ cell
    deg = rng.integers(1, 4, len(R)).astype(np.float64)
    deg = np.minimum(deg, deg[::-1])  # deg(-R) = deg(R)
    h = rng.standard_normal((len(R), 16, 16)) + 1j * rng.standard_normal((len(R), 16, 16))
    h *= np.exp(-np.linalg.norm(R, axis=1) / 4.0)[:, None, None]
    # H(-R) = H(R)^dagger makes H(k) Hermitian: R and -R are mirror indices
    h = 0.5 * (h + h[::-1].conj().transpose(0, 2, 1))
    recip = 2 * np.pi * np.linalg.inv(cell).T
    return R, deg, h, recip


def _compare(ks, R, deg, h0R):
    E0, U0 = P.el_bloch_diag(ks, R, deg, h0R)
    E1, U1 = P.el_bloch_diag(ks, R, deg, h0R, noalloc=True)
    scale = max(1.0, float(np.abs(E0).max()))
    de = float(np.abs(E1 - E0).max()) / scale
    du = float(np.abs(U1 - U0).max())
    say(f"  nk {len(ks)}, nw {h0R.shape[1]}, nR {len(R)}: max energy diff {de:.1e}, max eigenvector diff {du:.1e}, "
        f"bit-identical {np.array_equal(E0, E1) and np.array_equal(U0, U1)}")
    assert de <= 1e-12 and du <= 1e-12


def test_noalloc_bloch_diag_matches_allocating_entry_point():
    say("no-alloc Bloch diagonalization: tiny model")
    p = tiny_problem()
    rng = np.random.default_rng(31)
    for nk in (300, 17):  # second call smaller: the scratch keeps the first call's contents
        ks = (p.recip @ rng.uniform(-1, 1, (3, nk))).T
        _compare(ks, p.Rh, p.degh, p.h0R)
    say("no-alloc Bloch diagonalization: 16 bands, 343 lattice vectors")
    R, deg, h, recip = _model16(5)
    for nk in (6000, 1000):
        ks = (recip @ rng.uniform(-1, 1, (3, nk))).T
        _compare(ks, R, deg, h)


def test_noalloc_bloch_diag_rejects_calls_beyond_the_scratch():
    say("no-alloc Bloch diagonalization: oversized call")
    R, deg, h, recip = _model16(6)
    ks = np.zeros((20000, 3))  # 20000 x (16 x 16 x 16 B + ...) > 64 MiB
    with pytest.raises(P.PhxError, match="scratch"):
        P.el_bloch_diag(ks, R, deg, h, noalloc=True)
