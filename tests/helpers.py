"""Shared small problems and comparison helpers for the tests."""
import functools

import numpy as np

from phoebe_amd import synth


@functools.lru_cache(maxsize=None)
def tiny_problem(natoms=2, nw=4, rg_e=3, rg_p=2, mesh=4, seed=7):
    return synth.make_problem(natoms, nw, rg_e, rg_p, mesh, seed=seed, name="tiny")


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b), 1e-300)
    return np.abs(a - b) / scale


def projector_diff(V1, V2, W, tol=1e-6):
    """Max difference of the spectral projectors of degenerate groups."""
    n = len(W)
    worst = 0.0
    i = 0
    while i < n:
        j = i + 1
        while j < n and abs(W[j] - W[i]) < tol:
            j += 1
        P1 = V1[:, i:j] @ V1[:, i:j].conj().T
        P2 = V2[:, i:j] @ V2[:, i:j].conj().T
        worst = max(worst, np.abs(P1 - P2).max())
        i = j
    return worst
