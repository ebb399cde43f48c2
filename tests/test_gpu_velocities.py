"""GPU: band velocities (velocity operator in the band basis, degenerate
blocks diagonalized) for electrons and phonons of the reference's Si inputs,
against the oracle and the known answers of test/velocityTest.cpp."""
import os
import sys

import numpy as np
import pytest

from oracle import harmonic_oracle as HO
from phoebe_amd import harmonic as Hm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
FC2 = os.path.join(GOLD, "silicon.fc")
FC4 = os.path.join(GOLD, "444_silicon.fc")
TB = os.path.join(GOLD, "si_tb.dat")
TB6 = os.path.join(GOLD, "666_si_tb.dat")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _groups(E, tol):
    out, ib, n = [], 0, len(E)
    while ib < n:
        s = 1
        while ib + s < n and abs(E[ib] - E[ib + s]) <= tol:
            s += 1
        out.append(list(range(ib, ib + s)))
        ib += s
    return out


def _compare(E, v, Eo, vo, tol_deg, label):
    """Gauge-invariant comparison: band energies, the diagonal of every
    direction (degenerate blocks hold their sorted block eigenvalues) and,
    for every pair of different degenerate groups, the Frobenius norm of the
    block between them (invariant under rotations inside each group)."""
    scale = max(np.abs(vo).max(), 1e-300)
    assert np.abs(E - Eo).max() < 1e-10 * max(np.abs(Eo).max(), 1e-300), label
    for i in range(3):
        assert np.abs(np.diag(v[:, :, i]).real - np.diag(vo[:, :, i]).real).max() < 1e-6 * scale, (label, i)
    groups = _groups(Eo, tol_deg)
    for a in groups:
        for b in groups:
            if a is b:
                continue
            for i in range(3):
                na = np.linalg.norm(v[np.ix_(a, b, [i])])
                no = np.linalg.norm(vo[np.ix_(a, b, [i])])
                assert abs(na - no) < 1e-6 * scale, (label, a, b, i)


def test_electron_velocities_vs_oracle():
    say("electron velocities")
    ph = Hm.PhononH0(FC2, "simple")
    el = Hm.ElectronH0Wannier(TB, ph.cell)
    tb = HO.read_w90_tb(TB, ph.cell)
    rng = np.random.default_rng(11)
    ks = np.vstack([np.zeros(3), 0.2 * ph.recip[:, 0], 0.35 * ph.recip[:, 0],
                    (ph.recip @ rng.uniform(-0.5, 0.5, (3, 12))).T])
    E, U, v = el.diag_velocities(ks)
    tol = 1e-6 / HO.ENERGY_RY_TO_EV
    for j, k in enumerate(ks):
        Eo, Uo, vo = HO.electron_velocities(tb, k)
        if j == 0:
            assert not np.any(v[j])      # Gamma
            continue
        _compare(E[j], v[j], Eo, vo, tol, ("k", j))
    # the Gamma-L points hold a degenerate pair
    assert any(len(g) > 1 for g in _groups(E[1], tol))


def test_electron_velocity_known_answers_single_point_rule():
    """WannierH0.Velocity / Velocity2 (test/velocityTest.cpp) through the GPU
    with the single-point Gamma rule; the batched rule zeroes k = (0.00025, 0, 0)."""
    say("electron velocity known answers")
    el = Hm.ElectronH0Wannier(TB6)
    dk = 0.00025
    k = np.array([[0, 0, 0], [dk, 0, 0], [0.1, 0.2, 0.3]])
    E, _, v = el.diag_velocities(k, gamma_rule="point")
    mass1 = 2 * (E[1, 3] - E[0, 3]) / dk ** 2
    mass2 = 2 * v[1, 3, 3, 0].real / dk
    say("  mass1", mass1, "mass2", mass2)
    assert abs(abs(mass1) - 686.779) < 0.02
    assert abs((mass1 - mass2) / mass1) < 0.02
    for i in range(3):
        kp = k[2].copy()
        kp[i] += 1e-4
        Ep, _ = el.diag(kp[None, :])
        x = (Ep[0, 0] - E[2, 0]) / 1e-4
        assert abs((x - v[2, 0, 0, i].real) / v[2, 0, 0, i].real) < 0.001
    _, _, vb = el.diag_velocities(k[:2], gamma_rule="batched")
    assert not np.any(vb[1])


@pytest.mark.parametrize("path,rule", [(FC4, "simple"), (FC4, ""), (FC2, "simple")])
def test_phonon_velocities_vs_oracle(path, rule):
    say("phonon velocities", os.path.basename(path), rule or "no sum rule")
    ph = Hm.PhononH0(path, rule)
    d = HO.read_qe_fc(path)
    if rule:
        d = HO.sum_rule_simple(d)
    R, W, M = HO.ws_reorder(d)
    lr = (*HO.long_range_setup(d), d["eps"], d["born"], d["tau"])
    masses = np.repeat(d["species_masses"][d["species"]], 3)
    rng = np.random.default_rng(12)
    qs = np.vstack([np.zeros(3), ph.recip @ [1 / 40, 0, 0], 0.25 * ph.recip[:, 0],
                    (ph.recip @ rng.uniform(-0.5, 0.5, (3, 10))).T])
    om, ev, v = ph.diag_velocities(qs)
    tol = 1e-4 / (1.0e10 * (1.0 / (6.62607015e-34 / (2 * np.pi) / 4.359744650e-18 * 1.0e12) / 2.0 / (2 * np.pi)) / 299792458.0)
    for j, q in enumerate(qs):
        wo, vo = HO.phonon_velocities(R, W, M, q, masses, lr)
        if j == 0:
            assert not np.any(v[j])
            continue
        _compare(om[j], v[j], wo, vo, tol, ("q", j))
    # eigenvectors come back divided by sqrt(M) (kokkosBatchedScaleEigenvectors)
    om2, V2 = ph.diag(qs[1:3], mass_scaling=True)
    assert np.abs(om2 - om[1:3]).max() < 1e-12 * np.abs(om).max()
    for j in range(2):
        P1 = ev[1 + j] @ ev[1 + j].conj().T
        P2 = V2[j] @ V2[j].conj().T
        assert np.abs(P1 - P2).max() < 1e-9 * np.abs(P2).max()


def test_speed_of_sound_known_answer_on_gpu():
    """PhononH0.Velocity (test/velocityTest.cpp) through the GPU: 2200 +- 200 m/s."""
    ph = Hm.PhononH0(FC4, "simple")
    q = ph.recip @ np.array([1 / 40, 0, 0])
    om, _, v = ph.diag_velocities(q[None, :], gamma_rule="point")
    gv = np.array([[v[0, b, b, i].real for i in range(3)] for b in range(3)])
    sound = min(abs(gv[b].min()) for b in range(3)) * HO.VELOCITY_RY_TO_SI
    say("  speed of sound", sound)
    assert abs(sound - 2200) < 200
    for b in range(3):
        assert abs(abs(gv[b] @ q) / np.linalg.norm(q) / np.linalg.norm(gv[b]) - 1) < 0.04
