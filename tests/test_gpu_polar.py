"""GPU polar (Froehlich) term against the oracle: polarCorrectionPart1 on the
GPU, calcCouplingSquared with polar data, and the linewidth builder with the
term on (direct transform), on a two-species cell with a dielectric tensor
and Born charges."""
import dataclasses
import sys

import numpy as np
import pytest

import phoebe_amd as P
from oracle import oracle as O
from tests.helpers import rel_err, tiny_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def polar_problem(seed=9, mesh=4):
    """Two atoms of different masses, a dielectric tensor and neutral Born charges."""
    p = dataclasses.replace(tiny_problem(natoms=2, nw=4, rg_e=3, rg_p=3, mesh=mesh, seed=seed))
    p.masses = p.masses.copy()
    p.masses[3:] *= 1.7
    rng = np.random.default_rng(seed)
    Z = rng.standard_normal((2, 3, 3)) + 2.0 * np.eye(3)[None]
    Z -= Z.mean(axis=0)
    E = rng.standard_normal((3, 3)) * 0.4
    tau = np.array([[0.0, 0.0, 0.0], (p.A @ [0.25, 0.25, 0.25])])
    p.polar = dict(positions=tau, born=Z, dielectric=9.0 * np.eye(3) + E @ E.T,
                   volume=abs(np.linalg.det(p.A)), coarse_mesh=(3, 3, 3))
    return p


def test_polar_in_use_rule():
    eps = np.eye(3)
    assert P.polar_in_use(eps, 2) and not P.polar_in_use(eps, 1)
    assert not P.polar_in_use(np.zeros((3, 3)), 2)
    assert not P.polar_in_use(np.full((3, 3), 1e-6), 2)  # squared norm 9e-12 <= 1e-10


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-6)])
def test_coupling_with_polar_vs_oracle(precision, tol):
    say("calcCouplingSquared with the polar term", precision)
    p = polar_problem()
    pol = p.polar
    gW = p.gw_host()
    rng = np.random.default_rng(21)
    inter = P.InteractionElPhWan(gW, p.Re, p.degE, p.Rp, p.degP, precision)
    inter.setPolar(pol["positions"], pol["born"], pol["dielectric"], p.recip, pol["volume"],
                   pol["coarse_mesh"], num_species=2)
    assert inter.usePolarCorrection
    k1 = p.recip @ rng.uniform(0, 1, 3)
    _, U1 = O.el_diag(p, k1)
    U1 = U1[:, 1:3]
    inter.cacheElPh(U1, k1)
    nq = 29
    q3s = [p.recip @ rng.uniform(-1, 1, 3) for _ in range(nq - 1)] + [np.zeros(3)]
    U2s = [O.el_diag(p, k1 + q)[1][:, : 1 + j % p.nw] for j, q in enumerate(q3s)]
    e3s = [O.ph_diag(p, q)[1] for q in q3s]
    # Part1 on the GPU against the oracle
    xg = inter.polarCorrectionPart1Batch(q3s, e3s)
    xo = np.stack([O.polar_part1(pol, p.recip, q, e) for q, e in zip(q3s, e3s)])
    perr = np.abs(xg - xo).max() / np.abs(xo).max()
    say("  Part1 max err / max |x|", perr)
    assert perr < 1e-12
    inter.calcCouplingSquared(U1, U2s, e3s, q3s, list(xg))
    cached = O.cache_elph(p.nw, p.nm, p.Rp, p.Re, p.degE, gW, U1, k1)
    ref = O.coupling_sq_polar(p.nw, p.nm, p.Rp, p.degP, cached, U2s, e3s, q3s, U1, xo)
    plain = O.coupling_sq(p.nw, p.nm, p.Rp, p.degP, cached, U2s, e3s, q3s)
    worst, effect = 0.0, 0.0
    for j in range(nq):
        got = inter.getCouplingSquared(j)
        worst = max(worst, np.abs(got - ref[j]).max() / np.abs(ref[j]).max())
        effect = max(effect, np.abs(plain[j] - ref[j]).max() / np.abs(ref[j]).max())
    say("  max |g|^2 error / max |g|^2 =", worst, " polar effect", effect)
    assert effect > 1e-3
    assert worst < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-6)])
def test_linewidths_with_polar_vs_oracle(precision, tol):
    say("linewidths with the polar term", precision)
    p = polar_problem()
    pol = p.polar
    gW = p.gw_host()
    ref = O.Oracle(p, gW.reshape(-1)).linewidths()
    q = dataclasses.replace(p)  # the same system without the term
    plain = O.Oracle(q, gW.reshape(-1)).linewidths()
    lw = P.ElectronLinewidths(p, gW=gW, precision=precision, transform="direct", k1_batch=7)
    lw.set_polar(pol["positions"], pol["born"], pol["dielectric"], pol["volume"], pol["coarse_mesh"])
    lw.run()
    err = rel_err(lw.linewidths(), ref)
    effect = rel_err(plain, ref).max()
    say("  states", lw.num_states, "max rel err", err.max(), " polar effect", effect)
    assert effect > 1e-3
    assert err.max() < tol
    lw.set_polar()  # off again: the plain linewidths
    lw.reset()
    lw.run()
    assert rel_err(lw.linewidths(), plain).max() < tol
    s = P.ElectronLinewidths(p, gW=gW, precision=precision, transform="separable")
    with pytest.raises(P.PhxError):
        s.set_polar(pol["positions"], pol["born"], pol["dielectric"], pol["volume"], pol["coarse_mesh"])
