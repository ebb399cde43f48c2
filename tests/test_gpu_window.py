"""GPU window selection and degeneracy averaging on closed-form band models
(diagonal H(R): flat bands held exactly, nearest-neighbour cosine bands),
against the closed forms and the oracle."""
import dataclasses
import sys

import numpy as np
import pytest

import phoebe_amd as P
from oracle import oracle as O
from tests.helpers import rel_err
from tests.test_oracle_window_average import _closed_form_energies, _diagonal_model

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_window_bounds_exact_on_gpu():
    """Flat bands exactly on both window bounds stay out (strict
    inequalities); one ulp wider they join."""
    mesh = (5, 4, 6)
    p, eps, hop, nn = _diagonal_model(mesh)
    p = dataclasses.replace(p, win_min=-0.3, win_max=0.25)
    E = _closed_form_energies(p, eps, hop, nn)
    passing = (E > p.win_min) & (E < p.win_max)
    active = np.flatnonzero(passing.any(axis=1))
    lw = P.ElectronLinewidths(p, gW=p.gw_host(), precision="fp64")
    say("window: active", lw.num_active, "of", np.prod(mesh))
    assert np.array_equal(lw.active, active)
    assert np.array_equal(lw.band_lo, [np.flatnonzero(passing[i])[0] for i in active])
    assert np.array_equal(lw.band_hi, [np.flatnonzero(passing[i])[-1] for i in active])
    p2 = dataclasses.replace(p, win_min=np.nextafter(-0.3, -1.0), win_max=np.nextafter(0.25, 1.0))
    lw2 = P.ElectronLinewidths(p2, gW=p2.gw_host(), precision="fp64")
    assert lw2.num_active == np.prod(mesh) and set(lw2.band_lo) == {0}


@pytest.mark.parametrize("form", ["fp64", P.CHOSEN])
def test_degeneracy_average_chain_on_gpu(form):
    """Three flat bands in a non-transitive 1e-6 Ry chain and a cosine band:
    the GPU's averaged linewidths (row 0 averaged band by band, row 1 left
    alone) match the oracle's."""
    say("averaging chain", form)
    p, _, _, _ = _diagonal_model((4, 4, 4), eps=[0.0, 0.6e-6, 1.2e-6, 0.0], hop=[0.0, 0.0, 0.0, 0.01])
    p.temperatures = np.array([p.temperatures[0], 2.0 * p.temperatures[0]])
    p.chem_pots = np.array([0.0, 0.0])
    p.win_min, p.win_max = -0.2, 0.2
    gW = p.gw_host()
    orc = O.Oracle(p, gW.reshape(-1))
    ref = orc.linewidths(average=True)
    raw = orc.linewidths(average=False)
    lw = P.linewidths_form(p, form, gW=gW)
    lw.run()
    got = lw.linewidths(average=True)
    tol = 1e-10 if form == "fp64" else 1e-6
    err = rel_err(got, ref)
    say("  max rel err", err.max())
    assert err.max() < tol
    assert np.array_equal(got[1], lw.linewidths(average=False)[1])
    # the averaging changed row 0 (linewidths here are ~1e-15: no absolute tolerance)
    assert not np.allclose(raw[0], ref[0], rtol=1e-9, atol=0.0)
