"""The synthetic configurations of the measured workloads at their full sizes,
on the GPU against the FP64 oracle.

* cfg0 (Nw=8, 6 modes, WS(7^3) / WS(6^3), 12^3 meshes): every active outer
  point, plain FP64 at 1e-10 and the chosen form at 1e-6; window, q-set and
  pair count bit-exact.
* cfg2 (Nw=16, 12 modes, WS(10^3) / WS(9^3), 40^3): plain FP64 against the
  oracle on sampled outer points at 1e-10; the chosen form against plain FP64
  on two whole mesh lines at 1e-6.
* cfg4 (Nw=16, 18 modes, 60^3): plain FP64 against the oracle on a sampled
  outer point; the learned q-block screen against the exact screened run.
* cfg3 (uneven materials): 16 materials of the uneven draw, each against the
  oracle on sampled outer points at 1e-6.

The oracle's inputs (band structure, phonons, gW) are the same seeded
synthetic system; gW is copied back from the device or drawn by the
library's bit-identical host twin.
"""
import gc
import sys
import time

import numpy as np
import pytest

import phoebe_amd as P
from oracle import oracle as O
from phoebe_amd import synth
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def states_of(lw, k1):
    """State indices of the outer points k1 (active indices)."""
    nb = lw.band_hi - lw.band_lo + 1
    start = np.concatenate([[0], np.cumsum(nb)])
    return np.concatenate([np.arange(start[a], start[a + 1]) for a in k1])


def line_points(lw, m0, count, seed):
    line = np.asarray(lw.active) // m0
    pick = np.random.default_rng(seed).permutation(np.unique(line))[:count]
    return np.concatenate([np.flatnonzero(line == l) for l in pick]).astype(np.int32)


def check_window(lw, orc):
    assert np.array_equal(lw.active, orc.active)
    assert np.array_equal(lw.band_lo, orc.band_lo) and np.array_equal(lw.band_hi, orc.band_hi)
    _, _, set_o = orc.phonons()
    _, set_g = lw.phonons()
    assert np.array_equal(set_g, set_o)


@pytest.mark.timeout(1200)
def test_cfg0_full_size_vs_oracle():
    t0 = time.perf_counter()
    p = synth.config("cfg0")
    gw = P.synth_gw_host(p)
    orc = O.Oracle(p, gw)
    say(f"cfg0: nRe {p.nRe} nRp {p.nRp}, {orc.num_active} active points, {orc.num_states} states "
        f"({time.perf_counter() - t0:.1f}s)")
    ref = orc.linewidths()
    say(f"  oracle: {orc.last_pairs} pairs ({time.perf_counter() - t0:.1f}s)")
    for form, tol in (("fp64", 1e-10), (P.CHOSEN, 1e-6)):
        lw = P.linewidths_form(p, form)
        check_window(lw, orc)
        assert lw.run() == orc.last_pairs
        err = rel_err(lw.linewidths(), ref)
        say(f"  {form}: max rel err {err.max():.3e}")
        assert np.all(ref > 0)
        assert err.max() < tol
        del lw
        gc.collect()


@pytest.mark.timeout(1200)
def test_cfg2_full_size_vs_oracle():
    t0 = time.perf_counter()
    p = synth.config("cfg2")
    lw64 = P.linewidths_form(p, "fp64")
    say(f"cfg2: {lw64.num_active} active points ({time.perf_counter() - t0:.1f}s)")
    gw = lw64.copy_gw()
    orc = O.Oracle(p, gw)
    say(f"  oracle set-up ({time.perf_counter() - t0:.1f}s)")
    check_window(lw64, orc)
    k1 = np.sort(np.random.default_rng(202).choice(lw64.num_active, 2, replace=False)).astype(np.int32)
    ref = orc.linewidths(k1_list=k1, average=False)
    say(f"  oracle: {orc.last_pairs} pairs ({time.perf_counter() - t0:.1f}s)")
    assert lw64.run(k1_list=k1) == orc.last_pairs
    idx = states_of(lw64, k1)
    err = rel_err(lw64.linewidths(average=False)[:, idx], ref[:, idx])
    say(f"  fp64 vs oracle on {len(idx)} states: max rel err {err.max():.3e}")
    assert np.all(ref[:, idx] > 0) and err.max() < 1e-10
    del gw, orc
    gc.collect()
    # the chosen form against plain FP64 on two whole mesh lines
    pts = line_points(lw64, p.mesh[0], 2, seed=12345)
    lw64.reset()
    lw64.run(k1_list=pts)
    g64 = lw64.linewidths(average=False)
    del lw64
    gc.collect()
    lwc = P.linewidths_form(p, P.CHOSEN)
    lwc.run(k1_list=pts)
    gc_ = lwc.linewidths(average=False)
    idx = states_of(lwc, pts)
    err = rel_err(gc_[:, idx], g64[:, idx])
    say(f"  {P.CHOSEN} vs fp64 on {len(pts)} outer points, {len(idx)} states: max rel err {err.max():.3e} "
        f"({time.perf_counter() - t0:.1f}s)")
    assert np.all(g64[:, idx] > 0) and err.max() < 1e-6


@pytest.mark.timeout(1200)
def test_cfg4_system_vs_oracle_and_learned_screen():
    t0 = time.perf_counter()
    p = synth.config("cfg4")
    lw64 = P.linewidths_form(p, "fp64")
    say(f"cfg4: {p.nm} modes, {lw64.num_active} active points ({time.perf_counter() - t0:.1f}s)")
    gw = lw64.copy_gw()
    orc = O.Oracle(p, gw)
    say(f"  oracle set-up ({time.perf_counter() - t0:.1f}s)")
    check_window(lw64, orc)
    k1 = np.array([int(np.random.default_rng(404).integers(lw64.num_active))], dtype=np.int32)
    ref = orc.linewidths(k1_list=k1, average=False)
    assert lw64.run(k1_list=k1) == orc.last_pairs
    idx = states_of(lw64, k1)
    err = rel_err(lw64.linewidths(average=False)[:, idx], ref[:, idx])
    say(f"  fp64 vs oracle on {orc.last_pairs} pairs: max rel err {err.max():.3e} ({time.perf_counter() - t0:.1f}s)")
    assert np.all(ref[:, idx] > 0) and err.max() < 1e-10
    del gw, orc, lw64
    gc.collect()
    # learned screen: trained on other lines, checked against the exact screened run
    lw = P.linewidths_form(p, P.CHOSEN)
    ev = line_points(lw, p.mesh[0], 2, seed=7)
    tr = line_points(lw, p.mesh[0], 6, seed=8)
    tr = np.setdiff1d(tr, ev).astype(np.int32)
    lw.run(k1_list=ev)
    exact = lw.linewidths(average=False)
    lw.reset()
    ns = lw.screen_collect(k1_list=tr, sample_frac=0.02, seed=3)
    before, after = lw.screen_train(epochs=200, batch=512, lr=2e-3, seed=5)
    say(f"  learned screen: {ns} samples, loss {before:.3f} -> {after:.3f}")
    assert ns > 0 and after < before
    thr = 1e-8
    lw.set_learned_screen(True, threshold=thr, audit_frac=0.01, seed=9)
    lw.screen_reset_stats()
    lw.run(k1_list=ev)
    got = lw.linewidths(average=False)
    st = lw.screen_stats()
    idx = states_of(lw, ev)
    err = rel_err(got[:, idx], exact[:, idx])
    say(f"  stats {st}, max rel err vs exact {err.max():.3e} ({time.perf_counter() - t0:.1f}s)")
    assert st["blocks_in_screen"] > 0
    # a skipped block weighs below thr of its panel's predicted total; with the
    # audit's factor of 10 the error per linewidth stays within 1e-6
    assert err.max() < 1e-6


@pytest.mark.timeout(1200)
def test_cfg3_materials_vs_oracle():
    t0 = time.perf_counter()
    mats = synth.uneven_materials(16, seed=1003, max_gw_bytes=1.0e9)
    say(f"cfg3: {len(mats)} materials ({time.perf_counter() - t0:.1f}s)")
    forms = set()
    for i, p in enumerate(mats):
        gw = P.synth_gw_host(p)
        orc = O.Oracle(p, gw)
        lw = P.linewidths_auto(p)
        forms.add(lw.transform)
        check_window(lw, orc)
        rng = np.random.default_rng(300 + i)
        k1 = np.sort(rng.choice(lw.num_active, min(2, lw.num_active), replace=False)).astype(np.int32)
        ref = orc.linewidths(k1_list=k1, average=False)
        assert lw.run(k1_list=k1) == orc.last_pairs
        idx = states_of(lw, k1)
        got = lw.linewidths(average=False)
        err = rel_err(got[:, idx], ref[:, idx])
        say(f"  material {i}: Nw {p.nw}, {p.nm} modes, R {p.nRe}/{p.nRp}, mesh {p.mesh[0]}^3, "
            f"{lw.transform}, {len(idx)} states, max rel err {err.max():.3e} ({time.perf_counter() - t0:.1f}s)")
        assert np.all(ref[:, idx] > 0) and err.max() < 1e-6
        del lw, orc, gw
        gc.collect()
    say(f"  transforms used: {sorted(forms)}")
