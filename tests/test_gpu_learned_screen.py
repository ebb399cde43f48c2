"""Learned q-block screen: the GPU forward / backward kernels against the CPU
restatement, what the trained network gains over the exact screened path, and
the audit / fallback path."""
import sys

import numpy as np
import pytest

import phoebe_amd as P
from oracle import mlp_oracle as M
from phoebe_amd import synth
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [1, 37, 300, 1024])
def test_mlp_gradients_match_cpu_oracle(n):
    say("mlp gradients n =", n)
    rng = np.random.default_rng(n)
    params = M.random_params(11 + n).astype(np.float32)
    X = rng.standard_normal((n, M.IN)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    loss, g, y = P.mlp_gradients(params, X, t)
    loss_o, g_o, y_o = M.loss_and_grads(params, X, t, rnd=True)
    assert abs(loss - loss_o) <= 1e-3 * max(loss_o, 1e-6)
    assert np.abs(y - y_o).max() <= 2e-3 * max(np.abs(y_o).max(), 1e-6)
    for name, a, b in (("W1", "W1", "b1"), ("b1", "b1", "W2"), ("W2", "W2", "b2"), ("b2", "b2", "W3"),
                       ("W3", "W3", "b3")):
        ga, gb = g[M.OFF[a]:M.OFF[b]], g_o[M.OFF[a]:M.OFF[b]]
        err = np.abs(ga - gb).max() / max(np.abs(gb).max(), 1e-12)
        say(f"  {name}: max |g| {np.abs(gb).max():.3e}, rel err {err:.2e}")
        assert err <= 5e-3, name
    assert abs(g[M.OFF["b3"]] - g_o[M.OFF["b3"]]) <= 1e-4 * max(abs(g_o[M.OFF["b3"]]), 1e-6)


def _problem():
    return synth.make_problem(2, 6, 3, 3, (12, 12, 12), seed=61, name="screen")


def _handle(p, gW, precision="fp64", split=True):
    lw = P.ElectronLinewidths(p, gW=gW, precision=precision, transform="separable")
    lw.set_screen(7.0)
    lw.set_line_batching(1)
    lw.set_split_last_stage(split)  # FP32: x-stage and lean epilogue kernels, or the fused kernel
    return lw


def test_trained_screen_skips_blocks_within_its_error_bound():
    p = _problem()
    gW = p.gw_host()
    exact = _handle(p, gW)
    exact.run()
    g0 = exact.linewidths(average=False)
    lw = _handle(p, gW)
    ns = lw.screen_collect(sample_frac=0.2, seed=3)
    say("training samples", ns)
    assert ns >= 500
    before, after = lw.screen_train(epochs=300, batch=512, lr=2e-3, seed=5)
    say(f"standardized MSE {before:.3f} -> {after:.3f}")
    assert after < 0.5 * before
    assert np.array_equal(lw.linewidths(average=False), np.zeros_like(g0))  # collecting adds nothing
    thr = 1e-6
    lw.set_learned_screen(True, threshold=thr, audit_frac=0.05, seed=9)
    lw.run()
    st = lw.screen_stats()
    g1 = lw.linewidths(average=False)
    m = g0 != 0
    err = rel_err(g1[m], g0[m])
    gain = st["learned_skips"] / max(st["blocks_in_screen"], 1)
    say(f"stats {st}; skipped {gain:.1%} of the blocks inside the energy screen; max rel err {err.max():.2e}")
    assert st["blocks_in_screen"] > 0
    assert gain >= 0.05  # what the trained network saves over the exact screened path
    assert st["fallbacks"] == 0
    # every skipped block was predicted below thr x its panel's total: with at most
    # m1*m2 blocks per panel the linewidth error stays below that many thresholds
    assert err.max() <= thr * 144
    assert np.array_equal(g1 == 0, g0 == 0)


@pytest.mark.parametrize("precision,split", [("fp64", True), ("fp32", True), ("fp32", 3), ("fp32", False)])
def test_failed_audit_falls_back_to_the_exact_blocks(precision, split):
    """The three passes of a screened batch (computed blocks, audited blocks,
    the gated fallback over the skipped ones) reproduce the exact screened
    run, in the fused kernel and in the split x-stage + epilogue kernels
    (split mode 3: the electron rotation in the x-stage kernel)."""
    say("failed audit", precision, f"split {int(split)}" if split else "fused")
    p = _problem()
    gW = p.gw_host()
    exact = _handle(p, gW, precision, split)
    exact.run()
    g0 = exact.linewidths(average=False)
    lw = _handle(p, gW, precision, split)
    # y = 100 relu(f11) - 300 with f11 > 0 only for blocks with a row within 3 sigma:
    # every prediction is absurdly small, so the skipped tail blocks' exact
    # contributions dwarf the threshold and the audit must fail
    bad = np.zeros(M.NPARAMS, dtype=np.float32)
    bad[M.OFF["W1"] + 11 * M.HID + 0] = 1.0
    bad[M.OFF["W2"] + 0 * M.HID + 0] = 1.0
    bad[M.OFF["W3"] + 0] = 100.0
    bad[M.OFF["b3"]] = -300.0
    lw.screen_params(bad)
    lw.set_learned_screen(True, threshold=1e-8, audit_frac=0.2, seed=1)
    lw.run()
    st = lw.screen_stats()
    say("stats", st)
    assert st["learned_skips"] > 0
    assert st["fallbacks"] >= 1
    g1 = lw.linewidths(average=False)
    m = g0 != 0
    assert rel_err(g1[m], g0[m]).max() < 1e-12


def test_learned_screen_split_equals_fused():
    """FP32 with the trained network: the split last stage (which takes the
    screen's slot masks and gate), with the electron rotation in the epilogue
    or in the x-stage kernel, skips the same blocks and gives the fused
    kernel's linewidths."""
    p = _problem()
    gW = p.gw_host()
    lw = _handle(p, gW, "fp32", True)
    ns = lw.screen_collect(sample_frac=0.2, seed=3)
    lw.screen_train(epochs=200, batch=512, lr=2e-3, seed=5)
    params = lw.screen_params()
    out, stats = [], []
    for split in (1, 3, 0):
        h = _handle(p, gW, "fp32", split)
        h.screen_params(params)
        h.set_learned_screen(True, threshold=1e-6, audit_frac=0.05, seed=9)
        h.run()
        out.append(h.linewidths(average=False))
        st = h.screen_stats()
        stats.append((st["blocks_in_screen"], st["learned_skips"], st["audited"], st["fallbacks"]))
    say(f"{ns} samples; split 1, split 3, fused stats {stats}")
    assert stats[0] == stats[1] == stats[2] and stats[0][1] > 0
    nz = out[2] != 0
    for got in out[:2]:
        assert np.array_equal(got != 0, nz)
        assert rel_err(got[nz], out[2][nz]).max() < 1e-12


def test_learned_screen_argument_checks():
    p = synth.make_problem(1, 4, 2, 2, 4, seed=34, name="tiny")
    lw = P.ElectronLinewidths(p, gW=p.gw_host(), precision="fp64", transform="separable")
    with pytest.raises(P.PhxError):
        lw.screen_collect()                 # energy screen off
    lw.set_screen(7.0)
    with pytest.raises(P.PhxError):
        lw.set_learned_screen(True)         # no trained network
    with pytest.raises(P.PhxError):
        lw.screen_train()                   # no samples
