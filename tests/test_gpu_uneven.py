"""Uneven-material batching: prepared sweeps replayed as hipGraphs, several
materials concurrently on their own streams, against per-material sequential
runs (same kernels, so the linewidths are bit-identical), and the autotuner."""
import sys
import time

import numpy as np
import pytest

import phoebe_amd as P
from phoebe_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _handle(p):
    lw = None
    if p.nm * p.nw <= 192:
        try:  # the separable transform needs nm <= 32 and <= 8 windowed bands per point
            lw = P.ElectronLinewidths(p, gW=None, precision="fp32", transform="separable")
            lw.set_screen(6.0)
        except P.PhxError:
            lw = None
    if lw is None:
        lw = P.ElectronLinewidths(p, gW=None, precision="fp32", transform="direct")
    lw.set_line_batching(4)
    return lw


def test_graph_sweeps_and_concurrent_materials_match_sequential_runs():
    mats = synth.uneven_materials(5, seed=77, max_gw_bytes=4e7, max_mesh=10)
    say("materials", [(p.nw, p.nm, p.nRe, p.nRp, p.mesh[0]) for p in mats])
    seq, handles = [], []
    t_seq = 0.0
    for p in mats:
        lw = _handle(p)
        lw.sync()
        t0 = time.perf_counter()
        lw.run()
        lw.sync()
        t_seq += time.perf_counter() - t0
        seq.append(lw.linewidths(average=False))
        lw.reset()
        handles.append(lw)
    for lw in handles:
        assert lw.plan() >= 1
    # graph replay per material
    for lw, ref in zip(handles, seq):
        lw.plan_run(graph=True)
        lw.sync()
        assert np.array_equal(lw.linewidths(average=False), ref)
        lw.reset()
    # all materials at once, each on its own stream
    t0 = time.perf_counter()
    pairs = sum(lw.plan_run(graph=True) for lw in handles)
    for lw in handles:
        lw.sync()
    t_con = time.perf_counter() - t0
    for lw, ref in zip(handles, seq):
        assert np.array_equal(lw.linewidths(average=False), ref)
    say(f"pairs {pairs}: sequential {t_seq * 1e3:.1f} ms, concurrent graphs {t_con * 1e3:.1f} ms")
    # plan without graph gives the same too
    lw = handles[0]
    lw.reset()
    lw.plan_run(graph=False)
    lw.sync()
    assert np.array_equal(lw.linewidths(average=False), seq[0])


def test_autotuner_keeps_results():
    p = synth.make_problem(2, 6, 3, 3, 8, seed=91, name="tune")
    lw = P.ElectronLinewidths(p, gW=p.gw_host(), precision="fp64", transform="separable")
    lw.set_line_batching(4)
    lw.run()
    ref = lw.linewidths(average=False)
    lw.reset()
    k1 = np.arange(min(lw.num_active, 64), dtype=np.int32)
    best = lw.autotune(k1)
    say("autotuned", best)
    assert best["k1_batch"] in (8, 16, 32) and best["lines_per_batch"] in (1, 2, 4, 8, 16)
    assert np.array_equal(lw.linewidths(average=False), np.zeros_like(ref))  # restored
    lw.run()
    got = lw.linewidths(average=False)
    m = ref != 0
    assert np.abs(got[m] - ref[m]).max() <= 1e-12 * np.abs(ref[m]).max()
