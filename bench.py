#!/usr/bin/env python
"""el-ph (k,q) pairs/s on one MI355X: chosen form vs plain FP64.

A step is one mesh line of outer points k1 (the active points with equal j1,
j2, lines drawn from a seeded permutation) scattered against every active
inner point k2: cacheElPh + calcCouplingSquared + symmetrizeCoupling +
Gaussian-delta linewidth reduction, with all inputs already resident in HBM.
The plain FP64 form processes the same outer points in its own k1-batched
GEMM batches of --k1-batch points.
The workload is the Nw=16, 12-mode, 40^3 k/q mesh system (R_e on WS(10^3),
R_ph on WS(9^3)); band structures, phonons and the coupling tensor are built
on the device before the timed region.

Prints one JSON line.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from phoebe_amd import CHOSEN, FORMS  # noqa: E402  (the library itself loads lazily)

METRIC = "el-ph (k,q) pairs/sec, chosen vs plain FP64 on 1 GPU, at ≤1e-6 linewidth rel err"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3.0}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)




def run_form(P, prob, form, make_batches, warmup, k1_batch):
    t0 = time.perf_counter()
    precision, transform, screen, line_min = P.FORMS[form]
    lw = P.linewidths_form(prob, form, k1_batch=k1_batch)
    log(f"[{form}] setup {time.perf_counter() - t0:.1f}s: active {lw.num_active}, "
        f"states {lw.num_states}")
    k1_batches, warm_batches = make_batches(lw.active, line_min > 0, warmup)
    for b in warm_batches:
        lw.run(k1_list=b)
    lw.sync()
    lw.reset()
    lw.set_timing(True)
    lw.sync()
    t0 = time.perf_counter()
    pairs = 0
    for b in k1_batches:
        pairs += lw.run(k1_list=b)
    lw.sync()
    elapsed = time.perf_counter() - t0
    kt = lw.kernel_times()
    st = lw.stage_times()
    sst = lw.split_times() if transform == "separable" else None
    if sst is not None:
        sst.update(lw.split_bytes())
        sst["fused_y"] = lw.last_stage_fused_y()
    gam = lw.linewidths(average=False)
    log(f"[{form}] {len(k1_batches)} run calls in {elapsed:.3f}s, {pairs} pairs, "
        f"{pairs / elapsed:.4g} pairs/s; kernels {kt}")
    nb = lw.band_hi - lw.band_lo + 1
    k1all = np.concatenate(k1_batches)
    return lw, dict(pairs=pairs, seconds=elapsed, kernels=kt, stage=st, split=sst, linewidths=gam,
                    k1=int(sum(len(b) for b in k1_batches)), active=lw.num_active,
                    ncol=lw.prob.nm * lw.prob.nw, nm=lw.prob.nm, nb1_mean=float(nb[k1all].mean()),
                    nb2_mean=float(nb.mean()))


def traffic_per_dispatch(kernel, workload):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_dispatch:"):
            return float(line.split(":", 1)[1])
    return None


def roofline(res, form, workload):
    """Dominant kernel of the form: k_coupling (direct), on the FLOPs of its
    transform GEMMs against the MFMA peak; the slower of k_sep_xg and
    k_sep_ep2 (split separable last stage), on the bytes it streams against
    HBM peak; k_sep_xm (fused separable last stage), on the FLOPs of its
    x-stage GEMMs over the rows it computed."""
    precision, transform = FORMS[form][:2]
    peak = PEAK_TFLOPS[precision]
    if transform == "direct":
        kt = res["kernels"]
        ms, launches, flops = kt["coupling_ms"], max(kt["batches"], 1), kt["coupling_flops"]
        kern = "k_coupling"
    else:
        st, sp = res["stage"], res.get("split")
        if sp and sp["launches"] > 0:
            # split last stage: both kernels stream HBM.  The x-stage kernel reads
            # the x-stage slabs of the slots it computes and writes the kept rows'
            # g3 (complex FP32), or g' (complex FP64 per kept band) after reading
            # U2 when it does the electron rotation; the epilogue reads its output
            # back (and the per-pair reduction bytes are reported beside)
            n = sp["launches"]
            t_xg, t_ep = sp["xg_ms"] / n * 1e-3, sp["ep_ms"] / n * 1e-3
            g3 = sp["ep_bytes"] / n
            xg_bytes = sp["xg_bytes"] / n
            # per computed (k1, k2) pair the reduction reads |g|^2 (8 nb1 nb2 nm bytes), the
            # inner point's occupations (16 nb2) and the phonon data (16 nm); a computed
            # row is one (k1, b1, k2) triple, so pairs = rows / nb1
            nb1, nb2, nm = res["nb1_mean"], res["nb2_mean"], res["nm"]
            pair_bytes = 8.0 * nb1 * nb2 * nm + 16.0 * nb2 + 16.0 * nm
            pairs = st["rows"] / nb1 / n
            # with the y-stage inside the x-stage kernel (k_sep_yx) it reads T1 instead of the slabs
            xk = "k_sep_yx" if sp.get("fused_y") else "k_sep_xg"
            xg = {"kernel": xk, "avg_launch_ms": round(t_xg * 1e3, 4), "bytes_per_launch": xg_bytes,
                  "achieved_GBps": round(xg_bytes / t_xg / 1e9, 1),
                  "frac": round(xg_bytes / t_xg / 1e9 / HBM_PEAK_GBS, 4),
                  "traffic": traffic_per_dispatch(xk, workload),
                  "mfma_TFLOPs": round(st["flops"] / n / t_xg / 1e12, 3)}
            ep = {"kernel": "k_sep_ep2", "avg_launch_ms": round(t_ep * 1e3, 4), "bytes_per_launch": g3,
                  "achieved_GBps": round(g3 / t_ep / 1e9, 1), "frac": round(g3 / t_ep / 1e9 / HBM_PEAK_GBS, 4),
                  "traffic": traffic_per_dispatch("k_sep_ep2", workload),
                  "pair_bytes": round(pair_bytes, 1), "pairs_per_launch": pairs,
                  "frac_pair_bytes": round(pair_bytes * pairs / t_ep / 1e9 / HBM_PEAK_GBS, 4)}
            dom = xg if t_xg >= t_ep else ep
            return {"bound": "hbm", "kernel": dom["kernel"], "achieved": dom["achieved_GBps"],
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dom["frac"], "traffic": dom["traffic"],
                    "avg_launch_ms": dom["avg_launch_ms"], "launches": n,
                    "bytes_per_launch": dom["bytes_per_launch"],
                    "flop_model_note": "the per-pair FLOP model of the direct R_ph transform "
                                       "(8 N_Rph nm nb1 Nw + rotations) does not apply: the separable "
                                       "transform computes the same g3 with about 26x fewer operations",
                    xk: xg, "k_sep_ep2": ep}
        else:
            ms, launches, flops = st["ms"], max(st["launches"], 1), st["flops"]
            kern = "k_sep_xm"
    ach = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    return {"bound": "mfma", "kernel": kern, "achieved": round(ach, 3), "peak": peak,
            "unit": "TFLOP/s", "frac": round(ach / peak, 4),
            "traffic": traffic_per_dispatch(kern, workload),
            "avg_launch_ms": round(ms / launches, 4), "launches": launches,
            "flops_per_launch": flops / launches}


def cpu_baseline(lw64, prob, k1_batches, gam64, k1_count):
    """Oracle (FP64 restatement of the reference CPU path) on the host cores."""
    from oracle import oracle as O

    t0 = time.perf_counter()
    gw = lw64.copy_gw()
    log(f"[cpu] gW copied to host ({gw.nbytes / 1e9:.1f} GB) in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    orc = O.Oracle(prob, gw)
    log(f"[cpu] oracle setup {time.perf_counter() - t0:.1f}s")
    if not (np.array_equal(orc.active, lw64.active) and np.array_equal(orc.band_lo, lw64.band_lo)
            and np.array_equal(orc.band_hi, lw64.band_hi)):
        raise RuntimeError("window selection differs between oracle and GPU")
    k1s = list(k1_batches[0][:k1_count])
    times = []
    for _ in range(3):  # median of three runs
        t0 = time.perf_counter()
        ref = orc.linewidths(k1_list=k1s, average=False)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    pairs = orc.last_pairs
    # compare the sampled states with the GPU FP64 run
    nb = lw64.band_hi - lw64.band_lo + 1
    start = np.concatenate([[0], np.cumsum(nb)])
    idx = np.concatenate([np.arange(start[a], start[a + 1]) for a in k1s])
    err = np.abs(gam64[:, idx] - ref[:, idx]) / np.abs(ref[:, idx])
    log(f"[cpu] {pairs} pairs in {dt:.1f}s; GPU fp64 vs oracle max rel err {err.max():.3e}")
    del gw, orc
    return {"value": pairs / dt, "unit": "(k,q) pairs/s", "cores": int(O.lib().orc_num_threads()),
            "kind": "port",
            "sample": f"{len(k1s)} outer k1 x all {pairs // len(k1s)} inner k2, cfg2 inputs, "
                      "cacheElPh GEMV + g3/g4/gFinal loops + serial reduction (FP64 C++ restatement), "
                      "median of 3 runs",
            "seconds": round(dt, 2), "seconds_runs": [round(x, 3) for x in times],
            "gpu_fp64_vs_oracle_max_rel_err": float(err.max())}


def eig_main(args):
    """Batched Bloch build + Hermitian eigensolve (2^17 random points per family:
    H(k) 8x8 and 16x16, D(q) 6x6 and 18x18 over WS(8^3) R vectors), FP64, inputs
    resident in HBM, each kernel timed with HIP events.  Prints one JSON line."""
    import scipy.linalg.lapack as lapack

    import phoebe_amd as P
    from phoebe_amd import synth

    fams, total_m, total_s = {}, 0, 0.0
    jac_flops, jac_s = 0.0, 0.0
    for kind, n in (("el", 8), ("el", 16), ("ph", 6), ("ph", 18)):
        f = synth.eig_family(kind, n)
        ks = (f["recip"] @ np.random.default_rng(11 + n).uniform(0.0, 1.0, (3, args.points))).T
        W, V, mb, mj = P.bloch_diag_timed(kind, ks, f["R"], f["weight"], f["coef"], f["masses"],
                                          reps=max(args.steps, 1))
        # spot check against LAPACK on 32 points
        worst = 0.0
        for i in np.random.default_rng(1).choice(args.points, 32, replace=False):
            if kind == "el":
                Hk = np.tensordot(np.exp(1j * (f["R"] @ ks[i])) / f["weight"], f["coef"], axes=(0, 0))
                wl = lapack.zheev(Hk, compute_v=0, lower=1)[0]
            else:
                D = np.tensordot(np.exp(-1j * (f["R"] @ ks[i])) * f["weight"], f["coef"], axes=(0, 0))
                D = 0.5 * (D + D.conj().T) / np.sqrt(np.outer(f["masses"], f["masses"]))
                lam = lapack.zheev(D, compute_v=0, lower=1)[0]
                wl = np.sign(lam) * np.sqrt(np.abs(lam))
            worst = max(worst, float(np.abs(W[i] - wl).max() / max(1.0, np.abs(wl).max())))
        if worst > 1e-10:
            raise RuntimeError(f"{kind} {n}: eigenvalues differ from LAPACK by {worst:.2e}")
        nR = len(f["R"])
        ft = 8.0 * nR * n * n * args.points
        eig = 23.0 * n ** 3 * args.points
        t = (mb + mj) * 1e-3
        fams[f"{kind}{n}"] = {"N": n, "R_vectors": nR, "matrices": args.points,
                              "bloch_ms": round(mb, 4), "eig_ms": round(mj, 4),
                              "matrices_per_s": args.points / t,
                              "eig_matrices_per_s": args.points / (mj * 1e-3),
                              "bloch_TFLOPs": round(ft / (mb * 1e-3) / 1e12, 3),
                              "eig_TFLOPs_model": round(eig / (mj * 1e-3) / 1e12, 3),
                              "eig_frac_fp64_peak": round(eig / (mj * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"], 5),
                              "max_eig_rel_err_vs_lapack": worst}
        total_m += args.points
        total_s += t
        jac_flops += eig
        jac_s += mj * 1e-3
        log(f"[cfg1] {kind} {n}: bloch {mb:.3f} ms, eig {mj:.3f} ms, {args.points / t:.4g} matrices/s, "
            f"err {worst:.1e}")
    ach = jac_flops / jac_s / 1e12
    out = {"metric": "matrices/s, batched Bloch build + Hermitian eigensolve (FP64) on 1 GPU",
           "value": total_m / total_s, "unit": "matrices/s", "n_gpus": 1, "steps": max(args.steps, 1),
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"cfg1: {args.points} random points per family, H(k) 8x8 and 16x16, "
                                  "D(q) 6x6 and 18x18, R on WS(8^3)"},
           "families": fams,
           "roofline": {"bound": "mfma", "kernel": "k_jacobi", "achieved": round(ach, 4),
                        "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                        "frac": round(ach / PEAK_TFLOPS["fp64"], 5), "traffic": None,
                        "flop_model": "23 N^3 per matrix (Householder tridiagonalization, back-transform, QL)"}}
    print(json.dumps(out), flush=True)


def uneven_main(args):
    """Uneven-batching workload: many materials, per-material sequential runs
    (host-synchronized batches) against planned sweeps replayed as hipGraphs,
    all materials concurrently on their own streams.  Prints one JSON line."""
    import phoebe_amd as P
    from phoebe_amd import synth

    t0 = time.perf_counter()
    mats = synth.uneven_materials(args.materials, seed=1003, max_gw_bytes=args.max_gw_gb * 1e9)
    log(f"[cfg3] {len(mats)} materials generated ({time.perf_counter() - t0:.0f}s)")
    handles, seq_t, pairs = [], 0.0, 0
    for i, p in enumerate(mats):
        lw = P.linewidths_auto(p)
        sep = lw.transform == "separable"
        if args.autotune:
            lw.autotune(np.arange(min(lw.num_active, 2 * p.mesh[0]), dtype=np.int32))
        lw.run()  # warm
        lw.sync()
        lw.reset()
        t1 = time.perf_counter()
        pairs += lw.run()
        lw.sync()
        seq_t += time.perf_counter() - t1
        ref = lw.linewidths(average=False)
        lw.reset()
        lw.plan()
        lw.plan_run(graph=True)  # capture + warm replay
        lw.sync()
        if not np.array_equal(lw.linewidths(average=False), ref):
            raise RuntimeError(f"material {i}: graph replay differs from the sequential run")
        lw.reset()
        handles.append(lw)
        log(f"[cfg3] material {i}: Nw {p.nw}, {p.nm} modes, R {p.nRe}/{p.nRp}, mesh {p.mesh[0]}^3, "
            f"{'separable' if sep else 'direct'}")
    t1 = time.perf_counter()
    gp = sum(lw.plan_run(graph=True) for lw in handles)
    for lw in handles:
        lw.sync()
    con_t = time.perf_counter() - t1
    assert gp == pairs
    out = {"metric": "el-ph (k,q) pairs/sec, uneven materials, concurrent graph sweeps on 1 GPU",
           "value": pairs / con_t, "unit": "(k,q) pairs/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"cfg3: {len(mats)} materials, Nw 4..32, 1..12 atoms, R shells 5^3..9^3, "
                                  f"meshes 8^3..24^3, FP32 gW <= {args.max_gw_gb} GB each",
                      "autotune": bool(args.autotune)},
           "sequential_pairs_per_s": pairs / seq_t, "speedup_vs_sequential": seq_t / con_t,
           "pairs": pairs, "sequential_s": seq_t, "concurrent_s": con_t}
    print(json.dumps(out), flush=True)


def learned_main(args):
    """Learned-screen workload (cfg4: Nw=16, 18 modes, 60^3 mesh): exact screened
    linewidths on evaluation mesh lines, training samples from other lines,
    then the learned screen on the evaluation lines.  Prints one JSON line."""
    import phoebe_amd as P
    from phoebe_amd import synth

    t0 = time.perf_counter()
    prob = synth.config(args.config)
    log(f"problem {args.config}: nw {prob.nw} nm {prob.nm} nRe {prob.nRe} nRp {prob.nRp} mesh {prob.mesh} "
        f"({time.perf_counter() - t0:.1f}s)")
    lw = P.ElectronLinewidths(prob, gW=None, precision="fp32", transform="separable")
    lw.set_screen(6.0)
    lw.set_line_batching(8)
    line = np.asarray(lw.active) // prob.mesh[0]
    pick = np.random.default_rng(12345).permutation(np.unique(line))
    pts = lambda ls: np.concatenate([np.flatnonzero(line == l) for l in ls]).astype(np.int32)
    ev, tr, wu = pts(pick[:args.steps]), pts(pick[args.steps:args.steps + args.train_lines]), pts(pick[-1:])
    lw.run(k1_list=wu)
    lw.sync()
    lw.reset()

    def timed():
        lw.sync()
        t = time.perf_counter()
        pairs = lw.run(k1_list=ev)
        lw.sync()
        return pairs, time.perf_counter() - t
    pairs, t_exact = timed()
    g0 = lw.linewidths(average=False)
    lw.reset()
    t = time.perf_counter()
    ns = lw.screen_collect(k1_list=tr, sample_frac=args.sample, seed=3)
    t_collect = time.perf_counter() - t
    t = time.perf_counter()
    before, after = lw.screen_train(epochs=300, batch=512, lr=2e-3, seed=5)
    t_train = time.perf_counter() - t
    log(f"[cfg4] {ns} samples ({t_collect:.1f}s), trained {before:.3f} -> {after:.3f} ({t_train:.1f}s)")
    lw.set_learned_screen(True, threshold=args.threshold, audit_frac=args.audit, seed=9)
    lw.run(k1_list=wu)  # warm the screened path
    lw.sync()
    lw.reset()
    lw.screen_reset_stats()
    _, t_learned = timed()
    g1 = lw.linewidths(average=False)
    st = lw.screen_stats()
    m = g0 != 0
    err = np.abs(g1[m] - g0[m]) / np.abs(g0[m])
    log(f"[cfg4] exact {t_exact:.3f}s, learned {t_learned:.3f}s, stats {st}, max rel err {err.max():.3e}")
    out = {"metric": "el-ph (k,q) pairs/sec with the learned q-block screen on 1 GPU",
           "value": pairs / t_learned, "unit": "(k,q) pairs/s", "n_gpus": 1, "steps": args.steps,
           "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"{args.config}: Nw={prob.nw}, {prob.nm} modes, {prob.mesh[0]}^3 mesh, "
                                  "FP32 separable, 6 sigma screen, mesh-line batches",
                      "threshold": args.threshold, "audit_frac": args.audit, "train_sample": args.sample,
                      "train_lines": args.train_lines},
           "exact_screened_pairs_per_s": pairs / t_exact, "speedup_vs_exact_screened": t_exact / t_learned,
           "blocks_in_screen": st["blocks_in_screen"], "learned_skips": st["learned_skips"],
           "skip_fraction": st["learned_skips"] / max(st["blocks_in_screen"], 1),
           "audited": st["audited"], "fallback_batches": st["fallbacks"],
           "audit_max_ratio_to_threshold": st["audit_max_ratio"],
           "max_linewidth_rel_err_vs_exact": float(err.max()) if err.size else None,
           "train_samples": ns, "train_mse_before_after": [before, after], "margin_ln": st["margin"],
           "collect_s": t_collect, "train_s": t_train}
    print(json.dumps(out), flush=True)


def phel_main(args):
    """Phonon-linewidth workload (PhElScatteringMatrix::builder) on the cfg2
    system: Nw=16, 12 modes, R_e WS(10^3), R_ph WS(9^3), 40^3 k mesh, phonons
    on the 40^3 q mesh, FP64 coupling synthesized on the device.  A step is
    one batch of 32 outer points k1 against every q3.  Prints one JSON line."""
    import phoebe_amd as P
    from phoebe_amd import synth

    t0 = time.perf_counter()
    prob = synth.config("cfg2")
    qm = tuple(int(x) for x in prob.mesh)
    prec = args.precision
    inter = P.InteractionElPhWan(None, prob.Re, prob.degE, prob.Rp, prob.degP, prec,
                                 synthetic=(prob.nw, prob.nm, prob.gw_decay, prob.gw_seed))
    ph = P.PhononElectronLinewidths(prob, inter, qm)
    log(f"[phel] setup {time.perf_counter() - t0:.1f}s: active {ph.num_active}, q mesh {qm}")
    K = 32
    pts = np.random.default_rng(7).permutation(ph.num_active).astype(np.int32)
    warm = pts[:K]
    timed = [pts[K * (1 + i):K * (2 + i)] for i in range(args.steps)]
    for _ in range(max(args.warmup, 1)):
        ph.run(warm)
    ph.reset()
    ph.set_timing(True)  # events per batch; each run call ends with one synchronization anyway
    t0 = time.perf_counter()
    pairs = sum(ph.run(b) for b in timed)
    elapsed = time.perf_counter() - t0
    kt = ph.kernel_times()
    ph.set_timing(False)
    log(f"[phel] {pairs} pairs in {elapsed:.3f}s: {pairs / elapsed:.4g} pairs/s; kernels {kt}")
    # dominant kernel: the batched cacheElPh (one gW stream per batch) or the coupling
    gw_bytes = (16.0 if prec == "fp64" else 8.0) * prob.nRe * prob.nRp * prob.nm * prob.nw * prob.nw
    if kt["cache_ms"] >= kt["coupling_ms"]:
        n = max(kt["cache_launches"], 1)
        ach = gw_bytes / (kt["cache_ms"] / n * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": "k_cache", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_ms": round(kt["cache_ms"] / n, 4), "launches": n}
    else:
        n = max(kt["coupling_launches"], 1)
        ach = kt["coupling_flops"] / (kt["coupling_ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": "k_coupling", "achieved": round(ach, 3), "peak": PEAK_TFLOPS[prec],
                "unit": "TFLOP/s", "frac": round(ach / PEAK_TFLOPS[prec], 4), "traffic": None,
                "avg_launch_ms": round(kt["coupling_ms"] / n, 4), "launches": n,
                "flops_per_launch": kt["coupling_flops"] / n}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        k1 = int(pts[K * (1 + args.steps)])
        t1 = time.perf_counter()
        gw = P.synth_gw_host(prob)
        orc = O.Oracle(prob, gw)
        log(f"[phel/cpu] gW on the host and oracle set-up {time.perf_counter() - t1:.1f}s")
        ref = orc.phel_linewidths(qm, average=False, k1_list=[k1])
        cpairs, csec = orc.phel_pairs, orc.phel_seconds
        ph.reset()
        ph.run([k1])
        got = ph.linewidths(average=False)
        nz = np.abs(ref) > 1e-30 * np.abs(ref).max()
        err = float((np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])).max())
        log(f"[phel/cpu] {cpairs} pairs in {csec:.1f}s; GPU vs oracle max rel err {err:.3e}")
        del gw, orc
        cpu = {"value": cpairs / csec, "unit": "(k,q) pairs/s", "cores": int(O.lib().orc_num_threads()),
               "kind": "port", "sample": f"1 outer k1 x all {cpairs} q3 of the 40^3 mesh, cfg2 inputs "
                                         "(FP64 C++ restatement of PhElScatteringMatrix::builder)",
               "seconds": round(csec, 2), f"gpu_{prec}_vs_oracle_max_rel_err": err}
    out = {"metric": f"ph-el (k,q) pairs/sec, phonon linewidths (PhElScatteringMatrix builder), {prec.upper()} "
                     "coupling on 1 GPU",
           "value": pairs / elapsed, "unit": "(k,q) pairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": elapsed / max(len(timed), 1) * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": prec, "data": "synthetic",
           "config": {"workload": "phel on cfg2: Nw=16, 12 modes, R_e WS(10^3)=1163, R_ph WS(9^3)=857, "
                                  "40^3 k mesh, 40^3 q mesh, 0.5 eV window",
                      "k1_per_step": K, "q_per_k1": int(np.prod(qm))},
           "kernels": kt, "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--k1-batch", type=int, default=32)
    ap.add_argument("--chosen", default=CHOSEN, choices=sorted(k for k in FORMS if k != "fp64"))
    ap.add_argument("--cpu-k1", type=int, default=1, help="outer points in the CPU sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only-chosen", action="store_true",
                    help="skip the plain FP64 run and the CPU baseline (profiling runs)")
    ap.add_argument("--no-fp64-separable", action="store_true",
                    help="skip the FP64 separable form (line64s6) beside plain FP64")
    ap.add_argument("--points", type=int, default=2 ** 17, help="cfg1: wavevectors per family")
    ap.add_argument("--materials", type=int, default=64, help="cfg3: number of materials")
    ap.add_argument("--max-gw-gb", type=float, default=1.0, help="cfg3: FP32 gW budget per material")
    ap.add_argument("--autotune", action="store_true", help="cfg3: autotune each material's batching")
    ap.add_argument("--threshold", type=float, default=1e-8, help="cfg4: learned-screen threshold")
    ap.add_argument("--audit", type=float, default=0.01, help="cfg4: audited fraction of skipped blocks")
    ap.add_argument("--sample", type=float, default=0.02, help="cfg4: training sample of tail blocks")
    ap.add_argument("--train-lines", type=int, default=8, help="cfg4: mesh lines for training samples")
    ap.add_argument("--calcs", type=int, default=1,
                    help="(T, mu) calculations: the workload's T repeated at T, 1.5 T, 2 T, ...")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"], help="phel: coupling precision")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); sharding the outer "
              "points k1 across GPUs is not run here.", flush=True)
        sys.exit(2)

    if args.config == "cfg1":
        eig_main(args)
        return
    if args.config == "cfg3":
        uneven_main(args)
        return
    if args.config == "cfg4":
        learned_main(args)
        return
    if args.config == "phel":
        phel_main(args)
        return
    import phoebe_amd as P
    from phoebe_amd import synth

    t0 = time.perf_counter()
    prob = synth.config(args.config)
    if args.calcs > 1:
        prob.temperatures = prob.temperatures[:1].repeat(args.calcs) * (1 + 0.5 * np.arange(args.calcs))
        prob.chem_pots = prob.chem_pots[:1].repeat(args.calcs)
    log(f"problem {args.config}: nw {prob.nw} nm {prob.nm} nRe {prob.nRe} nRp {prob.nRp} "
        f"mesh {prob.mesh} gW {prob.gw_bytes() / 1e9:.1f} GB ({time.perf_counter() - t0:.1f}s)")

    B = args.k1_batch
    m0 = prob.mesh[0]

    def make_batches(active, by_line, warmup):
        """(timed, warmup) lists of outer points: the points of `steps` mesh
        lines (then of `warmup` further lines), as one list handed to the
        library's line batching, or as k1-batched lists of B points."""
        line = np.asarray(active) // m0
        pick = np.random.default_rng(12345).permutation(np.unique(line))
        def pts(ls):
            return np.concatenate([np.flatnonzero(line == l) for l in ls]).astype(np.int32)
        timed, warm = pts(pick[:args.steps]), pts(pick[args.steps:args.steps + max(warmup, 1)])
        if by_line:
            return [timed], ([warm] if warmup else [])
        return [timed[i:i + B] for i in range(0, len(timed), B)], [warm[i * B:(i + 1) * B] for i in range(warmup)]

    # plain FP64
    r64, rf64, cpu = None, None, None
    if not args.only_chosen:
        lw64, r64 = run_form(P, prob, "fp64", make_batches, args.warmup, B)
        batches, _ = make_batches(lw64.active, False, 0)
        if not args.no_cpu_baseline:
            try:
                cpu = cpu_baseline(lw64, prob, batches, r64["linewidths"], args.cpu_k1)
            except Exception as e:  # reported, not fatal for the GPU numbers
                log(f"[cpu] baseline failed: {e}")
                cpu = {"error": str(e)}
        rf64 = roofline(r64, "fp64", args.config)
        del lw64

    # the FP64 separable form: the exact algorithmic changes (mesh-line cacheElPh,
    # separable R_ph transform, 6-sigma screen) without the reduced precision
    rs64 = None
    if not args.only_chosen and not args.no_fp64_separable:
        lws, rs64 = run_form(P, prob, "line64s6", make_batches, args.warmup, B)
        del lws
    lwc, rc = run_form(P, prob, args.chosen, make_batches, args.warmup, B)
    del lwc
    # HBM reference on this GPU: a plain 4 GiB device copy (8 GiB of traffic)
    copy_ms = P.stream_copy_ms(4 << 30, reps=5)
    copy_gbps = 2 * (4 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"[copy] 4 GiB device copy {copy_ms:.3f} ms: {copy_gbps:.0f} GB/s")
    vc = rc["pairs"] / rc["seconds"]
    v64, relerr = None, None
    if r64 is not None:
        # per-linewidth error of the chosen form against FP64, over the processed outer states
        g64, gc = r64["linewidths"], rc["linewidths"]
        mask = g64 != 0
        relerr = np.abs(gc[mask] - g64[mask]) / np.abs(g64[mask])
        v64 = r64["pairs"] / r64["seconds"]
    sep64 = None
    if rs64 is not None:
        vs = rs64["pairs"] / rs64["seconds"]
        e_s = None
        if r64 is not None:
            g64, gs = r64["linewidths"], rs64["linewidths"]
            m = g64 != 0
            e_s = float((np.abs(gs[m] - g64[m]) / np.abs(g64[m])).max()) if m.any() else None
        sep64 = {"form": "line64s6", "value": vs, "ms_per_step": rs64["seconds"] * 1e3 / args.steps,
                 "speedup_vs_plain_fp64": None if v64 is None else vs / v64,
                 "chosen_speedup_over_it": vc / vs,
                 "max_linewidth_rel_err_vs_plain_fp64": e_s,
                 "roofline": roofline(rs64, "line64s6", args.config)}
    out = {
        "metric": METRIC,
        "value": vc,
        "unit": "(k,q) pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rc["seconds"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": FORMS[args.chosen][0],
        "data": "synthetic",
        "config": {"workload": f"{args.config}: Nw={prob.nw}, {prob.nm} modes, "
                               f"R_e WS(10^3)={prob.nRe}, R_ph WS(9^3)={prob.nRp}, "
                               f"{prob.mesh[0]}^3 k/q mesh, 0.5 eV window",
                   "k1_per_step": int(rc["k1"] // args.steps) if FORMS[args.chosen][3] else B,
                   "k2_per_k1": int(rc["pairs"] // rc["k1"]),
                   "calculations": int(len(prob.temperatures)),
                   "active_k": int(rc["active"]), "chosen_form": args.chosen,
                   "chosen_precision": FORMS[args.chosen][0],
                   "chosen_transform": FORMS[args.chosen][1],
                   "chosen_screen_sigma": FORMS[args.chosen][2],
                   "chosen_line_min_points": FORMS[args.chosen][3],
                   "lines_per_cache_stream": 16, "panel_chunk": "automatic (128 at cfg2)"},
        "plain_fp64": None if r64 is None else {
            "value": v64, "ms_per_step": r64["seconds"] * 1e3 / args.steps, "roofline": rf64},
        "fp64_separable": sep64,
        "speedup_vs_fp64": None if v64 is None else vc / v64,
        "max_linewidth_rel_err_vs_fp64": float(relerr.max()) if relerr is not None and relerr.size else None,
        "linewidths_checked": 0 if relerr is None else int(relerr.size),
        "roofline": roofline(rc, args.chosen, args.config),
        "hbm_copy_GBps": copy_gbps,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
