#!/usr/bin/env python3
"""Benchmark of the MI355X radiation hot path.

Step = one SynchrotronSpectrum_3D forward pass over 2^24 particles (float32
SoA u, B; |u| power law p = -2.5 on [1, 1e3], isotropic; B = (0,0,1) +
N(0, 0.3^2); unit weights, E = 0) into 256 log photon bins on [1e-2, 1e7] mc^2,
inputs resident in HBM, with the library's default method choice (at this size
the exact factorised form; the pair-by-pair kernel is timed beside it).
value = particle x bin pairs per second (n * nbins / step time).

The JSON line also carries the roofline of the step's longest kernel (timed
per launch with HIP events on its stream), the CPU baseline (the oracle's
faithful OpenMP restatement of the reference, timed on a bounded sample on this
host) and secondary measurements: forward+adjoint, the energy histogram (HBM
bound), the distribution form, batched cells and the batched contraction.
"""
from __future__ import annotations

import argparse
import json
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


def logspace(start, stop, num):
    """The reference's Logspace grid, from the `ragnar` module of this package."""
    import numpy as np

    from ragnar_amd import ragnar as rg

    return np.asarray(rg.Logspace(start, stop, num).as_array(), np.float32)


def generator(kind, params, e):
    """A reference generator ("plaw": p, emin, emax; "delta": energy0, denergy)
    evaluated at the energies e, from the `ragnar` module of this package."""
    import numpy as np

    from ragnar_amd import ragnar as rg

    gen = {"plaw": rg.PlawGenerator, "delta": rg.DeltaGenerator}[kind](*params)
    return np.asarray(gen.compute(rg.Bins(np.ascontiguousarray(e, np.float32))).as_array(), np.float32)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--profile-reps", type=int, default=10)
    ap.add_argument("--particles", type=int, default=1 << 24)
    ap.add_argument("--bins", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="only the headline step")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU work (thread-seconds) of the baseline sample")
    ap.add_argument("--method", choices=("auto", "direct", "binned", "compute"), default="auto",
                    help="emission method of --profile; with --summarize also \"compute\", the --profile-compute runs")
    ap.add_argument("--summarize", nargs=3, metavar=("TRACE_DIR", "FETCH_DIR", "WRITE_DIR"),
                    help="summarise three rocprofv3 --output-format csv runs of --profile "
                         "(--kernel-trace --stats, --pmc FETCH_SIZE, --pmc WRITE_SIZE) into "
                         "profiles/*.md and profiles/traffic_2p24.json, then exit")
    ap.add_argument("--profile", action="store_true",
                    help="dispatch each full-size kernel once (forward, adjoint, histogram) and a "
                         "1 GiB device copy that calibrates rocprofv3 byte counters, then exit")
    ap.add_argument("--profile-compute", action="store_true",
                    help="dispatch the compute-bound kernels (pair-by-pair forward, split-FP16 "
                         "contraction) for rocprofv3 VALU / MFMA counters, then exit")
    ap.add_argument("--summarize-bench", nargs=2, metavar=("TRACE_DIR", "BENCH_JSON"),
                    help="kernel durations of a rocprofv3 --kernel-trace run of the headline bench "
                         "command, beside an unprofiled bench line's event-timed stages")
    ap.add_argument("--summarize-compute", nargs=2, metavar=("TRACE_DIR", "PMC_DIR"),
                    help="summarise a --kernel-trace --stats run and a --pmc run of "
                         "--profile-compute into profiles/compute_units.md, then exit")
    return ap.parse_args()


def profile_run(args):
    """Single dispatches for rocprofv3 --pmc / --kernel-trace runs."""
    import torch

    from ragnar_amd import torch_ops as T
    from ragnar_amd import workloads

    n, nb = args.particles, args.bins
    u_h, b_h = workloads.particles(n)
    U = torch.from_numpy(u_h).cuda()
    B = torch.from_numpy(b_h).cuda()
    plan = T.SyncPlan(T.default_table(), logspace(1e-2, 1e7, nb))
    hplan = T.HistPlan(logspace(1e-2, 1e3, 200), True)
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
    dst = torch.empty_like(src)
    src.fill_(1.0)
    S = torch.empty(nb, device="cuda")
    ones = torch.ones(nb, device="cuda")
    reps = args.profile_reps
    m = args.method
    for _ in range(args.warmup):  # clocks ramp up before the counted dispatches
        T.sync_spectrum(plan, U, B, out=S, method=m)
    torch.cuda.synchronize()
    print(f"profile: forward ({m}) x{reps}", flush=True)
    for _ in range(reps):
        T.sync_spectrum(plan, U, B, out=S, method=m)
    print(f"profile: energy histogram x{reps}", flush=True)
    for _ in range(reps):
        T.energy_hist(hplan, U, True)
    print(f"profile: adjoint ({m}) x{reps}", flush=True)
    for _ in range(reps):
        T.sync_adjoint(plan, U, B, None, None, ones, need=("u", "b", "w"), method=m)
    print("profile: 1 GiB copy (calibration)", flush=True)
    dst.copy_(src)
    torch.cuda.synchronize()
    print("profile: done", flush=True)


# Counters of the --profile-compute pass: FP32 / FP64 VALU flops (MFMA excluded),
# F16 MFMA operations / 512, and rocprofv3's derived VALUBusy and MfmaUtil
# (5 SQ counters and GRBM_GUI_ACTIVE: one pass).
COMPUTE_COUNTERS = ("SQ_INSTS_VALU_FLOPS_FP32", "SQ_INSTS_VALU_FLOPS_FP64", "SQ_INSTS_VALU_MFMA_MOPS_F16",
                    "VALUBusy", "MfmaUtil")
PEAK_FP32_NONPACKED = 78.6e12  # flop/s: FP32 VALU without packed math (this build's kernels)
PEAK_FP64_VALU = 78.6e12
PEAK_MFMA_F16 = 2.5e15         # dense


def profile_compute_run(args):
    """Single dispatches of the compute-bound kernels for rocprofv3: the
    pair-by-pair forward at 2^24 x 256 (FP32 VALU), the batched contraction at
    65536 x 512 x 512 (MFMA): split FP16 with three and one products, split
    BF16 with six, and the rows generated in the kernel, and the
    inverse-Compton spectrum at 1000 x 1000 x 1000 triples (FP32/FP64 VALU)."""
    import torch

    from ragnar_amd import torch_ops as T
    from ragnar_amd import workloads

    n, nb = args.particles, args.bins
    u_h, b_h = workloads.particles(n)
    U = torch.from_numpy(u_h).cuda()
    B = torch.from_numpy(b_h).cuda()
    plan = T.SyncPlan(T.default_table(), logspace(1e-2, 1e7, nb))
    S = torch.empty(nb, device="cuda")
    nd, ng, nbc = 65536, 512, 512
    f, pidx, gcut = workloads.distributions(nd, ng)
    G = torch.from_numpy(logspace(1, 1e6, ng)).cuda()
    Fm = torch.from_numpy(f).cuda()
    Pi = torch.from_numpy(pidx.astype("float32")).cuda()
    Gc = torch.from_numpy(gcut.astype("float32")).cuda()
    cplan = T.SyncPlan(T.default_table(), logspace(1e-2, 1e7, nbc))
    out = torch.empty(nd, nbc, device="cuda")
    reps = args.profile_reps
    for _ in range(args.warmup):
        T.sync_spectrum(plan, U, B, out=S, method="direct")
        T.sync_fromdist_batched(cplan, G, Fm, True, 1, out=out)
    torch.cuda.synchronize()
    print(f"profile: forward (direct) x{reps}", flush=True)
    for _ in range(reps):
        T.sync_spectrum(plan, U, B, out=S, method="direct")
    for prec, name in ((1, "f16x3"), (2, "f16x1"), (3, "bf16x6")):
        print(f"profile: contraction {name} x{reps}", flush=True)
        for _ in range(reps):
            T.sync_fromdist_batched(cplan, G, Fm, True, prec, out=out)
    for prec, name in ((1, "f16x3"), (3, "bf16x6")):
        print(f"profile: contraction {name}, rows generated in the kernel x{reps}", flush=True)
        for _ in range(reps):
            T.sync_fromdist_batched_plaw(cplan, G, Pi, Gc, True, prec, out=out)
    gi = logspace(1e3, 1e7, 1000)
    si = logspace(1e-11, 1e-7, 1000)
    ic = [torch.from_numpy(a).cuda() for a in (gi, generator("plaw", [-1.5, 1e3, 1e7], gi), si,
                                                 generator("delta", [1e-8, 1e-9], si),
                                                 logspace(1e3, 1e7, 1000))]
    T.ic_spectrum(ic[0], ic[1], True, ic[2], ic[3], ic[4])
    torch.cuda.synchronize()
    print(f"profile: ic spectrum 1000^3 x{reps}", flush=True)
    for _ in range(reps):
        T.ic_spectrum(ic[0], ic[1], True, ic[2], ic[3], ic[4])
    print("profile: 1 GiB copy (calibration)", flush=True)
    del U, B, Fm
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    src.fill_(1.0)
    torch.empty_like(src).copy_(src)
    torch.cuda.synchronize()
    print("profile: done", flush=True)


def summarize_compute(trace_dir, pmc_dir):
    """Achieved FP32 VALU, FP64 VALU and F16 MFMA rates of the --profile-compute
    kernels against MI355X peaks, from rocprofv3 CSV output: counter values
    per dispatch (warm dispatches averaged, split by template instance via the
    dispatch order of the trace) over the kernel's average traced duration."""
    import csv
    import statistics

    trace = list(csv.DictReader(open(Path(trace_dir) / "run_kernel_trace.csv")))
    dur = {}
    for r in trace:
        dur.setdefault(r["Kernel_Name"], []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
    vals = {}
    for r in csv.DictReader(open(Path(pmc_dir) / "run_counter_collection.csv")):
        vals.setdefault(r["Kernel_Name"], {}).setdefault(r["Counter_Name"], {}).setdefault(
            int(r["Dispatch_Id"]), 0.0)
        vals[r["Kernel_Name"]][r["Counter_Name"]][int(r["Dispatch_Id"])] += float(r["Counter_Value"])

    def warm(v):
        v = [x for _, x in sorted(v.items())]
        return statistics.mean(v[1:] if len(v) > 1 else v)

    rows = ["| kernel | dispatches | avg duration (us) | FP32 VALU TFLOP/s | FP64 VALU TFLOP/s | "
            "F16 MFMA TFLOP/s (counter) | matrix TFLOP/s (products) | VALUBusy % | MfmaUtil % |",
            "|---|---|---|---|---|---|---|---|---|"]
    gemm = 2.0 * 65536 * 512 * 512  # one product pass of the profiled contraction
    report = {}
    for k, d in sorted(dur.items(), key=lambda kv: -statistics.mean(kv[1])):
        full = _demangle(k)
        short = _short_name(full)
        if short not in ("forward_kernel", "combine_kernel", "mfma_contract2", "mfma_contract_bf16",
                         "pack_kernel_matrix", "pack_kernel_matrix_bf16", "ic_kernel", "ic_combine"):
            continue
        c = vals.get(k, {})
        t = statistics.mean(d[1:] if len(d) > 1 else d)
        # the SQ VALU instruction counters count wave instructions: x 64 lanes
        # (as rocprofv3's own TOTAL_32_OPS does)
        f32 = (warm(c["SQ_INSTS_VALU_FLOPS_FP32"]) * 64 / t / 1e12
               if "SQ_INSTS_VALU_FLOPS_FP32" in c else float("nan"))
        f64 = (warm(c["SQ_INSTS_VALU_FLOPS_FP64"]) * 64 / t / 1e12
               if "SQ_INSTS_VALU_FLOPS_FP64" in c else float("nan"))
        mf = (warm(c["SQ_INSTS_VALU_MFMA_MOPS_F16"]) * 512 / t / 1e12
              if "SQ_INSTS_VALU_MFMA_MOPS_F16" in c else float("nan"))
        vb = warm(c["VALUBusy"]) if "VALUBusy" in c else float("nan")
        mu = warm(c["MfmaUtil"]) if "MfmaUtil" in c else float("nan")
        base = full.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").strip()
        tmpl = short + (base[base.index("<"):] if "<" in base else "")
        tmpl = tmpl.replace("rgn::", "")
        # products per element of the contraction kernels (template argument
        # NPROD of the split-FP16 kernels; six for split BF16)
        nprod = 6 if short == "mfma_contract_bf16" else (
            int(base[base.index("<") + 1:].split(",")[0]) if short.startswith("mfma_contract") and "<" in base else 0)
        mx = nprod * gemm / t / 1e12 if nprod else float("nan")
        rows.append(f"| `{tmpl}` | {len(d)} | {t * 1e6:.1f} | {f32:.1f} | {f64:.1f} | {mf:.0f} | {mx:.0f} | {vb:.0f} | "
                    f"{mu:.0f} |")
        report[tmpl] = {"us": t * 1e6, "fp32_tflops": f32, "fp64_tflops": f64, "mfma_f16_tflops": mf,
                        "matrix_tflops": mx, "valu_busy_pct": vb, "mfma_util_pct": mu}
    md = ["# VALU and MFMA rates of the compute-bound kernels (MI355X)", "",
          "Made by `python bench.py --summarize-compute <trace> <pmc>` from two rocprofv3 runs of",
          "`bench.py --profile-compute` (commands in profiles/README.md): the pair-by-pair",
          "forward at 2^24 particles x 256 bins, the batched contraction at 65536 x 512 x 512",
          "(split FP16 `mfma_contract2<3, 2, false>`: three products, f read as whole 64-g blocks,",
          "`<1, 2, false>`: one product; split BF16 `mfma_contract_bf16<2, false>`: six products;",
          "`mfma_contract2<3, 2, true>` and `mfma_contract_bf16<2, true>`: rows generated in the",
          "kernel) and the inverse-Compton",
          "spectrum at 1000 gamma x 1000 soft-photon x 1000 output bins (`ic_kernel`).  The",
          "matrix rate counts every product pass (2 x 65536 x 512 x 512 flop each) over the traced",
          "duration; the counter rate is F16 only.", "",
          f"Peaks: FP32 VALU {PEAK_FP32_NONPACKED / 1e12:.1f} TFLOP/s without packed math "
          f"({PEAK_FP32 / 1e12:.1f} with it; this build compiles with packed FP32 off), "
          f"FP64 VALU {PEAK_FP64_VALU / 1e12:.1f}, dense F16 MFMA {PEAK_MFMA_F16 / 1e12:.0f}.",
          "Counters: " + ", ".join(f"`{c}`" for c in COMPUTE_COUNTERS) + " (one --pmc pass).  The "
          "VALU flop counters count per wave instruction and are scaled by 64 lanes; MFMA flops = "
          "MOPS_F16 x 512; VALUBusy and MfmaUtil are rocprofv3's derived percentages (VALUBusy "
          "sums four SIMDs per CU, so it can exceed 100).", "",
          *rows, ""]
    (ROOT / "profiles" / "compute_units.md").write_text("\n".join(md))
    print("\n".join(md))
    return report


def _demangle(name):
    """Kernel name and integer / bool template arguments of a mangled
    _ZN...E name (what these kernels need; c++filt here cannot parse the
    _Float16 parameter types), else the name unchanged."""
    import re

    if not name.startswith("_ZN"):
        return name
    i, parts = 3, []
    while i < len(name) and name[i].isdigit():
        m = re.match(r"\d+", name[i:])
        n = int(m.group())
        i += len(m.group())
        parts.append(name[i:i + n])
        i += n
    args = []
    if i < len(name) and name[i] == "I":
        for kind, val in re.findall(r"L([ib])(\d+)E", name[i:name.find("EE", i) + 1]):
            args.append(("true" if val == "1" else "false") if kind == "b" else val)
    base = "::".join(p for p in parts if not p.startswith("_GLOBAL__N"))
    return "void " + base + (f"<{', '.join(args)}>" if args else "") + "()"


def _short_name(name):
    """Kernel name without namespaces, template arguments and parameters."""
    import re

    base = _demangle(name).replace("(anonymous namespace)::", "").split("(")[0]
    base = re.sub(r"<.*>", "", base)
    base = base.replace("void ", "").strip()
    return base.split("::")[-1] or name[:40]


def _kernel_label(name):
    """Kernel name with its template arguments, without namespaces and
    parameters (tells the template instances of one kernel apart)."""
    base = _demangle(name).replace("(anonymous namespace)::", "").split("(")[0]
    return base.replace("void ", "").replace("rgn::", "").strip() or name[:40]


def summarize(trace_dir, fetch_dir, write_dir, method, tag=None, title=None, source=None):
    """Per-kernel durations and HBM bytes from rocprofv3 CSV output of --profile.

    Byte counters (FETCH_SIZE, WRITE_SIZE, KiB per dispatch) are averaged over
    the warm dispatches of a kernel (all but the first) and each is scaled by
    the factor that makes the 1 GiB device copy of --profile read and write
    exactly 1 GiB.  Writes profiles/kernels_2p24_<method>.md and merges the
    kernels into profiles/traffic_2p24.json."""
    import csv
    import statistics

    # template instances apart for the compute kernels; short names for the
    # emission kernels (the keys the bench line's traffic lookup uses)
    label = _kernel_label if method == "compute" else _short_name
    stats = list(csv.DictReader(open(Path(trace_dir) / "run_kernel_stats.csv")))
    counters = {}
    for d, c in ((fetch_dir, "FETCH_SIZE"), (write_dir, "WRITE_SIZE")):
        for r in csv.DictReader(open(Path(d) / "run_counter_collection.csv")):
            key = (label(r["Kernel_Name"]), int(r["Grid_Size"]))
            counters.setdefault(key, {}).setdefault(c, []).append(float(r["Counter_Value"]))

    def warm(v):
        return statistics.mean(v[1:] if len(v) > 1 else v)

    copy = [v for (k, g), v in counters.items() if "copyBuffer" in k and g >= 131072]
    if not copy:
        raise SystemExit("no 1 GiB copy dispatch in the counter runs")
    gib_kib = 1024.0 * 1024.0
    ffac = gib_kib / warm(copy[0]["FETCH_SIZE"])
    wfac = gib_kib / warm(copy[0]["WRITE_SIZE"])
    lines = ["| kernel | dispatches | avg duration (us) | min (us) | max (us) |", "|---|---|---|---|---|"]
    for r in stats:
        lines.append(f"| `{label(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} |")
    tl = ["| kernel (grid threads) | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes / launch (scaled) |",
          "|---|---|---|---|"]
    traffic = {}
    for (k, g), v in sorted(counters.items()):
        if "FETCH_SIZE" not in v or "WRITE_SIZE" not in v:
            continue
        fk, wk = warm(v["FETCH_SIZE"]), warm(v["WRITE_SIZE"])
        b = (fk * ffac + wk * wfac) * 1024.0
        tl.append(f"| `{k}` ({g}) | {fk:.0f} | {wk:.0f} | {b / 1e6:.1f} MB |")
        if k not in traffic or b > traffic[k]["bytes_per_launch"]:
            traffic[k] = {"fetch_kib": fk, "write_kib": wk, "bytes_per_launch": b,
                          "fetch_factor": ffac, "write_factor": wfac, "grid_threads": g}
    tag = tag or f"2p24_{method}"
    title = title or f"2^24 particles x 256 bins, method {method}"
    source = source or f"--profile --method {method}"
    md = [f"# Kernel durations and HBM traffic, {title}", "",
          f"Made by `python bench.py --method {method} --summarize <trace> <fetch> <write>` from",
          f"the three `bench.py {source}` runs listed in profiles/README.md.", "",
          "## Durations (rocprofv3 --kernel-trace --stats)", "",
          *lines, "", "## HBM traffic (rocprofv3 --pmc, warm dispatches)", "",
          f"Scale factors from the 1 GiB copy: FETCH x{ffac:.3f}, WRITE x{wfac:.3f}.", "", *tl, ""]
    (ROOT / "profiles" / f"kernels_{tag}.md").write_text("\n".join(md))
    tpath = ROOT / "profiles" / "traffic_2p24.json"
    merged = json.loads(tpath.read_text()) if tpath.exists() else {}
    merged.update({k: v for k, v in traffic.items() if not k.startswith(("__amd", "vectorized", "at::"))})
    tpath.write_text(json.dumps(merged, indent=1, sort_keys=True))
    print("\n".join(md))


def summarize_bench(trace_dir, bench_json):
    """Durations of every dispatch of the headline step's kernels in a
    rocprofv3 --kernel-trace run of `bench.py --no-extras --no-cpu`, against
    the event-timed stages of an unprofiled bench line (bench_json: a file
    whose last line is that JSON).  Writes profiles/bench_headline_2p24.md."""
    import csv
    import statistics

    rows = list(csv.DictReader(open(Path(trace_dir) / "run_kernel_trace.csv")))
    line = json.loads(Path(bench_json).read_text().strip().splitlines()[-1])
    stages = line.get("stages_ms", {})
    out = ["| kernel | dispatches | traced avg (us) | median (us) | min (us) | max (us) | bench line, event-timed (us) |",
           "|---|---|---|---|---|---|---|"]
    for k in stages:
        d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rows
             if _short_name(r["Kernel_Name"]) == k]
        if not d:
            continue
        out.append(f"| `{k}` | {len(d)} | {statistics.mean(d):.1f} | {statistics.median(d):.1f} | "
                   f"{min(d):.1f} | {max(d):.1f} | {stages[k] * 1e3:.1f} |")
    rf = line.get("roofline", {})
    md = ["# Headline step under rocprofv3: 2^24 particles x 256 bins, default method", "",
          "Trace: `rocprofv3 --kernel-trace --stats --output-format csv -d <trace> -o run -- "
          "python3 bench.py --no-extras --no-cpu`;",
          "summary: `python bench.py --summarize-bench <trace> <bench line>`, the bench line being the",
          "same command run without the profiler, just before, on the same machine.  Every dispatch of the",
          "step's kernels in the traced run is counted (warm-up, timed steps and stage timing); an",
          "event-timed stage also holds the gap since the previous launch.", "",
          *out, "",
          f"Bench line: {line.get('ms_per_step', 0) * 1e3:.1f} us per step, roofline kernel "
          f"`{rf.get('kernel')}` at {rf.get('achieved', 0):.0f} {rf.get('unit', '')} "
          f"(frac {rf.get('frac', 0):.3f}).", ""]
    (ROOT / "profiles" / "bench_headline_2p24.md").write_text("\n".join(md))
    print("\n".join(md))


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


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


def read_traffic(kernel):
    """HBM bytes per launch of a kernel at the benchmark size, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, each scaled by what the
    counter reads for a 1 GiB device copy), or None."""
    p = ROOT / "profiles" / "traffic_2p24.json"
    if p.exists():
        try:
            return json.loads(p.read_text())[kernel]["bytes_per_launch"]
        except Exception:
            return None
    return None


def bench_fit(T, workloads, torch, table, bins):
    """2^26 particles (same generator as the headline step), weights
    W = |u|^(p + 2.5) exp(-|u| / c) at p = -2.2, c = 3e3, loss
    sum_e (log S_e - log S_e^target)^2 over the 256 bins: one step is the
    forward, the loss gradient dL/dS and the adjoint (dL/dw, dL/du, dL/dB),
    each with the library's default method."""
    n = 1 << 26
    nb = len(bins)
    u_h, b_h = workloads.particles(n, seed=321)
    w_h = workloads.fit_weights(u_h)
    U = torch.from_numpy(u_h).cuda()
    B = torch.from_numpy(b_h).cuda()
    w = torch.from_numpy(w_h).cuda()
    del u_h, b_h, w_h
    plan = T.SyncPlan(table, bins)
    target = (T.sync_spectrum(plan, U, B, w) * 1.05).clamp_min(1e-30).log()

    def fwd():
        return T.sync_spectrum(plan, U, B, w)

    def step():
        S = fwd()
        r = S.clamp_min(1e-30).log() - target
        gbar = torch.where(S > 0, 2.0 * r / S.clamp_min(1e-30), torch.zeros_like(S))
        T.sync_adjoint(plan, U, B, w, None, gbar, need=("w", "u", "b"))

    ms_f = event_time(fwd, 20, 3, torch)
    ms_fa = event_time(step, 20, 30, torch)
    evals = n * nb
    return {"particles": n, "ms_fwd": ms_f, "ms_fwd_adj": ms_fa, "ratio": ms_fa / ms_f,
            "fwd_adj_evals_per_s": evals / (ms_fa * 1e-3)}


def bench_cells(T, workloads, torch, table):
    """4096 cells, lognormal counts (median 4000, sigma 1): one persistent launch
    in a hipGraph against 4096 calls of the single-population form.  The launch
    shape is tuned afresh (a new cache file) on the first call; the chosen
    shape is reported."""
    import os
    import tempfile

    tune_file = os.path.join(tempfile.mkdtemp(prefix="ragnar_tune_"), "tune.json")
    os.environ["RAGNAR_TUNE_CACHE"] = tune_file
    u, b, offs, _, _ = workloads.cells(4096, 4000, 1.0)
    nb = 256
    plan = T.SyncPlan(table, logspace(1e-2, 1e7, nb))
    U = torch.from_numpy(u).cuda()
    B = torch.from_numpy(b).cuda()
    out = torch.empty(4096, nb, device="cuda")
    ms_graph = event_time(lambda: T.sync_cells(plan, U, B, offs, out=out), 10, 2, torch)
    views = [(U[:, a:z], B[:, a:z]) for a, z in zip(offs[:-1], offs[1:])]
    rows = [out[c] for c in range(4096)]

    def per_cell():
        for (uc, bc), o in zip(views, rows):
            T.sync_spectrum(plan, uc.contiguous() if not uc.is_contiguous() else uc, bc, out=o)

    # views of a (3, N) tensor are not contiguous: copy them once, outside the timing
    views = [(uc.contiguous(), bc.contiguous()) for uc, bc in views]
    ms_calls = event_time(per_cell, 2, 1, torch)
    total = int(offs[-1])
    shape = {}
    try:
        for line in open(tune_file):
            if '"' in line and ":" in line:
                key, val = line.strip().rstrip(",").rsplit(":", 1)
                shape[key.strip().strip('"').rsplit("/", 1)[-1]] = int(val)
    except OSError:
        pass
    k = 64  # the first cells, for the CPU leg
    sample = (u[:, :offs[k]].copy(), b[:, :offs[k]].copy(), offs[:k + 1].copy(), total, 4096)
    return {"particles": total, "ms_graph": ms_graph, "ms_per_cell_calls": ms_calls,
            "speedup": ms_calls / ms_graph, "evals_per_s": total * nb / (ms_graph * 1e-3),
            "tuned_shape": {"blocks_per_cu": shape.get("bpc"), "bin_split": shape.get("split")},
            "frac_fp32_peak": 6 * total * nb / (ms_graph * 1e-3) / PEAK_FP32}, sample


def bench_contraction(T, workloads, torch, table):
    """65536 power-law x cutoff distributions on 512 log gamma bins -> 512 photon
    bins: FP32 GEMM (rocBLAS) and the split-FP16 MFMA kernel (three and one
    products).  Returns the timings and, for the checker, 32 sampled rows of each
    output with their inputs."""
    nd, ng, nb = 65536, 512, 512
    f, pidx, gcut = workloads.distributions(nd, ng)
    g = logspace(1, 1e6, ng)
    bins = logspace(1e-2, 1e7, nb)
    plan = T.SyncPlan(table, bins)
    G = torch.from_numpy(g).cuda()
    Fm = torch.from_numpy(f).cuda()
    out = torch.empty(nd, nb, device="cuda")
    res = {}
    rows = list(range(0, nd, nd // 32))
    samples = {"g": g, "bins": bins, "f": f[rows], "out": {}}
    for name, prec in (("fp32", 0), ("f16x3", 1), ("f16x1", 2), ("bf16x6", 3)):
        ms = event_time(lambda: T.sync_fromdist_batched(plan, G, Fm, True, prec, out=out), 10, 2, torch)
        samples["out"][name] = out[rows].cpu().numpy()
        flop = 2.0 * nd * ng * nb
        res[name] = {"ms": ms, "TFLOP_per_s": flop / (ms * 1e-3) / 1e12,
                     "GB_per_s": (nd * ng + nd * nb) * 4 / (ms * 1e-3) / 1e9}
    # rows generated inside the kernel from (index, cutoff): 8 bytes per row in
    P = torch.from_numpy(pidx.astype("float32")).cuda()
    Cc = torch.from_numpy(gcut.astype("float32")).cuda()
    for name, prec in (("f16x3_generated", 1), ("f16x1_generated", 2), ("bf16x6_generated", 3)):
        ms = event_time(lambda: T.sync_fromdist_batched_plaw(plan, G, P, Cc, True, prec, out=out), 10, 2, torch)
        samples["out"][name] = out[rows].cpu().numpy()
        flop = 2.0 * nd * ng * nb
        res[name] = {"ms": ms, "TFLOP_per_s": flop / (ms * 1e-3) / 1e12,
                     "GB_per_s": (8 * nd + 4 * nd * nb) / (ms * 1e-3) / 1e9,
                     "hbm_floor_ms": (8 * nd + 4 * nd * nb) / PEAK_HBM * 1e3}
    samples["index"], samples["cutoff"] = pidx[rows], gcut[rows]
    # the plain per-call form: one SynchrotronSpectrumFromDist call per row
    # (8192 rows timed, scaled to all rows)
    k = 8192
    row_in = [Fm[d] for d in range(k)]
    row_out = [out[d] for d in range(k)]

    def per_row():
        for fr, o in zip(row_in, row_out):
            T.sync_fromdist(plan, G, fr, True, out=o)

    ms_rows = event_time(per_row, 2, 1, torch) * nd / k
    res["per_row_calls"] = {"ms": ms_rows, "rows_timed": k,
                            "speedup_f16x3": ms_rows / res["f16x3"]["ms"],
                            "speedup_fp32": ms_rows / res["fp32"]["ms"]}
    return res, samples


def stage_profile(T, plan, U, B, out, torch, reps):
    """Average per-launch times (HIP events on the launch stream) of `reps`
    default-method forward calls: kernels sorted by time, longest first."""
    acc = {}
    order = []
    for _ in range(3):
        T.spectrum_stages(plan, U, B, out=out)
    for _ in range(reps):
        for name, t in T.spectrum_stages(plan, U, B, out=out):
            if name not in acc:
                acc[name] = 0.0
                order.append(name)
            acc[name] += t
    kernels = sorted(((k, acc[k] / reps) for k in order), key=lambda kv: -kv[1])
    return {"kernels": kernels, "total_ms": sum(v for _, v in kernels)}


def roofline(prof, n, nb, plan):
    """Roofline of the step's longest kernel.

    forward_kernel (direct): 6 FP32 VALU flop per particle x bin eval (the
      inner loop: two FMAs forming u, one table-index conversion, the FMA of
      the interpolant and the accumulating FMA); bound: FP32 VALU.
    binned_bucket: reads u and B (24 B/particle) and writes each contributing
      source's (W, phi) (8 B/particle); its run descriptors (4 B per non-empty
      run, ~40 sources per run here) are left out; bound: HBM.
    binned_subhist: reads the sorted (W, phi) (8 B/particle) and writes its
      per-item moments (2 x (M + 1) x 8 B per item); bound: HBM.
    """
    name, ms = prof["kernels"][0]
    t = ms * 1e-3
    info = plan.info()
    if name == "forward_kernel":
        achieved = FLOP_PER_EVAL_FWD * n * nb / t
        return {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": read_traffic(name),
                "kernel": name, "kernel_ms": ms, "flop_per_eval": FLOP_PER_EVAL_FWD}
    if name == "binned_bucket":
        nbytes = 32 * n
    elif name == "binned_subhist":
        items = n // 16384 + info["NK"]
        nbytes = 8 * n + 16 * (info["M"] + 1) * items
    else:
        nbytes = 0
    achieved = nbytes / t
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": achieved / PEAK_HBM, "traffic": read_traffic(name), "kernel": name,
            "kernel_ms": ms, "algorithmic_bytes": nbytes}


def step_roofline(n, ms):
    """The whole step against the HBM roofline of its essential traffic: every
    particle's u and B read once (24 B/particle); the factorised form's
    intermediate (sorted sources, run tables, moments) is not counted."""
    nbytes = 24 * n
    achieved = nbytes / (ms * 1e-3)
    return {"bound": "hbm", "essential_bytes": nbytes, "achieved": achieved / 1e9, "unit": "GB/s",
            "peak": PEAK_HBM / 1e9, "frac": achieved / PEAK_HBM, "ms": ms}


def oracle_checks(u, b, bins, S_step, contraction_samples):
    """The checker: the timed path's outputs against the oracle.  S_step is the
    default-method spectrum of all the step's particles: its max relative error
    per bin (bins above 1e-12 of the peak) against the reference's float terms
    summed in FP64 (oracle EXACT) and against the FP64 model; contraction_samples
    are sampled rows of the batched contraction (or None)."""
    import numpy as np

    from oracle import oracle as orc

    out = {}
    table = orc.ftable()
    _, ex = orc.sync_spectrum(u, b, bins, mode=orc.EXACT, table=table, want_fp64=True)
    _, f64 = orc.sync_spectrum(u, b, bins, mode=orc.FP64, table=table, want_fp64=True)
    live = ex > 1e-12 * ex.max()
    out["parity_max_rel_err"] = float(np.max(np.abs(S_step[live] / ex[live] - 1)))
    out["accuracy_max_rel_err_vs_fp64"] = float(np.max(np.abs(S_step[live] / f64[live] - 1)))
    out["reference_fp32_max_rel_err_vs_fp64"] = float(np.max(np.abs(ex[live] / f64[live] - 1)))
    if contraction_samples is not None:
        g, cb, f = contraction_samples["g"], contraction_samples["bins"], contraction_samples["f"]
        errs = {}
        for name, rows in contraction_samples["out"].items():
            worst = 0.0
            for i in range(f.shape[0]):
                fi = f[i]
                if name.endswith("_generated"):
                    # the generated row: x^p exp(-x / xc) / trapezoid norm, p and xc as floats
                    gd = g.astype(np.float64)
                    v = np.exp(np.float64(np.float32(contraction_samples["index"][i])) * np.log(gd)
                               - gd / np.float64(np.float32(contraction_samples["cutoff"][i])))
                    fi = (v / np.trapezoid(v, gd)).astype(np.float32)
                _, f64 = orc.sync_fromdist(g, fi, True, cb, mode=orc.FP64, want_fp64=True)
                lv = f64 > 1e-12 * f64.max()
                worst = max(worst, float(np.max(np.abs(rows[i][lv] / f64[lv] - 1))))
            errs[name] = worst
        out["contraction_max_rel_err_vs_fp64_32rows"] = errs
    return out


def cpu_baseline(u, b, bins, target_cpu_seconds):
    """The oracle's faithful restatement (per-thread float accumulators, FP64
    per-pair OmegaSync, log10f interpolation), on a bounded sample."""
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


def _timed(fn, min_seconds):
    """(calls, wall seconds) of repeated fn() over at least min_seconds."""
    k, t = 0, time.perf_counter()
    while True:
        fn()
        k += 1
        dt = time.perf_counter() - t
        if dt >= min_seconds:
            return k, dt


def cpu_baselines(u, b, bins, cells_sample, contraction_samples, seconds=2.0):
    """The oracle's faithful restatement of the reference (per-thread float
    accumulators, per-pair FP64 OmegaSync, log10f interpolation, the F table
    rebuilt on every FromDist call as the reference does), on bounded samples
    of each workload, on this host's cores."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as orc

    threads = orc.num_threads()
    out = {}
    # distribution form, 200 x 200 (one call builds the table, then sums 40k pairs)
    g = orc.logspace(1, 100, 200)
    f = orc.generator("plaw", [-2.0, 1.0, 100.0], g)
    db = orc.logspace(0.01, 1e7, 200)
    k, dt = _timed(lambda: orc.sync_fromdist(g, f, True, db, mode=orc.FAITHFUL, table=orc.ftable()), seconds)
    out["fromdist_200x200"] = {"value": dt / k * 1e6, "unit": "us/call", "cores": 1, "kind": "port",
                               "sample": f"{k} calls, each building the F table as the reference does"}
    # energy histogram (12 B per particle read)
    hb = orc.logspace(1e-2, 1e3, 200)
    m = min(u.shape[1], 1 << 22)
    k, dt = _timed(lambda: orc.energy_hist(u[:, :m], hb, True, True, mode=orc.FAITHFUL), seconds)
    out["energy_hist"] = {"value": HIST_BYTES_PER_PARTICLE * m * k / dt / 1e9, "unit": "GB/s",
                          "cores": threads, "kind": "port",
                          "sample": f"{k} histograms of the first {m} particles of the workload"}
    # weighted forward of the fit (the reference has no adjoint)
    from ragnar_amd import workloads

    m = 1 << 18
    w = workloads.fit_weights(u[:, :m])
    tab = orc.ftable()
    k, dt = _timed(lambda: orc.sync_spectrum(u[:, :m], b[:, :m], bins, w=w, mode=orc.FAITHFUL, table=tab),
                   seconds)
    out["fit_forward"] = {"value": m * len(bins) * k / dt, "unit": "evals/s", "cores": threads, "kind": "port",
                          "sample": f"forward only (the reference has no adjoint), first {m} particles "
                                    "of the headline workload with the fit's weights x 256 bins"}
    # batched cells: per-cell calls, as the reference would make them
    if cells_sample is not None:
        cu, cb, offs, total, ncells = cells_sample
        cbins = orc.logspace(1e-2, 1e7, 256)

        def cells():
            for c in range(len(offs) - 1):
                orc.sync_spectrum(cu[:, offs[c]:offs[c + 1]], cb[:, offs[c]:offs[c + 1]], cbins,
                                  mode=orc.FAITHFUL, table=orc.ftable())

        k, dt = _timed(cells, seconds)
        n_s = int(offs[-1] - offs[0])
        rate = n_s * 256 * k / dt
        out["cells_4096"] = {"value": total * 256 / rate * 1e3, "unit": "ms (extrapolated)", "cores": threads,
                             "kind": "port", "evals_per_s": rate,
                             "sample": f"the first {len(offs) - 1} of {ncells} cells ({n_s} particles), one "
                                       "call per cell with its table build, scaled to all particles"}
    # batched contraction: one FromDist call per row, rows spread over the threads
    if contraction_samples is not None:
        cg, cbb, cf = contraction_samples["g"], contraction_samples["bins"], contraction_samples["f"]

        def row(i):
            orc.sync_fromdist(cg, cf[i % cf.shape[0]], True, cbb, mode=orc.FAITHFUL, table=orc.ftable())

        with ThreadPoolExecutor(threads) as ex:
            k, dt = _timed(lambda: list(ex.map(row, range(threads * 2))), seconds)
        rows_per_s = threads * 2 * k / dt
        out["contraction_65536x512x512"] = {
            "value": 65536 / rows_per_s * 1e3, "unit": "ms (extrapolated)", "cores": threads, "kind": "port",
            "sample": f"{threads * 2 * k} FromDist calls (512 x 512, table built per call) on {threads} "
                      "threads, scaled to 65536 rows"}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU sharding is designed but not launched)"}))
        sys.exit(2)

    if args.summarize:
        if args.method == "compute":
            summarize(*args.summarize, "compute", tag="compute_bound",
                      title="forward pair by pair at 2^24 x 256, contraction at 65536 x 512 x 512, "
                            "inverse Compton at 1000^3 (bench.py --profile-compute)", source="--profile-compute")
        else:
            summarize(*args.summarize, args.method)
        return
    if args.summarize_compute:
        summarize_compute(*args.summarize_compute)
        return
    if args.summarize_bench:
        summarize_bench(*args.summarize_bench)
        return
    if args.profile:
        profile_run(args)
        return
    if args.profile_compute:
        profile_compute_run(args)
        return

    import numpy as np
    import torch

    from ragnar_amd import torch_ops as T
    from ragnar_amd import workloads

    torch.cuda.init()
    dev = "cuda"
    n, nb = args.particles, args.bins
    u_h, b_h = workloads.particles(n)
    bins = logspace(1e-2, 1e7, nb)
    U = torch.from_numpy(u_h).to(dev)
    B = torch.from_numpy(b_h).to(dev)
    table = T.default_table()
    plan = T.SyncPlan(table, bins)
    out = torch.empty(nb, device=dev)
    evals = n * nb

    progress("inputs resident; timing both methods")
    # both methods, then the step the library runs by default (method "auto")
    method_ms = {}
    for m in ("direct", "binned"):
        method_ms[m] = event_time(lambda: T.sync_spectrum(plan, U, B, out=out, method=m),
                                  max(5, args.steps // 4), args.warmup, torch)

    def step():
        T.sync_spectrum(plan, U, B, out=out)

    ms = event_time(step, args.steps, args.warmup, torch)
    value = evals / (ms * 1e-3)
    auto = stage_profile(T, plan, U, B, out, torch, reps=20)
    # what a plain device copy reaches here (1 GiB read + 1 GiB written): the
    # practical HBM ceiling the roofline fraction can be read against
    csrc = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
    cdst = torch.empty_like(csrc)
    copy_ms = event_time(lambda: cdst.copy_(csrc), 5, 2, torch)
    copy_gbps = 2 * csrc.numel() * 4 / (copy_ms * 1e-3) / 1e9
    del csrc, cdst
    auto_method = "binned" if auto["kernels"][0][0].startswith("binned") else "direct"

    # the default method's output, for the checker (CPU leg)
    S_step = T.sync_spectrum(plan, U, B).double().cpu().numpy()
    S_full = {m: T.sync_spectrum(plan, U, B, method=m).double().cpu().numpy() for m in ("direct", "binned")}
    live = S_full["direct"] > 1e-12 * S_full["direct"].max()
    methods_agree = float(np.max(np.abs(S_full["binned"][live] / S_full["direct"][live] - 1)))

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
                   "particles": n, "photon_bins": nb, "method": auto_method},
        "roofline": dict(roofline(auto, n, nb, plan), step=step_roofline(n, ms), device_copy_GBps=copy_gbps),
        "methods_ms": method_ms,
        "methods_max_rel_diff": methods_agree,
        "stages_ms": {name: t for name, t in auto["kernels"]},
    }

    rf = res["roofline"]
    if rf.get("unit") == "GB/s":
        rf["frac_of_device_copy"] = rf["achieved"] / copy_gbps
    progress(f"headline step {ms:.4f} ms")
    extras = {}
    if not args.no_extras:
        # forward + adjoint (differentiable fit step at the same size)
        w = torch.from_numpy(workloads.fit_weights(u_h)).to(dev)
        target = T.sync_spectrum(plan, U, B, w).clamp_min(1e-30).log()

        def fwd_adj():
            S = T.sync_spectrum(plan, U, B, w)
            r = S.clamp_min(1e-30).log() - target
            gbar = torch.where(S > 0, 2.0 * r / S.clamp_min(1e-30), torch.zeros_like(S))
            T.sync_adjoint(plan, U, B, w, None, gbar, need=("w", "u", "b"))

        # the first ~60 forward + adjoint calls of a process run up to 12%
        # slower (445 -> 396 us per call over calls 11-70 on one box); the
        # untimed calls reach the steady state
        ms_fa = event_time(fwd_adj, max(3, args.steps // 2), 60, torch)
        # the same weighted forward alone: the ratio compares like with like
        ms_fw = event_time(lambda: T.sync_spectrum(plan, U, B, w), max(3, args.steps // 2), 3, torch)
        gones = torch.ones(nb, device=dev)
        adj_ms = {m: event_time(lambda: T.sync_adjoint(plan, U, B, w, None, gones, need=("w", "u", "b"),
                                                       method=m), 10, 3, torch)
                  for m in ("direct", "binned")}
        extras["fwd_adj"] = {"ms": ms_fa, "evals_per_s": evals / (ms_fa * 1e-3), "ms_fwd_weighted": ms_fw,
                             "ratio_to_fwd": ms_fa / ms_fw, "ratio_to_unweighted_fwd": ms_fa / ms,
                             "adjoint_ms_by_method": adj_ms}
        progress("fwd+adjoint done; fit at 2^26")
        # differentiable fit at 2^26 particles (fwd + adjoint of the reweighting loss)
        try:
            extras["fit_2p26"] = bench_fit(T, workloads, torch, table, bins)
        except Exception as exc:  # reported, the headline line still prints
            extras["fit_2p26"] = {"error": repr(exc)}
        progress("fit done; histogram, FromDist, IC")
        # energy histogram (HBM bound: 12 B per particle)
        hbins = logspace(1e-2, 1e3, 200)
        hplan = T.HistPlan(hbins, True)
        ms_h = event_time(lambda: T.energy_hist(hplan, U, True), args.steps, 2, torch)
        extras["energy_hist"] = {"ms": ms_h, "GB_per_s": HIST_BYTES_PER_PARTICLE * n / (ms_h * 1e-3) / 1e9,
                                 "frac_hbm_peak": HIST_BYTES_PER_PARTICLE * n / (ms_h * 1e-3) / PEAK_HBM}
        # distribution form, 200 x 200 (launch-latency bound)
        g = logspace(1, 100, 200)
        f = generator("plaw", [-2.0, 1.0, 100.0], g)
        dplan = T.SyncPlan(table, logspace(0.01, 1e7, 200))
        G, Fd = torch.from_numpy(g).to(dev), torch.from_numpy(f).to(dev)
        ms_d = event_time(lambda: T.sync_fromdist(dplan, G, Fd, True), 50, 5, torch)
        extras["fromdist_200x200"] = {"us": ms_d * 1e3}
        # inverse-Compton spectrum, the reference test's 200 x 200 x 200 triples
        gi = logspace(1e3, 1e7, 200)
        si = logspace(1e-11, 1e-7, 200)
        ti = [torch.from_numpy(a).to(dev) for a in (gi, generator("plaw", [-1.5, 1e3, 1e7], gi), si,
                                                     generator("delta", [1e-8, 1e-9], si),
                                                     logspace(1e3, 1e7, 200))]
        ms_i = event_time(lambda: T.ic_spectrum(ti[0], ti[1], True, ti[2], ti[3], ti[4]), 50, 5, torch)
        extras["ic_200x200x200"] = {"us": ms_i * 1e3, "triples_per_s": 200 ** 3 / (ms_i * 1e-3)}
        progress("cells and contraction")
        for name, fn in (("cells_4096", lambda: bench_cells(T, workloads, torch, table)),
                         ("contraction_65536x512x512",
                          lambda: bench_contraction(T, workloads, torch, table))):
            try:
                extras[name] = fn()
            except Exception as exc:  # reported, the headline line still prints
                extras[name] = {"error": repr(exc)}
    samples = None
    c = extras.get("contraction_65536x512x512")
    if isinstance(c, tuple):
        extras["contraction_65536x512x512"], samples = c
    cells_sample = None
    c = extras.get("cells_4096")
    if isinstance(c, tuple):
        extras["cells_4096"], cells_sample = c
    res["extras"] = extras

    # CPU leg: the oracle times the reference algorithm and checks the outputs
    if not args.no_cpu:
        progress("CPU leg: oracle checks and baselines")
        checks = oracle_checks(u_h, b_h, bins, S_step, samples)
        for key in ("parity_max_rel_err", "accuracy_max_rel_err_vs_fp64", "reference_fp32_max_rel_err_vs_fp64"):
            res[key] = checks[key]
        for name, err in checks.get("contraction_max_rel_err_vs_fp64_32rows", {}).items():
            extras["contraction_65536x512x512"][name]["max_rel_err_vs_fp64_32rows"] = err
        res["cpu_baseline"] = cpu_baseline(u_h, b_h, bins, args.cpu_seconds)
        extras["cpu_baselines"] = cpu_baselines(u_h, b_h, bins, cells_sample, samples)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
