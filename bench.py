"""Throughput benchmark of the MI355X ERI engine (one JSON line on stdout).

Workload: the 64-atom, all-classes-up-to-d system (jittered 3-bohr cubic
grid; per atom s(6) s(3) p(3) core and s(1) p(1) d(1) valence shells), every
unique shell quartet passing a Schwarz screen of 1e-12, FP64.  One step is
one pass over all of those quartets through the device-resident batched path
(simint_hip_compute_eri_prefix), integrals written to HBM (a reused output
buffer: the full set is ~0.85 TB of integrals).  Inputs are resident before
the timed region.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--natom 64] [--no-cpu]

Besides the metric the line carries:
  roofline      for the kernel with the largest share of device time:
                algorithmic FP64 FLOPs per launch (FMA = 2; per quartet
                K (f_setup + f_vrr + f_boys + bcont) + f_hrr, the reference
                generator's operation counts) over its mean launch time,
                measured with HIP events on the library stream in the timed
                region, against the 78.6 TFLOP/s FP64 peak;
  cpu_baseline  the CPU oracle (a C restatement of the reference algorithm,
                one core) on a random sample of the same quartets.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "simint-generator_amd")

METRIC = "shell quartets/s at ≤1e-10 abs err, mixed vs FP64, 1 GPU; p50/p99 latency"
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` as recorded in the rocprofv3 summaries
    under profiles/ (FETCH_SIZE + WRITE_SIZE, calibrated against a copy
    kernel), or None if that kernel has not been measured."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for f in sorted(os.listdir(pdir)):
        if not f.endswith(".md"):
            continue
        for line in open(os.path.join(pdir, f)):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return float(parts[2])
    return None


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print("[bench %.1fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


T_START = time.perf_counter()
BATCH_STREAMS = 16   # side streams the batched stream's classes are spread over (4 and 8 measured equal, 16 0-2 % faster)


def kernel_symbol(info, cls):
    """Name of the kernel the class runs now (the family selected at run
    time; the split kernels carry their group count)."""
    import simint
    name = "_".join("spdfghik"[x] for x in cls)
    fam = simint.lib().simint_hip_class_kernel(*cls)
    if fam == 1:
        return "eri_lpq_%s" % name
    if fam in (3, 4):
        return "eri_" + info[name]["split_kernels"][fam - 3]
    return "eri_wpq_%s" % name


def list_kernel_symbol(cls):
    """Name of the quartet-list kernel the class runs now."""
    import simint
    name = "_".join("spdfghik"[x] for x in cls)
    return ("eri_lpql_%s" if simint.lib().simint_hip_class_list_kernel(*cls) == 1 else "eri_wpql_%s") % name


def cpu_baseline(ss, seconds=12.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    sa = oracle.ShellArrays(ss.shells)
    oracle.normalize(sa)
    done = 0
    t_used = 0.0
    chunk = 200
    seed = 0
    while t_used < seconds:
        qs = ss.sample(chunk, seed=seed)
        seed += 1
        t0 = time.perf_counter()
        oracle.eri(sa, qs)
        t_used += time.perf_counter() - t0
        done += len(qs)
        if t_used < seconds / 4:
            chunk = min(chunk * 2, 20000)
    return {"value": done / t_used, "unit": "shell quartets/s", "cores": 1, "kind": "port",
            "sample": "%d quartets drawn uniformly from the screened 64-atom set, %.1f s of oracle time"
                      % (done, t_used)}


def batched_leg(steps, warmup, nquartet=10 ** 6):
    """The batched stream: 1e6 quartets over every class up to f (pool of
    4096 shells, depths 1-6), one graph launch per pass of 256 quartet-list
    class launches, inputs and the 5 GB output resident in HBM."""
    import torch
    import batch
    import simint
    import workloads
    info = json.load(open(os.path.join(PKG, "csrc", "gen", "classes.json")))
    t0 = time.perf_counter()
    shells, q = workloads.config3(nquartet=nquartet, seed=0)
    basis = simint.Basis(shells)
    b = batch.QuartetBatch(basis, q)
    setup = time.perf_counter() - t0
    progress("batched stream set up (%.1f s)" % setup)
    flops = 0.0
    for g in b.groups:
        c = info["_".join("spdfghik"[x] for x in g.cls)]
        flops += g.prim_quartets * (c["f_setup"] + c["f_vrr"] + c["f_boys"] + c["bcont"]) + g.n * c["f_hrr"]
    out = torch.empty(b.out_doubles, dtype=torch.float64, device="cuda")
    stream = torch.cuda.ExternalStream(simint.lib().simint_hip_stream())
    # one eager pass with per-class events (class breakdown)
    ev = []

    def rec(g, start):
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream)
        ev.append((g, e))
    b.run(out.data_ptr(), events=rec)
    simint.synchronize()
    progress("batched: eager pass with per-class events done")
    per = []
    for k in range(0, len(ev), 2):
        g, e0 = ev[k]
        _, e1 = ev[k + 1]
        per.append((e0.elapsed_time(e1), g))
    gr = b.capture(out.data_ptr(), streams=BATCH_STREAMS)
    for _ in range(warmup):
        gr.launch()
    simint.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t1 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        gr.launch()
    e1.record(stream)
    simint.synchronize()
    wall = time.perf_counter() - t1
    dev_s = e0.elapsed_time(e1) / 1000.0
    gr.free()
    per.sort(key=lambda x: -x[0])
    res = {"workload": "%d quartets of 4 independent uniform draws from a pool of 4096 shells (1024 per AM "
                       "0..3, depth U{1..6}, exponents log-uniform [0.05, 1e4], 30-bohr cube), all 256 classes, "
                       "FP64" % nquartet,
           "value": nquartet * steps / wall, "unit": "shell quartets/s",
           "ms_per_pass": 1000.0 * wall / steps, "device_ms_per_pass": 1000.0 * dev_s / steps,
           "tflops": flops * steps / dev_s / 1e12, "frac_fp64_peak": flops * steps / dev_s / PEAK_FP64,
           "flops_per_pass": flops, "output_bytes": 8.0 * b.out_doubles,
           "launch": "one hipGraph per pass holding %d quartet-list class launches on %d forked streams"
                     % (len(b.groups), BATCH_STREAMS),
           "setup_s": setup,
           "top_classes": [{"class": "_".join("spdfghik"[x] for x in g.cls), "ms": ms, "quartets": g.n,
                            "kernel": list_kernel_symbol(g.cls)}
                           for ms, g in per[:8]]}
    fam = {}
    for ms, g in per:
        k = "wpql" if "wpql" in list_kernel_symbol(g.cls) else "lpql"
        fam[k] = fam.get(k, 0.0) + ms
    res["eager_class_ms"] = {"total": sum(ms for ms, _ in per), "by_family": fam,
                             "note": "classes one after another on one stream, per-class events"}
    b.free()
    return res


def mixed_leg(natom=64, threshold=1e-4):
    """A system of the 64-atom recipe (64 atoms by default, 256 on request): every screened quartet
    (Schwarz 1e-12) in FP64, then again with quartets whose bound is below
    `threshold` run through the FP32-VRR/HRR kernels; one pass each (timed
    on the library stream), then an untimed chunk-by-chunk comparison of the
    two outputs for the max absolute error per class."""
    import torch
    import quartets
    import simint
    import workloads
    shells = workloads.config2(natom=natom, seed=0)
    t0 = time.perf_counter()
    ss = quartets.ScreenedSet(shells, tol=1e-12)
    plan64 = ss.launches
    ss.mixed = threshold
    ss._plan()
    planm = ss.launches
    setup = time.perf_counter() - t0
    progress("%d-atom set: %d quartets (setup %.1f s)" % (natom, ss.total_quartets, setup))
    out = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
    ref = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
    stream = torch.cuda.ExternalStream(simint.lib().simint_hip_stream())

    def timed(plan):
        ss.launches = plan
        ss.run_pass(out.data_ptr(), chunk=0)      # first use of the kernels
        simint.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        e0.record(stream)
        ss.run_pass(out.data_ptr())
        e1.record(stream)
        simint.synchronize()
        return time.perf_counter() - t, e0.elapsed_time(e1) / 1000.0
    w64, d64 = timed(plan64)
    progress("FP64 pass %.1f s" % w64)
    wm, dm = timed(planm)
    progress("mixed pass %.1f s" % wm)
    nmixed = sum(ln.nq for ln in planm if ln.precision == 1)
    # error of the mixed pass against the FP64 pass, chunk by chunk (a
    # chunk holds one class; FP64 launches of both plans are identical)
    errs = {}
    t_last = time.perf_counter()
    ss.launches = planm
    for c in range(ss.nchunk):
        if time.perf_counter() - t_last > 20.0:
            progress("error check: chunk %d of %d" % (c, ss.nchunk))
            t_last = time.perf_counter()
        if not any(ln.precision == 1 for ln in ss.chunk_launches(c)):
            continue
        n = ss.chunk_doubles(c)
        ss.launches = plan64
        ss.run_pass(ref.data_ptr(), chunk=c)
        ss.launches = planm
        ss.run_pass(out.data_ptr(), chunk=c)
        simint.synchronize()
        e = float((out[:n] - ref[:n]).abs().max())
        cls = ss.chunk_launches(c)[0].cls
        name = "_".join("spdfghik"[x] for x in cls)
        errs[name] = max(errs.get(name, 0.0), e)
    ss.launches = planm
    ss.free()
    return {"workload": "%d-atom grid of the 64-atom recipe (%d shells), Schwarz 1e-12; quartets with "
                        "sqrt((ab|ab)(cd|cd)) < %g (scaled per class by its HRR sides) run the FP32 VRR"
                        % (natom, len(shells), threshold),
            "quartets": ss.total_quartets, "fp32_quartets": nmixed,
            "mixed_value": ss.total_quartets / wm, "fp64_value": ss.total_quartets / w64,
            "speedup": w64 / wm, "mixed_device_s": dm, "fp64_device_s": d64, "unit": "shell quartets/s",
            "max_abs_err_vs_fp64": max(errs.values()) if errs else 0.0,
            "max_abs_err_per_class": errs, "bound": 1e-10, "setup_s": setup}


def deriv_leg(natom=64, tol=1e-12):
    """First derivatives of every screened quartet of the 64-atom system
    (12 values per Cartesian), one pass through a 1 GiB chunk buffer on the
    device (simint_hip_compute_eri_deriv1_range per launch)."""
    import torch
    import quartets
    import simint
    import workloads
    shells = workloads.config2(natom=natom, seed=0)
    ss = quartets.ScreenedSet(shells, tol=tol, out_doubles=1 << 27, deriv=1)
    work = torch.empty(ss.deriv_worksize(), dtype=torch.float64, device="cuda")
    out = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
    stream = torch.cuda.ExternalStream(simint.lib().simint_hip_stream())
    ss.run_pass(out.data_ptr(), chunk=0, work_ptr=work.data_ptr())
    simint.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t = time.perf_counter()
    e0.record(stream)
    t_note = t
    for c in range(ss.nchunk):
        ss.run_pass(out.data_ptr(), chunk=c, work_ptr=work.data_ptr())
        if time.perf_counter() - t_note > 20.0:
            progress("deriv: chunk %d of %d" % (c, ss.nchunk))
            t_note = time.perf_counter()
    e1.record(stream)
    simint.synchronize()
    wall = time.perf_counter() - t
    res = {"workload": "first derivatives of the %d-atom system, Schwarz %g" % (natom, tol),
           "quartets": ss.total_quartets, "value": ss.total_quartets / wall, "unit": "shell quartets/s",
           "pass_s": wall, "device_s": e0.elapsed_time(e1) / 1000.0,
           "output_GBps": ss.total_bytes / wall / 1e9}
    ss.free()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--natom", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched all-classes stream")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-deriv", action="store_true", help="skip the first-derivative leg")
    ap.add_argument("--mixed-natom", type=int, default=64,
                    help="atoms of the mixed-precision leg (the 256-atom system takes ~7 minutes)")
    ap.add_argument("--tuning", default=None, help="apply this kernel-family tuning file before running")
    ap.add_argument("--only-class", default=None,
                    help="profiling aid: run only this class's launches (e.g. d_p_d_p)")
    ap.add_argument("--force-family", choices=["lpq", "wpq", "lpqs_a", "lpqs_b"], default=None,
                    help="profiling aid: use this kernel family for every class that has it")
    ap.add_argument("--copy-gib", type=float, default=None,
                    help="profiling aid: run one device copy of this many GiB and exit "
                         "(calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(shell quartets are independent; sharding them over GPUs is described in DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_gib:
        n = int(args.copy_gib * (1 << 30)) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 8 * n}))
        return
    sys.path.insert(0, PKG)
    import quartets
    import simint
    import workloads

    shells = workloads.config2(natom=args.natom, seed=0)
    if args.tuning:
        simint.apply_tuning(args.tuning)
    if args.force_family:
        fam = simint.FAMILIES[args.force_family]
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    for d in range(3):
                        simint.lib().simint_hip_set_class_kernel(a, b, c, d, fam)
    t_setup = time.perf_counter()
    ss = quartets.ScreenedSet(shells, tol=1e-12)
    t_setup = time.perf_counter() - t_setup
    progress("64-atom set: %d quartets in %d launches (setup %.1f s)" % (ss.total_quartets, len(ss.launches),
                                                                         t_setup))
    if args.only_class:
        ss.launches = [ln for ln in ss.launches if quartets.simint_class_name(ln.cls) == args.only_class]
        ss.total_quartets = sum(ln.nq for ln in ss.launches)
        ss.total_flops = sum(ln.flops for ln in ss.launches)
        ss.total_bytes = sum(8.0 * ln.nout for ln in ss.launches)
    out = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
    stream = torch.cuda.ExternalStream(simint.lib().simint_hip_stream())

    for _ in range(args.warmup):
        ss.run_pass(out.data_ptr())
    simint.synchronize()
    torch.cuda.synchronize()
    progress("warmup done")

    events = []

    def rec(ln, start):
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream)
        events.append((ln, start, e))

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ss.run_pass(out.data_ptr(), events=rec)
    simint.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    progress("timed steps done: %.1f ms per step" % ms_per_step)
    value = ss.total_quartets * args.steps / elapsed

    # per-class device time from the events
    per = {}
    for k in range(0, len(events), 2):
        ln, _, e0 = events[k]
        _, _, e1 = events[k + 1]
        ms = e0.elapsed_time(e1)
        c = per.setdefault(ln.cls, {"ms": 0.0, "flops": 0.0, "bytes": 0.0, "launches": 0, "quartets": 0})
        c["ms"] += ms
        c["flops"] += ln.flops
        c["bytes"] += 8.0 * ln.nout
        c["launches"] += 1
        c["quartets"] += ln.nq
    dom = max(per, key=lambda k: per[k]["ms"])
    d = per[dom]
    mean_s = d["ms"] / 1000.0 / d["launches"]
    flops_per_launch = d["flops"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    t_flop = flops_per_launch / PEAK_FP64
    t_byte = bytes_per_launch / PEAK_HBM
    if t_flop >= t_byte:
        achieved = flops_per_launch / mean_s / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / (PEAK_FP64 / 1e12), "traffic": None}
    else:
        achieved = bytes_per_launch / mean_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": achieved / (PEAK_HBM / 1e9), "traffic": None}
    kname = kernel_symbol(ss.info, dom)
    roof["traffic"] = measured_traffic(kname)
    roof.update({"kernel": kname,
                 "algorithmic_bytes_per_launch": bytes_per_launch,
                 "compute": "FP64 VALU (no MFMA); peak is the FP64 vector rate",
                 "launches": d["launches"], "mean_launch_ms": 1000.0 * mean_s,
                 "flops_per_launch": flops_per_launch, "share_of_device_time":
                     d["ms"] / sum(v["ms"] for v in per.values())})
    whole = {"tflops": ss.total_flops * args.steps / elapsed / 1e12,
             "frac_fp64_peak": ss.total_flops * args.steps / elapsed / PEAK_FP64,
             "output_GBps": ss.total_bytes * args.steps / elapsed / 1e9}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "shell quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "%d-atom jittered 3-bohr grid, s6 s3 p3 s1 p1 d1 per atom, all unique "
                               "quartets with Schwarz bound >= 1e-12, plain FP64 (l <= 2)" % args.natom,
                   "natom": args.natom, "nshell": len(shells), "quartets_per_step": ss.total_quartets,
                   "launches_per_step": len(ss.launches), "path": "device-resident batched",
                   "setup_s": t_setup},
        "roofline": roof,
        "whole_job": whole,
    }
    if not args.no_batched and not args.only_class:
        ss.free()
        line["batched"] = batched_leg(args.steps, args.warmup)
        progress("batched leg done")
    if not args.no_mixed and not args.only_class:
        line["mixed"] = mixed_leg(args.mixed_natom)
        progress("mixed leg done")
    if not args.no_deriv and not args.only_class:
        line["deriv1"] = deriv_leg()
        progress("deriv leg done")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ss)
        progress("cpu baseline done")
    # device time by (class, primitive quartets per quartet)
    perk = {}
    for k in range(0, len(events), 2):
        ln, _, e0 = events[k]
        _, _, e1 = events[k + 1]
        key = (ln.cls, int(ln.bra.nprim * ln.ket.nprim))
        c = perk.setdefault(key, [0.0, 0])
        c[0] += e0.elapsed_time(e1)
        c[1] += ln.nq
    line["top_class_K"] = [{"class": quartets.simint_class_name(k[0]), "K": k[1], "ms_per_step": v[0] / args.steps,
                            "quartets_per_step": v[1] // args.steps}
                           for k, v in sorted(perk.items(), key=lambda kv: -kv[1][0])[:12]]
    top = sorted(per.items(), key=lambda kv: -kv[1]["ms"])[:8]
    line["top_classes"] = [{"class": quartets.simint_class_name(k), "ms_per_step": v["ms"] / args.steps,
                            "tflops": v["flops"] / (v["ms"] / 1000.0) / 1e12,
                            "quartets_per_step": v["quartets"] // args.steps,
                            "kernel": kernel_symbol(ss.info, k)} for k, v in top]
    print(json.dumps(line))
    if args.no_batched or args.only_class:
        ss.free()


if __name__ == "__main__":
    main()
