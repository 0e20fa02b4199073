"""Throughput benchmark of the MI355X ERI engine (one JSON line on stdout).

Main workload (`value`): the 256-atom system (jittered 3-bohr 8x8x4 grid; per
atom s(6) s(3) p(3) core and s(1) p(1) d(1) valence shells; 1536 shells),
every unique shell quartet passing a Schwarz screen of 1e-12 (1.2e11
quartets), through the device-resident path in mixed precision: quartets
whose Schwarz bound is below 1e-4 (divided by 400 with one HRR side and by
1e5 with two) run the VRR in FP32, with the Boys function, prefactors,
contraction and HRR in FP64; the others run FP64.  The same set through the
plain FP64 path is timed beside it, and the maximum absolute error of the
mixed path is measured per class against the FP64 GPU path (every quartet
of a strided sample of output chunks) and against the CPU oracle (a located
sample of FP32 quartets).

A step is one share of that quartet set: the launch plan is cut into --steps
consecutive shares of about equal quartet count, so --steps steps are one
full pass whatever --steps is.  Integrals go to a reused 4 GiB output buffer
in HBM (the whole set is ~70 TB of integrals); inputs are resident before
the timed region, which is bracketed by device synchronisations.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--no-batched]
                    [--no-serving] [--no-deriv] [--no-fp64-64]

Besides the metric the line carries:
  fp64 / mixed      quartets/s of both paths, their ratio, per-class error;
  roofline          the kernel with the largest share of device time in the
                    timed mixed pass: algorithmic FP64 FLOPs (FMA = 2; per
                    quartet K (f_setup + f_vrr + f_boys + bcont) + f_hrr, the
                    reference generator's operation counts) or output bytes,
                    whichever bounds it, over its mean launch time (HIP events
                    on the library stream in the timed region);
  whole_job         sum over launches of max(F64/P64 + F32/P32, bytes/BW)
                    over the measured time (the job's roofline fraction);
  serving           a bounded Poisson trace through the continuous-batching
                    server: p50/p99 latency and quartets/s;
  batched           1e6 quartets of all 256 classes up to f in one graph;
  deriv1            first derivatives of the 64-atom system;
  fp64_64atom       the 64-atom system in FP64 (the all-classes-to-d config),
                    the same pass with the Coulomb/exchange digestion
                    after every launch (fp64_64atom.jk), and with the
                    primitive Schwarz screen at 1e-13 (fp64_64atom.prim_screen,
                    with its largest change over every integral);
  dropin_pcie       the host-buffer drop-in call on one large batch (its
                    rate includes the PCIe copies, which `value` does not);
  cpu_baseline      the CPU oracle (a C restatement of the reference
                    algorithm) on every host core the box gives this job,
                    on a sample of the same quartets.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "simint-generator_amd")

METRIC = "shell quartets/s at ≤1e-10 abs err, mixed vs FP64, 1 GPU; p50/p99 latency"
PEAK_FP64 = 78.6e12
PEAK_FP32 = 157.3e12
PEAK_HBM = 8.0e12
T_START = time.perf_counter()
BATCH_STREAMS = 16   # side streams the batched stream's classes are spread over


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print("[bench %.1fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


def cname(cls):
    return "_".join("spdfghik"[x] for x in cls)


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


def kernel_symbol(info, cls, precision=0):
    """Name of the kernel a range launch of the class runs now (the family
    selected at run time; the split kernels carry their group count)."""
    import simint
    name = cname(cls)
    fam = simint.lib().simint_hip_class_kernel(*cls)
    if fam == 1:
        return ("eri_lpqm_%s" if precision == 1 else "eri_lpq_%s") % name
    if fam in (3, 4):
        return "eri_" + info[name]["split_kernels"][fam - 3]
    if fam in (5, 6):
        return "eri_" + info[name]["split_kernels_f"][fam - 5]
    return "eri_wpq_%s" % name


def list_kernel_symbol(cls):
    import simint
    return ("eri_lpql_%s" if simint.lib().simint_hip_class_list_kernel(*cls) == 1 else "eri_wpql_%s") % cname(cls)


def shares(launches, k):
    """The plan cut into k consecutive shares of about equal quartet count."""
    tot = max(sum(ln.nq for ln in launches), 1)
    out = [[] for _ in range(k)]
    acc = 0
    for ln in launches:
        out[min(k - 1, acc * k // tot)].append(ln)
        acc += ln.nq
    return out


def launch_split_flops(info, ln):
    """(FP64 flops, FP32 flops) of a range launch: the mixed kernels run
    the VRR in FP32, everything else FP64."""
    c = info[cname(ln.cls)]
    K = ln.bra.nprim * ln.ket.nprim
    if ln.precision == 1:
        f32 = float(ln.nq) * K * c["f_vrr"]
        return ln.flops - f32, f32
    return ln.flops, 0.0


class Timer:
    """Per-launch HIP events on the library stream."""

    def __init__(self, stream):
        self.stream = stream
        self.ev = []

    def __call__(self, ln, start):
        import torch
        e = torch.cuda.Event(enable_timing=True)
        e.record(self.stream)
        self.ev.append((ln, e))

    def per_launch(self):
        return [(self.ev[k][0], self.ev[k][1].elapsed_time(self.ev[k + 1][1]) / 1000.0)
                for k in range(0, len(self.ev), 2)]


def timed_pass(ss, plan, steps, warmup, out, stream):
    """Warm up on the first `warmup` shares, then time all `steps` shares of
    the plan; returns (wall s, device s, per-launch seconds)."""
    import torch
    import simint
    ss.launches = plan
    sh = shares(plan, steps)
    for s in range(min(warmup, steps)):
        ss.run_launches(out.data_ptr(), sh[s])
    simint.synchronize()
    torch.cuda.synchronize()
    tm = Timer(stream)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for s in range(steps):
        ss.run_launches(out.data_ptr(), sh[s], events=tm)
    e1.record(stream)
    simint.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, e0.elapsed_time(e1) / 1000.0, tm.per_launch()


def error_check(ss, plan64, planm, out, ref, max_chunks=96, per_class=40, seed=0):
    """Max absolute error of the mixed path per class: (a) against the FP64
    GPU path over every quartet of the sampled output chunks (the first two
    chunks of each class holding FP32 quartets plus a stride over all such
    chunks, at most max_chunks), (b) against the CPU oracle on `per_class`
    located FP32 quartets of each class drawn from those chunks."""
    import numpy as np
    import oracle
    import simint
    ss.launches = planm
    fp32_chunks = {}
    for ln in planm:
        if ln.precision == 1:
            fp32_chunks.setdefault(ln.chunk, set()).add(ln.cls)
    chunks = sorted(fp32_chunks)
    pick = set()
    seen = {}
    for c in chunks:
        for cls in fp32_chunks[c]:
            if seen.get(cls, 0) < 2:
                pick.add(c)
                seen[cls] = seen.get(cls, 0) + 1
    stride = max(1, len(chunks) // max(1, max_chunks - len(pick)))
    pick |= set(chunks[::stride])
    pick = sorted(pick)
    dev_err = {}
    probes = {}
    mixed_launches = [ln for c in pick for ln in ss.chunk_launches(c) if ln.precision == 1]
    bycls = {}
    for ln in mixed_launches:
        bycls.setdefault(ln.cls, []).append(ln)
    for k, (cls, lns) in enumerate(sorted(bycls.items())):
        for c, off, nc, q in ss.located_sample(per_class, seed=seed + k, launches=lns):
            probes.setdefault(c, []).append((off, nc, q, cls))
    got = []
    t_note = time.perf_counter()
    for i, c in enumerate(pick):
        if time.perf_counter() - t_note > 20.0:
            progress("error check: chunk %d of %d" % (i, len(pick)))
            t_note = time.perf_counter()
        n = ss.chunk_doubles(c)
        ss.launches = plan64
        ss.run_pass(ref.data_ptr(), chunk=c)
        ss.launches = planm
        ss.run_pass(out.data_ptr(), chunk=c)
        simint.synchronize()
        for ln in ss.chunk_launches(c):
            if ln.precision != 1:
                continue
            # the FP32 launch's own region of the chunk
            lo = np.zeros_like(ln.nket) if ln.lo is None else ln.lo
            nc = ln.nout // ln.nq
            a = int(ln.off[0])
            b = int(ln.off[-1]) + (int(ln.nket[-1]) - int(lo[-1])) * nc
            e = float((out[a:b] - ref[a:b]).abs().max()) if b > a else 0.0
            dev_err[cname(ln.cls)] = max(dev_err.get(cname(ln.cls), 0.0), e)
        for off, nc, q, cls in probes.get(c, []):
            got.append((cls, q, out[off:off + nc].cpu().numpy()))
    shells = ss.shells
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    orc_err = {}
    for cls in sorted({g[0] for g in got}):
        qs = [g[1] for g in got if g[0] == cls]
        vals = np.concatenate([g[2] for g in got if g[0] == cls])
        r, _, _ = oracle.eri(sa, qs)
        orc_err[cname(cls)] = float(np.abs(vals - r).max())
    ss.launches = planm
    cover = sum(ln.nq for c in pick for ln in ss.chunk_launches(c) if ln.precision == 1)
    return dev_err, orc_err, {"chunks_compared": len(pick), "fp32_quartets_compared": int(cover),
                              "oracle_quartets": len(got)}


def mixed_system_leg(args):
    """The metric: the 256-atom system, mixed precision vs plain FP64."""
    import torch
    import quartets
    import simint
    import workloads
    shells = workloads.grid_molecule(natom=args.natom, seed=0)
    t0 = time.perf_counter()
    ss = quartets.ScreenedSet(shells, tol=1e-12)
    plan64 = ss.launches
    ss.mixed = args.mixed_threshold
    ss._plan()
    planm = ss.launches
    if args.only_class:
        plan64 = [ln for ln in plan64 if cname(ln.cls) == args.only_class]
        planm = [ln for ln in planm if cname(ln.cls) == args.only_class]
    setup = time.perf_counter() - t0
    nq = sum(ln.nq for ln in planm)
    progress("%d-atom set: %d quartets, %d mixed launches, %d FP64 launches (setup %.1f s)"
             % (args.natom, nq, len(planm), len(plan64), setup))
    out = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
    stream = torch.cuda.ExternalStream(simint.lib().simint_hip_stream())
    wm, dm, perm = timed_pass(ss, planm, args.steps, args.warmup, out, stream)
    progress("mixed: %d steps in %.1f s" % (args.steps, wm))
    w64, d64, per64 = timed_pass(ss, plan64, args.steps, args.warmup, out, stream)
    progress("FP64: %d steps in %.1f s" % (args.steps, w64))
    res = {"wall_mixed": wm, "wall_fp64": w64, "dev_mixed": dm, "dev_fp64": d64, "quartets": nq,
           "setup_s": setup, "per_launch": perm, "per_launch64": per64,
           "fp32_quartets": sum(ln.nq for ln in planm if ln.precision == 1)}
    if not args.no_errcheck:
        ref = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
        dev_err, orc_err, cov = error_check(ss, plan64, planm, out, ref)
        progress("error check done")
        res.update({"dev_err": dev_err, "orc_err": orc_err, "err_cover": cov})
        del ref
    res["info"] = ss.info
    res["ss"] = ss
    return res


def roofline_of(per, info):
    """Roofline object for the kernel with the largest device time among
    per-launch timings.  Launches are grouped by the kernel they run: the
    FP32-VRR launches of a class whose selected kernel has no mixed variant
    run the same FP64 kernel as its FP64 launches and count with them."""
    agg = {}
    for ln, s in per:
        key = kernel_symbol(info, ln.cls, ln.precision)
        a = agg.setdefault(key, {"s": 0.0, "f64": 0.0, "f32": 0.0, "bytes": 0.0, "n": 0, "cls": ln.cls,
                                 "mixed": False})
        a["mixed"] = a["mixed"] or (ln.precision == 1 and key.startswith("eri_lpqm_"))
        f64, f32 = launch_split_flops(info, ln)
        a["s"] += s
        a["f64"] += f64
        a["f32"] += f32
        a["bytes"] += 8.0 * ln.nout
        a["n"] += 1
    tot = sum(a["s"] for a in agg.values())
    key = max(agg, key=lambda k: agg[k]["s"])
    a = agg[key]
    mean_s = a["s"] / a["n"]
    f64, f32, by = a["f64"] / a["n"], a["f32"] / a["n"], a["bytes"] / a["n"]
    t_flop = f64 / PEAK_FP64 + f32 / PEAK_FP32
    t_byte = by / PEAK_HBM
    kname = key
    if t_flop >= t_byte:
        # FP64 vector (VALU) bound: no MFMA is involved; "mfma" is the
        # compute-bound label of the line format
        ach = (f64 + f32) / mean_s / 1e12
        pk = (f64 + f32) / (t_flop * 1e12)
        roof = {"bound": "mfma", "achieved": ach, "peak": pk, "unit": "TFLOP/s", "frac": ach / pk}
    else:
        ach = by / mean_s / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": ach / (PEAK_HBM / 1e9)}
    roof.update({"traffic": measured_traffic(kname), "kernel": kname, "class": cname(a["cls"]),
                 "precision": "mixed" if a["mixed"] else "fp64",
                 "compute": "vector ALU (no MFMA); compute peak is the FP64 vector rate (78.6 TFLOP/s) with "
                            "FP32 flops priced at 157.3 TFLOP/s",
                 "algorithmic_bytes_per_launch": by, "flops_per_launch": f64 + f32, "fp32_flops_per_launch": f32,
                 "launches": a["n"], "mean_launch_ms": 1000.0 * mean_s, "share_of_device_time": a["s"] / tot})
    return roof, agg


def whole_job(per, info, wall):
    t_min = 0.0
    fl = 0.0
    for ln, _ in per:
        f64, f32 = launch_split_flops(info, ln)
        t_min += max(f64 / PEAK_FP64 + f32 / PEAK_FP32, 8.0 * ln.nout / PEAK_HBM)
        fl += f64 + f32
    return {"roofline_time_s": t_min, "measured_s": wall, "frac": t_min / wall, "tflops": fl / wall / 1e12,
            "note": "sum over launches of max(F64/78.6T + F32/157.3T, bytes/8TB/s) over the measured time"}


def cpu_baseline(ss, seconds=12.0):
    """The oracle on every core this job may use, on uniform samples of the
    screened set (ctypes releases the interpreter lock during each oracle
    call, so the threads run in parallel)."""
    import oracle
    sa = oracle.ShellArrays(ss.shells)
    oracle.normalize(sa)
    cores = min(16, os.cpu_count() or 1)
    oracle.eri(sa, ss.sample(8, seed=99))      # tables and plans built once
    lock = threading.Lock()
    tot = {"done": 0}
    stop = time.perf_counter() + seconds

    def work(t):
        seed = 1000 * t
        while time.perf_counter() < stop:
            with lock:
                qs = ss.sample(400, seed=seed)
            seed += 1
            oracle.eri(sa, qs)
            with lock:
                tot["done"] += len(qs)
    t0 = time.perf_counter()
    th = [threading.Thread(target=work, args=(t,)) for t in range(cores)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    el = time.perf_counter() - t0
    model, flags = "unknown", ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            if line.startswith("flags") and not flags:
                flags = line
    except OSError:
        pass
    isa = [x for x in ("avx512f", "avx2", "fma") if (" %s " % x) in flags + " "]
    return {"value": tot["done"] / el, "unit": "shell quartets/s", "cores": cores, "kind": "port",
            "cpu": model, "isa": isa,
            "sample": "%d quartets drawn uniformly from the screened %d-atom set, %.1f s wall on %d threads "
                      "of the oracle (gcc -O2 scalar C)" % (tot["done"], len(ss.shells) // 6, el, cores)}


def serving_leg(nreq=10, rate=0.5, natom=(10, 48), seed=0):
    """A bounded Poisson trace through the continuous-batching server."""
    import serving
    reqs = serving.request_stream(nreq, rate, seed, natom)
    srv = serving.Server(hbm_budget=64 << 30, slots=4, tol=1e-10)
    srv.serve(reqs)
    res = serving.summarize(reqs, srv.wall)
    res["max_concurrent"] = srv.max_concurrent
    res["trace"] = ("%d requests, Poisson arrivals at %.2f/s, molecules of %d-%d atoms of the 64-atom recipe "
                    "(own seeds), 30%% first derivatives, Schwarz 1e-10, 4 slots, 64 GiB admission budget"
                    % (nreq, rate, natom[0], natom[1]))
    res["latencies_s"] = [None if r.t_done is None else r.t_done - r.arrival for r in reqs]
    return res


def batched_leg(steps, warmup, nquartet=10 ** 6):
    """The batched stream: 1e6 quartets over every class up to f, one graph
    launch per pass of 256 quartet-list class launches, inputs and the 5 GB
    output resident in HBM."""
    import torch
    import batch
    import simint
    import workloads
    info = json.load(open(os.path.join(PKG, "csrc", "gen", "classes.json")))
    t0 = time.perf_counter()
    shells, q = workloads.all_class_stream(nquartet=nquartet, seed=0)
    basis = simint.Basis(shells)
    b = batch.QuartetBatch(basis, q)
    setup = time.perf_counter() - t0
    progress("batched stream set up (%.1f s)" % setup)
    flops = 0.0
    for g in b.groups:
        c = info[cname(g.cls)]
        flops += g.prim_quartets * (c["f_setup"] + c["f_vrr"] + c["f_boys"] + c["bcont"]) + g.n * c["f_hrr"]
    out = torch.empty(b.out_doubles, dtype=torch.float64, device="cuda")
    stream = torch.cuda.ExternalStream(simint.lib().simint_hip_stream())
    ev = []

    def rec(g, start):
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream)
        ev.append((g, e))
    b.run(out.data_ptr(), events=rec)
    simint.synchronize()
    per = [(ev[k][1].elapsed_time(ev[k + 1][1]), ev[k][0]) for k in range(0, len(ev), 2)]
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
           "top_classes": [{"class": cname(g.cls), "ms": ms, "quartets": g.n, "kernel": list_kernel_symbol(g.cls)}
                           for ms, g in per[:8]]}
    fam = {}
    for ms, g in per:
        k = "wpql" if "wpql" in list_kernel_symbol(g.cls) else "lpql"
        fam[k] = fam.get(k, 0.0) + ms
    res["eager_class_ms"] = {"total": sum(ms for ms, _ in per), "by_family": fam,
                             "note": "classes one after another on one stream, per-class events"}
    b.free()
    basis.free()
    return res


def deriv_leg(natom=64, tol=1e-12):
    """First derivatives of every screened quartet of the 64-atom system
    (12 values per Cartesian), one pass through an 8 GiB chunk buffer: each
    derivative call runs a class kernel per (a+1)/(a-1) term and the
    assembly, so a call carries ~0.5 ms of launch ramp-up and tail, and
    larger chunks (fewer calls) amortise it (11.2 s per pass with 1 GiB
    chunks, 8.8 s with 8 GiB; the term scratch grows to 6 GB)."""
    import torch
    import quartets
    import simint
    import workloads
    shells = workloads.grid_molecule(natom=natom, seed=0)
    ss = quartets.ScreenedSet(shells, tol=tol, out_doubles=1 << 30, deriv=1)
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


def dropin_leg():
    """The host-buffer drop-in call (simint_compute_eri: pairs packed and
    copied to the device, integrals copied back over PCIe) on one large
    batch: P = Q = every (p p) shell pair of 16 atoms of the 64-atom recipe.
    Its rate includes both host<->device copies; `value` does not."""
    import simint
    import workloads
    shells = workloads.grid_molecule(natom=16, seed=0)
    basis = simint.Basis(shells)
    ps = [i for i, s in enumerate(shells) if s[0] == 1]
    pairs = [(a, b) for ia, a in enumerate(ps) for b in ps[:ia + 1]]
    P = simint.Pairs.from_pairs(basis, pairs)
    simint.compute_eri(P, P)          # warm-up: device set-up and staging buffers
    t = time.perf_counter()
    rc, out = simint.compute_eri(P, P)
    wall = time.perf_counter() - t
    nq = len(pairs) ** 2
    res = {"what": "simint_compute_eri, host buffers, P = Q = %d (pp) pairs of 16 atoms of the 64-atom recipe"
                   % len(pairs), "quartets": nq, "returned": int(rc), "s": wall, "value": nq / wall,
           "unit": "shell quartets/s", "GB_to_host": 8.0 * out.size / 1e9}
    P.free()
    basis.free()
    return res


def fp64_64atom_leg():
    """The 64-atom all-classes-to-d system in plain FP64, one full pass."""
    import torch
    import quartets
    import simint
    import workloads
    ss = quartets.ScreenedSet(workloads.grid_molecule(natom=64, seed=0), tol=1e-12)
    out = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
    ss.run_launches(out.data_ptr(), shares(ss.launches, 4)[0])
    simint.synchronize()
    t = time.perf_counter()
    ss.run_pass(out.data_ptr())
    simint.synchronize()
    wall = time.perf_counter() - t
    res = {"workload": "64-atom jittered 3-bohr grid, s6 s3 p3 s1 p1 d1 per atom, every unique quartet with "
                       "Schwarz bound >= 1e-12, plain FP64", "quartets": ss.total_quartets,
           "value": ss.total_quartets / wall, "unit": "shell quartets/s", "pass_s": wall,
           "tflops": ss.total_flops / wall / 1e12}
    # the same pass with the Coulomb/exchange digestion after every launch
    nbf = ss.nbf()
    g = torch.Generator(device="cpu").manual_seed(5)
    C = torch.rand((nbf, nbf // 4), generator=g, dtype=torch.float64) - 0.5
    D = (C @ C.T / nbf).cuda()
    J = torch.zeros((nbf, nbf), dtype=torch.float64, device="cuda")
    K = torch.zeros_like(J)
    ss.run_jk(out.data_ptr(), D.data_ptr(), J.data_ptr(), K.data_ptr(), shares(ss.launches, 4)[0])
    J.zero_()
    K.zero_()
    simint.synchronize()
    t = time.perf_counter()
    ss.run_jk(out.data_ptr(), D.data_ptr(), J.data_ptr(), K.data_ptr())
    simint.jk_symmetrize(nbf, J.data_ptr(), K.data_ptr())
    simint.synchronize()
    wjk = time.perf_counter() - t
    res["jk"] = {"what": "Coulomb and exchange matrices (%d basis functions) of a random density, digested after "
                         "every class launch (simint_hip_compute_jk_range)" % nbf,
                 "pass_s": wjk, "value": ss.total_quartets / wjk, "unit": "shell quartets/s",
                 "digestion_share": 1.0 - wall / wjk,
                 "energy_check": {"tr_DJ": float((D * J).sum()), "tr_DK": float((D * K).sum())}}
    del J, K, D
    res["prim_screen"] = prim_screen_leg(ss, out)
    ss.free()
    return res


def prim_screen_leg(ss, out, prim_tol=1e-13):
    """The same 64-atom pass with simint's primitive screen: pairs built
    with SIMINT_SCREEN_SCHWARZ (a bound per primitive pair) and screen_tol
    applied in the primitive loop; the change against the unscreened pass
    is measured over every integral."""
    import torch
    import quartets
    import simint
    import workloads
    s2 = quartets.ScreenedSet(workloads.grid_molecule(natom=64, seed=0), tol=1e-12, prim_tol=prim_tol)
    if [(ln.cls, ln.b0, ln.nq) for ln in s2.launches] != [(ln.cls, ln.b0, ln.nq) for ln in ss.launches]:
        raise RuntimeError("the screened set's plan differs from the plain one")
    out2 = torch.empty(s2.out_doubles, dtype=torch.float64, device="cuda")
    s2.run_launches(out2.data_ptr(), shares(s2.launches, 4)[0])
    simint.synchronize()
    t = time.perf_counter()
    s2.run_pass(out2.data_ptr())
    simint.synchronize()
    wall = time.perf_counter() - t
    change = 0.0
    for c in range(ss.nchunk):
        ss.run_pass(out.data_ptr(), chunk=c)
        s2.run_pass(out2.data_ptr(), chunk=c)
        simint.synchronize()
        n = ss.chunk_doubles(c)
        change = max(change, float((out[:n] - out2[:n]).abs().max()))
    res = {"what": "the 64-atom FP64 pass with the primitive Schwarz screen (SIMINT_SCREEN_SCHWARZ pairs, "
                   "screen_tol %g in the primitive loop)" % prim_tol,
           "value": s2.total_quartets / wall, "unit": "shell quartets/s", "pass_s": wall,
           "max_abs_change_vs_unscreened": change, "setup_s": s2.setup_s}
    del out2
    s2.free()
    return res


def make_parser():
    """The command line (also parsed by the CPU tests against the commands
    the profile summaries give)."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="shares of the quartet set timed (together one pass)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed shares run first")
    ap.add_argument("--natom", type=int, default=256)
    ap.add_argument("--mixed-threshold", type=float, default=1e-4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched all-classes stream")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving trace")
    ap.add_argument("--no-deriv", action="store_true", help="skip the first-derivative leg")
    ap.add_argument("--no-fp64-64", action="store_true", help="skip the 64-atom FP64 leg")
    ap.add_argument("--no-errcheck", action="store_true", help="profiling aid: skip the error measurement")
    ap.add_argument("--tuning", default=None, help="apply this kernel-family tuning file before running")
    ap.add_argument("--only-class", default=None,
                    help="profiling aid: run only this class's launches (e.g. d_p_d_p)")
    ap.add_argument("--copy-gib", type=float, default=None,
                    help="profiling aid: run one device copy of this many GiB and exit "
                         "(calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--calib-flops", type=int, default=None,
                    help="profiling aid: symmetrise one N x N FP64 matrix with the library's kernel "
                         "(exactly N(N-1) FP64 flops: one add and one multiply per off-diagonal pair) "
                         "and exit (calibrates the SQ_INSTS_VALU_FLOPS_FP64 counter)")
    return ap


def main():
    args = make_parser().parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (shell quartets are independent; a "
              "multi-GPU run would shard the screened set by bra pairs with replicated shell pairs)",
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
    if args.calib_flops:
        sys.path.insert(0, PKG)
        import simint
        simint.lib().simint_init()
        n = args.calib_flops
        M = torch.rand((n, n), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        simint.jk_symmetrize(n, M.data_ptr(), None)
        simint.synchronize()
        print(json.dumps({"calib_flops_fp64": n * (n - 1)}))
        return
    sys.path.insert(0, PKG)
    # the CPU oracle (error check and cpu_baseline only): oracle/oracle.py
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import simint
    if args.tuning:
        simint.apply_tuning(args.tuning)

    m = mixed_system_leg(args)
    info = m["info"]
    value = m["quartets"] / m["wall_mixed"]
    fp64_value = m["quartets"] / m["wall_fp64"]
    roof, agg = roofline_of(m["per_launch"], info)
    per_class = {}
    for ln, s in m["per_launch"]:
        pc = per_class.setdefault(cname(ln.cls), {"ms_mixed": 0.0})
        pc["ms_mixed"] += 1000.0 * s
    for ln, s in m["per_launch64"]:
        pc = per_class.setdefault(cname(ln.cls), {"ms_mixed": 0.0})
        pc["ms_fp64"] = pc.get("ms_fp64", 0.0) + 1000.0 * s
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "shell quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * m["wall_mixed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "%d-atom jittered 3-bohr grid (8x8x4), s6 s3 p3 s1 p1 d1 per atom, every unique "
                               "quartet with Schwarz bound >= 1e-12, mixed precision: FP32 VRR for quartets with "
                               "bound < %g (/400 with one HRR side, /1e5 with two), Boys, contraction and HRR "
                               "FP64" % (args.natom, args.mixed_threshold),
                   "natom": args.natom, "nshell": 6 * args.natom, "quartets_per_pass": m["quartets"],
                   "fp32_vrr_quartets": m["fp32_quartets"], "step": "1/%d of the quartet set" % args.steps,
                   "path": "device-resident screened ranges (simint_hip_compute_eri_range)",
                   "setup_s": m["setup_s"]},
        "mixed": {"value": value, "device_s": m["dev_mixed"], "wall_s": m["wall_mixed"]},
        "fp64": {"value": fp64_value, "device_s": m["dev_fp64"], "wall_s": m["wall_fp64"],
                 "ms_per_step": 1000.0 * m["wall_fp64"] / args.steps},
        "mixed_over_fp64": value / fp64_value,
        "roofline": roof,
        "whole_job": whole_job(m["per_launch"], info, m["wall_mixed"]),
        "whole_job_fp64": whole_job(m["per_launch64"], info, m["wall_fp64"]),
    }
    if "dev_err" in m:
        line["error"] = {"bound": 1e-10,
                         "max_abs_err_vs_oracle": max(m["orc_err"].values()) if m["orc_err"] else 0.0,
                         "max_abs_err_vs_fp64_gpu": max(m["dev_err"].values()) if m["dev_err"] else 0.0,
                         "per_class_vs_oracle": m["orc_err"], "per_class_vs_fp64_gpu": m["dev_err"],
                         "sample": m["err_cover"],
                         "note": "classes absent here run FP64 throughout (no quartet under the threshold, or "
                                 "their kernel has no FP32 variant)"}
    top = sorted(per_class.items(), key=lambda kv: -kv[1]["ms_mixed"])[:10]
    line["top_classes"] = [dict(cls=k, **v) for k, v in top]
    ss = m["ss"]
    if not args.no_serving and not args.only_class:
        progress("serving trace")
        line["serving"] = serving_leg()
        line["p50_latency_s"] = line["serving"]["p50_latency_s"]
        line["p99_latency_s"] = line["serving"]["p99_latency_s"]
        progress("serving leg done")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ss)
        progress("cpu baseline done")
    ss.free()
    del m
    if not args.no_batched and not args.only_class:
        line["batched"] = batched_leg(3, 1)
        progress("batched leg done")
    if not args.no_deriv and not args.only_class:
        line["deriv1"] = deriv_leg()
        progress("deriv leg done")
    if not args.no_fp64_64 and not args.only_class:
        line["fp64_64atom"] = fp64_64atom_leg()
        progress("64-atom FP64 leg done")
        line["dropin_pcie"] = dropin_leg()
        progress("drop-in leg done")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
