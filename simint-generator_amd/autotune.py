"""Per-class kernel-family autotuner (runs on the GPU).

For every class, times each kernel family built for it -- lane-per-quartet
(registers), wave-per-quartet (LDS-staged, batched), the split
lane-per-quartet kernels with two group counts and, for bra-heavy d
classes, their ket-top-group variants -- on the screened quartets
of a synthetic system built with the 64-atom benchmark recipe at a smaller
size, and writes the fastest family per class to tuning.json (read by the
build to set each class's default):

    python simint-generator_amd/autotune.py [--natom 16] [--reps 3] [--out PATH]
    python simint-generator_amd/autotune.py --list [--split]   # quartet-list kernels

--list --split also times the lanes per quartet of the lane-per-quartet
list launches per class and records the counts that beat the automatic
choice ("list_split"; batch.QuartetBatch applies them).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--natom", type=int, default=16)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    ap.add_argument("--list", action="store_true",
                    help="tune the quartet-list kernels on the batched all-classes stream instead "
                         "(merges 'list_family' into the output file)")
    ap.add_argument("--nquartet", type=int, default=200000)
    ap.add_argument("--split", action="store_true",
                    help="with --list: also time 1-16 lanes per quartet for the classes whose list kernel is "
                         "lane-per-quartet (merges 'list_split' into the output file)")
    a = ap.parse_args()
    if a.list:
        tune_lists(a)
        if a.split:
            tune_split(a)
        return None
    import torch
    import quartets
    import simint
    import workloads
    L = simint.lib()
    ss = quartets.ScreenedSet(workloads.grid_molecule(natom=a.natom, seed=0), tol=1e-12)
    out = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
    stream = torch.cuda.ExternalStream(L.simint_hip_stream())
    by = {}
    for ln in ss.launches:
        by.setdefault(ln.cls, []).append(ln)
    family = {}
    timings = {}
    for cls in sorted(by):
        name = quartets.simint_class_name(cls)
        prev = L.simint_hip_class_kernel(*cls)
        res = {}
        for fam, fname in ((1, "lpq"), (2, "wpq"), (3, "lpqs_a"), (4, "lpqs_b"), (5, "lpqsf_a"), (6, "lpqsf_b")):
            if L.simint_hip_set_class_kernel(*cls, fam) < 0:
                continue
            best = None
            for rep in range(a.reps + 1):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for ln in by[cls]:
                    simint.compute_prefix(ln.bra.dev, ln.ket.dev, ln.b0, ln.nket, ln.off, out.data_ptr(), 0.0)
                e1.record(stream)
                simint.synchronize()
                ms = e0.elapsed_time(e1)
                if rep > 0:
                    best = ms if best is None else min(best, ms)
            res[fname] = best
        L.simint_hip_set_class_kernel(*cls, prev)
        if res:
            family[name] = min(res, key=res.get)
            timings[name] = res
            print("%-8s %s" % (name, "  ".join("%s %.3f ms" % kv for kv in sorted(res.items()))), flush=True)
    doc = json.load(open(a.out)) if os.path.exists(a.out) else {}
    doc.update({"workload": "64-atom recipe, %d atoms, Schwarz 1e-12, quartets of each class" % a.natom,
                "family": family, "timings_ms": timings,
                "generated": time.strftime("%Y-%m-%d %H:%M:%S")})
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print("wrote", a.out)


def tune_lists(a):
    """Quartet-list family per class, timed on a batched stream of the
    all-classes recipe (a 4096-shell pool, depths 1-6)."""
    import torch
    import batch
    import simint
    import workloads
    L = simint.lib()
    shells, q = workloads.all_class_stream(nquartet=a.nquartet, seed=1)
    basis = simint.Basis(shells)
    b = batch.QuartetBatch(basis, q)
    out = torch.empty(b.out_doubles, dtype=torch.float64, device="cuda")
    stream = torch.cuda.ExternalStream(L.simint_hip_stream())
    family, timings = {}, {}
    for g in b.groups:
        name = "_".join("spdfghik"[x] for x in g.cls)
        prev = L.simint_hip_class_list_kernel(*g.cls)
        res = {}
        for fam, fname in ((1, "lpq"), (2, "wpq")):
            if L.simint_hip_set_class_list_kernel(*g.cls, fam) < 0:
                continue
            best = None
            for rep in range(a.reps + 1):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                simint.compute_list(g.P, g.Q, g.n, g.idx.data_ptr(), g.idx.data_ptr(), g.qout.data_ptr(),
                                    out.data_ptr())
                e1.record(stream)
                simint.synchronize()
                ms = e0.elapsed_time(e1)
                if rep > 0:
                    best = ms if best is None else min(best, ms)
            res[fname] = best
        L.simint_hip_set_class_list_kernel(*g.cls, prev)
        if res:
            family[name] = min(res, key=res.get)
            timings[name] = res
            print("%-8s %s" % (name, "  ".join("%s %.3f ms" % kv for kv in sorted(res.items()))), flush=True)
    doc = json.load(open(a.out)) if os.path.exists(a.out) else {}
    doc["list_family"] = family
    doc["list_timings_ms"] = timings
    doc["list_workload"] = "%d quartets of the all-classes stream recipe (seed 1)" % a.nquartet
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print("wrote", a.out)


def tune_split(a):
    """Lanes per quartet of the lane-per-quartet list launches, per class,
    on the same batched stream: every class whose list kernel (after the
    family tuning) is lane-per-quartet is timed with the automatic choice
    (0: about 65536 lanes per launch) and with 1, 2, 4, 8 and 16 lanes; a
    class keeps a fixed count only where it beats the automatic one by more
    than 2 %."""
    import torch
    import batch
    import simint
    import workloads
    L = simint.lib()
    if os.path.exists(a.out):
        simint.apply_tuning(a.out)
    shells, q = workloads.all_class_stream(nquartet=a.nquartet, seed=1)
    basis = simint.Basis(shells)
    b = batch.QuartetBatch(basis, q)
    out = torch.empty(b.out_doubles, dtype=torch.float64, device="cuda")
    stream = torch.cuda.ExternalStream(L.simint_hip_stream())
    split, timings = {}, {}
    for g in b.groups:
        if L.simint_hip_class_list_kernel(*g.cls) != 1:
            continue
        name = "_".join("spdfghik"[x] for x in g.cls)
        res = {}
        for lanes in (0, 1, 2, 4, 8, 16):
            prev = L.simint_hip_set_list_split(lanes)
            best = None
            for rep in range(a.reps + 1):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                simint.compute_list(g.P, g.Q, g.n, g.idx.data_ptr(), g.idx.data_ptr(), g.qout.data_ptr(),
                                    out.data_ptr())
                e1.record(stream)
                simint.synchronize()
                ms = e0.elapsed_time(e1)
                if rep > 0:
                    best = ms if best is None else min(best, ms)
            L.simint_hip_set_list_split(prev)
            res[lanes] = best
        timings[name] = res
        fastest = min(res, key=res.get)
        if fastest != 0 and res[fastest] < 0.98 * res[0]:
            split[name] = fastest
        print("%-8s %s" % (name, "  ".join("%d: %.3f ms" % kv for kv in sorted(res.items()))), flush=True)
    doc = json.load(open(a.out)) if os.path.exists(a.out) else {}
    doc["list_split"] = split
    doc["list_split_timings_ms"] = {k: {str(l): t for l, t in v.items()} for k, v in timings.items()}
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print("wrote", a.out, "(%d classes with a fixed lane count)" % len(split))


if __name__ == "__main__":
    main()
