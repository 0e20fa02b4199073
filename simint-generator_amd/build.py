"""Build libsimint.so (host runtime + generated gfx950 kernels) in-tree.

    python simint-generator_amd/build.py [--maxam N] [--jobs J]

Steps: emit the generated sources (emitter/generate.py), compile every HIP
source for gfx950 and the host C++ sources, keep each lane-per-quartet
kernel only if the compiler reports no scratch spills (otherwise the class
uses its wave-per-quartet kernel), and link.  Objects are cached under
_build/ keyed by a hash of their inputs.
"""
import argparse
import concurrent.futures as cf
import glob
import hashlib
import json
import os
import re
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(CSRC, "gen")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libsimint.so")

DEFAULT_MAXAM = 3
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-I" + INCLUDE, "-I" + CSRC,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-I" + INCLUDE, "-I" + CSRC, "-Wall"]

sys.path.insert(0, PKG)


def _header_hash(kernels_only=False):
    """Hash of the headers a source depends on: the generated class kernels
    include only the device layout, the Boys code and its tables; the
    runtime includes everything."""
    h = hashlib.sha1()
    if kernels_only:
        files = sorted([os.path.join(CSRC, "device_types.h"), os.path.join(CSRC, "boys_device.h")]
                       + glob.glob(os.path.join(GEN, "*.inc")))
    else:
        files = sorted(glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(GEN, "*.inc"))
                       + glob.glob(os.path.join(INCLUDE, "simint", "**", "*.h*"), recursive=True))
    for f in files:
        h.update(f.encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def _compile(src, hdr_hash, want_report):
    base = os.path.basename(src)
    obj = os.path.join(BUILD, base + ".o")
    keyf = obj + ".key"
    repf = obj + ".res"
    is_hip = src.endswith(".hip")
    cmd = [HIPCC] + HIP_FLAGS if is_hip else ["g++"] + CXX_FLAGS
    if is_hip and want_report:
        cmd = cmd + ["-Rpass-analysis=kernel-resource-usage"]
    cmd = cmd + ["-c", src, "-o", obj]
    h = hashlib.sha1()
    with open(src, "rb") as f:
        h.update(f.read())
    h.update(hdr_hash.encode())
    h.update(" ".join(cmd).encode())
    key = h.hexdigest()
    if os.path.exists(obj) and os.path.exists(keyf) and open(keyf).read() == key:
        rep = open(repf).read() if os.path.exists(repf) else ""
        return obj, rep, False
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), p.stderr[-4000:]))
    with open(repf, "w") as f:
        f.write(p.stderr)
    with open(keyf, "w") as f:
        f.write(key)
    return obj, p.stderr, True


def _scratch(report, kernel=None):
    """Scratch bytes per lane from the compiler's resource report: of the
    named kernel (mangled name contains it), or the largest in the file."""
    if kernel is None:
        m = re.findall(r"ScratchSize \[bytes/lane\]: (\d+)", report)
        return max(int(x) for x in m) if m else 0
    cur = None
    for line in report.splitlines():
        fm = re.search(r"Function Name: (\S+)", line)
        if fm:
            cur = fm.group(1)
            continue
        sm = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if sm and cur and re.search(r"\d%s\d" % re.escape(kernel), cur):
            return int(sm.group(1))
    return 0


def build(maxam=DEFAULT_MAXAM, jobs=None, verbose=True):
    from emitter import generate as gen_mod
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 4)
    okf = os.path.join(GEN, "lpq_ok.json")
    lpq_ok = None
    if os.path.exists(okf):
        d = json.load(open(okf))
        if d.get("maxam") == maxam:
            lpq_ok = set(tuple(x) for x in d["ok"])
    lpq, wpq, sel, lpqs = gen_mod.generate(PKG, INCLUDE, maxam, lpq_ok)
    hdr = _header_hash()
    srcs = sorted(glob.glob(os.path.join(GEN, "eri_*.hip")))
    # drop stale generated kernels of classes outside this build
    from emitter.hip_lpq import class_name
    want = set("eri_lpq_%s.hip" % class_name(a) for a in lpq) | set("eri_wpq_%s.hip" % class_name(a) for a in wpq)
    from emitter.hip_lpqs import kernel_name
    # split kernels: keys (a,b,c,d) for the default kind, ("f", a,b,c,d) for
    # the ket-top variants of bra-heavy classes
    want |= set("eri_%s.hip" % (kernel_name(a[1:], G, "f") if a[0] == "f" else kernel_name(a, G))
                for a, Gs in lpqs.items() for G in Gs)
    srcs = [s for s in srcs if os.path.basename(s) in want]
    rt = [os.path.join(CSRC, "runtime.hip"), os.path.join(CSRC, "osoei.hip"), os.path.join(CSRC, "frontend.hip"),
          os.path.join(CSRC, "jk.hip"),
          os.path.join(CSRC, "shell.cpp"),
          os.path.join(CSRC, "shell_screen.cpp")]
    reports = {}
    objs = {}
    nbuilt = 0
    with cf.ThreadPoolExecutor(jobs) as ex:
        khdr = _header_hash(kernels_only=True)
        futs = {ex.submit(_compile, s, khdr if s in srcs else hdr, "eri_lpq_" in s): s for s in srcs + rt}
        for f in cf.as_completed(futs):
            s = futs[f]
            obj, rep, built = f.result()
            objs[s] = obj
            reports[s] = rep
            nbuilt += built
    # keep LPQ kernels that compiled without scratch spills
    ok = set()
    for am in lpq:
        s = os.path.join(GEN, "eri_lpq_%s.hip" % class_name(am))
        if _scratch(reports[s], "eri_lpq_" + class_name(am)) == 0:
            ok.add(am)
    with open(okf, "w") as f:
        json.dump({"maxam": maxam, "ok": sorted(list(a) for a in ok)}, f)
    lpq, wpq, sel, lpqs = gen_mod.generate(PKG, INCLUDE, maxam, ok)
    hdr = _header_hash()
    dobj, _, built = _compile_dispatch(hdr)
    objs["dispatch"] = dobj
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + sorted(objs.values())
    p = subprocess.run(link, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stderr[-4000:])
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        nl = sum(1 for v in sel.values() if v == 1)
        print("libsimint.so: maxam %d, %d classes (%d lane-per-quartet, %d wave-per-quartet), %d objects rebuilt"
              % (maxam, len(sel), nl, len(sel) - nl, nbuilt))
    return LIB


def _compile_dispatch(hdr):
    # dispatch.cpp includes HIP headers: compile with hipcc as host code
    src = os.path.join(GEN, "dispatch.cpp")
    obj = os.path.join(BUILD, "dispatch.cpp.o")
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-I" + INCLUDE, "-I" + CSRC, "-x", "hip",
           "--offload-arch=gfx950", "-c", src, "-o", obj]
    h = hashlib.sha1(open(src, "rb").read() + hdr.encode() + " ".join(cmd).encode()).hexdigest()
    keyf = obj + ".key"
    if os.path.exists(obj) and os.path.exists(keyf) and open(keyf).read() == h:
        return obj, "", False
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), p.stderr[-4000:]))
    open(keyf, "w").write(h)
    return obj, "", True


def build_oracle(verbose=True):
    """Build the CPU oracle (test infrastructure) and, where the reference
    tree is present, the reference pieces under oracle/_ref."""
    od = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", od], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", od, "ref", "-j8"], check=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--maxam", type=int, default=DEFAULT_MAXAM)
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.maxam, a.jobs)
    build_oracle()
