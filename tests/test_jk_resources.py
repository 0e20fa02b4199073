"""The J/K digestion kernels, compiled for gfx950 on the CPU: no instance
of the specialised kernels spills to scratch memory.  Their register arrays
are sized by a budget (doubles per lane, the second ket shell taken in
chunks) meant to keep every instance within the 256 VGPRs plus AGPRs of a
lane; a budget that let the compiler spill would turn every digestion of
that shape into scratch traffic."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_jk_digestion_kernels_do_not_spill(tmp_path):
    src = os.path.join(ROOT, "simint-generator_amd", "csrc", "jk.hip")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "simint-generator_amd", "csrc"), "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "--cuda-device-only", "-c", src, "-o", str(tmp_path / "jk.o"),
           "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    kern, scratch = None, {}
    for line in r.stderr.splitlines():
        m = re.search(r"remark:\s+Function Name: (\S+)", line)
        if m:
            kern = m.group(1)
            continue
        m = re.search(r"remark:\s+ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and kern:
            scratch[kern] = int(m.group(1))
    digest = {k: v for k, v in scratch.items() if "jk_digest_rows" in k or "jk_digest_t" in k}
    assert len(digest) >= 50, sorted(scratch)      # every (second bra, ket, ket) shape up to f
    spilled = {k: v for k, v in digest.items() if v > 0}
    assert not spilled, spilled
